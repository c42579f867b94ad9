"""adgpu_eval_poses_noalloc through the C ABI: the same energies and gradients,
bit for bit, as adgpu_eval_poses (same kernels, staging in the library's fixed
per-device area instead of allocated buffers), on two calls in a row (the
second with fewer poses, so stale contents of the shared area would show),
and sizes over the limit refused."""
import ctypes as C

import numpy as np
import pytest

import adgpu
import helpers
pytestmark = pytest.mark.gpu

_fp = C.POINTER(C.c_float)
ADGPU_OK, ADGPU_ERR_LIMIT = 0, 2       # adgpu_status (include/adgpu.h)


def _noalloc(ctx, mid, lid, mode, poses):
    g = np.ascontiguousarray(poses, np.float32)
    n = g.shape[0]
    e = np.full(n, np.nan, np.float32)
    gr = np.full((n, 64), np.nan, np.float32)
    rc = ctx.lib.adgpu_eval_poses_noalloc(ctx.h, mid, lid, int(mode), g.ctypes.data_as(_fp), n,
                                          e.ctypes.data_as(_fp), gr.ctypes.data_as(_fp))
    return rc, e, gr


def test_noalloc_matches_eval_poses_on_two_calls(gpu_ctx):
    case = helpers.synthetic_case(30, 6, 7, 11, npts=60)
    mid = gpu_ctx.maps_upload(case.packed, case.size, case.maps.num_map_atypes, case.spacing)
    lid = gpu_ctx.ligand_upload(case.tables, case.ep)
    big = helpers.poses_for(case, 3000, seed=71)
    small = helpers.poses_for(case, 700, seed=72)
    for mode in (adgpu.EVAL_ENERGRAD, adgpu.EVAL_ENERGY, adgpu.EVAL_GRADIENT):
        for poses in (big, small):
            rc, e, g = _noalloc(gpu_ctx, mid, lid, mode, poses)
            assert rc == ADGPU_OK, gpu_ctx.lib.adgpu_last_error(gpu_ctx.h)
            e_ref, g_ref = gpu_ctx.eval_poses(mid, lid, poses, mode)
            if mode != adgpu.EVAL_GRADIENT:
                assert np.array_equal(e, e_ref), mode
            if mode != adgpu.EVAL_ENERGY:
                assert np.array_equal(g, g_ref), mode
    # a ligand of another size on the same staging area
    case2 = helpers.synthetic_case(60, 12, 4, 9, npts=60)
    mid2 = gpu_ctx.maps_upload(case2.packed, case2.size, case2.maps.num_map_atypes, case2.spacing)
    lid2 = gpu_ctx.ligand_upload(case2.tables, case2.ep)
    p2 = helpers.poses_for(case2, 500, seed=73)
    rc, e, g = _noalloc(gpu_ctx, mid2, lid2, adgpu.EVAL_ENERGRAD, p2)
    assert rc == ADGPU_OK
    e_ref, g_ref = gpu_ctx.eval_poses(mid2, lid2, p2, adgpu.EVAL_ENERGRAD)
    assert np.array_equal(e, e_ref) and np.array_equal(g, g_ref)
    helpers.check_energies(case2.orc, p2, e)
    for c_, m_, l_ in ((gpu_ctx, mid, lid), (gpu_ctx, mid2, lid2)):
        c_.ligand_release(l_)
        c_.maps_release(m_)


def test_noalloc_refuses_sizes_over_the_limit(gpu_ctx):
    case = helpers.synthetic_case(12, 0, 3, 5, npts=40)
    mid = gpu_ctx.maps_upload(case.packed, case.size, case.maps.num_map_atypes, case.spacing)
    lid = gpu_ctx.ligand_upload(case.tables, case.ep)
    poses = helpers.poses_for(case, 8193, seed=74)
    rc, _, _ = _noalloc(gpu_ctx, mid, lid, adgpu.EVAL_ENERGY, poses)
    assert rc == ADGPU_ERR_LIMIT
    assert b"ADGPU_NOALLOC_MAX_POSES" in gpu_ctx.lib.adgpu_last_error(gpu_ctx.h)
    # the largest accepted size works
    rc, e, _ = _noalloc(gpu_ctx, mid, lid, adgpu.EVAL_ENERGY, poses[:8192])
    assert rc == ADGPU_OK
    e_ref, _ = gpu_ctx.eval_poses(mid, lid, poses[:8192], adgpu.EVAL_ENERGY)
    assert np.array_equal(e, e_ref)
    gpu_ctx.ligand_release(lid)
    gpu_ctx.maps_release(mid)
