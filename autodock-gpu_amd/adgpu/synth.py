"""Seeded synthetic docking inputs: ligands with random torsion trees and
Gaussian-well receptor maps.

The generator builds the ligand's device tables directly, following the
reference's ligand preparation rules (host/src/processligand.cpp), and can
also write the same ligand as pdbqt plus the maps as AutoGrid .fld/.map
files, so the reference's own parser reads identical inputs.

Rules restated here (each cited at its use):
  * AD4.1 per-type parameters and combining rules -- get_VWpars,
    processligand.cpp:749-1066;
  * H-bond pairs -- is_H_bond / is_H_acceptor, processligand.cpp:1505-1540;
  * intramolecular pairs -- get_intraE_contributors, processligand.cpp:250-457;
  * rigid fragments and rotatable-bond order (ENDBRANCH order) --
    parse_liganddata, processligand.cpp:1226-1290;
  * moving / unit vectors -- get_moving_and_unit_vectors,
    processligand.cpp:1068-1122;
  * centring and grid scaling -- gen_initpop_and_reflig,
    getparameters.cpp:2624-2630.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np

# AD4.1 parameters, in the order of the reference's atom_names table
# (processligand.cpp:866-896): name, Rii (sum of vdW radii of two like atoms),
# epsii, Rij_hb, epsij_hb, volume, solpar.
_AD4 = [
    ("H", 2.00, 0.020, 0.0, 0.0, 0.0000, 0.00051),
    ("HD", 2.00, 0.020, 0.0, 1.0, 0.0000, 0.00051),
    ("HS", 2.00, 0.020, 0.0, 1.0, 0.0000, 0.00051),
    ("C", 4.00, 0.150, 0.0, 0.0, 33.5103, -0.00143),
    ("A", 4.00, 0.150, 0.0, 0.0, 33.5103, -0.00052),
    ("N", 3.50, 0.160, 0.0, 0.0, 22.4493, -0.00162),
    ("NA", 3.50, 0.160, 1.9, 5.0, 22.4493, -0.00162),
    ("NS", 3.50, 0.160, 1.9, 5.0, 22.4493, -0.00162),
    ("OA", 3.20, 0.200, 1.9, 5.0, 17.1573, -0.00251),
    ("OS", 3.20, 0.200, 1.9, 5.0, 17.1573, -0.00251),
    ("F", 3.09, 0.080, 0.0, 0.0, 15.4480, -0.00110),
    ("MG", 1.30, 0.875, 0.0, 0.0, 1.5600, -0.00110),
    ("P", 4.20, 0.200, 0.0, 0.0, 38.7924, -0.00110),
    ("SA", 4.00, 0.200, 2.5, 1.0, 33.5103, -0.00214),
    ("S", 4.00, 0.200, 0.0, 0.0, 33.5103, -0.00214),
    ("CL", 4.09, 0.276, 0.0, 0.0, 35.8235, -0.00110),
    ("CA", 1.98, 0.550, 0.0, 0.0, 2.7700, -0.00110),
    ("MN", 1.30, 0.875, 0.0, 0.0, 2.1400, -0.00110),
    ("FE", 1.30, 0.010, 0.0, 0.0, 1.8400, -0.00110),
    ("ZN", 1.48, 0.550, 0.0, 0.0, 1.7000, -0.00110),
    ("BR", 4.33, 0.389, 0.0, 0.0, 42.5661, -0.00110),
    ("I", 4.72, 0.550, 0.0, 0.0, 55.0585, -0.00110),
    ("SI", 4.10, 0.200, 0.0, 0.0, 35.8235, -0.00143),
    ("B", 3.84, 0.155, 0.0, 0.0, 29.6478, -0.00152),
]
AD4_INDEX = {row[0]: i for i, row in enumerate(_AD4)}
_ACCEPTORS = {"NA", "NS", "OA", "OS", "SA"}
_HDONOR_H = {"HD", "HS"}

# AD4.2 "bound same as unbound" free-energy coefficients
# (unbound_models[0], getparameters.h:55-57).
ELEC_SCALE_FACTOR = 332.06363
AD4_COEFF_VDW = 0.1662
AD4_COEFF_HB = 0.1209
AD4_COEFF_ELEC = ELEC_SCALE_FACTOR * 0.1406
AD4_COEFF_DESOLV = 0.1322


def is_hbond(t1: str, t2: str) -> bool:
    """is_H_bond, processligand.cpp:1521-1538."""
    return (t1 in _HDONOR_H and t2 in _ACCEPTORS) or (t2 in _HDONOR_H and t1 in _ACCEPTORS)


@dataclass
class Ligand:
    """A ligand in pdbqt atom order plus its device tables (numpy arrays)."""

    names: list
    types: list                      # per atom AD4 type name
    xyz: np.ndarray                  # (N,3) Angstrom, pdbqt frame
    charges: np.ndarray              # (N,)
    bonds: list                      # list of (i,j)
    rotbonds: np.ndarray             # (T,2) parent-side atom, child-side atom
    atom_rotbonds: np.ndarray        # (N,T) uint8
    rigid_id: np.ndarray             # (N,)
    pdbqt_lines: list = field(default_factory=list)
    type_names: list = field(default_factory=list)   # ligand type table, first appearance order

    @property
    def num_atoms(self) -> int:
        return len(self.types)

    @property
    def num_rotbonds(self) -> int:
        return int(self.rotbonds.shape[0])

    def pdbqt(self) -> str:
        return "\n".join(self.pdbqt_lines) + "\n"


# ----------------------------------------------------------------------------
# geometry
# ----------------------------------------------------------------------------
def _rand_unit(rng):
    v = rng.normal(size=3)
    return v / np.linalg.norm(v)


def _place(rng, pos, nbrs, parent, length, min_13=2.35, min_far=2.9, tries=400):
    """Direction from `parent` at ~tetrahedral angles to its bonds, keeping
    non-bonded distances above the reference's 2.1 A bond cutoff
    (get_bonds, processligand.cpp:576-600)."""
    p = pos[parent]
    bvec = [pos[k] - p for k in nbrs[parent]]
    bvec = [b / np.linalg.norm(b) for b in bvec]
    first13 = set(nbrs[parent])
    for _ in range(tries):
        d = _rand_unit(rng)
        if any(np.dot(d, b) > math.cos(math.radians(100.0)) for b in bvec):
            continue
        x = p + length * d
        ok = True
        for k, q in enumerate(pos):
            if k == parent:
                continue
            dist = np.linalg.norm(x - q)
            lim = min_13 if k in first13 else min_far
            if dist < lim:
                ok = False
                break
        if ok:
            return x
    return None


def make_ligand(num_atoms: int = 30, num_torsions: int = 6, seed: int = 0,
                charge_range: float = 0.5) -> Ligand:
    """Tree-shaped ligand of `num_atoms` atoms (types C, A, N, OA, HD), with
    `num_torsions` rotatable bonds chosen among bonds between two non-terminal
    heavy atoms; charges uniform in [-charge_range, charge_range]."""
    rng = np.random.default_rng(seed)
    for attempt in range(200):
        lig = _try_make(rng, num_atoms, num_torsions, charge_range)
        if lig is not None:
            return lig
    raise RuntimeError("could not build ligand %d/%d (seed %d)" % (num_atoms, num_torsions, seed))


def _try_make(rng, n_atoms, n_tors, crange):
    n_hd = 0 if n_atoms < 6 else max(1, int(round(0.1 * n_atoms)))
    n_heavy = n_atoms - n_hd
    pos = [np.zeros(3)]
    nbrs = [[]]
    for i in range(1, n_heavy):
        cands = [k for k in range(len(pos)) if len(nbrs[k]) < 3]
        # chain-like growth: prefer recently added atoms
        w = np.array([1.0 + 0.15 * k for k in cands])
        placed = None
        for _ in range(30):
            par = int(rng.choice(cands, p=w / w.sum()))
            placed = _place(rng, pos, nbrs, par, 1.52)
            if placed is not None:
                break
        if placed is None:
            return None
        pos.append(placed)
        nbrs.append([par])
        nbrs[par].append(i)
    deg = [len(n) for n in nbrs]
    # heavy types respecting valence (OA <= 2 bonds incl. H, N <= 3)
    # Heteroatoms bond only to carbons: the reference's distance rule accepts
    # N-O bonds only up to 1.50 A (get_bonds, processligand.cpp:592-595).
    types = []
    for k in range(n_heavy):
        r = rng.random()
        hetero_nb = any(types[b] in ("N", "OA") for b in nbrs[k] if b < k)
        if not hetero_nb and deg[k] <= 1 and r < 0.30:
            types.append("OA")
        elif not hetero_nb and deg[k] <= 2 and r < 0.45:
            types.append("N")
        elif r < 0.75:
            types.append("C")
        else:
            types.append("A")
    for k in range(n_heavy):
        if types[k] in ("N", "OA") and any(types[b] in ("N", "OA") for b in nbrs[k]):
            types[k] = "C"
    # hydrogen donors on N / OA with a free valence
    sites = [k for k in range(n_heavy) if (types[k] == "OA" and deg[k] < 2) or (types[k] == "N" and deg[k] < 3)]
    rng.shuffle(sites)
    while len(sites) < n_hd:
        leaves = [k for k in range(n_heavy) if deg[k] == 1 and types[k] in ("C", "A") and k not in sites
                  and not any(types[b] in ("N", "OA") for b in nbrs[k])]
        if not leaves:
            return None
        k = int(rng.choice(leaves))
        types[k] = "OA"
        sites.append(k)
    for s in sites[:n_hd]:
        x = _place(rng, pos, nbrs, s, 1.02, min_13=1.5, min_far=2.0)
        if x is None:
            return None
        pos.append(x)
        nbrs.append([s])
        nbrs[s].append(len(pos) - 1)
        types.append("HD")
    n = len(pos)
    bonds = sorted({(min(a, b), max(a, b)) for a in range(n) for b in nbrs[a]})
    heavy_deg = [sum(1 for b in nbrs[k] if types[b] != "HD") for k in range(n)]
    internal = [(a, b) for (a, b) in bonds
                if types[a] != "HD" and types[b] != "HD" and heavy_deg[a] >= 2 and heavy_deg[b] >= 2]
    if len(internal) < n_tors:
        # fall back to any heavy-heavy bond whose smaller side has >= 2 atoms
        internal = [(a, b) for (a, b) in bonds if types[a] != "HD" and types[b] != "HD"
                    and len(nbrs[a]) >= 2 and len(nbrs[b]) >= 2]
        if len(internal) < n_tors:
            return None
    pick = rng.choice(len(internal), size=n_tors, replace=False) if n_tors else []
    rot = {internal[int(i)] for i in pick}
    charges = rng.uniform(-crange, crange, size=n)
    return _torsion_tree(np.array(pos), types, charges, nbrs, bonds, rot, rng)


def _torsion_tree(pos, types, charges, nbrs, bonds, rot, rng):
    """Orders atoms as a pdbqt torsion tree and derives rigid fragments and
    the rotatable-bond tables (parse_liganddata, processligand.cpp:1226-1290)."""
    n = len(types)
    rotset = set(rot)

    def is_rot(a, b):
        return (min(a, b), max(a, b)) in rotset

    # rigid fragments
    frag = [-1] * n
    nf = 0
    for s in range(n):
        if frag[s] >= 0:
            continue
        stack = [s]
        frag[s] = nf
        while stack:
            a = stack.pop()
            for b in nbrs[a]:
                if frag[b] < 0 and not is_rot(a, b):
                    frag[b] = nf
                    stack.append(b)
        nf += 1
    members = [[a for a in range(n) if frag[a] == f] for f in range(nf)]
    # root = fragment minimising the deepest branch depth
    fadj = {f: [] for f in range(nf)}
    for (a, b) in rot:
        fadj[frag[a]].append((a, b))
        fadj[frag[b]].append((b, a))

    def depth(root):
        seen = {root}
        best = 0
        stack = [(root, 0)]
        while stack:
            f, d = stack.pop()
            best = max(best, d)
            for (x, y) in fadj[f]:
                if frag[y] not in seen:
                    seen.add(frag[y])
                    stack.append((frag[y], d + 1))
        return best

    root = min(range(nf), key=lambda f: (depth(f), -len(members[f]), f))
    order = []          # new index -> old index
    lines = ["REMARK  synthetic ligand", "ROOT"]
    events = []         # ('atom', old) / ('branch', pa_old, ch_old) / ('end', ...)

    def emit(f, entry_atom, parent_frag):
        atoms = members[f]
        if entry_atom is not None:
            atoms = [entry_atom] + [a for a in atoms if a != entry_atom]
        for a in atoms:
            events.append(("atom", a))
        if entry_atom is None:
            events.append(("endroot",))
        kids = sorted([(x, y) for (x, y) in fadj[f] if frag[y] != parent_frag], key=lambda e: e[1])
        for (x, y) in kids:
            events.append(("branch", x, y))
            emit(frag[y], y, f)
            events.append(("end", x, y))

    emit(root, None, -1)
    newidx = {}
    for ev in events:
        if ev[0] == "atom":
            newidx[ev[1]] = len(order)
            order.append(ev[1])
    names = []
    cnt = {}
    for old in order:
        t = types[old]
        el = "H" if t == "HD" else ("O" if t == "OA" else ("N" if t == "N" else "C"))
        cnt[el] = cnt.get(el, 0) + 1
        names.append("%s%d" % (el, cnt[el]))
    for ev in events:
        if ev[0] == "atom":
            i = newidx[ev[1]]
            x, y, z = pos[ev[1]]
            lines.append("%-6s%5d %-4s %3s %1s%4d    %8.3f%8.3f%8.3f%6.2f%6.2f    %6.3f %-2s" % (
                "HETATM", i + 1, names[i], "LIG", "A", 1, x, y, z, 1.0, 0.0, charges[ev[1]], types[ev[1]]))
        elif ev[0] == "endroot":
            lines.append("ENDROOT")
        elif ev[0] == "branch":
            lines.append("BRANCH %3d %3d" % (newidx[ev[1]] + 1, newidx[ev[2]] + 1))
        else:
            lines.append("ENDBRANCH %3d %3d" % (newidx[ev[1]] + 1, newidx[ev[2]] + 1))
    lines.append("TORSDOF %d" % len(rot))
    # rotatable bonds in ENDBRANCH order; atoms rotate while their branch is open
    rotbonds, open_br, rb_of_branch = [], [], {}
    rig = np.zeros(n, dtype=np.int32)
    cur_id = 1
    reserved = 1
    atom_rot_tmp = np.zeros((n, len(rot)), dtype=np.uint8)
    branches = []
    ai = 0
    for ev in events:
        if ev[0] == "atom":
            for bi, st in enumerate(open_br):
                if st:
                    atom_rot_tmp[ai, bi] = 1
            rig[ai] = cur_id
            ai += 1
        elif ev[0] == "branch":
            branches.append((newidx[ev[1]], newidx[ev[2]]))
            open_br.append(True)
            reserved += 1
            cur_id = reserved
        elif ev[0] == "end":
            key = (newidx[ev[1]], newidx[ev[2]])
            rotbonds.append(key)
            open_br[branches.index(key)] = False
            cur_id -= 1
    T = len(rotbonds)
    atom_rotbonds = np.zeros((n, T), dtype=np.uint8)
    for i, key in enumerate(rotbonds):
        atom_rotbonds[:, i] = atom_rot_tmp[:, branches.index(key)]
    new_pos = pos[order]
    new_types = [types[o] for o in order]
    new_q = charges[order]
    new_bonds = sorted({(min(newidx[a], newidx[b]), max(newidx[a], newidx[b])) for (a, b) in bonds})
    type_names = []
    for t in new_types:
        if t not in type_names:
            type_names.append(t)
    # charges as the pdbqt text carries them (3 decimals)
    new_q = np.array([float("%6.3f" % q) for q in new_q])
    new_pos = np.array([[float("%8.3f" % c) for c in p] for p in new_pos])
    return Ligand(names=names, types=new_types, xyz=new_pos, charges=new_q, bonds=new_bonds,
                  rotbonds=np.array(rotbonds, dtype=np.int32).reshape(-1, 2),
                  atom_rotbonds=atom_rotbonds, rigid_id=rig, pdbqt_lines=lines, type_names=type_names)


# ----------------------------------------------------------------------------
# device tables
# ----------------------------------------------------------------------------
def intra_pairs(lig: Ligand) -> np.ndarray:
    """get_intraE_contributors (processligand.cpp:250-457) for ligands
    without CG/G0/W atoms."""
    n = lig.num_atoms
    contrib = np.ones((n, n), dtype=bool)
    np.fill_diagonal(contrib, False)
    rig = lig.rigid_id
    same = rig[:, None] == rig[None, :]
    contrib[same] = False
    adj = np.zeros((n, n), dtype=bool)
    for a, b in lig.bonds:
        adj[a, b] = adj[b, a] = True
    # 1-2, 1-3, 1-4 neighbours (the reference expands neighbours twice)
    reach = adj.copy()
    for _ in range(2):
        reach = reach | ((reach.astype(np.int32) @ adj.astype(np.int32)) > 0)
    contrib[reach] = False
    for (a, b) in lig.rotbonds:
        contrib[rig == rig[a], b] = False
        contrib[b, rig == rig[a]] = False
        contrib[rig == rig[b], a] = False
        contrib[a, rig == rig[b]] = False
    T = lig.num_rotbonds
    for i in range(T - 1):
        for j in range(i + 1, T):
            a, b = lig.rotbonds[i]
            a2, b2 = lig.rotbonds[j]
            if rig[a] == rig[a2]:
                contrib[b, b2] = contrib[b2, b] = False
            if rig[a] == rig[b2]:
                contrib[b, a2] = contrib[a2, b] = False
            if rig[b] == rig[a2]:
                contrib[a, b2] = contrib[b2, a] = False
            if rig[b] == rig[b2]:
                contrib[a, a2] = contrib[a2, a] = False
    iu = np.triu_indices(n, 1)
    sel = contrib[iu]
    return np.stack([iu[0][sel], iu[1][sel]], axis=1).astype(np.int32)


def vdw_tables(type_names, coeff_vdw=AD4_COEFF_VDW, coeff_hb=AD4_COEFF_HB):
    """get_VWpars (processligand.cpp:930-1043) and the AC/BD selection of
    prepare_const_fields_for_gpu (calcenergy.cpp:171-206)."""
    nt = len(type_names)
    reqm = np.zeros(nt, np.uint32)
    vw_exp = np.zeros((nt, nt), np.uint16)
    reqm_ab = np.zeros((nt, nt), np.float64)
    ac = np.zeros((nt, nt), np.float64)
    bd = np.zeros((nt, nt), np.float64)
    ds_s = np.zeros(nt, np.float64)
    ds_v = np.zeros(nt, np.float64)
    for i, t in enumerate(type_names):
        reqm[i] = AD4_INDEX[t]
        ds_v[i] = _AD4[AD4_INDEX[t]][5]
        ds_s[i] = _AD4[AD4_INDEX[t]][6]
    for i, ti in enumerate(type_names):
        for j, tj in enumerate(type_names):
            pi, pj = _AD4[AD4_INDEX[ti]], _AD4[AD4_INDEX[tj]]
            if is_hbond(ti, tj):
                eps = coeff_hb * pi[4] * pj[4]
                r = pi[3] + pj[3]
                vw_exp[i, j] = (12 << 8) + 10
                reqm_ab[i, j] = r
                ac[i, j] = 5 * eps * r ** 12
                bd[i, j] = 6 * eps * r ** 10
            else:
                eps = coeff_vdw * math.sqrt(pi[2] * pj[2])
                r = 0.5 * (pi[1] + pj[1])
                vw_exp[i, j] = (12 << 8) + 6
                reqm_ab[i, j] = r
                ac[i, j] = eps * r ** 12
                bd[i, j] = 2 * eps * r ** 6
    return reqm, vw_exp, reqm_ab, ac, bd, ds_s, ds_v


@dataclass
class LigandTables:
    """Arrays backing an adgpu_ligand struct (see include/adgpu.h)."""

    num_atoms: int
    true_ligand_atoms: int
    num_rotbonds: int
    num_atypes: int
    charges: np.ndarray
    atom_types: np.ndarray
    atom_map: np.ndarray
    ignore_inter: np.ndarray
    ref_coords: np.ndarray
    rotbond_moving: np.ndarray
    rotbond_unit: np.ndarray
    rotbond_atoms: np.ndarray
    atom_rotbonds: np.ndarray
    pairs: np.ndarray
    atype_reqm: np.ndarray
    vw_exp: np.ndarray
    reqm_ab: np.ndarray
    vw_ac: np.ndarray
    vw_bd: np.ndarray
    ds_s: np.ndarray
    ds_v: np.ndarray

    @property
    def num_pairs(self) -> int:
        return int(self.pairs.shape[0])

    @property
    def num_genes(self) -> int:
        return self.num_rotbonds + 6


def ligand_tables(lig: Ligand, spacing: float, map_types) -> LigandTables:
    """Device tables as docking_with_gpu hands them to the kernels: centred
    at the ligand's geometric centre and scaled to grid units
    (getparameters.cpp:2624-2630), vectors per calcenergy.cpp / processligand.cpp."""
    n = lig.num_atoms
    xyz = lig.xyz.astype(np.float64)
    xyz = (xyz - xyz.mean(axis=0)) / spacing
    T = lig.num_rotbonds
    mov = np.zeros((T, 3))
    uni = np.zeros((T, 3))
    for t, (a, b) in enumerate(lig.rotbonds):
        A, B = xyz[a], xyz[b]
        d = B - A
        u = d / math.sqrt(float(np.dot(d, d)))
        u = np.where(u >= 1, 0.999999, u)
        uni[t] = u
        tt = float(np.dot(d, -A)) / float(np.dot(d, d))
        mov[t] = tt * d + A
    reqm, vw_exp, reqm_ab, ac, bd, ds_s, ds_v = vdw_tables(lig.type_names)
    tindex = {t: i for i, t in enumerate(lig.type_names)}
    atom_types = np.array([tindex[t] for t in lig.types], np.int32)
    atom_map = np.array([list(map_types).index(t) for t in lig.types], np.int32)
    return LigandTables(
        num_atoms=n, true_ligand_atoms=n, num_rotbonds=T, num_atypes=len(lig.type_names),
        charges=lig.charges.astype(np.float32), atom_types=atom_types, atom_map=atom_map,
        ignore_inter=np.zeros(n, np.uint8), ref_coords=xyz.astype(np.float32).reshape(-1),
        rotbond_moving=mov.astype(np.float32).reshape(-1), rotbond_unit=uni.astype(np.float32).reshape(-1),
        rotbond_atoms=lig.rotbonds.astype(np.int32).reshape(-1),
        atom_rotbonds=np.ascontiguousarray(lig.atom_rotbonds.astype(np.uint8)).reshape(-1),
        pairs=intra_pairs(lig).reshape(-1), atype_reqm=reqm,
        vw_exp=vw_exp.reshape(-1), reqm_ab=reqm_ab.astype(np.float32).reshape(-1),
        vw_ac=ac.astype(np.float32).reshape(-1), vw_bd=bd.astype(np.float32).reshape(-1),
        ds_s=ds_s.astype(np.float32), ds_v=ds_v.astype(np.float32))


# ----------------------------------------------------------------------------
# receptor maps
# ----------------------------------------------------------------------------
@dataclass
class Maps:
    size: tuple                  # points per axis
    spacing: float
    center: tuple
    type_names: list             # type maps in order; e and d follow
    values: np.ndarray           # (nmaps, nz, ny, nx) float32, nmaps = ntypes + 2

    @property
    def num_map_atypes(self) -> int:
        return len(self.type_names)

    def packed(self) -> np.ndarray:
        """Packed-4 layout of get_gridvalues (processgrid.cpp:234-249)."""
        return pack_maps(self.values)

    @property
    def size_angstr(self):
        return tuple((s - 1) * self.spacing for s in self.size)


def pack_maps(values: np.ndarray) -> np.ndarray:
    nm, nz, ny, nx = values.shape
    out = np.zeros((nm, nz, ny, nx, 4), np.float32)
    out[..., 0] = values
    out[:, :, :-1, :, 1] = values[:, :, 1:, :]
    out[:, :-1, :, :, 2] = values[:, 1:, :, :]
    out[:, :-1, :-1, :, 3] = values[:, 1:, 1:, :]
    return out.reshape(-1)


def make_maps(type_names, npts: int = 60, spacing: float = 0.375, seed: int = 0,
              n_wells: int = 200, n_cores: int = 60, clip: float = 1e5) -> Maps:
    """Per map: a seeded sum of `n_wells` Gaussian wells (depth U[-1.0,-0.1]
    kcal/mol, width U[0.8, 2.0] A) plus `n_cores` repulsive r^-12 cores,
    clipped at `clip`.  The e map is a signed screened potential, the d map a
    small positive desolvation field."""
    rng = np.random.default_rng(seed)
    n = (npts, npts, npts) if np.isscalar(npts) else tuple(npts)
    box = np.array([(k - 1) * spacing for k in n])
    ax = [np.arange(k, dtype=np.float32) * spacing for k in n]
    X = ax[0][None, None, :]
    Y = ax[1][None, :, None]
    Z = ax[2][:, None, None]
    nmaps = len(type_names) + 2
    vals = np.zeros((nmaps, n[2], n[1], n[0]), np.float32)
    cores = rng.uniform(0.0, 1.0, size=(n_cores, 3)) * box
    core_r = rng.uniform(1.4, 2.0, size=n_cores)
    core_field = np.zeros((n[2], n[1], n[0]), np.float32)
    for c, r0 in zip(cores, core_r):
        d2 = (X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2
        d2 = np.maximum(d2, 1e-6)
        core_field += (r0 * r0 / d2) ** 6
    for m in range(nmaps):
        wells = rng.uniform(0.0, 1.0, size=(n_wells, 3)) * box
        depth = rng.uniform(-1.0, -0.1, size=n_wells)
        width = rng.uniform(0.8, 2.0, size=n_wells)
        f = np.zeros_like(core_field)
        for c, dp, w in zip(wells, depth, width):
            lo = np.maximum(((c - 3 * w) / spacing).astype(int), 0)
            hi = np.minimum(((c + 3 * w) / spacing).astype(int) + 2, np.array(n))
            sx, sy, sz = slice(lo[0], hi[0]), slice(lo[1], hi[1]), slice(lo[2], hi[2])
            d2 = (X[:, :, sx] - c[0]) ** 2 + (Y[:, sy, :] - c[1]) ** 2 + (Z[sz, :, :] - c[2]) ** 2
            f[sz, sy, sx] += (dp * np.exp(-d2 / (2 * w * w))).astype(np.float32)
        if m < len(type_names):
            v = f + core_field * rng.uniform(0.5, 1.5)
        elif m == len(type_names):
            v = f * rng.choice([-1.0, 1.0], size=1)[0] * 2.0 + 0.05 * core_field
        else:
            v = -0.3 * f + 0.01 * core_field
        v = np.minimum(v, clip)
        # AutoGrid writes 3 decimals
        vals[m] = np.round(v, 3)
    center = tuple(float(x) for x in (box / 2.0))
    return Maps(size=n, spacing=spacing, center=center, type_names=list(type_names), values=vals)


def write_maps(maps: Maps, directory: str, base: str = "synth") -> str:
    """AutoGrid-style .fld + .map files readable by get_gridinfo /
    get_gridvalues (processgrid.cpp:28-253).  Returns the .fld path."""
    os.makedirs(directory, exist_ok=True)
    nx, ny, nz = maps.size
    labels = ["%s-affinity" % t for t in maps.type_names] + ["Electrostatics", "Desolvation"]
    files = ["%s.%s.map" % (base, t) for t in maps.type_names] + ["%s.e.map" % base, "%s.d.map" % base]
    cx, cy, cz = maps.center
    hdr = ["GRID_PARAMETER_FILE synth.gpf", "GRID_DATA_FILE %s.maps.fld" % base,
           "MACROMOLECULE synth_receptor.pdbqt", "SPACING %.3f" % maps.spacing,
           "NELEMENTS %d %d %d" % (nx - 1, ny - 1, nz - 1), "CENTER %.3f %.3f %.3f" % (cx, cy, cz)]
    for m, fn in enumerate(files):
        with open(os.path.join(directory, fn), "w") as fh:
            fh.write("\n".join(hdr) + "\n")
            flat = maps.values[m].reshape(-1)
            fh.write("\n".join("%.3f" % v for v in flat))
            fh.write("\n")
    fld = ["# AVS field file", "#SPACING %.3f" % maps.spacing,
           "#NELEMENTS %d %d %d" % (nx - 1, ny - 1, nz - 1), "#CENTER %.3f %.3f %.3f" % (cx, cy, cz),
           "#MACROMOLECULE synth_receptor.pdbqt", "#GRID_PARAMETER_FILE synth.gpf", "#",
           "ndim=3", "dim1=%d" % nx, "dim2=%d" % ny, "dim3=%d" % nz, "nspace=3",
           "veclen=%d" % len(files), "data=float", "field=uniform"]
    fld += ["label=%s" % lb for lb in labels]
    fld += ["variable %d file=%s filetype=ascii skip=6" % (i + 1, fn) for i, fn in enumerate(files)]
    path = os.path.join(directory, "%s.maps.fld" % base)
    with open(path, "w") as fh:
        fh.write("\n".join(fld) + "\n")
    return path


def make_flexres(num_atoms: int = 12, num_torsions: int = 4, seed: int = 0, offset=(0.0, 0.0, 0.0),
                 charge_range: float = 0.3, resname: str = "FLX", resnum: int = 1) -> str:
    """One flexible residue as the reference's flexres pdbqt reader takes it
    (parse_liganddata, processligand.cpp:1124-1300: each ROOT starts a
    residue; its first two atoms get ignore_inter): a synthetic tree of
    `num_atoms` atoms with `num_torsions` rotatable bonds, translated by
    `offset` (A, receptor frame)."""
    lig = make_ligand(num_atoms, num_torsions, seed=seed, charge_range=charge_range)
    out = ["BEGIN_RES %s A%4d" % (resname, resnum)]
    for line in lig.pdbqt_lines:
        if line.startswith(("REMARK", "TORSDOF")):
            continue
        if line.startswith("HETATM"):
            x, y, z = float(line[30:38]) + offset[0], float(line[38:46]) + offset[1], float(line[46:54]) + offset[2]
            line = "ATOM  " + line[6:17] + "%3s A%4d    %8.3f%8.3f%8.3f" % (resname, resnum, x, y, z) + line[54:]
        out.append(line)
    out.append("END_RES %s A%4d" % (resname, resnum))
    return "\n".join(out) + "\n"


def flexres_types(text: str) -> list:
    """AD4 atom types of a flexres pdbqt text, in order of appearance."""
    seen = []
    for line in text.splitlines():
        if line.startswith("ATOM"):
            t = line[77:].split()[0]
            if t not in seen:
                seen.append(t)
    return seen


def write_ligand(lig: Ligand, path: str) -> str:
    with open(path, "w") as fh:
        fh.write(lig.pdbqt())
    return path


# ----------------------------------------------------------------------------
# genotypes
# ----------------------------------------------------------------------------
def random_poses(n: int, num_genes: int, grid_size, seed: int = 0) -> np.ndarray:
    """n random poses: translation uniform over the grid (grid units),
    orientation a uniform random quaternion converted to (phi, theta, alpha)
    as gen_initpop_and_reflig does (getparameters.cpp:2594-2619), torsions
    uniform in [-180, 180) degrees.  Returns (n, 64) float32."""
    rng = np.random.default_rng(seed)
    g = np.zeros((n, 64), np.float32)
    for k in range(3):
        g[:, k] = rng.uniform(0.0, grid_size[k] - 1, size=n)
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    qw = np.sqrt(1 - u1) * np.sin(2 * np.pi * u2)
    qx = np.sqrt(1 - u1) * np.cos(2 * np.pi * u2)
    qy = np.sqrt(u1) * np.sin(2 * np.pi * u3)
    qz = np.sqrt(u1) * np.cos(2 * np.pi * u3)
    rot = 2 * np.arccos(np.clip(qw, -1, 1))
    s = np.sqrt(np.maximum(1 - qw * qw, 0))
    small = s < 1e-3
    s = np.where(small, 1.0, s)
    x, y, z = qx / s, qy / s, qz / s
    theta = np.arccos(np.clip(z, -1, 1))
    phi = np.arctan2(y, x)
    deg = 0.0174533
    g[:, 3] = phi / deg
    g[:, 4] = theta / deg
    g[:, 5] = rot / deg
    nt = num_genes - 6
    if nt > 0:
        g[:, 6:6 + nt] = rng.uniform(-180.0, 180.0, size=(n, nt))
    return g
