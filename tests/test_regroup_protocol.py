"""The ADADELTA regrouping protocol of k_ls_ad.hip as a step-level model (CPU).

Regrouping starts once a launch's queue and hand-off ring are drained: every
team then either runs one search to its end or has none.  Teams sit in
wavefronts of W (64 / team width) that move in lockstep; the model runs one
wavefront iteration at a time, picked by a scheduler, and splits the one
place where other wavefronts can act inside an iteration that matters here:
between a hand-over's seat claim and its compare-and-swap on the seat entry.
Per wavefront iteration, in the kernel's order:

* every seated team looks at its seat entry as it read it before its
  previous evaluation (an entry is written once, so a stale read only
  delays); a search there is taken (its saved iteration count restored), and
  with no other team of the wavefront running a search, an empty seat is
  given up by compare-and-swap (which takes the search instead when one
  arrived first);
* every idle, unseated team beside a team with a search takes a seat ticket
  (counter [5]) below the seat capacity;
* the wavefront ends when no team has a search or a seat;
* every team with a search runs one iteration; a search that is the only
  work of its wavefront and has run `rk` iterations there claims the oldest
  seat ticket nobody claimed (counter [6] < [5]), if that seat is older than
  every seat its own wavefront holds, and, in a later step, hands itself
  over by compare-and-swap, or keeps running when the seat was given up
  first.  A search that came from a seat does not move again.  The age rule makes two half-idle wavefronts merge toward the older
  seat instead of trading their searches.

A team that started its search before the queue and the ring ran out has not
seen them drained; it finds out when it next looks (after its search ends or
moves), and a wavefront ends only when every team has looked.

Checked for wavefronts of 2 and 4 teams, several search-length mixes and
schedules (uniform; hand-overs completed only in 3 % of the steps; one
wavefront starved until the others wait): every search runs exactly its
length once, every search handed to a seat is taken exactly once, seat and
ticket indices stay below their bounds, the model finishes within a bound,
and (uniform schedule) regrouping leaves fewer wavefront steps with idle
teams than no regrouping.  A seat given up with a plain store instead of a
compare-and-swap loses searches; a seated team that never gives up its seat,
or a team whose search moved and that never looks for work again, waits
forever: all three fail these checks.
"""
import random

import pytest

UNSET, CANCELLED = -1, -3


class Team:
    def __init__(self, t, it):
        self.t = t                 # search in progress (None: idle)
        self.iters = it
        self.done = t is None      # found the queue and the ring used up (set when it looks)
        self.seated = False
        self.seat = None
        self.since = 0
        self.moved = False         # the search came from a seat: it does not move again
        self.peek = UNSET          # the seat entry as read in the team's previous iteration
        self.swap = None           # claimed seat ticket whose entry is still to be swapped


def cas(arr, i, expect, new):
    """Compare-and-swap; returns the entry's value afterwards."""
    if arr[i] == expect:
        arr[i] = new
        return new
    return arr[i]


def simulate(lengths, start, W, rk, pick, seed, cancel_by_cas=True, give_up=True, acquire_after_move=True,
             max_steps=1_000_000):
    """start[k] = (search, iterations already run) or None for team k."""
    rng = random.Random(seed)
    ntot = len(lengths)
    cap = 2 * ntot
    seats = [UNSET] * cap
    q5 = q6 = 0
    saved = {}
    ran = [0] * ntot
    finished = [0] * ntot
    taken = []
    teams = [Team(*(s if s is not None else (None, 0))) for s in start]
    for tm in teams:
        if tm.t is not None:
            ran[tm.t] += tm.iters
    waves = [teams[w * W:(w + 1) * W] for w in range(len(teams) // W)]
    live = set(range(len(waves)))
    idle_steps = steps = regroups = 0
    for _ in range(max_steps):
        if not live:
            break
        w = pick(rng, waves, sorted(live))
        wv = waves[w]
        # a hand-over claimed in the previous iteration completes first; so
        # does a give-up by plain store (the mutation: read, then store later)
        for tm in wv:
            if getattr(tm, "giving_up", False):
                tm.giving_up = False
                seats[tm.seat] = CANCELLED
                tm.seated = False
            if tm.swap is not None:
                tk, tm.swap = tm.swap, None
                saved[tm.t] = tm.iters
                if cas(seats, tk, UNSET, tm.t) == tm.t:
                    regroups += 1
                    tm.t = None
                    if acquire_after_move:
                        tm.done = True     # its next acquisition finds everything used up
                else:
                    del saved[tm.t]
        active = [tm.t is not None for tm in wv]
        for k, tm in enumerate(wv):
            others = any(a for j, a in enumerate(active) if j != k)
            if tm.seated:
                v = tm.peek                       # read one iteration earlier (entries are written once)
                if v == UNSET and not others and give_up:
                    if cancel_by_cas:
                        v = cas(seats, tm.seat, UNSET, CANCELLED)
                    else:
                        tm.giving_up = True
                        continue
                if v >= 0:
                    assert v in saved, "a search taken twice or never saved"
                    tm.seated = False
                    tm.t, tm.iters = v, saved.pop(v)
                    tm.since = 0
                    tm.moved = True
                    taken.append(v)
                elif v == CANCELLED:
                    tm.seated = False
            elif tm.t is None and tm.done and others:
                h = q5
                q5 += 1
                if h < cap:
                    tm.seated, tm.seat, tm.peek = True, h, UNSET
        if not any(tm.t is not None or tm.seated or not tm.done for tm in wv):
            live.discard(w)
            continue
        if not any(tm.t is not None for tm in wv):
            continue
        for tm in wv:
            if tm.seated:
                tm.peek = seats[tm.seat]
        steps += 1
        if any(tm.t is None for tm in wv):
            idle_steps += 1
        busy = [tm.t is not None or not tm.done for tm in wv]
        # a search moves only to a seat older than every seat its wavefront holds
        lim = min([tm.seat for tm in wv if tm.seated] or [cap])
        for k, tm in enumerate(wv):
            if tm.t is None:
                continue
            alone = not any(b for j, b in enumerate(busy) if j != k)
            tm.iters += 1
            ran[tm.t] += 1
            if tm.iters == lengths[tm.t]:
                finished[tm.t] += 1
                tm.t = None
                tm.done = True
            elif rk > 0 and alone and not tm.moved:
                tm.since += 1
                if tm.since >= rk and q6 < min(q5, cap, lim):
                    tm.swap = q6
                    q6 += 1
    else:
        raise AssertionError("model did not finish: a wait that never ends")
    assert all(tm.t is None and not tm.seated for tm in teams)
    assert finished == [1 if lengths[i] > 0 else 0 for i in range(ntot)]
    assert ran == list(lengths), "a search lost or run twice"
    assert not saved and len(taken) == regroups   # each hand-over taken once (a search may move again)
    assert q6 <= cap
    return regroups, idle_steps, steps


def uniform(rng, waves, live):
    return rng.choice(live)


def swaps_last(rng, waves, live):
    """A wavefront with a hand-over in flight moves only in 3 % of the steps."""
    others = [w for w in live if not any(tm.swap is not None for tm in waves[w])]
    return rng.choice(others) if others and rng.random() < 0.97 else rng.choice(live)


def starve_first(rng, waves, live):
    """Wavefront 0 moves only when every other live wavefront has no search."""
    busy = [w for w in live if w != 0 and any(tm.t is not None for tm in waves[w])]
    return rng.choice(busy) if busy else rng.choice(live)


def _case(nwaves, W, seed):
    rng = random.Random(seed)
    n = nwaves * W
    lengths, start = [], []
    for k in range(n):
        if rng.random() < 0.75:
            total = 300 if rng.random() < 0.1 else rng.randint(2, 200)
            lengths.append(total)
            start.append((len(lengths) - 1, rng.randint(0, total - 1)))
        else:
            start.append(None)
    return lengths, start


@pytest.mark.parametrize("pick", [uniform, swaps_last, starve_first])
@pytest.mark.parametrize("nwaves,W,rk", [(40, 2, 8), (25, 4, 8), (64, 2, 1), (8, 2, 32), (1, 2, 8), (30, 4, 3)])
def test_every_search_runs_once_and_every_wait_ends(pick, nwaves, W, rk):
    lengths, start = _case(nwaves, W, nwaves * 31 + W)
    regroups, idle, steps = simulate(lengths, start, W, rk, pick, seed=nwaves + rk)
    base = simulate(lengths, start, W, 0, pick, seed=nwaves + rk)
    assert base[0] == 0
    if nwaves > 1:
        assert regroups > 0                          # the regrouping path was exercised
        if pick is uniform:
            assert idle < base[1]                    # fewer wavefront steps beside idle teams


def test_plain_store_give_up_loses_searches():
    lengths, start = _case(40, 2, 7)
    with pytest.raises(AssertionError):
        for seed in range(20):
            simulate(lengths, start, 2, 8, swaps_last, seed=seed, cancel_by_cas=False)


def test_moved_team_that_never_looks_again_waits_forever():
    lengths, start = _case(10, 2, 3)
    with pytest.raises(AssertionError, match="never ends"):
        simulate(lengths, start, 2, 8, uniform, seed=0, acquire_after_move=False, max_steps=200_000)


def test_seat_never_given_up_waits_forever():
    lengths, start = _case(25, 4, 5)
    with pytest.raises(AssertionError, match="never ends"):
        simulate(lengths, start, 4, 8, uniform, seed=0, give_up=False, max_steps=200_000)
