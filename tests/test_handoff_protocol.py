"""The ADADELTA hand-off protocol of k_ls_ad.hip as a step-level model (CPU).

The kernel's teams run asynchronously: each team's loop iteration is split
here into the steps that other teams can interleave with (a queue claim, the
ring write that follows it, one search iteration, one look at a ticket's
entry), and a scheduler picks which team moves next.  The model follows the
kernel's rules:

* need: claim the next unstarted slot (counter [0]); a claim below ntot is
  followed by the ring write "none" for it, then the search starts;
* a search that reaches the hand-off point while the team still sees
  unstarted slots claims one; on success the team saves the search, writes it
  into the ring entry of its claim (a later step) and starts the claimed slot;
  a search is handed off at most once, a resumed one never again;
* with no unstarted slot left, a team takes tickets over the ring in claim
  order (counter [2]); an unwritten entry is looked at again in the team's
  next step, "none" moves on to the next ticket, a search is resumed; a
  ticket at or past ntot ends the team.

Checked for several sizes, hand-off points and schedules (uniformly random;
teams about to write a ring entry moved in only 3 % of the steps; one team
starved until the others wait): every search runs exactly its length once, every handed-off
search is resumed exactly once, every ring and ticket index stays below ntot,
every wait ends (the model finishes within a bound), and no ticket is taken
before every claim has been made, which is what makes each wait end.  A
forgotten ring write leaves a wait that never ends, and tickets that skip
entries leave searches unfinished: both fail these checks.
"""
import random

import pytest

UNSET, NONE = -1, -2


class Team:
    def __init__(self):
        self.state = "claim"
        self.fresh = True
        self.t = None          # search in progress
        self.iters = 0
        self.may_yield = False
        self.pending = None    # claim whose ring entry this team still has to write
        self.ticket = None
        self.polls = 0


def simulate(lengths, n_teams, yq, pick, seed, max_steps=2_000_000):
    rng = random.Random(seed)
    ntot = len(lengths)
    ring = [UNSET] * ntot
    saved = {}
    q_claim = q_ticket = 0
    done_iters = [0] * ntot
    finished = [0] * ntot
    handed = set()
    resumed = []
    teams = [Team() for _ in range(n_teams)]
    max_polls = 0
    for step in range(max_steps):
        live = [k for k, tm in enumerate(teams) if tm.state != "done"]
        if not live:
            break
        k = pick(rng, teams, live)
        tm = teams[k]
        if tm.state == "claim":
            c = q_claim
            q_claim += 1
            if c < ntot:
                tm.pending = c
                tm.state = "write_none"
            else:
                tm.fresh = False
                tm.state = "ticket" if yq > 0 else "done"
        elif tm.state == "write_none":
            assert ring[tm.pending] == UNSET
            ring[tm.pending] = NONE
            tm.t, tm.iters, tm.may_yield = tm.pending, 0, yq > 0
            tm.pending = None
            tm.state = "run"
        elif tm.state == "run":
            tm.iters += 1
            done_iters[tm.t] += 1
            if tm.iters == lengths[tm.t]:
                finished[tm.t] += 1
                tm.t = None
                tm.state = "claim" if tm.fresh else ("ticket" if yq > 0 else "done")
            elif tm.may_yield and tm.iters == yq:
                tm.may_yield = False
                if tm.fresh:
                    c = q_claim
                    q_claim += 1
                    if c < ntot:
                        tm.pending = c
                        tm.state = "save"
                    else:
                        tm.fresh = False
        elif tm.state == "save":
            assert tm.t not in handed
            handed.add(tm.t)
            saved[tm.t] = tm.iters
            assert ring[tm.pending] == UNSET
            ring[tm.pending] = tm.t
            tm.t, tm.iters, tm.may_yield = tm.pending, 0, True
            tm.pending = None
            tm.state = "run"
        elif tm.state == "ticket":
            if tm.ticket is None:
                assert q_claim >= ntot           # every claim precedes the first ticket
                tm.ticket = q_ticket
                q_ticket += 1
                tm.polls = 0
            if tm.ticket >= ntot:
                tm.state = "done"
                continue
            v = ring[tm.ticket]
            if v == UNSET:
                tm.polls += 1
                max_polls = max(max_polls, tm.polls)
                continue
            tm.ticket = None
            if v == NONE:
                continue
            assert 0 <= v < ntot and v in saved
            resumed.append(v)
            tm.t, tm.iters, tm.may_yield = v, saved.pop(v), False
            tm.state = "run"
    else:
        raise AssertionError("model did not finish: a wait that never ends")
    assert finished == [1] * ntot
    assert done_iters == list(lengths)
    assert sorted(resumed) == sorted(handed) and not saved
    return len(handed), max_polls


def uniform(rng, teams, live):
    return rng.choice(live)


def writers_last(rng, teams, live):
    """A team holding a claim whose ring entry it has not written yet moves
    only in 3 % of the steps: tickets meet unwritten entries often."""
    others = [k for k in live if teams[k].state not in ("write_none", "save")]
    return rng.choice(others) if others and rng.random() < 0.97 else rng.choice(live)


def starve_first(rng, teams, live):
    """Team 0 moves only once every other live team is waiting on a ticket."""
    busy = [k for k in live if k != 0 and teams[k].state != "ticket"]
    if busy:
        return rng.choice(busy)
    return rng.choice(live)


@pytest.mark.parametrize("pick", [uniform, writers_last, starve_first])
@pytest.mark.parametrize("ntot,teams,yq", [(300, 64, 32), (120, 100, 1), (500, 32, 8), (64, 80, 32), (1, 4, 1),
                                           (200, 1, 16)])
def test_every_search_runs_once_and_every_wait_ends(pick, ntot, teams, yq):
    rng = random.Random(ntot * 7 + teams)
    # lengths like ADADELTA's: most searches end early, a tenth at the 300 cap
    lengths = [300 if rng.random() < 0.1 else rng.randint(1, 200) for _ in range(ntot)]
    n_handed, _ = simulate(lengths, teams, yq, pick, seed=ntot + teams + yq)
    if teams < ntot and yq < 200:
        assert n_handed > 0                      # the hand-off path was exercised


def test_no_handoff_when_off():
    lengths = [50] * 100
    n_handed, polls = simulate(lengths, 16, 0, uniform, seed=1)
    assert n_handed == 0 and polls == 0
