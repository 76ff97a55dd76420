"""CPU simulation of the second radix pass's per-tile barrier schedule
(k_sortchunk_g in hashjoin-scm_amd/csrc/hj_radix.hip).

The kernel sorts one tile after another with LDS digit counters that are
cleared by their owning thread right after it reads them, and the next tile's
counting has no barrier of its own before it.  The protocol per tile is:

  count   every thread: rank = atomicAdd(&cnt[digit], 1) for its items
  B1      barrier
  own     the owner of digit p reads cnt[p] (twice when it owns several
          digits: once for its block-scan input, once after the scan) and
          clears it; the block scan has barriers of its own
  B4      barrier
  place   every thread: tile[base[digit] + rank] = item
  B5      barrier
  store   every thread writes out tile positions t, t + NT, ...
  stage   the next tile's staging (a barrier only for tiles covered by few
          runs), then the next tile's count

This test runs that program for several threads under random interleavings
(any runnable thread may take the next step; a barrier releases when every
thread has arrived) and checks that every owner reads exactly its digit's
count of the current tile and that every tile comes out sorted by digit with
all its items.  A planted variant that clears the counters at the start of
the next tile instead shows the race the schedule avoids: a fast thread's
count of the next tile lands before a slow owner's clear and is lost."""
import random

import pytest

BARRIER = object()


def _program(tid, nt, tiles, per, plant, few):
    """Steps of thread `tid` (closures over the shared state `st`)."""
    steps = []
    for h, items in enumerate(tiles):
        mine = [(k, items[k]) for k in range(tid, len(items), nt)]
        owned = [tid * per + q for q in range(per)]
        if plant and h > 0:
            # planted mistake: clear at the start of the next tile, no barrier
            for p in owned:
                steps.append(lambda st, p=p: st["cnt"].__setitem__(p, 0))
        for k, d in mine:
            def count(st, k=k, d=d, h=h):
                st["rank"][(h, k)] = st["cnt"][d]
                st["cnt"][d] += 1  # one atomic step
            steps.append(count)
        steps.append(BARRIER)  # B1
        if per > 1:
            # block-scan input: the sum of the owned counters, then the scan's
            # barriers, then a second read of each counter (and its clear)
            steps.append(lambda st, h=h: st["sum"].__setitem__((h, tid), sum(st["cnt"][p] for p in owned)))
            steps.append(BARRIER)
            steps.append(BARRIER)
        for p in owned:
            def own(st, p=p, h=h):
                st["seen"][(h, p)] = st["cnt"][p]
                if not plant:
                    st["cnt"][p] = 0
            steps.append(own)
        steps.append(BARRIER)  # end of the scan
        steps.append(lambda st, h=h: st["base"].__setitem__(h, _bases(st, h, nt * per)) if tid == 0 else None)
        steps.append(BARRIER)  # B4
        for k, d in mine:
            def place(st, k=k, d=d, h=h):
                st["tile"][st["base"][h][d] + st["rank"][(h, k)]] = (d, k)
            steps.append(place)
        steps.append(BARRIER)  # B5
        for u in range(tid, len(items), nt):
            steps.append(lambda st, u=u, h=h: st["out"].setdefault(h, {}).__setitem__(u, st["tile"][u]))
        if few[h]:
            steps.append(BARRIER)  # the cooperative staging of a few-run tile
    return steps


def _bases(st, h, nd):
    run, out = 0, []
    for p in range(nd):
        out.append(run)
        run += st["seen"][(h, p)]
    return out


def _simulate(seed, nt=8, per=1, ntiles=6, plant=False):
    rng = random.Random(seed)
    nd = nt * per
    tiles = []
    for _ in range(ntiles):
        n = rng.randint(1, 3 * nt)
        skew = rng.random() < 0.4
        tiles.append([0 if skew and rng.random() < 0.8 else rng.randrange(nd) for _ in range(n)])
    few = [rng.random() < 0.3 for _ in range(ntiles)]
    progs = [_program(t, nt, tiles, per, plant, few) for t in range(nt)]
    st = {"cnt": [0] * nd, "rank": {}, "seen": {}, "sum": {}, "base": {}, "tile": [None] * (3 * nt), "out": {}}
    pc = [0] * nt
    waiting = set()
    while True:
        runnable = [t for t in range(nt) if t not in waiting and pc[t] < len(progs[t])]
        if not runnable:
            if len(waiting) == nt:  # every thread has the same barriers
                for t in waiting:
                    pc[t] += 1  # release the barrier
                waiting.clear()
                continue
            assert not waiting, "threads stuck at a barrier"
            break
        t = rng.choice(runnable)
        step = progs[t][pc[t]]
        if step is BARRIER:
            waiting.add(t)
            continue
        try:
            step(st)
        except (IndexError, TypeError, KeyError):
            return False  # a lost or doubled count sent an item outside the tile
        pc[t] += 1
    try:
        ok = True
        for h, items in enumerate(tiles):
            for p in range(nd):
                ok &= st["seen"][(h, p)] == items.count(p)
            got = [st["out"][h][u] for u in range(len(items))]
            ok &= None not in got and sorted(got) == sorted((d, k) for k, d in enumerate(items))
            ok &= [d for d, _ in got] == sorted(items)
        return ok
    except (TypeError, KeyError):
        return False


@pytest.mark.parametrize("per", [1, 2])
def test_counter_clearing_is_race_free(per):
    assert all(_simulate(seed, per=per) for seed in range(400))


def test_planted_late_clear_races():
    """Clearing at the start of the next tile without a barrier loses counts
    in some interleaving: the simulation can see the race it rules out."""
    assert not all(_simulate(seed, plant=True) for seed in range(400))
