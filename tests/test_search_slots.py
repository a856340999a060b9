"""The donor search's candidate list and selection, restated on the host (no
GPU).

Listing (k_collect / k_find): each wave (one per workgroup) gathers its
lanes' (rank | nv << 28, query) pairs in an LDS buffer of WBUF pairs, one
step adding at most one pair per lane, and whenever more than WBUF - 64
pairs are held, and at the end, appends them by cell type: for each type
present, one atomic add of its count to the counter of region (type,
b % NREG) (b: its workgroup), then the type's pairs in the order held.
Region (t, r) holds slots tbase[t] + r * rcap[t] .. + rcap[t] - 1; pairs that
land at or past rcap[t] are dropped, and the host then sees a count above
it, grows that type's regions and lists again before any test runs.  The
evaluation kernels read region (t, r)'s slots below min(used, rcap[t]).
(The model runs each wave as one workgroup, as k_collect does; k_find's
work items of one wave may run in different workgroups, so there a query's
pairs may span regions, which the order-free selection below does not
mind.)

Selection (k_eval's per-query words, k_select): the reference walks the ADT
in preorder, tests every candidate (an element whose padded box holds the
point) in increasing node id, and stops at the first clean hit; its result is
that hit's cell, or else the outcome of the last candidate tested (its cell
when accepted with a near-zero weight, status 1, else -1).  The device folds
the outcomes, in any order, into best = the lowest rank of a clean hit, then
for the queries without one into last = the highest (rank << 2 | status)
among candidates inside their element box (status OUTSIDE = 3 takes no
part), and picks best's cell, else last's cell when its status is 1, else
-1.

The model replays the listing with random candidate counts, interleavings and
capacities, and checks that every slot read later was written once by this
launch, and that with room for everything each query's pairs are exactly its
candidates.  It then checks the order-free selection against the sequential
rule on shuffled lists with random outcomes, and that a planted mistake (the
highest clean hit instead of the lowest) is caught."""
import numpy as np
import pytest

WAVE = 64
WBUF = 512
NTYPE = 4
OUTSIDE = 3


def ntype(v):
    """the cell-type bucket of a listed value (the model's stand-in for nv)"""
    return v % NTYPE


def replay(lists, nreg, rcap, rng):
    """lists[q]: the candidate values of query q (queries in wave order; wave
    w runs as workgroup w); rcap[t]: the region capacity of type t.  Returns
    the slots (None = never written), write counts, the regions' counts
    used[t][r] and the types' first slots."""
    nq = len(lists)
    nwave = (nq + WAVE - 1) // WAVE
    tbase = np.concatenate([[0], np.cumsum([nreg * c for c in rcap])]).astype(int)
    slots = [None] * int(tbase[-1])
    writes = np.zeros(int(tbase[-1]), np.int64)
    used = [[0] * nreg for _ in range(NTYPE)]
    # each wave: its steps, then its flushes; waves interleave at flushes
    pending = []
    for w in range(nwave):
        lanes = range(w * WAVE, min((w + 1) * WAVE, nq))
        steps = max((len(lists[q]) for q in lanes), default=0)
        buf, flushes = [], []
        for k in range(steps):
            for q in lanes:  # lane order inside a step (the ballot prefix)
                if k < len(lists[q]):
                    buf.append((q, lists[q][k]))
            if len(buf) > WBUF - WAVE:
                flushes.append(buf)
                buf = []
        if buf:
            flushes.append(buf)
        pending.append(flushes)
    # the waves' flushes reach their regions' atomics in any interleaving
    order = [w for w in range(nwave) for _ in pending[w]]
    rng.shuffle(order)
    nxt = [0] * nwave
    for w in order:
        buf = pending[w][nxt[w]]
        nxt[w] += 1
        r = w % nreg
        for t in range(NTYPE):
            mine = [qv for qv in buf if ntype(qv[1]) == t]  # in the order held
            base = used[t][r]
            used[t][r] += len(mine)
            for j, v in enumerate(mine):
                if base + j < rcap[t]:
                    k = int(tbase[t]) + r * rcap[t] + base + j
                    slots[k] = v
                    writes[k] += 1
    return slots, writes, used, tbase


def held(used, rcap, tbase):
    """The slots the evaluation kernels read: each region's first
    min(used, rcap[t])."""
    return [int(tbase[t]) + r * rcap[t] + k for t in range(NTYPE) for r, u in enumerate(used[t])
            for k in range(min(u, rcap[t]))]


@pytest.mark.parametrize("seed", range(10))
def test_every_read_slot_was_written_by_this_launch(seed):
    rng = np.random.default_rng(seed)
    nq = int(rng.integers(1, 800))
    counts = rng.choice([0, 1, 2, 3, 5, 8, 30, 90, 700], size=nq,
                        p=[.2, .2, .15, .15, .1, .1, .05, .03, .02])
    lists = [[int(q) * 10000 + k for k in range(int(c))] for q, c in enumerate(counts)]
    total = int(counts.sum())
    for nreg in (1, 4, 32):
        for cap in sorted({1, 7, max(1, total // (5 * nreg)), max(1, total // nreg), total + 10}):
            rcap = [cap, max(1, cap // 2), cap, 2 * cap]  # uneven type capacities
            slots, writes, used, tbase = replay(lists, nreg, rcap, rng)
            assert sum(map(sum, used)) == total
            read = held(used, rcap, tbase)
            assert all(slots[j] is not None for j in read), (seed, nreg, cap)
            assert (writes[read] == 1).all(), (seed, nreg, cap)
            assert writes.max(initial=0) <= 1
            fits = all(max(used[t]) <= rcap[t] for t in range(NTYPE))
            if fits:  # each query's pairs are exactly its candidates, in step order within a type
                assert len(read) == total
                got = {}
                for j in read:
                    q, v = slots[j]
                    got.setdefault((q, ntype(v)), []).append(v)
                for q, lst in enumerate(lists):
                    for t in range(NTYPE):
                        assert got.get((q, t), []) == [v for v in lst if ntype(v) == t], (seed, q)
            else:  # the host grows each overflowing type to its largest count and lists again
                grown = [max(used[t]) + max(used[t]) // 4 if max(used[t]) > rcap[t] else rcap[t]
                         for t in range(NTYPE)]
                slots, writes, used, tbase = replay(lists, nreg, grown, rng)
                assert all(max(used[t]) <= grown[t] for t in range(NTYPE))
                assert len(held(used, grown, tbase)) == total


def sequential(cands):
    """The reference's walk over its candidates in preorder (rank order):
    cands = [(rank, cell, status)], OUTSIDE entries never tested."""
    res = -1
    for rank, cell, st in sorted(cands):
        if st == OUTSIDE:
            continue
        res = cell if st else -1
        if st == 2:
            break
    return res


def order_free(cands, lowest=True):
    best, last = None, None
    for rank, cell, st in cands:
        if st == OUTSIDE:
            continue
        if st == 2 and (best is None or (rank < best[0] if lowest else rank > best[0])):
            best = (rank, cell)
        if last is None or rank > last[0]:
            last = (rank, cell, st)
    if best is not None:
        return best[1]
    if last is not None and last[2] == 1:
        return last[1]
    return -1


@pytest.mark.parametrize("seed", range(6))
def test_order_free_selection_is_the_sequential_rule(seed):
    rng = np.random.default_rng(100 + seed)
    planted_caught = False
    for _ in range(3000):
        n = int(rng.integers(0, 12))
        ranks = rng.choice(1000, size=n, replace=False)
        cells = rng.integers(0, 10 ** 6, size=n)
        st = rng.choice([0, 1, 2, OUTSIDE], size=n, p=[.55, .1, .2, .15])
        cands = [(int(r), int(c), int(s)) for r, c, s in zip(ranks, cells, st)]
        want = sequential(cands)
        rng.shuffle(cands)  # the listing order is free
        assert order_free(cands) == want, cands
        if order_free(cands, lowest=False) != want:
            planted_caught = True
    assert planted_caught  # the checks see a selection that takes the wrong clean hit
