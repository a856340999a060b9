"""The donor search's candidate list and selection, restated on the host (no
GPU).

Listing (k_collect / k_find): each wave (one per workgroup) gathers its
lanes' (rank | nv << 28, query) pairs in an LDS buffer of WBUF pairs, one
step adding at most one pair per lane, and appends the buffer to region
b % NREG of the list (b: its workgroup) with one atomic add of its length to
that region's counter whenever more than WBUF - 64 pairs are held, and at
the end; the copy sorts the pairs by lane, keeping each lane's step order.
Region r holds slots r * rcap .. (r + 1) * rcap - 1; pairs that land at or
past rcap are dropped, and the host then sees a count above rcap, grows
every region and lists again before any test runs.  The binning and
evaluation kernels read region r's slots below min(used_r, rcap).  (The
model runs each wave as one workgroup, as k_collect does; k_find's work
items of one wave may run in different workgroups, so there a query's pairs
may span regions, which the order-free selection below does not mind.)

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
OUTSIDE = 3


def replay(lists, nreg, rcap, rng):
    """lists[q]: the candidate values of query q (queries in wave order; wave
    w runs as workgroup w).  Returns the slots (None = never written), write
    counts and each region's count."""
    nq = len(lists)
    nwave = (nq + WAVE - 1) // WAVE
    slots = [None] * (nreg * rcap)
    writes = np.zeros(nreg * rcap, np.int64)
    used = [0] * nreg
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
    # the waves' flushes reach their region's atomic in any interleaving
    order = [w for w in range(nwave) for _ in pending[w]]
    rng.shuffle(order)
    nxt = [0] * nwave
    for w in order:
        buf = pending[w][nxt[w]]
        nxt[w] += 1
        r = w % nreg
        base = used[r]
        used[r] += len(buf)
        buf = sorted(buf, key=lambda qv: qv[0] % WAVE)  # stable: by lane, step order kept
        for j, v in enumerate(buf):
            if base + j < rcap:
                slots[r * rcap + base + j] = v
                writes[r * rcap + base + j] += 1
    return slots, writes, used


def held(slots, used, rcap):
    """The slots the binning and evaluation kernels read: each region's first
    min(used_r, rcap)."""
    return [j for r, u in enumerate(used) for j in range(r * rcap, r * rcap + min(u, rcap))]


@pytest.mark.parametrize("seed", range(10))
def test_every_read_slot_was_written_by_this_launch(seed):
    rng = np.random.default_rng(seed)
    nq = int(rng.integers(1, 800))
    counts = rng.choice([0, 1, 2, 3, 5, 8, 30, 90, 700], size=nq,
                        p=[.2, .2, .15, .15, .1, .1, .05, .03, .02])
    lists = [[int(q) * 10000 + k for k in range(int(c))] for q, c in enumerate(counts)]
    total = int(counts.sum())
    for nreg in (1, 4, 32):
        for rcap in sorted({1, 7, max(1, total // (5 * nreg)), max(1, total // nreg), total, total + 10}):
            slots, writes, used = replay(lists, nreg, rcap, rng)
            assert sum(used) == total
            read = held(slots, used, rcap)
            assert all(slots[j] is not None for j in read), (seed, nreg, rcap)
            assert (writes[read] == 1).all(), (seed, nreg, rcap)
            assert writes.max(initial=0) <= 1
            if max(used) <= rcap:  # room for everything: each query's pairs are exactly its candidates
                assert len(read) == total
                got = {}
                for j in read:
                    q, v = slots[j]
                    got.setdefault(q, []).append(v)
                for q, lst in enumerate(lists):
                    assert got.get(q, []) == lst, (seed, q)  # a query's pairs keep its step order
            else:  # the host lists again with every region grown to the largest count
                grown = max(used) + max(used) // 4
                slots, writes, used = replay(lists, nreg, grown, rng)
                assert max(used) <= grown and len(held(slots, used, grown)) == total


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
