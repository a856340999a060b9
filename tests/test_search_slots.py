"""The slot accounting of the donor search's candidate list, restated on the
host (no GPU).

k_collect gives each wave of 64 queries one chunk of the flat candidate list
(an atomic add of the wave's total to the `used` counter; each lane's slots
follow by an exclusive prefix sum).  A lane's first KSTAGE candidates are
staged in LDS; the rest go to spill chunks of SPILL_K entries per lane that
the wave takes from a global pool (one atomic per chunk, at most SPILL_MAX
per wave), and are copied into the lane's slots with the staged ones.  When
the pool runs out, or a wave would need more than SPILL_MAX chunks, the
wave's lanes with more than KSTAGE candidates take no slots there and are
listed by k_collect_more (one atomic add each).  A list that does not fit
below the capacity is not stored; the slots it reserved below the capacity
are marked empty.  The binning and evaluation kernels then read every slot
below min(used, cap).

This model replays that allocation with random candidate counts, random
atomic orders, spill pools from empty to ample, and capacities far below the
total, and checks that every slot the later kernels read was written by this
search (a candidate of the query that owns it, or the empty mark), that a
query whose list fits finds exactly its own candidates at [qoff, qoff +
count) in order, and that spill chunks are never shared between waves."""
import numpy as np
import pytest

WAVE = 64
KSTAGE = 24
SPILL_K = 16
SPILL_MAX = 4
EMPTY = -1


def replay(counts, cap, pool, rng):
    """counts[q]: candidates of query q (queries in wave order); pool: spill
    chunks.  Returns the slot contents (None = never written), write counts,
    qoff, used, and the queries listed by k_collect_more."""
    nq = counts.size
    nwave = (nq + WAVE - 1) // WAVE
    slots = [None] * cap
    writes = np.zeros(cap, np.int64)
    qoff = np.full(nq, -1, np.int64)
    used = 0
    taken = 0  # the pool's chunk counter
    owner = {}  # spill chunk -> wave
    ovf = []

    def write(j, v):
        slots[j] = v
        writes[j] += 1

    # waves reach their atomics in any order
    for w in rng.permutation(nwave):
        lanes = np.arange(w * WAVE, min((w + 1) * WAVE, nq))
        c = counts[lanes]
        # the walk: a chunk is taken when the longest list reaches entry
        # SPILL_K * nchunk past the stage; a failed take ends spilling
        need = int(-(-max(int(c.max()) - KSTAGE, 0) // SPILL_K))
        chunks, ok = [], True
        for _ in range(need):
            if len(chunks) == SPILL_MAX:
                ok = False
                break
            cid = taken
            taken += 1
            if cid >= pool:
                ok = False
                break
            assert cid not in owner
            owner[cid] = w
            chunks.append(cid)
        spill = {}
        if ok:  # spilled entries, lane-major inside each chunk
            for lane, q in enumerate(lanes):
                for e in range(max(int(counts[q]) - KSTAGE, 0)):
                    spill[(chunks[e // SPILL_K] * SPILL_K + e % SPILL_K) * WAVE + lane] = (int(q), KSTAGE + e)
        held = ok | (c <= KSTAGE)
        k = np.where(held, c, 0)
        pre = np.concatenate([[0], np.cumsum(k)[:-1]])
        base = used
        used += int(k.sum())
        for lane, (q, kq, p, h) in enumerate(zip(lanes, k, pre, held)):
            if not h:
                ovf.append(q)
                continue
            off = base + int(p)
            qoff[q] = off
            if off + kq > cap:
                for j in range(off, min(cap, off + kq)):
                    write(j, EMPTY)
                continue
            for j in range(min(kq, KSTAGE)):
                write(off + j, (int(q), j))  # from the stage
            for j in range(KSTAGE, kq):
                e = j - KSTAGE
                write(off + j, spill[(chunks[e // SPILL_K] * SPILL_K + e % SPILL_K) * WAVE + lane])
    # k_collect_more: the queued queries in any order
    for q in rng.permutation(np.array(ovf, np.int64)) if ovf else []:
        cnt = int(counts[q])
        off = used
        used += cnt
        qoff[q] = off
        if off + cnt > cap:
            for j in range(off, min(cap, off + cnt)):
                write(j, EMPTY)
            continue
        for j in range(cnt):
            write(off + j, (int(q), j))
    return slots, writes, qoff, used, ovf


@pytest.mark.parametrize("seed", range(12))
def test_every_read_slot_was_written_by_this_search(seed):
    rng = np.random.default_rng(seed)
    nq = int(rng.integers(1, 700))
    # mostly short lists, some empty, some above KSTAGE
    counts = rng.choice([0, 1, 2, 3, 5, 8, 23, 24, 25, 40, 41, 90, 200], size=nq,
                        p=[.2, .2, .15, .1, .1, .08, .04, .03, .03, .03, .02, .01, .01]).astype(np.int64)
    total = int(counts.sum())
    caps = sorted({1, 7, max(1, total // 5), max(1, total // 2), max(1, total - 1), total, total + 10})
    for cap, pool in [(c, p) for c in caps for p in (0, 1, 3, 10 ** 6)]:
        slots, writes, qoff, used, ovf = replay(counts, cap, pool, rng)
        assert used == total
        if pool == 0:  # nothing spills: exactly the lists past the stage are re-walked
            assert sorted(ovf) == sorted(np.nonzero(counts > KSTAGE)[0].tolist())
        n = min(used, cap)
        # every slot the binning reads was written exactly once by this search
        assert all(s is not None for s in slots[:n]), (seed, cap, pool)
        assert (writes[:n] == 1).all(), (seed, cap, pool)
        # a query whose list fits finds exactly its own candidates, in order
        for q in range(nq):
            c = int(counts[q])
            if c == 0:
                continue
            if qoff[q] + c <= cap:
                assert slots[qoff[q]:qoff[q] + c] == [(q, j) for j in range(c)], (seed, cap, pool, q)
        # with room for everything, no slot is marked empty
        if cap >= total:
            assert EMPTY not in slots[:n]
