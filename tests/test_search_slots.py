"""The slot accounting of the donor search's candidate list, restated on the
host (no GPU).

k_collect gives each wave of 64 queries one chunk of the flat candidate list
(an atomic add of the wave's total to the `used` counter; each lane's slots
follow by an exclusive prefix sum), lanes with more than KSTAGE candidates
take none there and are listed by k_collect_more (one atomic add each).
A list that does not fit below the capacity is not stored; the slots it
reserved below the capacity are marked empty.  The binning and evaluation
kernels then read every slot below min(used, cap).

This model replays that allocation with random candidate counts, random
atomic orders and capacities far below the total, and checks that every
slot the later kernels read was written by this search (a candidate of the
query that owns it, or the empty mark), and that a query whose list fits
finds exactly its own candidates at [qoff, qoff + count)."""
import numpy as np
import pytest

WAVE = 64
KSTAGE = 32
EMPTY = -1


def replay(counts, cap, rng):
    """counts[q]: candidates of query q (queries in wave order).  Returns the
    slot contents (None = never written), qoff, used."""
    nq = counts.size
    nwave = (nq + WAVE - 1) // WAVE
    slots = [None] * cap
    writes = np.zeros(cap, np.int64)
    qoff = np.full(nq, -1, np.int64)
    used = 0
    ovf = []

    def write(j, v):
        slots[j] = v
        writes[j] += 1

    # waves reach their atomic in any order
    for w in rng.permutation(nwave):
        lanes = np.arange(w * WAVE, min((w + 1) * WAVE, nq))
        k = np.where(counts[lanes] <= KSTAGE, counts[lanes], 0)
        pre = np.concatenate([[0], np.cumsum(k)[:-1]])
        total = int(k.sum())
        base = used
        used += total
        for q, kq, p in zip(lanes, k, pre):
            if counts[q] > KSTAGE:
                ovf.append(q)
                continue
            off = base + int(p)
            qoff[q] = off
            if off + kq > cap:
                for j in range(off, min(cap, off + kq)):
                    write(j, EMPTY)
                continue
            for j in range(kq):
                write(off + j, (q, j))
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
            write(off + j, (q, j))
    return slots, writes, qoff, used


@pytest.mark.parametrize("seed", range(12))
def test_every_read_slot_was_written_by_this_search(seed):
    rng = np.random.default_rng(seed)
    nq = int(rng.integers(1, 700))
    # mostly short lists, some empty, some above KSTAGE
    counts = rng.choice([0, 1, 2, 3, 5, 8, 31, 32, 33, 40, 90], size=nq,
                        p=[.2, .2, .15, .1, .1, .1, .04, .03, .03, .03, .02]).astype(np.int64)
    total = int(counts.sum())
    for cap in sorted({1, 7, max(1, total // 5), max(1, total // 2), max(1, total - 1), total, total + 10}):
        slots, writes, qoff, used = replay(counts, cap, rng)
        assert used == total
        n = min(used, cap)
        # every slot the binning reads was written exactly once by this search
        assert all(s is not None for s in slots[:n]), (seed, cap)
        assert (writes[:n] == 1).all(), (seed, cap)
        # a query whose list fits finds exactly its own candidates, in order
        for q in range(nq):
            c = int(counts[q])
            if c == 0:
                continue
            if qoff[q] + c <= cap:
                assert slots[qoff[q]:qoff[q] + c] == [(q, j) for j in range(c)], (seed, cap, q)
        # with room for everything, no slot is marked empty
        if cap >= total:
            assert EMPTY not in slots[:n]
