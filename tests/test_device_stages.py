"""Device stages against the oracle at sizes the golden cases do not reach.

The ADT builder splits segments longer than 16384 elements across many
workgroups and builds several trees in one level loop; the test runs it
through the C ABI on trees of mixed sizes (long, medium, short, single,
empty) whose boxes have heavy ties along every cut dimension and signed
zeros, and compares every tree node for node with the oracle's restatement
of ADT.C / buildADTrecursion.c / median.F90.  The coincident-point dedupe
likewise splits octree nodes of more than 65536 points across workgroups
and takes several point sets at once; its representatives and raised
resolutions are compared with the oracle's uniquenodes_octree."""
import ctypes as C

import numpy as np
import pytest

from oracle.oracle import lib as oracle_lib, ptr

pytestmark = pytest.mark.gpu


def boxes(n, rng):
    c = np.round(rng.uniform(-1, 1, (n, 3)) * 64) / 64  # coarse grid: many equal coordinates
    c[rng.uniform(size=(n, 3)) < 0.02] = 0.0
    h = np.round(rng.uniform(0.001, 0.05, (n, 3)) * 512) / 512
    lo, hi = c - h, c + h
    z = rng.uniform(size=(n, 3)) < 0.01
    lo[z] = -0.0  # signed zeros tie with +0.0 under <
    hi[rng.uniform(size=(n, 3)) < 0.01] = 0.0
    hi = np.maximum(hi, lo)
    return np.ascontiguousarray(np.concatenate([lo, hi], 1))


def test_device_adt_batch_matches_oracle():
    from tioga_amd.tioga import load

    L = load()
    L.tioga_amd_device_build_adts.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.tioga_amd_device_build_adts.restype = C.c_int
    O = oracle_lib()
    rng = np.random.default_rng(7)
    sizes = [300000, 70000, 5000, 129, 128, 1, 0, 2, 66000, 512, 513, 2048, 2049, 17, 16, 9000]
    ebs = [boxes(n, rng) for n in sizes]
    for batch in ([0], list(range(len(sizes)))):
        ns = np.array([sizes[t] for t in batch], np.int32)
        eb = np.ascontiguousarray(np.concatenate([ebs[t] for t in batch]).reshape(-1)) if ns.sum() else np.zeros(6)
        tot = int(ns.sum())
        ai = np.zeros(4 * max(tot, 1), np.int32)
        ar = np.zeros(6 * max(tot, 1))
        ext = np.zeros(6 * len(batch))
        print(f"[device adt] trees {ns.tolist()}", flush=True)
        rc = L.tioga_amd_device_build_adts(len(batch), ptr(ns), ptr(eb), ptr(ai), ptr(ar), ptr(ext))
        assert rc == 0
        off = 0
        for k, t in enumerate(batch):
            n = sizes[t]
            wi = np.zeros(4 * max(n, 1), np.int32)
            wr = np.zeros(6 * max(n, 1))
            we = np.zeros(6)
            O.orc_build_adt(n, ptr(ebs[t]), ptr(wi), ptr(wr), ptr(we))
            assert np.array_equal(ai[4 * off:4 * (off + n)], wi[:4 * n]), f"tree {t} adtIntegers"
            assert np.array_equal(ar[6 * off:6 * (off + n)].view(np.int64), wr[:6 * n].view(np.int64)), \
                f"tree {t} adtReals"
            if n:
                assert np.array_equal(ext[6 * k:6 * k + 6], we), f"tree {t} extents"
            off += n


def points(n, rng, ndup=0.3):
    """Points on a coarse grid (ties at the octree's mid planes) with groups
    of up to 4 coincident copies, some of them under another mesh tag."""
    m = max(1, int(n * (1 - ndup)))
    base = np.round(rng.uniform(-1, 1, (m, 3)) * 256) / 256
    pick = np.concatenate([np.arange(m), rng.integers(0, m, n - m)])
    rng.shuffle(pick)
    x = np.ascontiguousarray(base[pick])
    tag = rng.integers(1, 3, n).astype(np.int32)
    rtag = rng.uniform(0, 1, n)
    return x, tag, rtag


def test_device_dedupe_batch_matches_oracle():
    from tioga_amd.tioga import load

    L = load()
    L.tioga_amd_device_unique_nodes.argtypes = [C.c_int] + [C.c_void_p] * 5
    L.tioga_amd_device_unique_nodes.restype = C.c_int
    O = oracle_lib()
    rng = np.random.default_rng(11)
    sizes = [400000, 70000, 25, 5, 0, 21, 120000]
    sets = [points(n, rng) if n else (np.zeros((0, 3)), np.zeros(0, np.int32), np.zeros(0)) for n in sizes]
    # one set of 500 identical points: a node that cannot be split
    sets.append((np.ascontiguousarray(np.full((500, 3), 0.25)), np.ones(500, np.int32), rng.uniform(0, 1, 500)))
    sizes.append(500)
    # sets at the subtree threshold, and a graded one (geometric spacing
    # towards a wall, duplicated) whose octree is deep and lopsided
    for m in (1024, 1025, 3000):
        sets.append(points(m, rng))
        sizes.append(m)
    g = 0.5 * 0.8 ** np.arange(450)
    xg = np.stack([g, 0.3 * g, np.full_like(g, 0.1)], 1)
    xg = np.ascontiguousarray(np.concatenate([xg, xg[::3]]))
    sets.append((xg, np.ones(xg.shape[0], np.int32), rng.uniform(0, 1, xg.shape[0])))
    sizes.append(xg.shape[0])
    for batch in ([0], list(range(len(sizes)))):
        ns = np.array([sizes[k] for k in batch], np.int32)
        x = np.ascontiguousarray(np.concatenate([sets[k][0] for k in batch]).reshape(-1))
        tag = np.ascontiguousarray(np.concatenate([sets[k][1] for k in batch]))
        rtag = np.ascontiguousarray(np.concatenate([sets[k][2] for k in batch]))
        xtag = np.zeros(max(int(ns.sum()), 1), np.int32)
        print(f"[device dedupe] sets {ns.tolist()}", flush=True)
        assert L.tioga_amd_device_unique_nodes(len(batch), ptr(ns), ptr(x), ptr(tag), ptr(rtag), ptr(xtag)) == 0
        off = 0
        for k in batch:
            n = sizes[k]
            xs, ts, rs = sets[k]
            wr = rs.copy()
            wi = np.zeros(max(n, 1), np.int32)
            if n:
                O.orc_unique_nodes_octree(ptr(np.ascontiguousarray(xs)), ptr(ts), ptr(wr), ptr(wi), n)
            assert np.array_equal(xtag[off:off + n], wi[:n]), f"set {k} xtag"
            assert np.array_equal(rtag[off:off + n], wr[:n]), f"set {k} rtag"
            off += n


def obb_sets(rng):
    """Point sets for the device OBB: sizes around the special branches and
    the 512-point segments, sums that stay in one binade (offset clusters),
    sums that keep returning to zero (centred data: every segment stepped
    term by term), dyadic coordinates (quantisation ties), signed zeros and
    a plane (extents tied at +-0), large offsets, tiny and subnormal scales,
    and a set of millions of points (many segments, long prefix scans).
    (Sets so small that the covariance underflows give all-NaN boxes whose
    NaN sign bits are the compiler's choice, so none is included.)"""
    sets = [rng.normal(size=(n, 3)) * [3.0, 1.0, 0.2] for n in (1, 2, 3, 4, 5, 511, 512, 513, 1025)]
    c = rng.normal(size=(300000, 3))
    sets.append(np.array([0.3, -0.5, 0.7]) + 0.05 * c / np.linalg.norm(c, axis=1)[:, None])
    sets.append(rng.uniform(-1, 1, (20000, 3)))
    g = np.round(rng.uniform(-1, 1, (200000, 3)) * 8) / 8
    g[rng.uniform(size=g.shape) < 0.05] = -0.0
    sets.append(g)
    flat = np.zeros((70000, 3))
    flat[:, :2] = rng.uniform(-1, 1, (70000, 2))
    sets.append(flat)
    sets.append(1e6 + 1e-3 * rng.normal(size=(100000, 3)))
    sets.append(1e-60 * (0.5 + rng.uniform(size=(50000, 3))))
    # a run of subnormal and tiny points before O(1) ones: running sums below
    # the smallest binade the segment step takes
    t = np.concatenate([1e-310 * rng.uniform(size=(1500, 3)), 1e-300 * rng.uniform(size=(1500, 3)),
                        rng.normal(size=(3000, 3))])
    sets.append(t)
    i = np.arange(40)
    sets.append(np.stack(np.meshgrid(i, i, i, indexing="ij"), -1)[..., ::-1].reshape(-1, 3) * (2.0 / 39) - 1.0)
    sets.append(np.array([-0.2, 0.4, 0.1]) + 0.08 * rng.uniform(-1, 1, (2500000, 3)))
    return [np.ascontiguousarray(s, dtype=np.float64) for s in sets]


def test_device_obbs_match_oracle():
    """The device OBB batch (segmented exact running sums, order-carrying
    extents) gives the oracle's findOBB bits for every set."""
    from tioga_amd.tioga import load

    L = load()
    L.tioga_amd_device_find_obbs.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.tioga_amd_device_find_obbs.restype = C.c_int
    O = oracle_lib()
    sets = obb_sets(np.random.default_rng(12))
    want = []
    for x in sets:
        xc, dxc, vec = np.zeros(3), np.zeros(3), np.zeros(9)
        O.orc_find_obb(ptr(x), ptr(xc), ptr(dxc), ptr(vec), x.shape[0])
        want.append(np.concatenate([xc, dxc, vec]))
    for batch in ([len(sets) - 1], [10], list(range(len(sets)))):
        ns = np.array([sets[k].shape[0] for k in batch], np.int32)
        x = np.ascontiguousarray(np.concatenate([sets[k] for k in batch]).reshape(-1))
        out = np.zeros(15 * len(batch))
        stepped = np.zeros(1, np.int64)
        print(f"[device obb] sets {ns.tolist()}", flush=True)
        assert L.tioga_amd_device_find_obbs(len(batch), ptr(ns), ptr(x), ptr(out), ptr(stepped)) == 0
        print(f"[device obb] segments stepped term by term: {int(stepped[0])}", flush=True)
        for a, k in enumerate(batch):
            got = out[15 * a:15 * a + 15]
            assert np.array_equal(got.view(np.int64), want[k].view(np.int64)), (k, sets[k].shape[0], got, want[k])
