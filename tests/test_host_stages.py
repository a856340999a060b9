"""The engine's host-side sequential stages against the oracle, on the CPU
(no device calls): the median quickselect that shapes the ADT (sequential and
the list-based rounds the device build runs), the principal-axis OBB, the ADT
build and the coincident-point dedupe.  Inputs are chosen with many equal
keys, where the quickselect's exact swap order matters."""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "tioga_amd", "libtioga_amd.so")
P = C.c_void_p


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        from tioga_amd import build

        build.build()
    L = C.CDLL(LIB)
    L.tioga_amd_host_median.argtypes = [P, P, C.c_int, C.c_int]
    L.tioga_amd_host_find_obb.argtypes = [P, C.c_int, P]
    L.tioga_amd_host_build_adt.argtypes = [C.c_int, P, P, P, P]
    L.tioga_amd_host_build_adt.restype = C.c_int
    L.tioga_amd_host_unique_nodes.argtypes = [P, P, P, P, C.c_int]
    return L


def ptr(a):
    return a.ctypes.data_as(P)


@pytest.mark.parametrize("lists", [0, 1])
def test_median_matches_oracle(lib, lists):
    rng = np.random.default_rng(5)
    ol = O.lib()
    for trial in range(400):
        n = int(rng.integers(1, 300)) if trial % 10 else int(rng.integers(300, 5000))
        nvals = int(rng.integers(1, max(2, n)))
        x = rng.integers(0, nvals, n).astype(np.float64) * 0.25
        ix = np.arange(n, dtype=np.int32)
        x1, ix1 = x.copy(), ix.copy()
        x2, ix2 = x.copy(), ix.copy()
        lib.tioga_amd_host_median(ptr(ix1), ptr(x1), n, lists)
        ol.orc_median(ptr(ix2), ptr(x2), n)
        assert np.array_equal(ix1, ix2), (trial, n)
        assert np.array_equal(x1, x2)


def test_find_obb_matches_oracle(lib):
    rng = np.random.default_rng(6)
    ol = O.lib()
    for n in (1, 2, 3, 4, 17, 1000):
        x = np.ascontiguousarray(rng.normal(size=(n, 3)) * [3.0, 1.0, 0.2])
        a = np.zeros(15)
        lib.tioga_amd_host_find_obb(ptr(x), n, ptr(a))
        xc, dxc, vec = np.zeros(3), np.zeros(3), np.zeros(9)
        ol.orc_find_obb(ptr(x), ptr(xc), ptr(dxc), ptr(vec), n)
        assert np.array_equal(a, np.concatenate([xc, dxc, vec])), n


def test_build_adt_matches_oracle(lib):
    rng = np.random.default_rng(7)
    ol = O.lib()
    for n in (1, 2, 3, 7, 64, 1000, 5000):
        lo = np.round(rng.uniform(-1, 1, (n, 3)) * 8) / 8  # many equal keys
        box = np.ascontiguousarray(np.concatenate([lo, lo + 0.125], axis=1))
        ai, ar, ext = np.zeros(4 * n, np.int32), np.zeros(6 * n), np.zeros(6)
        lib.tioga_amd_host_build_adt(n, ptr(box), ptr(ai), ptr(ar), ptr(ext))
        bi, br, bext = np.zeros(4 * n, np.int32), np.zeros(6 * n), np.zeros(6)
        ol.orc_build_adt(n, ptr(box), ptr(bi), ptr(br), ptr(bext))
        assert np.array_equal(ai, bi) and np.array_equal(ar, br) and np.array_equal(ext, bext), n


def test_unique_nodes_matches_oracle(lib):
    rng = np.random.default_rng(8)
    ol = O.lib()
    n = 3000
    base = rng.uniform(-1, 1, (n // 3, 3))
    x = np.ascontiguousarray(np.concatenate([base, base, base + 1e-12])[rng.permutation(n)])
    tag = rng.integers(1, 3, n).astype(np.int32)
    r1 = rng.uniform(0, 1, n)
    r2 = r1.copy()
    t1, t2 = np.zeros(n, np.int32), np.zeros(n, np.int32)
    lib.tioga_amd_host_unique_nodes(ptr(x), ptr(tag), ptr(r1), ptr(t1), n)
    ol.orc_unique_nodes_octree(ptr(x), ptr(tag), ptr(r2), ptr(t2), n)
    assert np.array_equal(t1, t2) and np.array_equal(r1, r2)
    assert (t1 != np.arange(n)).sum() > n // 3  # duplicates were found

