"""The allocation-free dataUpdate against the existing one, through the C ABI.

tioga_amd_dataupdate_mb_noalloc updates the caller's device solutions in
place with every staging buffer in one fixed device array; it must give the
bits tioga_dataupdate_mb_ gives on the registered host arrays (the update
of tioga.C:336-524 / dataUpdate.C:27-119).  Each case runs two calls in a
row on the same connectivity, the second with a smaller input (fewer
fields, the other layout), in the fused form (no receptor is a donor
vertex: the tiny case) and in the staged form (some receptors are donor
vertices: three overlapping blocks, and the 6-neighbour hole fill).  An
update whose staged values would not fit the fixed array, and a call before
any connectivity, are rejected with an error."""
import ctypes as C

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

P = C.c_void_p


def _lib(t):
    L = t.L
    L.tioga_amd_dataupdate_mb_noalloc.argtypes = [C.POINTER(C.c_int), C.c_char_p, C.POINTER(P)]
    L.tioga_amd_dataupdate_mb_noalloc.restype = C.c_int
    L.tioga_amd_dataupdate_noalloc_max_bytes.restype = C.c_long
    return L


def _connected(name):
    from tioga_amd.tioga import Tioga

    c = cases.build(name)
    t = Tioga()
    t.set_params(c.isym, c.nfringe, c.mexclude)
    for b in c.blocks:
        t.register_grid(b)
    t.preprocess()
    t.connect()
    return c, t


def _hip():
    """The HIP runtime the engine itself links (device buffers for the test)."""
    H = C.CDLL("libamdhip64.so.7")
    H.hipMalloc.argtypes = [C.POINTER(P), C.c_size_t]
    H.hipMemcpy.argtypes = [P, P, C.c_size_t, C.c_int]
    H.hipFree.argtypes = [P]
    return H


def _noalloc(t, qs, nvar, itype):
    """Device copies of qs, updated by the allocation-free entry point
    (the test allocates them; the entry point allocates nothing)."""
    H = _hip()
    L = _lib(t)
    qs = [np.ascontiguousarray(q, np.float64) for q in qs]
    dev = []
    try:
        for q in qs:
            p = P()
            assert H.hipMalloc(C.byref(p), max(q.nbytes, 8)) == 0
            dev.append(p)
            assert H.hipMemcpy(p, q.ctypes.data_as(P), q.nbytes, 1) == 0  # host to device
        ptrs = (P * len(dev))(*dev)
        rc = L.tioga_amd_dataupdate_mb_noalloc(C.byref(C.c_int(nvar)), itype, ptrs)
        msg = C.create_string_buffer(1024)
        status = L.tioga_amd_status(msg, 1024)
        L.tioga_amd_clear_status()
        out = []
        for p, q in zip(dev, qs):
            o = np.empty_like(q)
            assert H.hipMemcpy(o.ctypes.data_as(P), p, q.nbytes, 2) == 0  # device to host
            out.append(o)
    finally:
        for p in dev:
            H.hipFree(p)
    return rc, status, msg.value.decode(), out


@pytest.mark.parametrize("name", ["tiny", "three", "tiny_isym0"])
def test_noalloc_update_equals_dataupdate(name):
    print(f"\n[no-alloc dataUpdate] case {name}", flush=True)
    c, t = _connected(name)
    try:
        rng = np.random.default_rng(90)
        for nvar, itype, interptype in ((5, b"row", 0), (2, b"column", 1)):
            qs = [rng.uniform(-1, 1, b.nnodes * nvar) for b in c.blocks]
            want = t.data_update_copies(qs, nvar, interptype)
            rc, status, msg, got = _noalloc(t, qs, nvar, itype)
            assert rc == 0 and status == 0, msg
            for b, g, w, q0 in zip(c.blocks, got, want, qs):
                assert np.array_equal(g, w), f"{name} nvar {nvar}: block {b.tag} differs"
            # the update changed something (receptors exist)
            assert any(not np.array_equal(w, q0) for w, q0 in zip(want, qs))
    finally:
        t.close()


def test_noalloc_update_rejects_what_does_not_fit():
    """Three blocks (some receptors are donor vertices, so the values are
    staged): enough fields that the staged values exceed the fixed array."""
    print("\n[no-alloc dataUpdate] staging larger than the fixed array", flush=True)
    c, t = _connected("three")
    try:
        L = _lib(t)
        maxb = L.tioga_amd_dataupdate_noalloc_max_bytes()
        assert maxb == 64 << 20
        nrec = sum(t.donor_info(b.tag)[0].size // 4 for b in c.blocks)
        nvar = maxb // (8 * nrec) + 8
        qs = [np.zeros(b.nnodes * nvar) for b in c.blocks]
        rc, status, msg, _ = _noalloc(t, qs, nvar, b"row")
        assert rc == -1 and status != 0
        assert "more than" in msg, msg
        # and the entry point still works at a size that fits
        rng = np.random.default_rng(3)
        qs = [rng.uniform(-1, 1, b.nnodes * 3) for b in c.blocks]
        want = t.data_update_copies(qs, 3, 0)
        rc, status, msg, got = _noalloc(t, qs, 3, b"row")
        assert rc == 0, msg
        assert all(np.array_equal(g, w) for g, w in zip(got, want))
    finally:
        t.close()


def test_noalloc_update_needs_connectivity():
    from tioga_amd.tioga import Tioga

    c = cases.build("tiny")
    t = Tioga()
    try:
        for b in c.blocks:
            t.register_grid(b)
        t.preprocess()
        qs = [np.zeros(b.nnodes * 2) for b in c.blocks]
        rc, status, msg, _ = _noalloc(t, qs, 2, b"row")
        assert rc == -1 and status != 0 and "performconnectivity" in msg, msg
    finally:
        t.close()
