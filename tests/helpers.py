"""Shared test helpers: run a parity case through the oracle in the exact
call sequence of oracle/ref_driver.cpp and compare against the fixtures."""
from __future__ import annotations

import hashlib
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def load_golden(name: str):
    with open(os.path.join(GOLDEN, f"{name}.json")) as f:
        meta = json.load(f)
    full = dict(np.load(os.path.join(GOLDEN, f"{name}.npz")))
    return meta, full


def run_oracle_case(case):
    """Every array ref_driver.cpp dumps, computed by the oracle."""
    from oracle.oracle import Oracle

    o = Oracle(case.blocks, case.isym, case.nfringe, case.mexclude)
    out = {}
    try:
        for s in range(len(case.moves) + 1):
            if s > 0:
                for bi, xb in enumerate(case.moves[s - 1]):
                    if xb is not None:
                        o.set_coords(bi, xb)
            o.profile()
            o.connect()
            pre = "" if s == 0 else f"s{s}."
            for k, v in o.results(case.qs, case.nvar, case.interptype).items():
                out[pre + k] = v
            if getattr(case, "reduce", False):
                o.reduce_fringes()
                for k, v in o.results(case.qs, case.nvar, case.interptype).items():
                    out[pre + "rf." + k] = v
        out["weight_aborts"] = o.L.orc_weight_aborts()
    finally:
        o.close()
    return out


def compare_to_golden(name: str, got: dict, skip=(), skip_prefix=()):
    """Returns a list of mismatch descriptions (empty when bit-exact)."""
    meta, full = load_golden(name)
    bad = []
    for k, m in meta.items():
        if any(k.endswith(s) for s in skip) or any(k.startswith(s) for s in skip_prefix):
            continue
        if k not in got:
            bad.append(f"{k}: missing")
            continue
        a = np.ascontiguousarray(got[k]).astype(m["dtype"])
        if a.size != m["size"]:
            bad.append(f"{k}: size {a.size} != {m['size']}")
            continue
        if digest(a) != m["sha256"]:
            if k in full:
                d = np.nonzero(full[k] != a)[0]
                bad.append(f"{k}: {d.size} of {a.size} differ, first at {d[:4]}")
            else:
                bad.append(f"{k}: digest differs")
    return bad


def set_test_callbacks(t, case):
    """Hand the engine the high-order callbacks of tests/cxx/ho_callbacks.c
    (the ones the reference driver sets), on block 0 as the receptor mesh and
    block 1 as the donor mesh; returns what must stay alive."""
    import ctypes as C

    lib = C.CDLL(os.path.join(os.path.dirname(os.path.abspath(__file__)), "cxx", "libho_callbacks.so"))
    r, d = case.blocks[0].normalized(), case.blocks[1].normalized()
    keep = [lib, r, d]
    P = C.c_void_p
    lib.ho_cb_set.argtypes = [P, P, C.c_int, P, P, C.c_int, C.c_int, C.c_int]
    lib.ho_cb_set(r.x.ctypes.data, r.types[0][1].ctypes.data, r.types[0][0], d.x.ctypes.data,
                  d.types[0][1].ctypes.data, d.types[0][0], d.nnodes, case.nvar)
    fns = [C.cast(getattr(lib, n), P) for n in
           ("ho_nodes_per_cell", "ho_receptor_nodes", "ho_inclusion", "ho_frac", "ho_modal")]
    t.L.tioga_set_highorder_callback_.argtypes = [P] * 5
    t.L.tioga_set_highorder_callback_(*fns)
    t._check()
    return keep


def run_engine_case(case, batched: bool = True):
    """Every array ref_driver.cpp dumps, computed by the HIP engine through
    its C ABI, in the driver's call sequence (batched or block-by-block form)."""
    from tioga_amd.tioga import Tioga

    t = Tioga()
    out = {}
    try:
        t.set_batching(batched)
        t.set_params(case.isym, case.nfringe, case.mexclude)
        ranked = getattr(case, "ranks", None) is not None
        ranks = case.ranks if ranked else [0] * len(case.blocks)
        gids = any(b.node_gid is not None for b in case.blocks)
        for b, r in zip(case.blocks, ranks):
            t.register_grid(b, rank=r)
        callbacks = getattr(case, "points", 0) == 2
        keep = set_test_callbacks(t, case) if callbacks else None
        if ranked:
            t.record_packets(True)
        conn_ex = ["boxes", "search", "donors1", "donors2", "donors3", "donors4"] + (["receptor_info"] if gids else [])
        for s in range(len(case.moves) + 1):
            if s > 0:
                for bi, xb in enumerate(case.moves[s - 1]):
                    if xb is not None:
                        t.set_coords(bi, xb)
            t.preprocess()
            t.connect()
            pre = "" if s == 0 else f"s{s}."
            # with callbacks the C ABI's dataupdate is the point update: no node update here
            for k, v in t.results(None if callbacks else case.qs, case.nvar, case.interptype).items():
                out[pre + k] = v
            if ranked:  # what the exchanges sent (the dataUpdate ran inside results)
                for ex in conn_ex + ["update"]:
                    out.update(t.packets(ex, pre + ex))
            if getattr(case, "points", False):  # performConnectivityHighOrder, point update
                t.connect_points()
                for k, v in t.points_results(case.qs, case.nvar, case.interptype).items():
                    out[pre + "hp." + k] = v
            if getattr(case, "reduce", False):
                t.reduce_fringes()
                for k, v in t.results(case.qs, case.nvar, case.interptype).items():
                    out[pre + "rf." + k] = v
                if ranked:
                    for ex in ["rf.donors3", "donors4"] + (["receptor_info"] if gids else []) + ["update"]:
                        out.update(t.packets(ex, pre + ("rf." + ex if not ex.startswith("rf.") else ex)))
        out["weight_aborts"] = t.weight_aborts()
        del keep
    finally:
        t.close()
    return out


def run_oracle_amr_case(case):
    """The arrays ref_driver.cpp dumps in AMR mode, computed by the oracle."""
    from oracle.oracle import Oracle

    o = Oracle(case.blocks)
    try:
        o.register_amr(case.grid, case.patch_q)
        o.profile()
        rc = o.connect_amr()
        out = o.amr_results(case.qs, case.nvar, case.interptype)
        out["status"] = rc
    finally:
        o.close()
    return out


def run_engine_amr_case(case):
    """The arrays ref_driver.cpp dumps in AMR mode, computed by the HIP
    engine through the C ABI."""
    from tioga_amd.tioga import Tioga

    t = Tioga()
    full = getattr(case, "full", False)
    try:
        for b in case.blocks:
            t.register_grid(b)
        t.register_amr(case.grid, case.patch_q)
        t.preprocess()
        if full:  # the hole maps the Cartesian donor processing reads
            t.connect()
        t.connect_amr()
        out = t.amr_results(case.qs, case.nvar, case.interptype, full=full)
    finally:
        t.close()
    return out
