"""Serving: requests that move one body rigidly over a resident background.

Each request's result is the two-block connectivity of the moved body with
the background.  The tests check that a result is the same when the request
is served alone and when it is served concurrently with others (four slots),
and that it equals a fresh two-block connectivity of the moved coordinates
by the engine and by the oracle; that the resident-tree mode gives the same
results wherever it reports no tie-band query; and that the memory budget
is obeyed (a budget below any request serialises them, and a request's
charge bounds what its slot held once that block has been seen)."""
import numpy as np
import pytest

from tioga_amd import workloads as W

pytestmark = pytest.mark.gpu


def setup_blocks():
    rng = np.random.default_rng(71)
    shells = []
    for t in range(1, 4):
        shells.append(W.sphere_shell(float(rng.uniform(800, 3000)), float(rng.uniform(0.08, 0.14)), t,
                                     rng.uniform(-0.4, 0.4, 3), W.shoemake_rotation(rng.uniform(size=3))))
    bg = W.cart_box(18, tag=4, jitter=0.1, seed=72)
    reqs = W.motion_requests(10, 3, seed=73)
    return shells, bg, reqs


def serve_all(shells, bg, reqs, nslots, together, budget_gb=16.0, resident=False):
    from tioga_amd.tioga import Tioga

    t = Tioga()
    try:
        for b in shells + [bg]:
            t.register_grid(b)
        t.preprocess()
        t.serve_open(bg.tag, nslots=nslots, budget_gb=budget_gb, resident_tree=resident)
        cen = {s.tag: s.x.mean(axis=0) for s in shells}
        size = {s.tag: s.nnodes for s in shells}
        out = []
        if together:
            ids = [t.serve_submit(tag, R, cen[tag], sh) for tag, R, sh in reqs]
            out = [t.serve_wait(i, size[tag]) for i, (tag, _, _) in zip(ids, reqs)]
        else:
            for tag, R, sh in reqs:
                out.append(t.serve_wait(t.serve_submit(tag, R, cen[tag], sh), size[tag]))
        t.serve_close()
    finally:
        t.close()
    return out


def test_serving_same_alone_and_in_mix():
    shells, bg, reqs = setup_blocks()
    print("[serving] alone", flush=True)
    alone = serve_all(shells, bg, reqs, 1, False)
    print("[serving] mix of 4 slots", flush=True)
    mix = serve_all(shells, bg, reqs, 4, True)
    for a, m in zip(alone, mix):
        assert np.array_equal(a["shell_iblank"], m["shell_iblank"])
        assert np.array_equal(a["bg_nodes"], m["bg_nodes"])
        assert np.array_equal(a["bg_vals"], m["bg_vals"])
        assert a["nsearch"] == m["nsearch"]
    assert any(len(a["bg_nodes"]) > 0 for a in alone)


def test_serving_equals_two_block_connectivity():
    from oracle.oracle import Oracle
    from tioga_amd.tioga import Tioga

    shells, bg, reqs = setup_blocks()
    served = serve_all(shells, bg, reqs[:3], 2, True)
    byt = {s.tag: s for s in shells}
    for k, (tag, R, sh) in enumerate(reqs[:3]):
        s0 = byt[tag]
        moved = W.Block(tag, W.rigid_apply(s0.x, R, s0.x.mean(axis=0), sh), s0.types, s0.wbc, s0.obc)
        print(f"[serving] request {k} vs a fresh two-block engine", flush=True)
        t = Tioga()
        try:
            t.register_grid(moved)
            t.register_grid(bg)
            t.preprocess()
            t.connect()
            ib_shell, ib_bg = t.iblank[0].copy(), t.iblank[1].copy()
        finally:
            t.close()
        nz = np.nonzero(ib_bg != 1)[0]
        assert np.array_equal(served[k]["shell_iblank"], ib_shell)
        assert np.array_equal(served[k]["bg_nodes"], nz)
        assert np.array_equal(served[k]["bg_vals"], ib_bg[nz])
        if k == 0:
            o = Oracle([moved, bg])
            o.profile()
            o.connect()
            assert np.array_equal(o.iblank[0], ib_shell)
            assert np.array_equal(o.iblank[1], ib_bg)
            o.close()


def test_serving_resident_tree_same_outside_tie_band():
    """Requests against one tree over all background cells equal the
    per-request trees' results wherever the mode reports no tie-band query
    (and the tie band is a small part of the searches)."""
    shells, bg, reqs = setup_blocks()
    print("[serving] per-request trees", flush=True)
    exact = serve_all(shells, bg, reqs, 2, True)
    print("[serving] resident background tree", flush=True)
    res = serve_all(shells, bg, reqs, 2, True, resident=True)
    clean = 0
    for a, r in zip(exact, res):
        assert a["tieband"] == 0  # no resident tree, nothing counted
        assert r["tieband"] <= 0.01 * r["nsearch"] + 1
        if r["tieband"] == 0:
            clean += 1
            assert np.array_equal(a["shell_iblank"], r["shell_iblank"])
            assert np.array_equal(a["bg_nodes"], r["bg_nodes"])
            assert np.array_equal(a["bg_vals"], r["bg_vals"])
        assert a["nsearch"] == r["nsearch"]
    assert clean >= len(reqs) // 2, [r["tieband"] for r in res]


def test_serving_budget_below_one_request_serialises():
    """A budget smaller than any single request: every request still runs
    (one always runs), one at a time, with the same results."""
    shells, bg, reqs = setup_blocks()
    print("[serving] budget below one request, 4 slots", flush=True)
    tight = serve_all(shells, bg, reqs, 4, True, budget_gb=1e-6)
    spans = sorted((r["t_start"], r["t_done"]) for r in tight)
    for (s0, d0), (s1, d1) in zip(spans, spans[1:]):
        assert s1 >= d0 - 1e-6, "two requests ran at once under a budget below one request"
    ref = serve_all(shells, bg, reqs, 1, False)
    for a, m in zip(ref, tight):
        assert np.array_equal(a["shell_iblank"], m["shell_iblank"])
        assert np.array_equal(a["bg_nodes"], m["bg_nodes"])


def test_serving_charge_bounds_what_a_slot_holds():
    """Once a block has been served, a request on it is charged what its
    slot then holds to within 1% (the slot's buffers only grow, so buffers
    kept from other blocks can add a little); under a budget of two such
    charges the requests still complete with the same results."""
    shells, bg, reqs = setup_blocks()
    print("[serving] measured charges", flush=True)
    warm = serve_all(shells, bg, reqs, 1, False)
    first = {}
    for (tag, _, _), r in zip(reqs, warm):
        assert r["peak_bytes"] > 0 and r["est_bytes"] > 0
        if tag in first:  # the block was seen before: its charge bounds the slot's peak
            assert r["est_bytes"] >= 0.99 * r["peak_bytes"], (tag, r["est_bytes"], r["peak_bytes"])
        first.setdefault(tag, r)
    budget = 2.0 * max(r["peak_bytes"] for r in warm) / 1e9
    two = serve_all(shells, bg, reqs, 4, True, budget_gb=budget)
    # whenever two or more requests ran at once, their charges fit the budget
    for r in two:
        t = r["t_start"] + 1e-7
        running = [q for q in two if q["t_start"] <= t < q["t_done"]]
        if len(running) > 1:
            assert sum(q["est_bytes"] for q in running) <= budget * 1e9 * (1 + 1e-9)
    for a, m in zip(warm, two):
        assert np.array_equal(a["shell_iblank"], m["shell_iblank"])
