"""The headline search path (query sets) against the oracle, bit for bit.

A query set is searched the way MeshBlock::search (search.C:107-269)
searches it: the donor block's cells that overlap the set's OBB go into an
ADT (ADT.C:34-135) and every point walks it (searchADTrecursion.C:26-101),
with checkContainment's isoparametric test (checkContainment.C:27-96).
The engine runs that through tioga_amd_query_set_create / _search and the
timed passes of tioga_amd_time_query_sets (no host capacity check between
the walk and the evaluation).  The oracle builds the same tree from the
same set's OBB and searches every point sequentially.

Workloads: the mixed tet/pyramid/prism/hex shell of the large plain-method
configuration in a jittered hex box, at three sizes, in both directions
(every box node against the shell, every shell node against the box);
trees of more than 16384 cells are included.  Edge cases: points exactly on
cell vertices and faces (every node of a block searched against that same
block, so most points lie in several cells and the result is decided by the
walk order), sets of 1, 2, 3 and 5 points (the n < 4 branches of findOBB),
a set entirely outside the tree, and a candidate list that starts far too
small (the regrow path of the checked search)."""
import numpy as np
import pytest

from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

SIZES = {
    # (shell nphi, ntheta, box cells per side)
    "small": (24, 12, 20),
    "mid": (64, 32, 64),      # shell tree 122 880 cells, box tree about 57 000
    "large": (128, 64, 128),  # shell tree 491 520 cells, box tree about 450 000
}


def _blocks(size):
    from tioga_amd import workloads as W

    nphi, nth, nbox = SIZES[size]
    return [W.mixed_shell(nphi, nth, tag=1), W.cart_box(nbox, tag=2, jitter=0.1, seed=2)]


def _threads():
    import os

    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def _oracle(blocks):
    o = Oracle(blocks)
    o.profile()
    return o


def _engine(blocks):
    from tioga_amd.tioga import Tioga

    t = Tioga()
    for b in blocks:
        t.register_grid(b)
    t.preprocess()
    return t


def _check_sets(t, o, sets, label):
    """Search every set on the device (checked path), then time two passes
    (unchecked path) and read the donors they left; both equal the oracle."""
    ids = [t.query_set(btag, xq) for btag, xq in sets]
    want, nelem = [], []
    for qid, (btag, xq) in zip(ids, sets):
        d_ref, ne_ref, _ = o.search_points_obb(btag, xq, xq, _threads())
        assert t.query_nelem(qid) == ne_ref, f"{label}: tree of set {qid} has {t.query_nelem(qid)} cells, oracle {ne_ref}"
        got = t.query_search(qid, xq.shape[0])
        bad = np.nonzero(got != d_ref)[0]
        assert bad.size == 0, (f"{label}: set {qid} (block {btag}, {xq.shape[0]} points): {bad.size} donors differ, "
                               f"first {bad[:5]} got {got[bad[:5]]} want {d_ref[bad[:5]]}")
        want.append(d_ref)
        nelem.append(ne_ref)
    t.time_query_sets(2)
    for qid, (btag, xq), d_ref in zip(ids, sets, want):
        got = t.query_donors(qid, xq.shape[0])
        assert np.array_equal(got, d_ref), f"{label}: timed pass of set {qid} left different donors"
    t.query_clear()
    return want, nelem


@pytest.mark.parametrize("size", ["small", "mid", "large"])
def test_query_sets_match_oracle(size):
    """Every box node against the shell and every shell node against the
    box, checked and timed passes, bit for bit; the found counts are not
    trivial."""
    print(f"\n[query sets] mixed shell + jittered box, size {size}", flush=True)
    blocks = _blocks(size)
    o = _oracle(blocks)
    t = _engine(blocks)
    try:
        sets = [(blocks[0].tag, blocks[1].x), (blocks[1].tag, blocks[0].x)]
        want, nelem = _check_sets(t, o, sets, size)
        assert (want[0] >= 0).sum() > 0 and (want[1] >= 0).sum() > 0
        if size != "small":  # both trees above the one-workgroup segment size of the ADT build
            assert min(nelem) > 16384, nelem
    finally:
        t.close()
        o.close()


def test_query_sets_points_on_vertices_and_faces():
    """Nodes of a block searched against the same block: every point is a
    vertex of up to 8 cells (and lies on their faces), so which cell the
    reference reports depends on its preorder walk and first-hit rule.  An
    unjittered box and the mixed shell (tets, pyramids and prisms share
    faces there)."""
    from tioga_amd import workloads as W

    print("\n[query sets] nodes against their own block (vertex and face ties)", flush=True)
    blocks = [W.mixed_shell(32, 16, tag=1), W.cart_box(24, tag=2)]
    o = _oracle(blocks)
    t = _engine(blocks)
    try:
        sets = [(blocks[0].tag, blocks[0].x), (blocks[1].tag, blocks[1].x)]
        # the face centres of the box's cells as well
        m = 25
        h = 2.0 / 24
        g = -1.0 + h * (np.arange(24) + 0.5)
        gi = -1.0 + h * np.arange(m)
        fx = np.stack(np.meshgrid(gi, g, g, indexing="ij"), -1).reshape(-1, 3)
        sets.append((blocks[1].tag, np.ascontiguousarray(fx)))
        want, _ = _check_sets(t, o, sets, "ties")
        assert all((w >= 0).mean() > 0.5 for w in want)
    finally:
        t.close()
        o.close()


def test_query_sets_tiny_and_outside():
    """Sets of 1, 2, 3 and 5 points (findOBB's n < 4 branches and the
    smallest trees) and a set that lies entirely outside the donor block."""
    print("\n[query sets] tiny sets and a set outside the block", flush=True)
    blocks = _blocks("small")
    o = _oracle(blocks)
    t = _engine(blocks)
    rng = np.random.default_rng(7)
    try:
        sets = []
        for n in (1, 2, 3, 5):
            sets.append((blocks[1].tag, rng.uniform(-0.9, 0.9, (n, 3))))
            sets.append((blocks[0].tag, rng.uniform(-0.5, 0.5, (n, 3))))
        sets.append((blocks[0].tag, rng.uniform(2.0, 3.0, (1000, 3))))
        want, _ = _check_sets(t, o, sets, "tiny")
        assert (want[-1] == -1).all()
    finally:
        t.close()
        o.close()


def test_query_sets_empty_set_rejected():
    """A set of no points is an error of the host call (the reference's
    search returns at once for nsearch == 0; there is nothing to search)."""
    from tioga_amd.tioga import TiogaError

    blocks = _blocks("small")
    t = _engine(blocks)
    try:
        with pytest.raises(TiogaError):
            t.query_set(blocks[0].tag, np.zeros((0, 3)))
    finally:
        t.close()


def test_query_sets_candidate_list_regrows():
    """The checked search starting from a candidate list of 1000 entries
    (far below the ~10^5 candidates of these sets): walks that do not fit
    mark their reserved slots empty, the host grows the list and walks
    again, and the donors equal the oracle's, in the checked and in the
    timed passes."""
    print("\n[query sets] candidate list regrow from capacity 1000", flush=True)
    blocks = _blocks("mid")
    o = _oracle(blocks)
    t = _engine(blocks)
    try:
        t.search_capacity(1000)
        assert t.search_capacity() == 1000
        sets = [(blocks[1].tag, blocks[0].x), (blocks[0].tag, blocks[1].x)]
        _check_sets(t, o, sets, "regrow")
        assert t.search_capacity() > 1000
    finally:
        t.close()
        o.close()


@pytest.mark.parametrize("tuning", [dict(morton_bits=0, seg=256, eval_grid=1024, bins=0),
                                    dict(morton_bits=30, seg=1024, eval_grid=8192, bins=1),
                                    dict(morton_bits=9, seg=64, eval_grid=64, bins=64),
                                    dict(morton_bits=21, seg=16, eval_grid=4096, bins=4096),
                                    dict(morton_bits=0, seg=1024, eval_grid=8192, bins=4096)])
def test_query_sets_any_tuning(tuning):
    """The search's parameters (candidates from bin grids of any density or
    from the tree walk, bin lists cut into work items of any length, Morton
    bits, evaluation grid) change no donor; very coarse bins cut into short
    items make many items per wave, and cut into long ones leave most of
    every list to the wave-wide scan of the longer lists."""
    print(f"\n[query sets] tuning {tuning}", flush=True)
    blocks = _blocks("mid")
    o = _oracle(blocks)
    t = _engine(blocks)
    try:
        got = t.search_tuning(**tuning)
        assert got == tuning
        sets = [(blocks[0].tag, blocks[1].x), (blocks[1].tag, blocks[0].x)]
        _check_sets(t, o, sets, f"tuning {tuning}")
        st = t.search_stats()
        assert st["listing"] == ("bins" if tuning["bins"] else "walk")
        if tuning["bins"] == 4096:
            assert st["listing_items"] > 2 * (st["walking_queries"] // 64 + 1), st
    finally:
        t.close()
        o.close()


def test_query_sets_graph_replays():
    """The timed passes replay the search launch from a HIP graph recorded
    after the checked pass: several replays of one graph, and graphs
    recorded again with eager searches in between, all leave the oracle's
    donors; so does the kernel-by-kernel form."""
    print("\n[query sets] graph replays", flush=True)
    blocks = _blocks("mid")
    o = _oracle(blocks)
    t = _engine(blocks)
    try:
        sets = [(blocks[0].tag, blocks[1].x), (blocks[1].tag, blocks[0].x)]
        ids = [t.query_set(btag, xq) for btag, xq in sets]
        want = [o.search_points_obb(btag, xq, xq, _threads())[0] for btag, xq in sets]
        for rep in range(3):
            t.time_query_sets(4)
            for qid, (_, xq), d in zip(ids, sets, want):
                assert np.array_equal(t.query_donors(qid, xq.shape[0]), d), f"replay round {rep}, set {qid}"
            assert np.array_equal(t.query_search(ids[rep % 2], sets[rep % 2][1].shape[0]), want[rep % 2])
        t.search_graph(False)
        t.time_query_sets(2)
        for qid, (_, xq), d in zip(ids, sets, want):
            assert np.array_equal(t.query_donors(qid, xq.shape[0]), d), f"kernel by kernel, set {qid}"
        t.search_graph(True)
    finally:
        t.close()
        o.close()
