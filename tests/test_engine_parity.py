"""The HIP engine against the reference, bit for bit.

Each case runs through the C ABI in the sequence of the reference driver
(register, preprocess, performconnectivity, dataupdate; moving cases repeat
preprocess/performconnectivity after rigid motions) and every array is
compared with the fixtures written from the compiled reference: donorId,
iblank, iblank_cell, receptor lists, weights and updated q bit-exact, plus
every intermediate array (resolutions, OBBs, search arrays, xtag, ADT, hole
maps)."""
import numpy as np
import pytest

import cases
from helpers import compare_to_golden, run_engine_amr_case, run_engine_case, run_oracle_amr_case, run_oracle_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("batched", [True, False], ids=["batched", "per_block"])
@pytest.mark.parametrize("name", cases.ALL_CASES + cases.POINT_CASES + cases.RANK_CASES)
def test_engine_matches_reference(name, batched):
    print(f"\n[engine parity] case {name} batched={batched}", flush=True)
    got = run_engine_case(cases.build(name), batched)
    assert got["weight_aborts"] == 0
    bad = compare_to_golden(name, got)
    assert not bad, "\n".join(bad[:20])


def test_engine_batched_equals_per_block_many_shells():
    """Eight overlapping shells of uneven sizes in a jittered background:
    the batched form and the block-by-block form give the same arrays."""
    from tioga_amd import workloads as W

    rng = np.random.default_rng(61)
    blocks = []
    for t in range(1, 9):
        rot = W.shoemake_rotation(rng.uniform(size=3))
        blocks.append(W.sphere_shell(float(rng.uniform(300, 6000)), float(rng.uniform(0.05, 0.12)), t,
                                     rng.uniform(-0.5, 0.5, 3), rot))
    blocks.append(W.cart_box(20, tag=9, jitter=0.1, seed=62))
    c = cases.Case("shells8", blocks, nvar=2)
    print("[engine parity] 8 shells, batched vs per block", flush=True)
    a = run_engine_case(c, True)
    b = run_engine_case(c, False)
    for k in b:
        assert k in a, k
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k
    o = run_oracle_case(c)
    for k in o:
        if k == "weight_aborts":
            continue
        assert np.array_equal(np.asarray(a[k]), np.asarray(o[k])), k


def test_engine_matches_oracle_random_seeds():
    """Cases with no fixture: random jittered background and shifted body,
    engine vs the oracle."""
    from tioga_amd import workloads as W

    for seed in (21, 22):
        rng = np.random.default_rng(seed)
        c = cases.Case(f"rand{seed}", [W.latlong_shell(12, 8, 4, 0.2, 0.5, 1.08, tag=1,
                                                       center=tuple(rng.uniform(-0.2, 0.2, 3))),
                                       W.cart_box(9, tag=2, jitter=0.15, seed=seed)], nvar=2)
        print(f"[engine parity] random seed {seed}", flush=True)
        a = run_engine_case(c)
        b = run_oracle_case(c)
        for k in b:
            if k == "weight_aborts":
                continue
            assert k in a, k
            assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k


@pytest.mark.parametrize("name", cases.AMR_CASES + cases.AMR_FULL_CASES)
def test_engine_amr_matches_reference(name):
    """AMR gather direction: receptors, donor patches, patch interpolation
    records (list order, stencils, weights), iblanks and updated q; in the
    full cases also the Cartesian-receptor direction: the near-body search
    of the patch cells, patch iblanks (holes, receptors), cancelled donors,
    interpListCart and the updated patch solutions."""
    print(f"\n[engine parity] AMR case {name}", flush=True)
    got = run_engine_amr_case(cases.build_amr(name))
    bad = compare_to_golden(name, got)
    assert not bad, "\n".join(bad[:20])


def test_engine_amr_random_hierarchy_vs_oracle():
    """No fixture: a shuffled 3-level hierarchy with 8^3-cell patches and a
    randomly placed shell whose every node is a receptor, engine vs oracle."""
    from tioga_amd import workloads as W

    rng = np.random.default_rng(51)
    grid = W.amr_hierarchy((-1.0, -1.0, -1.0), 2.0 / 32,
                           [((0, 0, 0), (32, 32, 32)), ((8, 16, 8), (56, 48, 56)), ((40, 40, 40), (88, 88, 72))],
                           patch=8, nf=1, shuffle_seed=52, local_seed=53)
    body = W.all_receptors(W.latlong_shell(20, 10, 6, 0.2, 0.45, 1.1, tag=1,
                                           center=tuple(rng.uniform(-0.3, 0.3, 3))))
    c = cases.AmrCase("rand_amr", [body], grid, nvar=4)
    print("[engine parity] random AMR hierarchy", flush=True)
    a = run_engine_amr_case(c)
    b = run_oracle_amr_case(c)
    for k in b:
        if k == "status":
            continue
        assert k in a, k
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k
