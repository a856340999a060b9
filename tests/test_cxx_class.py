"""The engine's C++ class TIOGA::tioga (include/tioga.h), called directly.

tests/cxx/tioga_class_driver (built by __graft_entry__.build()) reads the
case files the golden fixtures were made from and drives the class through
the reference's call sequence (registerGridData with global ids where the
case has them, profile, performConnectivity, registerSolution, dataUpdate,
getDonorCount/getDonorInfo, getReceptorInfo, reduce_fringes); every array it
writes must equal the compiled reference's, bit for bit: iblank,
iblank_cell, donor records, weights, receptor info (64-bit donor cell ids
packed in two ints, tioga.C:572-671) and the updated fields.  The global-id
cases were written by the reference built with TIOGA_HAS_NODEGID, whose
coincident-point dedupe is the node-id map of search.C:48-74."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import cases
from helpers import digest, load_golden

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cxx", "tioga_class_driver")

CASES = ["tiny", "userres", "moving", "three", "tiny_rf", "tiny_gid", "seam_gid", "seam_gid_rf"]


@pytest.mark.parametrize("name", CASES)
def test_cxx_class_matches_reference(name):
    from oracle import oracle as O

    assert os.path.exists(DRIVER), "tests/cxx/tioga_class_driver is built by __graft_entry__.build()"
    print(f"\n[C++ class] case {name}", flush=True)
    c = cases.build(name)
    meta, full = load_golden(name)
    with tempfile.TemporaryDirectory() as td:
        cp, op = os.path.join(td, "case.bin"), os.path.join(td, "out.bin")
        O.write_case(cp, c.blocks, c.qs, c.nvar, c.interptype, c.isym, c.nfringe, c.mexclude, c.moves,
                     reduce=c.reduce)
        r = subprocess.run([DRIVER, cp, op], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr[-2000:]
        got = O.read_result(op)
    assert got, "the driver wrote nothing"
    bad = []
    for k, a in got.items():
        assert k in meta, f"{k}: not in the reference's dump"
        m = meta[k]
        a = np.ascontiguousarray(a).astype(m["dtype"])
        if a.size != m["size"] or digest(a) != m["sha256"]:
            d = np.nonzero(full[k] != a)[0] if k in full and full[k].size == a.size else []
            bad.append(f"{k}: {len(d)} of {a.size} differ" if len(d) else f"{k}: differs")
    assert not bad, "\n".join(bad[:20])
    if "gid" in name:
        assert any(k.endswith("receptor_info") for k in got)
