"""Writes the golden fixtures under tests/golden/ from the reference TIOGA
compiled from its own sources (oracle/Makefile, target ``ref``), run on one
MPI rank by oracle/ref_driver.cpp.

For each case of cases.py (the AMR gather cases run the driver's AMR mode),
``<case>.npz`` holds every array that the driver
dumps whose size is at most ``FULL_LIMIT`` elements, and ``<case>.json``
holds the SHA-256 digest, dtype and size of every array (the hole maps and
config-1 arrays are pinned by digest only).  Run in a container where
/root/reference exists:

    make -C oracle ref && python tests/golden/make_golden.py [case ...]
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

from oracle import oracle as O  # noqa: E402
import cases  # noqa: E402

FULL_LIMIT = 20000


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main(names):
    for name in names:
        with tempfile.TemporaryDirectory() as td:
            cp, op = os.path.join(td, "case.bin"), os.path.join(td, "out.bin")
            if name in cases.AMR_CASES:
                c = cases.build_amr(name)
                O.write_case(cp, c.blocks, c.qs, c.nvar, c.interptype, amr=(c.grid, c.patch_q))
            else:
                c = cases.build(name)
                O.write_case(cp, c.blocks, c.qs, c.nvar, c.interptype, c.isym, c.nfringe, c.mexclude,
                             c.moves, reduce=c.reduce)
            res = O.run_reference(cp, op, cwd=td)
        meta = {}
        full = {}
        for k, a in sorted(res.items()):
            if k.endswith("time_connectivity"):
                continue
            meta[k] = {"dtype": str(a.dtype), "size": int(a.size), "sha256": digest(a)}
            if a.size <= FULL_LIMIT and name != "config1":
                full[k] = a
        with open(os.path.join(HERE, f"{name}.json"), "w") as f:
            json.dump(meta, f, indent=0, sort_keys=True)
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **full)
        print(name, len(meta), "arrays,", len(full), "stored in full")


if __name__ == "__main__":
    main(sys.argv[1:] or cases.ALL_CASES + cases.AMR_CASES)
