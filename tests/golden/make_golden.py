"""Writes the golden fixtures under tests/golden/ from the reference TIOGA
compiled from its own sources (oracle/Makefile, target ``ref``), run by
oracle/ref_driver.cpp on one MPI rank (the multi-rank cases under mpiexec,
with the packets every rank sends: ``pk.<exchange>.<src>.<dst>.*``).

For each case of cases.py (the AMR gather cases run the driver's AMR mode),
``<case>.npz`` holds every array that the driver
dumps whose size is at most ``FULL_LIMIT`` elements, and ``<case>.json``
holds the SHA-256 digest, dtype and size of every array (the hole maps and
hex48_pair arrays are pinned by digest only).  Run in a container where
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


def packet_stages(gids: bool, nmoves: int, reduce: bool):
    """Names of the sendRecvPackets calls of ref_driver.cpp's call sequence,
    in order: per step the connectivity (exchangeBoxes, exchangeSearchData,
    the four rounds of exchangeDonors), getReceptorInfo when node ids are
    registered, dataUpdate; after reduce_fringes its two rounds, then
    getReceptorInfo and dataUpdate again."""
    out = []
    for st in range(nmoves + 1):
        pre = "" if st == 0 else f"s{st}."
        seq = ["boxes", "search", "donors1", "donors2", "donors3", "donors4"]
        seq += (["receptor_info"] if gids else []) + ["update"]
        if reduce:
            seq += ["rf.donors3", "rf.donors4"] + (["rf.receptor_info"] if gids else []) + ["rf.update"]
        out += [pre + x for x in seq]
    return out


def packet_arrays(packets, gids: bool, nmoves: int, reduce: bool):
    """pk.<stage>.<src>.<dst>.{ints,reals,counts} for every packet of the
    run.  exchangeBoxes sends 18 doubles per pair of which it fills the first
    six (exchangeBoxes.C:184-238); only those are kept."""
    names = packet_stages(gids, nmoves, reduce)
    out = {}
    for src, calls in enumerate(packets):
        if not any(calls):  # a rank without partners sends nothing (and skips some exchanges)
            continue
        if len(calls) != len(names):
            raise RuntimeError(f"rank {src}: {len(calls)} exchanges, expected {len(names)}")
        for nm, call in zip(names, calls):
            for dst, (ints, reals, ni, nr) in call.items():
                p = f"pk.{nm}.{src}.{dst}."
                if nm.endswith("boxes"):
                    reals = reals[:2 * ni]
                out[p + "ints"] = ints.astype(np.int32)
                out[p + "reals"] = reals.astype(np.float64)
                out[p + "counts"] = np.array([ni, nr], np.int32)
    return out


def main(names):
    for name in names:
        with tempfile.TemporaryDirectory() as td:
            cp, op = os.path.join(td, "case.bin"), os.path.join(td, "out.bin")
            if name in cases.AMR_CASES + cases.AMR_FULL_CASES:
                c = cases.build_amr(name)
                O.write_case(cp, c.blocks, c.qs, c.nvar, c.interptype, amr=(c.grid, c.patch_q), amr_full=c.full)
            else:
                c = cases.build(name)
                O.write_case(cp, c.blocks, c.qs, c.nvar, c.interptype, c.isym, c.nfringe, c.mexclude,
                             c.moves, reduce=c.reduce, ranks=c.ranks, points=c.points)
            if getattr(c, "ranks", None) is not None:
                gids = any(b.node_gid is not None for b in c.blocks)
                res, packets = O.run_reference(cp, op, cwd=td, nranks=max(c.ranks) + 1, gids=gids)
                res.update(packet_arrays(packets, gids, len(c.moves), c.reduce))
            else:
                res = O.run_reference(cp, op, cwd=td, gids=any(b.node_gid is not None for b in c.blocks))
        meta = {}
        full = {}
        for k, a in sorted(res.items()):
            if k.endswith("time_connectivity"):
                continue
            meta[k] = {"dtype": str(a.dtype), "size": int(a.size), "sha256": digest(a)}
            if a.size <= FULL_LIMIT and name != "hex48_pair":
                full[k] = a
        with open(os.path.join(HERE, f"{name}.json"), "w") as f:
            json.dump(meta, f, indent=0, sort_keys=True)
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **full)
        print(name, len(meta), "arrays,", len(full), "stored in full")


if __name__ == "__main__":
    main(sys.argv[1:] or cases.ALL_CASES + cases.AMR_CASES + cases.AMR_FULL_CASES + cases.POINT_CASES + cases.RANK_CASES)
