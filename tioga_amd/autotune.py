"""Launch parameters of the donor search, chosen by measurement.

The search's parameters -- how candidates are listed (bin grids of about
`bins` elements per bin, or the tree walk when 0), the Morton bits of the
query order (the sort's passes against the listing's coherence), the
candidates each lane stages in LDS (occupancy against spill traffic), and
the workgroups of each containment-test kernel -- change no result, only
speed.  :func:`tune` measures the registered query sets of an
engine over a grid of values by coordinate descent (each parameter swept
with the others held, twice round) and returns the fastest; the table
``search_tuning.json`` beside this file keeps the chosen values per device
architecture and workload, and :func:`apply` sets them on an engine.

    python -m tioga_amd.autotune [--out table.json]   # (on a GPU) tunes the
                                                      # large plain-method
                                                      # workload
"""
from __future__ import annotations

import json
import os
import time
from typing import Dict, List, Optional, Tuple

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "search_tuning.json")

SPACE = {
    "bins": [0, 1, 2, 4, 8],
    "morton_bits": [0, 9, 12, 13, 14, 15, 21, 30],
    "seg": [64, 128, 256, 512, 1024],
    "eval_grid": [1024, 2048, 4096, 8192],
}
DEFAULT = {"morton_bits": 14, "seg": 256, "eval_grid": 8192, "bins": 8}


def arch() -> str:
    """The device architecture name the table is keyed by (gfx950 here), as
    the engine's HIP runtime reports it."""
    import ctypes as C

    from tioga_amd.tioga import load as load_lib

    L = load_lib()
    L.tioga_amd_device_arch.argtypes = [C.c_char_p, C.c_int]
    L.tioga_amd_device_arch.restype = C.c_int
    buf = C.create_string_buffer(64)
    if L.tioga_amd_device_arch(buf, 64) != 0:
        L.tioga_amd_clear_status()
        return "unknown"
    return buf.value.decode()


def _measure(tg, params: Dict[str, int], steps: int) -> float:
    """Device time of one whole pass (ordering, listing, tests, selection)."""
    tg.search_tuning(**params)
    tg.time_query_sets(1)  # warm-up at these parameters
    pass_ms, _, _ = tg.time_query_sets(steps)
    return pass_ms


def tune(tg, steps: int = 5, rounds: int = 2, log=None) -> Tuple[Dict[str, int], List[dict]]:
    """Coordinate descent over SPACE on the engine's registered query sets;
    returns (best parameters, every measurement as {params..., "ms"})."""
    best = dict(tg.search_tuning())
    seen: Dict[Tuple[int, ...], float] = {}
    trials: List[dict] = []

    def cost(p):
        key = tuple(p[k] for k in sorted(SPACE))
        if key not in seen:
            seen[key] = _measure(tg, p, steps)
            trials.append(dict(p, ms=round(seen[key], 4)))
            if log:
                log(f"autotune {p}: {seen[key]:.4f} ms")
        return seen[key]

    best_ms = cost(best)
    for _ in range(rounds):
        changed = False
        for name, values in SPACE.items():
            for v in values:
                p = dict(best, **{name: v})
                c = cost(p)
                if c < best_ms * 0.995:  # a change must win by more than the noise
                    best, best_ms, changed = p, c, True
        if not changed:
            break
    tg.search_tuning(**best)
    return best, trials


def load(workload: str, device: Optional[str] = None) -> Optional[Dict[str, int]]:
    """The table's parameters for this device and workload, if any."""
    try:
        with open(TABLE) as f:
            t = json.load(f)
    except OSError:
        return None
    entry = t.get(device or arch(), {}).get(workload)
    return None if entry is None else {k: int(entry.get(k, DEFAULT[k])) for k in DEFAULT}


def apply(tg, workload: str, device: Optional[str] = None) -> Dict[str, int]:
    """Sets the table's parameters (the defaults when the table has none)."""
    p = load(workload, device) or DEFAULT
    return tg.search_tuning(**p)


def save(workload: str, params: Dict[str, int], trials: List[dict], device: Optional[str] = None,
         path: str = TABLE) -> None:
    try:
        with open(path) as f:
            t = json.load(f)
    except OSError:
        t = {}
    dev = device or arch()
    t.setdefault(dev, {})[workload] = dict(params, measured=time.strftime("%Y-%m-%d"),
                                           trials=sorted(trials, key=lambda d: d["ms"]))
    with open(path, "w") as f:
        json.dump(t, f, indent=1, sort_keys=True)


def main() -> int:
    import argparse
    import sys

    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=TABLE, help="table to update (default: the package's)")
    args = ap.parse_args()
    sys.path.insert(0, os.path.dirname(HERE))
    from tioga_amd import workloads as W
    from tioga_amd.tioga import Tioga

    blocks = W.hex256_mixed_pair()
    tg = Tioga()
    for b in blocks:
        tg.register_grid(b)
    tg.preprocess()
    for btag, xq in ((blocks[0].tag, blocks[1].x), (blocks[1].tag, blocks[0].x)):
        tg.query_set(btag, xq)
    best, trials = tune(tg, log=lambda m: print(m, flush=True))
    tg.close()
    save("hex256_mixed_pair", best, trials, path=args.out)
    print(json.dumps({"hex256_mixed_pair": best}), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
