"""TEST INFRASTRUCTURE ONLY: times the compiled reference TIOGA on the CPU.

Builds the timers variant of the reference (oracle/Makefile target `timers`:
tioga.C with -DTIOGA_ENABLE_TIMERS, whose myTimer prints "End <phase>, time
taken=<s>" around each phase, tioga.C:97-151, 715-735), writes a case file
for a golden case (default: hex48_pair, the 48^3 hex background and the
64x32x24 lat-long shell) and runs oracle/_ref/timers/ref_driver on it on one
MPI rank, several times, printing per-phase times (the median over the runs)
and the driver's own wall times of profile, performConnectivity and
dataUpdate.  The reference's myTimer keeps one start time for all phases, so
its line for performConnectivity covers only the last phase; the driver's
wall time is the whole call.  Needs the reference sources and this image's
MPICH (as `make -C oracle ref`); it is a cross-check on this machine's CPU,
never part of a GPU run.

    python oracle/time_reference.py [--case hex48_pair] [--runs 5]
"""
import argparse
import os
import re
import statistics
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--case", default="hex48_pair")
    ap.add_argument("--runs", type=int, default=5)
    args = ap.parse_args()
    r = subprocess.run(["make", "-C", HERE, "timers"], capture_output=True, text=True)
    if r.returncode != 0:
        print(r.stdout[-2000:], r.stderr[-2000:])
        return 1
    # this file's folder first on the path would make `oracle` this folder's module
    sys.path[:] = [p for p in sys.path if os.path.abspath(p or ".") != HERE]
    sys.path.insert(0, ROOT)
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    import cases
    from oracle import oracle as O

    c = cases.build(args.case)
    drv = os.path.join(HERE, "_ref", "timers", "ref_driver")
    phases, walls = {}, {}
    with tempfile.TemporaryDirectory() as td:
        cp, op = os.path.join(td, "case.bin"), os.path.join(td, "out.bin")
        O.write_case(cp, c.blocks, c.qs, c.nvar, c.interptype, c.isym, c.nfringe, c.mexclude, c.moves,
                     reduce=c.reduce)
        for _ in range(args.runs):
            r = subprocess.run([drv, cp, op], capture_output=True, text=True,
                               env=dict(os.environ, REF_DRIVER_TIMES="1"), timeout=600)
            if r.returncode != 0:
                print(r.stderr[-2000:])
                return 1
            for name, t in re.findall(r"End (\S+), time taken=([0-9.eE+-]+)", r.stdout + r.stderr):
                phases.setdefault(name, []).append(float(t))
            for line in r.stderr.splitlines():
                if line.startswith("ref_driver step 0:"):
                    for name, t in re.findall(r"(\w+) ([0-9.]+) s", line):
                        walls.setdefault(name, []).append(float(t))
    nodes = sum(b.nnodes for b in c.blocks)
    cells = sum(b.ncells for b in c.blocks)
    print(f"case {args.case}: {len(c.blocks)} blocks, {nodes} nodes, {cells} cells; {args.runs} runs, "
          f"median seconds, 1 MPI rank on {os.cpu_count()} host cores")
    for name, v in phases.items():
        print(f"  reference timer {name:32s} {statistics.median(v):.6f}")
    for name, v in walls.items():
        print(f"  driver wall     {name:32s} {statistics.median(v):.6f}")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
