"""Donor-search throughput of the MI355X overset-connectivity engine.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload hex256_mixed_pair|hex48_pair]

Workload (default ``hex256_mixed_pair``, the large plain-method configuration): a 256^3
hex background on [-1,1]^3 with interior nodes jittered by +-10 % of the
spacing (16.97 M nodes) and a 3.87 M-cell mixed-element shell (40 % tet, 30 %
prism, 5 % pyramid, 25 % hex) from an ellipsoid wall (semi-axes 0.3, 0.2,
0.2) to radius 0.6; FP64.  Both blocks are registered through the C ABI,
preprocessed and connected once (performconnectivity, timed end to end).
The measured step then searches about 19 M receptor points, as the
configuration states: every background node against the near-body block and
every near-body node against the background block, each against the search
structure MeshBlock::search builds for that query set (cells overlapping the
query set's OBB, ADT).  A step is one pass of the search kernels over both
query sets with all inputs resident in HBM.

Printed: one JSON line with searches/s, the search kernel's HBM roofline
fraction (algorithmic bytes per launch over its HIP-event duration; `traffic`
is the HBM bytes per launch that rocprofv3 counted, from the committed
summary under profiles/), the CPU baseline (the oracle's sequential search on
a bounded sample, 1 core), the end to end connectivity time, dataUpdate
bandwidth, the AMR gather leg (`amr_gather`: 4096 patches of 16^3 cells on
3 levels, every node of the mixed-element shell a receptor, 1000
device-resident dataUpdate_AMR gathers of 8 fields), the batching leg
(`batching`: 64 shells and a 192^3 background, batched against
block-by-block performconnectivity) and the serving leg (`serving`: moving
shells over a resident 256^3 background, p50/p99 latency under Poisson
arrivals, against the plain form of the same requests).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def workload(name: str):
    from tioga_amd import workloads as W

    if name == "hex256_mixed_pair":
        return W.hex256_mixed_pair(), ("hex256_mixed_pair: 256^3 jittered hex background + 3.87M-cell mixed tet/pyramid/prism/hex "
                             "shell (ellipsoid wall 0.3,0.2,0.2 to r=0.6), FP64; all background nodes vs the "
                             "near-body block and all near-body nodes vs the background block")
    if name == "hex48_pair":
        return W.hex48_pair(), ("hex48_pair: 48^3 hex background + 64x32x24 lat-long hex shell (r 0.25..0.6, "
                             "stretch 1.1), FP64; all background nodes vs the near-body block and all "
                             "near-body nodes vs the background block")
    raise SystemExit(f"unknown workload {name}")


def search_bytes(block, donor: np.ndarray, nelem: int) -> float:
    """Algorithmic bytes of one search launch: 40 B per query (xyz 24, res 8,
    id 4 read, donor 4 write), 64 B per ADT cell (6-double box + 4-int node
    record), and the cheaper of the found cells' vertex data (28 B per vertex
    per found query: a 4-B connectivity entry and 24-B coordinates) and one
    read of the whole block's connectivity and coordinates."""
    nv_of = np.concatenate([np.full(c.shape[0], nv, np.int64) for nv, c in block.types])
    found = donor[donor >= 0]
    q = donor.size
    per_found = 28.0 * float(nv_of[found].sum())
    whole = 4.0 * float(nv_of.sum()) + 24.0 * block.nnodes
    return 40.0 * q + 64.0 * nelem + min(per_found, whole)


def dataupdate_bytes(tg, blocks, nvar: int) -> float:
    """Algorithmic bytes of one dataUpdate pass: per live receptor 4 B of
    receptor id, 12 B per weight (4-B node id, 8-B weight) and 8*nvar B of
    result, plus one read of nvar fields at every distinct donor node."""
    tot = 0.0
    for b in blocks:
        rec, ind, _ = tg.donor_info(b.tag)
        nw = rec[3::4]
        starts = np.concatenate([[0], np.cumsum(nw + 1)[:-1]])
        donor_nodes = np.concatenate([ind[s:s + n] for s, n in zip(starts, nw)]) if nw.size else np.zeros(0)
        tot += nw.size * (4 + 8 * nvar) + 12.0 * float(nw.sum()) + 8.0 * nvar * np.unique(donor_nodes).size
    return tot


def profiled_traffic(summary: str, key: str):
    """HBM bytes per launch that rocprofv3 --pmc measured for this workload
    (FETCH_SIZE and WRITE_SIZE, each calibrated on a plain device copy); the
    committed summary under profiles/ names the runs.  None when absent."""
    try:
        with open(os.path.join(ROOT, "profiles", summary)) as f:
            for line in f:
                if line.startswith(key + ":"):
                    return int(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def amr_gather_leg(shell, nvar: int = 8, updates: int = 1000):
    """The AMR gather workload: the mixed-element shell, every node a
    receptor, against 4096 patches of 16^3 cells on 3 levels (8 FP64 fields
    per patch); `updates` device-resident dataUpdate_AMR gathers with fixed
    donor weights, timed with HIP events on the engine's stream."""
    from tioga_amd import workloads as W
    from tioga_amd.tioga import Tioga

    t0 = time.time()
    body = W.all_receptors(shell)
    grid = W.amr3_grid()
    pq = grid.patch_fields(nvar, seed=5)
    log(f"AMR workload built in {time.time() - t0:.1f} s: {grid.ngrids} patches, "
        f"{sum(q.size for q in pq) * 8 / 2**30:.2f} GiB of patch fields")
    tg = Tioga()
    tg.register_grid(body)
    tg.register_amr(grid, pq)
    tg.preprocess()
    t0 = time.time()
    tg.connect_amr()
    t_conn = time.time() - t0
    q = np.random.default_rng(6).uniform(-1, 1, body.nnodes * nvar)
    tg.register_solution(body.tag, q)
    ms, nrec = tg.time_data_update_amr(nvar, 0, updates)
    uniq = int(tg.amr_array(0, "rec_unique_cells")[0])
    nrcp = int(tg.amr_array(body.tag, "receptorIdCart").size)
    tg.close()
    # per record: patch, node, block|stencil bits, base index (4 B each), six
    # basis values (48 B) read and nvar results (8 B each) written; per
    # distinct donor cell its nvar field values read once
    algo = nrec * (16 + 48 + 8 * nvar) + 8 * nvar * uniq
    gbps = algo / (ms / 1000) / 1e9 if ms > 0 else None
    log(f"AMR gather: {nrec} records, {uniq} donor cells, {ms:.4f} ms per update")
    return {"workload": "3 levels (ratio 2), 4096 patches of 16^3 cells, one ghost layer, pdegree 0, "
                        f"{nvar} FP64 fields; receptors: every node of the mixed-element shell",
            "receptors": nrcp, "records": int(nrec), "unique_donor_cells": uniq, "updates": updates,
            "kernel": "k_amr_gather", "connectivity_amr_ms": round(1000 * t_conn, 2),
            "ms_per_update": round(ms, 5), "algorithmic_bytes_per_update": algo,
            "GBps": round(gbps, 2) if gbps else None,
            "frac": round(gbps / HBM_PEAK_GBPS, 4) if gbps else None,
            "traffic": profiled_traffic("amr_gather_amr3_mixed_shell.md", "traffic_bytes_per_update")}


def host_cpu():
    """(search threads: the CPUs this process may run on, at most 16; those
    CPUs; the CPU model name)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return max(1, min(16, n)), n, model


def cpu_baseline(blocks, sets, gpu_donors, budget_s: float):
    """The oracle's donor search (tioga_oracle.c, restating
    search.C:107-233, searchADTrecursion.C:26-101 and checkContainment.C:27-96)
    on a bounded random sample of each query set, searched against the tree
    the GPU searches (the cells overlapping the OBB of the whole set), once
    on 1 core and once on all the cores this process may use (POSIX threads
    over contiguous chunks); the search loops alone are timed.  The sampled
    donors are compared with the GPU's donors of the timed passes."""
    from oracle.oracle import Oracle

    nthr, nproc, model = host_cpu()
    o = Oracle(blocks)
    o.profile()
    rng = np.random.default_rng(0)
    tot_q, t1, tall, mism = 0, 0.0, 0.0, 0
    for (btag, xq), gd in zip(sets, gpu_donors):
        # calibrate the per-query cost on a small sample, then size the sample
        n0 = min(20000, xq.shape[0])
        idx = np.sort(rng.choice(xq.shape[0], n0, replace=False))
        _, _, t = o.search_points_obb(btag, xq, xq[idx], 1)
        per = t / n0
        n = int(min(xq.shape[0], max(n0, budget_s / len(sets) / max(per, 1e-9))))
        idx = np.sort(rng.choice(xq.shape[0], n, replace=False))
        xs = np.ascontiguousarray(xq[idx])
        d1, _, ts = o.search_points_obb(btag, xq, xs, 1)
        dn, _, tp = o.search_points_obb(btag, xq, xs, nthr)
        mism += int((d1 != gd[idx]).sum()) + int((dn != d1).sum())
        tot_q += n
        t1 += ts
        tall += tp
        log(f"cpu baseline: {n} queries vs block {btag}: 1 core {ts:.2f} s, {nthr} threads {tp:.2f} s")
    o.close()
    return {"one_core": tot_q / t1, "all_cores": tot_q / tall, "threads": nthr, "nproc": nproc, "cpu_model": model,
            "sampled": tot_q, "seconds_one_core": t1, "seconds_all_cores": tall, "mismatches": mism}


def batching_leg(steps: int, warmup: int, batched_only: bool = False) -> dict:
    """64 near-body shells + 192^3 background: end-to-end performconnectivity
    steps (all inputs resident, coordinates unchanged between steps) in the
    batched form (every stage one launch for all blocks) and the
    block-by-block form; donor searches per second of connectivity, and
    whether the two forms give identical iblank and donor records."""
    import torch
    from tioga_amd import workloads as W
    from tioga_amd.tioga import Tioga

    t0 = time.time()
    blocks = W.shells64_hex192()
    log(f"shells64_hex192 generated in {time.time() - t0:.1f} s: {len(blocks)} blocks, "
        f"{sum(b.ncells for b in blocks)} cells, {sum(b.nnodes for b in blocks)} nodes")
    tg = Tioga()
    for b in blocks:
        tg.register_grid(b)
    t0 = time.time()
    tg.preprocess()
    log(f"shells64_hex192 preprocess {time.time() - t0:.2f} s")

    def run(batched: bool):
        tg.set_batching(batched)
        for _ in range(warmup):
            tg.connect()
        torch.cuda.synchronize()
        times, phases = [], []
        for _ in range(steps):
            t1 = time.perf_counter()
            tg.connect()
            times.append(time.perf_counter() - t1)
            phases.append(tg.phase_times())
        log(f"shells64_hex192 connectivity ({'batched' if batched else 'block by block'}): "
            f"{', '.join(f'{1000 * t:.1f}' for t in times)} ms")
        ib = [a.copy() for a in tg.iblank]
        rec = [tg.donor_info(b.tag) for b in blocks]
        return times, phases, ib, rec

    t_b, ph_b, ib_b, rec_b = run(True)
    if not batched_only:
        t_p, ph_p, ib_p, rec_p = run(False)
        same = all(np.array_equal(a, b) for a, b in zip(ib_p, ib_b)) and all(
            all(np.array_equal(x, y) for x, y in zip(r1, r2)) for r1, r2 in zip(rec_p, rec_b))
    tg.close()
    nsearch = ph_b[-1]["nsearch"]
    ms = 1000 * float(np.median(t_b))
    out = {"workload": "shells64_hex192: 64 hex/prism shells (1e4..1e6 cells, log-uniform) + 192^3 background, "
                       "performconnectivity with all inputs resident",
           "blocks": len(blocks), "searches_per_step": int(nsearch), "steps": steps,
           "batched": {"ms_per_step": round(ms, 3), "p99_ms": round(1000 * float(np.percentile(t_b, 99)), 3),
                       "searches_per_s": round(nsearch / (ms / 1000), 1),
                       "phases_ms": {k: round(float(np.median([p[k] for p in ph_b])), 3) for k in ph_b[-1]}}}
    if not batched_only:
        ms_p = 1000 * float(np.median(t_p))
        out["per_block"] = {"ms_per_step": round(ms_p, 3), "searches_per_s": round(nsearch / (ms_p / 1000), 1),
                            "phases_ms": {k: round(float(np.median([p[k] for p in ph_p])), 3) for k in ph_p[-1]}}
        out["batched_vs_per_block"] = round(ms_p / ms, 3)
        out["identical_iblank_and_donors"] = bool(same)
    return out


def batching_bench(args) -> int:
    leg = batching_leg(args.steps, args.warmup, args.batched_only)
    b = leg["batched"]
    out = {"metric": "donor searches/s of performconnectivity (64 shells + background), batched form",
           "value": b["searches_per_s"], "unit": "searches/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": b["ms_per_step"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": leg["workload"], "blocks": leg["blocks"],
                      "searches_per_step": leg["searches_per_step"]}, "batching": leg}
    print(json.dumps(out), flush=True)
    return 0


def serving_leg(slots: int, budget_gb: float, rates, seconds: float, max_requests: int, nplain: int = 6) -> dict:
    """16 resident shells over the resident 256^3 background: requests that
    move one shell rigidly, served by `slots` concurrent slots.  Reports the
    latency of requests served alone (one at a time), p50/p99 latency and
    throughput under Poisson arrivals at each rate, the throughput at
    saturation (a burst of requests submitted at once), and the plain GPU
    form of the same requests: one engine per shell holding that shell and
    the background, each request moving the shell, preprocessing that block
    alone and running performconnectivity, one request after the other.
    Latency = done - arrival on one steady clock."""
    import threading

    from tioga_amd import workloads as W
    from tioga_amd.tioga import Tioga

    t0 = time.time()
    shells = W.serve16_shells()
    bg = W.cart_box(256, tag=17, jitter=0.1, seed=2)
    log(f"serve16_hex256 generated in {time.time() - t0:.1f} s: shell nodes {sorted(s.nnodes for s in shells)}")
    tg = Tioga()
    for b in shells + [bg]:
        tg.register_grid(b)
    t0 = time.time()
    tg.preprocess()
    log(f"serve16_hex256 preprocess {time.time() - t0:.2f} s")
    cen = {s.tag: s.x.mean(axis=0) for s in shells}
    size = {s.tag: s.nnodes for s in shells}
    tg.serve_open(bg.tag, nslots=slots, budget_gb=budget_gb)
    for s_ in shells:  # warm-up: every shell once
        tg.serve_wait(tg.serve_submit(s_.tag, np.eye(3), cen[s_.tag], np.zeros(3)), size[s_.tag])
    alone, alone_s = [], 0
    for tag, R, sh in W.motion_requests(16, len(shells), seed=40):
        r = tg.serve_wait(tg.serve_submit(tag, R, cen[tag], sh), size[tag])
        alone.append(1000 * (r["t_done"] - r["t_submit"]))
        alone_s += r["nsearch"]
    log(f"serving alone: p50 {np.median(alone):.1f} ms, max {max(alone):.1f} ms")

    def stream(reqs, arrive):
        n = len(reqs)
        ids = [None] * n
        start = time.perf_counter()

        def submitter():
            for i, (tag, R, sh) in enumerate(reqs):
                d = start + arrive[i] - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                ids[i] = tg.serve_submit(tag, R, cen[tag], sh)

        th = threading.Thread(target=submitter)
        th.start()
        res = []
        for i in range(n):
            while ids[i] is None:
                time.sleep(0.0005)
            res.append(tg.serve_wait(ids[i], size[reqs[i][0]]))
        th.join()
        lat = np.array([1000 * (r["t_done"] - r["t_submit"]) for r in res])
        span = max(r["t_done"] for r in res) - min(r["t_submit"] for r in res)
        return {"requests": n, "p50_ms": round(float(np.percentile(lat, 50)), 3),
                "p99_ms": round(float(np.percentile(lat, 99)), 3), "throughput_req_s": round(float(n / span), 2),
                "searches_per_s": round(float(sum(r["nsearch"] for r in res) / span), 1)}

    per_rate = []
    for k, lam in enumerate(rates):
        n = int(min(max_requests, max(40, lam * seconds)))
        reqs = W.motion_requests(n, len(shells), seed=41 + k)
        arrive = np.cumsum(np.random.default_rng(4 + k).exponential(1.0 / lam, n))
        per_rate.append(dict(rate_req_s=lam, **stream(reqs, arrive)))
        log(f"serving at {lam}/s: {per_rate[-1]}")
    burst = stream(W.motion_requests(64, len(shells), seed=60), np.zeros(64))
    log(f"serving burst of 64: {burst}")
    tg.serve_close()
    # resident-tree mode: one tree over all background cells, built at open
    t0 = time.perf_counter()
    tg.serve_open(bg.tag, nslots=slots, budget_gb=budget_gb, resident_tree=True)
    open_ms = 1000 * (time.perf_counter() - t0)
    for s_ in shells:
        tg.serve_wait(tg.serve_submit(s_.tag, np.eye(3), cen[s_.tag], np.zeros(3)), size[s_.tag])
    r_alone, r_s, r_tie, r_clean = [], 0, 0, 0
    for tag, R, sh in W.motion_requests(16, len(shells), seed=40):
        r = tg.serve_wait(tg.serve_submit(tag, R, cen[tag], sh), size[tag])
        r_alone.append(1000 * (r["t_done"] - r["t_submit"]))
        r_s += r["nsearch"]
        r_tie += r["tieband"]
        r_clean += r["tieband"] == 0
    r_burst = stream(W.motion_requests(64, len(shells), seed=60), np.zeros(64))
    tg.serve_close()
    tg.close()
    resident = {"open_ms": round(open_ms, 1), "alone_p50_ms": round(float(np.median(r_alone)), 3),
                "alone_max_ms": round(float(max(r_alone)), 3),
                "alone_searches_per_s": round(float(r_s / (sum(r_alone) / 1000)), 1),
                "saturation_burst": r_burst, "tieband_queries": int(r_tie), "searches": int(r_s),
                "requests_without_tieband": f"{r_clean}/16",
                "what": "one tree over all background cells built at open; a request with tieband 0 "
                        "has the per-request tree's results"}
    log(f"resident-tree serving: {resident}")
    # the plain GPU form of the same kind of requests
    plain, engines = [], {}
    for tag, R, sh in W.motion_requests(nplain, len(shells), seed=40):
        s0 = shells[tag - 1]
        if tag not in engines:
            moving = W.Block(tag, s0.x.copy(), s0.types, s0.wbc, s0.obc)
            e = Tioga()
            e.register_grid(moving)
            e.register_grid(bg)
            e.preprocess()
            engines[tag] = e
        e = engines[tag]
        e.set_coords(0, W.rigid_apply(s0.x, R, cen[tag], sh))
        t1 = time.perf_counter()
        e.preprocess_block(tag)
        e.connect()
        plain.append((1000 * (time.perf_counter() - t1), e.phase_times()["nsearch"]))
    for e in engines.values():
        e.close()
    plain_ms = float(np.mean([p[0] for p in plain]))
    plain_sps = sum(p[1] for p in plain) / (sum(p[0] for p in plain) / 1000)
    log(f"plain form: {plain_ms:.1f} ms per request, {plain_sps:.3e} searches/s")
    stable = [r for r in per_rate if r["throughput_req_s"] >= 0.95 * r["rate_req_s"]] or per_rate[:1]
    head = max(stable, key=lambda r: r["rate_req_s"])
    return {"workload": "serve16_hex256: 16 hex/prism shells (1e4..2e6 cells) moving rigidly over a resident 256^3 "
                        "jittered hex background; Poisson arrivals", "slots": slots, "budget_gb": budget_gb,
            "alone_p50_ms": round(float(np.median(alone)), 3), "alone_max_ms": round(float(max(alone)), 3),
            "alone_searches_per_s": round(alone_s / (sum(alone) / 1000), 1),
            "per_rate": per_rate, "p99_ms_at_highest_sustained_rate": head["p99_ms"],
            "highest_sustained_rate_req_s": head["rate_req_s"], "saturation_burst": burst,
            "plain_form": {"ms_per_request": round(plain_ms, 3), "searches_per_s": round(plain_sps, 1),
                           "requests": len(plain),
                           "what": "one engine per shell (shell + background), move, preprocess the shell "
                                   "alone, performconnectivity; sequential"},
            "served_vs_plain_searches": round(burst["searches_per_s"] / plain_sps, 3),
            "resident_tree": resident}


def serving_bench(args) -> int:
    rates = [float(v) for v in args.rates.split(",")]
    leg = serving_leg(args.slots, args.budget_gb, rates, args.seconds, args.max_requests)
    out = {"metric": "p99 connectivity latency under Poisson load (serving), donor searches/s",
           "value": leg["p99_ms_at_highest_sustained_rate"], "unit": "ms",
           "at_rate_req_s": leg["highest_sustained_rate_req_s"], "n_gpus": 1,
           "steps": sum(r["requests"] for r in leg["per_rate"]), "warmup": 16, "ms_per_step": leg["alone_p50_ms"],
           "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": leg["workload"], "slots": args.slots, "budget_gb": args.budget_gb},
           "serving": leg}
    print(json.dumps(out), flush=True)
    return 0


def search_only(args, tg, blocks, desc) -> int:
    """The headline passes alone (what the committed kernel profiles run)."""
    sets = [(blocks[0].tag, blocks[1].x), (blocks[1].tag, blocks[0].x)]
    qids = [tg.query_set(btag, xq) for btag, xq in sets]
    from tioga_amd import autotune

    autotune.apply(tg, args.workload)
    tg.search_graph(not args.no_graph)
    for i, (_, xq) in zip(qids, sets):
        tg.query_search(i, xq.shape[0])
    for _ in range(args.warmup):
        tg.time_query_sets(1)
    pass_ms, nq, kern_ms = tg.time_query_sets(args.steps)
    stats = tg.search_stats()
    tg.close()
    print(json.dumps({"workload": desc, "steps": args.steps, "queries_per_step": int(nq),
                      "search_pass_ms_device": round(pass_ms, 4), "search_kernel_ms": round(kern_ms, 4),
                      "search_launch": "kernel by kernel" if args.no_graph else "HIP graph replay",
                      "search_stats": stats}),
          flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="hex256_mixed_pair")
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of CPU search for the baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true",
                    help="timed passes launch the search kernels one by one instead of replaying a HIP graph")
    ap.add_argument("--no-amr", action="store_true", help="skip the AMR gather leg")
    ap.add_argument("--search-only", action="store_true",
                    help="only the headline query-set passes (profiling runs): no performconnectivity, dataUpdate, "
                         "CPU baseline, reduce_fringes or AMR leg")
    ap.add_argument("--batched-only", action="store_true", help="shells64_hex192: skip the block-by-block form")
    ap.add_argument("--no-batching", action="store_true", help="skip the batching leg (64 shells, both forms)")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving leg (moving shells)")
    ap.add_argument("--slots", type=int, default=8, help="serving slots (serve16_hex256)")
    ap.add_argument("--budget-gb", type=float, default=64.0, help="serving HBM budget (serve16_hex256)")
    ap.add_argument("--rates", default="50,100,200,500,1000,2000", help="request rates per second (serve16_hex256)")
    ap.add_argument("--seconds", type=float, default=2.0, help="arrival window per rate (serve16_hex256)")
    ap.add_argument("--max-requests", type=int, default=300, help="requests per rate at most (serve16_hex256)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run 5 device copies of MIB MiB (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs one rank on one GPU; multi-GPU runs are not part of "
                                   "this build; the path shards by mesh partition, one rank per GPU", "n_gpus": args.gpus}))
        return 2

    import torch

    if not torch.cuda.is_available():
        raise SystemExit("bench.py needs a GPU")
    if args.calibrate_copy:
        n = args.calibrate_copy * 2**20 // 8
        a = torch.rand(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": 8 * n, "copies": 5}), flush=True)
        return 0
    from tioga_amd.tioga import Tioga

    if args.workload == "shells64_hex192":
        return batching_bench(args)
    if args.workload == "serve16_hex256":
        return serving_bench(args)
    t0 = time.time()
    blocks, desc = workload(args.workload)
    log(f"workload generated in {time.time() - t0:.1f} s: " +
        ", ".join(f"tag {b.tag}: {b.nnodes} nodes {b.ncells} cells" for b in blocks))
    tg = Tioga()
    for b in blocks:
        tg.register_grid(b)
    t0 = time.time()
    tg.preprocess()
    torch.cuda.synchronize()
    t_pre = time.time() - t0
    if args.search_only:
        return search_only(args, tg, blocks, desc)
    t0 = time.time()
    tg.connect()
    torch.cuda.synchronize()
    t_conn = time.time() - t0
    phases = tg.phase_times()
    log(f"preprocess {t_pre:.2f} s, first performconnectivity {t_conn:.2f} s ({phases})")
    # performconnectivity warm: the same blocks again, scratch already sized;
    # each step ends with iblank and iblank_cell written to the host arrays
    for _ in range(args.warmup):
        tg.connect()
    conn_ms, conn_ph = [], []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        tg.connect()
        conn_ms.append(1000 * (time.perf_counter() - t0))
        conn_ph.append(tg.phase_times())
    conn_warm = {"steps": args.steps, "p50_ms": round(float(np.median(conn_ms)), 3),
                 "p99_ms": round(float(np.percentile(conn_ms, 99)), 3),
                 "searches": int(conn_ph[-1]["nsearch"]),
                 "searches_per_s": round(conn_ph[-1]["nsearch"] / (float(np.median(conn_ms)) / 1000), 1),
                 "phases_ms": {k: round(float(np.median([p[k] for p in conn_ph])), 3) for k in conn_ph[-1]}}
    log(f"performconnectivity warm: {conn_warm}")

    # dataUpdate over the connectivity's receptors, 5 FP64 fields, q resident
    nvar = 5
    qs = [np.random.default_rng(1 + i).uniform(-1, 1, b.nnodes * nvar) for i, b in enumerate(blocks)]
    for b, q in zip(blocks, qs):
        tg.register_solution(b.tag, q)
    du_ms, nrec = tg.time_data_update(nvar, 0, 20)

    # the measured workload: two query sets
    nb_tag, bg_tag = blocks[0].tag, blocks[1].tag
    sets = [(nb_tag, blocks[1].x), (bg_tag, blocks[0].x)]
    t0 = time.time()
    qids = [tg.query_set(btag, xq) for btag, xq in sets]
    log(f"query sets built in {time.time() - t0:.1f} s")
    # the search's launch parameters from the autotuner's table (they change
    # no result; tioga_amd/autotune.py measured them on this workload)
    from tioga_amd import autotune

    tuning = autotune.apply(tg, args.workload)
    tuning["source"] = "tioga_amd/search_tuning.json" if autotune.load(args.workload) else "defaults"
    tg.search_graph(not args.no_graph)
    donors = [tg.query_search(i, xq.shape[0]) for i, (_, xq) in zip(qids, sets)]

    for _ in range(args.warmup):
        tg.time_query_sets(1)
    torch.cuda.synchronize()
    t0 = time.time()
    pass_ms, nq, kern_ms = tg.time_query_sets(args.steps)
    torch.cuda.synchronize()
    wall = time.time() - t0
    ms_per_step = 1000.0 * wall / args.steps
    value = nq / (ms_per_step / 1000.0)
    # the donors the timed passes left (the path the rate is measured on)
    timed = [tg.query_donors(i, xq.shape[0]) for i, (_, xq) in zip(qids, sets)]
    stats = tg.search_stats()
    timed_same = all(np.array_equal(a, b) for a, b in zip(timed, donors))
    # queries inside their tree's root box (the others are rejected by one box test)
    in_root = 0
    for i, (_, xq) in zip(qids, sets):
        e = tg.query_extents(i)
        m = np.ones(xq.shape[0], bool)
        for d in range(3):
            m &= (xq[:, d] >= e[2 * d] - 1e-10) & (xq[:, d] <= e[2 * d + 1] + 1e-10)
        in_root += int(m.sum())

    byblk = {b.tag: b for b in blocks}
    nelem_of = [tg.query_nelem(i) for i in qids]
    algo = sum(search_bytes(byblk[btag], d, ne) for (btag, _), d, ne in zip(sets, timed, nelem_of))
    achieved = algo / (kern_ms / 1000.0) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 5),
            "traffic": profiled_traffic("search_pipeline_hex256_mixed_pair.md", "traffic_bytes_per_launch")
            if args.workload == "hex256_mixed_pair" else None,
            "traffic_source": "profiles/search_pipeline_hex256_mixed_pair.md (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE)",
            "kernel": "search launch (k_find_cost, k_find / k_collect, k_eval<4,5,6,8>, "
                      "k_eval_long<5,6,8>, k_last, k_select)",
            "algorithmic_bytes_per_step": algo, "launches_per_step": 1,
            "avg_launch_ms": round(kern_ms, 4)}

    cpu, parity = None, None
    if not args.no_cpu:
        c = cpu_baseline(blocks, sets, timed, args.cpu_budget)
        cpu = {"value": round(c["all_cores"], 1), "unit": "searches/s", "cores": c["threads"], "kind": "port",
               "sample": f"{c['sampled']} random queries drawn from both query sets, each searched against the "
                         f"tree the GPU searches (cells overlapping the OBB of the whole set); search loops only: "
                         f"{c['seconds_all_cores']:.1f} s on {c['threads']} threads, "
                         f"{c['seconds_one_core']:.1f} s on 1 core",
               "one_core": round(c["one_core"], 1), "nproc": c["nproc"], "cpu_model": c["cpu_model"]}
        parity = {"sampled": c["sampled"], "mismatches": c["mismatches"],
                  "against": "oracle donors of the sampled queries (1 core and threaded) vs the GPU's timed passes"}

    bdu = dataupdate_bytes(tg, blocks, nvar)
    # tioga::reduce_fringes on this connectivity (last: it changes the
    # interpolation records), iblank written back to the host arrays included
    nrec_before = int(sum(int((ib == -1).sum()) for ib in tg.iblank))
    torch.cuda.synchronize()
    t0 = time.time()
    tg.reduce_fringes()
    t_rf = time.time() - t0
    rf = {"ms": round(1000 * t_rf, 3), "receptors_before": nrec_before,
          "receptors_after": int(sum(int((ib == -1).sum()) for ib in tg.iblank)),
          "holes_after": int(sum(int((ib == 0).sum()) for ib in tg.iblank)), "nfringe": 1}
    log(f"reduce_fringes {rf}")
    tg.close()
    amr = None if args.no_amr else amr_gather_leg(blocks[0])
    batching = None if args.no_batching else batching_leg(3, 1)
    serving = None if args.no_serving else serving_leg(8, 64.0, [50.0, 100.0], 1.5, 150)
    out = {
        "metric": "donor searches/s on 1 GPU",
        "value": round(value, 1),
        "unit": "searches/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": desc, "queries_per_step": int(nq),
                   "blocks": {str(b.tag): {"nodes": b.nnodes, "cells": b.ncells} for b in blocks}},
        "roofline": roof,
        "cpu_baseline": cpu,
        "parity": parity,
        "timed_donors_equal_checked": bool(timed_same),
        "search_tuning": tuning,
        "search_launch": "kernel by kernel" if args.no_graph else "HIP graph replay (recorded once after the checked pass)",
        "queries_in_root_box": in_root,
        "search_stats": stats,
        "searches_per_s_in_root_box": round(in_root / (ms_per_step / 1000.0), 1),
        "search_kernel_ms": round(kern_ms, 4),
        "search_pass_ms_device": round(pass_ms, 4),
        "found": int(sum(int((d >= 0).sum()) for d in timed)),
        "connectivity_first_ms": round(1000 * t_conn, 2),
        "connectivity_first_phases_ms": {k: round(v, 3) for k, v in phases.items()},
        "connectivity_warm": conn_warm,
        "preprocess_ms": round(1000 * t_pre, 2),
        "dataupdate": {"ms": round(du_ms, 4), "receptors": int(nrec), "nvar": nvar,
                       "algorithmic_bytes": bdu,
                       "GBps": round(bdu / (du_ms / 1000) / 1e9, 2) if du_ms > 0 else None,
                       "frac": round(bdu / (du_ms / 1000) / 1e9 / HBM_PEAK_GBPS, 4) if du_ms > 0 else None,
                       "kernel": "k_update<0> (fused gather and scatter)",
                       "traffic": profiled_traffic("dataupdate_hex256_mixed_pair.md", "traffic_bytes_per_update")
                       if args.workload == "hex256_mixed_pair" else None},
        "reduce_fringes": rf,
        "amr_gather": amr,
        "batching": batching,
        "serving": serving,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
