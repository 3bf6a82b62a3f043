#!/usr/bin/env python3
"""Benchmark: QDB-banded scanning CYK + Inside on one MI355X.

Workload (the GPU FastCYKScan + FastFInsideScan configuration): a covariance
model scanned over a seeded synthetic target of --mb megabases (second-order
Markov background at 40% GC with --homologs emitted homolog fragments of
0.5-1.0x model length planted at seeded positions).  The target is cut into
database blocks of --block-mb megabases (each after the first starts W-1
residues early, the overlap cmsearch gives its sequence windows).  One step =
one block: FastCYKScan with the tight bands (QDB1, beta 1e-7), then
FastFInsideScan with the loose bands (QDB2, beta 1e-15), both over the whole
block ("plain method": Inside on every position, no filter); steps cycle
through the blocks.

value = DP cells of both scans / wall time of the timed steps (Gcells/s),
with the target resident in device memory (pinned once before the steps);
cells are (v, j, d) cells inside the QDB bands, E states excluded, counted
exactly as the reference scanners compute them.  Each wall time includes the
host's hit replay.

A complete JSON line is printed after every timed step (the last one is the
result); no timed step is begun after --budget-s seconds from the start, and
`steps` is the number actually timed.  The CPU baseline runs before the steps.
Optional legs run after the steps and print one more line with their records.

The JSON line also carries
  roofline:     the dominant kernel's algorithmic VALU ops (2*cnum+1 per
                non-B cell, 2 per bifurcation split, FLogsum counted as 8)
                per second of its HIP-event kernel time, against MI355X's
                published FP32 peak and its non-packed FP32 VALU rate, and
  cpu_baseline: the C oracle (a restatement of the reference scanners,
                gcc -O3, no FMA) timed on this host on a bounded sample.
"""
import argparse
import concurrent.futures as cf
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

BG_SEED = 20261203   # the benchmark target's background chain and sample
VALU_PEAK_TOPS = 256 * 4 * 16 * 2.4e9 / 1e12   # 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz, non-packed FP32
FP32_PEAK_TFLOPS = 157.3                        # MI355X published FP32 vector peak (packed FMA)
WORKLOAD_TAG = "multiloop350_block12p5mb"       # names the traffic summary of this workload in profiles/


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def ops_per_res(cm, qdbidx, inside):
    """Algorithmic VALU ops per residue at steady state (bands at jp >= W):
    CYK 2*cnum (+1 emitting) per non-B cell and 2 per bifurcation split;
    Inside counts every max reduction as an 8-op log-sum."""
    cells, ops, red = cm.work_per_res(qdbidx)
    return ops + 7.0 * red if inside else ops


def traffic_rate(bytes_per_launch, kernel_ms):
    """HBM bytes of one launch over its kernel time, against MI355X's 8 TB/s."""
    rate = bytes_per_launch / (kernel_ms / 1e3) / 1e9
    return {"achieved": round(rate, 1), "peak": 8000.0, "unit": "GB/s", "frac": round(rate / 8000.0, 4)}


def host_cpu():
    """The host's CPU model and logical CPU count (for the baseline's record)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count()}


def cpu_baseline(cm, codes, threads, sample_len, W):
    """Oracle CYK (QDB1) + FInside (QDB2) on `threads` windows of sample_len+W-1."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import infernal_amd as ia
    import oracle_bind as ob
    ob.load()
    Lw = sample_len + W - 1
    rng = np.random.default_rng(99)
    starts = rng.integers(1, len(codes) - Lw, threads)
    dsq = ia.dsq_from_codes(codes)

    def one(algo, q, s):
        t = time.perf_counter()
        ob.scan(cm, algo, q, dsq, int(s), int(s) + Lw - 1, 1e30, want_rows=False)
        return time.perf_counter() - t

    res = {}
    for name, algo, q in (("cyk", ob.ORACLE_CYK, 1), ("finside", ob.ORACLE_FINSIDE, 2)):
        t0 = time.perf_counter()
        with cf.ThreadPoolExecutor(threads) as ex:
            list(ex.map(lambda s: one(algo, q, s), starts))
        wall = time.perf_counter() - t0
        cells = 0.0
        smx = cm.c.smx
        for s in starts:
            cells += _cells(cm, q, int(s), int(s) + Lw - 1)
        res[name] = (cells, wall)
    # one core: the first window alone
    s0 = int(starts[0])
    for name, algo, q in (("cyk", ob.ORACLE_CYK, 1), ("finside", ob.ORACLE_FINSIDE, 2)):
        res[name + "_1core"] = (_cells(cm, q, s0, s0 + Lw - 1), one(algo, q, s0))
    return res


def cyk_1500(args, ia, synth, threads):
    """Bit-exact FastCYKScan (QDB1) of the 1,500-nt synthetic model over
    --cyk1500-mb Mb of the same kind of target, on the GPU and (threads not
    None) with the oracle on the host, for the GPU/CPU Gcells/s ratio."""
    path = os.path.join(ROOT, "tests", "golden", "cm", "synth_nested1500.cm")
    cm = ia.CM.read(path)
    cm.configure(local=True)
    n = int(args.cyk1500_mb * 1e6)
    bg = synth.markov2_background(n, gc=0.40, seed=1500)
    frags = synth.homologs(ia.CM.read(path), max(1, n // 20000), seed=15)
    codes, _ = synth.plant(bg, frags, seed=16)
    dsq = ia.dsq_from_codes(codes)
    # first use (profile build, device buffers sized for this target), untimed, as the benchmark's
    # warm-up steps are
    ia.FastCYKScan(cm, 1, dsq, 1, n, 0.0, want_hits=True)
    t0 = time.perf_counter()
    r = ia.FastCYKScan(cm, 1, dsq, 1, n, 0.0, want_hits=True)
    wall = time.perf_counter() - t0
    tm = ia.last_call_timing()
    log("1500-nt CYK: %.2f s wall, %.2f s kernel, %d hits, best %.2f" % (wall, tm["kernel_ms"] / 1e3, len(r["hits"]), r["sc"]))
    out = {"cm": "%s (clen %d, M %d, W %d)" % (cm.c.name.decode(), cm.clen, cm.M, cm.W),
           "target_mb": n / 1e6, "homologs": len(frags), "hits": len(r["hits"]),
           "gpu_gcells_per_s": round(tm["cells"] / wall / 1e9, 3),
           "gpu_kernel_gcells_per_s": round(tm["cells"] / (tm["kernel_ms"] / 1e3) / 1e9, 3)}
    if threads:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_bind as ob
        ob.load()
        Lw = 2000 + cm.W - 1
        starts = np.random.default_rng(98).integers(1, n - Lw, threads)

        def one(s):
            ob.scan(cm, ob.ORACLE_CYK, 1, dsq, int(s), int(s) + Lw - 1, 1e30, want_rows=False)

        t0 = time.perf_counter()
        with cf.ThreadPoolExecutor(threads) as ex:
            list(ex.map(one, starts))
        cw = time.perf_counter() - t0
        cells = sum(_cells(cm, 1, int(s), int(s) + Lw - 1) for s in starts)
        out["cpu_gcells_per_s"] = round(cells / cw / 1e9, 4)
        out["cpu_cores"] = threads
        s0 = int(starts[0])
        t1 = time.perf_counter()
        one(s0)
        c1 = time.perf_counter() - t1
        out["cpu_1core_gcells_per_s"] = round(_cells(cm, 1, s0, s0 + Lw - 1) / c1 / 1e9, 4)
        out["cpu_sample"] = "oracle FastCYKScan, %d windows of %d nt in %d threads, %.1f s" % (threads, Lw, threads, cw)
        out["gpu_over_cpu"] = round(out["gpu_gcells_per_s"] / out["cpu_gcells_per_s"], 1)
        out["gpu_over_cpu_1core"] = round(out["gpu_gcells_per_s"] / out["cpu_1core_gcells_per_s"], 1)
    return out


def batch_leg(args, ia, synth):
    """Many models in one launch.  Models: every .cm / .cm.gz file of
    --batch-cmdir (default: the shipped models under tests/golden/cm; the
    many-models workload uses the 200 synthetic models that
    tests/golden/make_batch_cms.py builds), local mode, FastCYKScan QDB1.
    Target: --batch-mb Mb of 2nd-order Markov background with two emitted
    homologs of every model planted.  Windows: --batch-windows per model, of
    1.2 W each, at seeded disjoint positions.  Reported: one
    infx_FastScanBatch launch for all models (direct, then replayed from a
    captured graph) against one infx_FastScanBatch launch per model (cells/s,
    per-model latency p50/p99: every model of the batched launch finishes
    with the launch), hits identical between the two."""
    import gzip
    cmdir = args.batch_cmdir or os.path.join(ROOT, "tests", "golden", "cm")
    names = sorted(f for f in os.listdir(cmdir) if f.endswith(".cm") or f.endswith(".cm.gz"))

    def read(nm):
        p = os.path.join(cmdir, nm)
        if nm.endswith(".gz"):
            with gzip.open(p, "rt") as fh:
                return ia.CM.from_text(fh.read())
        return ia.CM.read(p)

    t0 = time.perf_counter()
    cms = []
    for nm in names:
        cm = read(nm)
        cm.configure(local=True)
        cms.append(cm)
    rng = np.random.default_rng(404)
    frags = []
    for nm in names:
        frags += synth.homologs(read(nm), 2, seed=int(rng.integers(1 << 30)))
    n = int(args.batch_mb * 1e6)
    bg = synth.markov2_background(n, gc=0.45, seed=405)
    codes, _ = synth.plant(bg, frags, seed=406)
    dsq = ia.dsq_from_codes(codes)
    wins, per = [], []
    for m, cm in enumerate(cms):
        wl = int(1.2 * cm.W)
        starts = np.sort(rng.choice(np.arange(1, n - wl, wl + 1), args.batch_windows, replace=False))
        ws = [(int(a), int(a) + wl - 1) for a in starts]
        per.append(ws)
        wins += [(i, j, m) for (i, j) in ws]
    log("many models: %d models (W %d-%d), %d windows, set up in %.1f s" % (len(cms), min(c.W for c in cms), max(c.W for c in cms),
                                                                          len(wins), time.perf_counter() - t0))
    cut = [0.0] * len(cms)
    ia.pin_target(dsq)
    try:
        t0 = time.perf_counter()
        ia.ScanBatch("cyk", cms, 1, dsq, [per[m][0] + (m,) for m in range(len(cms))], cut)   # every model's set-up
        log("many models: first use (profiles, device tables) %.1f s" % (time.perf_counter() - t0))
        t0 = time.perf_counter()
        got = ia.ScanBatch("cyk", cms, 1, dsq, wins, cut)
        tb = time.perf_counter() - t0
        cells = ia.last_call_timing()["cells"]
        kb = ia.last_call_timing()["kernel_ms"] / 1e3
        log("batched: one launch %.2f s (kernel %.2f s), %.3g cells, %.1f Gcells/s" % (tb, kb, cells, cells / tb / 1e9))
        tg = []
        ia.set_batch_graph(True)
        try:
            for rep in range(2):              # capture, then replay
                t0 = time.perf_counter()
                gg = ia.ScanBatch("cyk", cms, 1, dsq, wins, cut)
                tg.append(time.perf_counter() - t0)
                log("batched: %s %.2f s" % ("graph capture and run" if rep == 0 else "graph replay", tg[-1]))
        finally:
            ia.set_batch_graph(False)
        key = lambda h: (h[0], h[1])
        same = all(sorted(a["hits"], key=key) == sorted(b["hits"], key=key) for a, b in zip(got, gg))
        tuned = None
        if args.batch_tune:
            from infernal_amd import autotune
            best, trials = autotune.tune("cyk", cms, 1, dsq, wins, cut, sample=0.1)
            with autotune.apply(best, max(cm.W for cm in cms)):
                t0 = time.perf_counter()
                gt = ia.ScanBatch("cyk", cms, 1, dsq, wins, cut)
                tt = time.perf_counter() - t0
            same = same and all(sorted(a["hits"], key=key) == sorted(b["hits"], key=key) for a, b in zip(got, gt))
            tuned = {"config": best, "wall_s": round(tt, 3), "gcells_per_s": round(cells / tt / 1e9, 3),
                     "trials": [[c, round(t, 4)] for c, t in trials]}
            log("many models, tuned %s: %.2f s" % (best, tt))
        lat = []
        for m, cm in enumerate(cms):
            ia.ScanBatch("cyk", [cm], 1, dsq, [(i, j, 0) for (i, j) in per[m][:1]], [0.0])
            t1 = time.perf_counter()
            r = ia.ScanBatch("cyk", [cm], 1, dsq, [(i, j, 0) for (i, j) in per[m]], [0.0])[0]
            lat.append(time.perf_counter() - t1)
            if (m + 1) % 25 == 0:
                log("per-model launches: %d of %d models, %.1f s" % (m + 1, len(cms), sum(lat)))
            same = same and sorted(r["hits"], key=key) == sorted(got[m]["hits"], key=key) and \
                np.float32(r["sc"]).view(np.uint32) == np.float32(got[m]["sc"]).view(np.uint32)
    finally:
        ia.unpin_target()
    log("many models: one launch %.2f s, graph replay %.2f s, one launch per model %.2f s" % (tb, tg[-1], sum(lat)))
    return {"models": len(cms), "model_dir": os.path.relpath(cmdir, ROOT) if cmdir.startswith(ROOT) else cmdir,
            "W_range": [min(c.W for c in cms), max(c.W for c in cms)],
            "clen_range": [min(c.clen for c in cms), max(c.clen for c in cms)],
            "windows": len(wins), "windows_per_model": args.batch_windows, "window_len": "1.2 W", "target_mb": args.batch_mb,
            "cells": cells,
            "batched": {"wall_s": round(tb, 3), "kernel_s": round(kb, 3), "gcells_per_s": round(cells / tb / 1e9, 3),
                        "latency_s": round(tb, 3)},
            "batched_graph_replay": {"wall_s": round(tg[-1], 3), "gcells_per_s": round(cells / tg[-1] / 1e9, 3)},
            "per_model_launches": {"wall_s": round(sum(lat), 3), "gcells_per_s": round(cells / sum(lat) / 1e9, 3),
                                   "latency_p50_s": round(float(np.percentile(lat, 50)), 4),
                                   "latency_p99_s": round(float(np.percentile(lat, 99)), 4)},
            "tuned": tuned, "hits_identical": bool(same)}


def latency_leg(args, ia, synth):
    """Single-call latency of the drop-in on short targets: one FastCYKScan
    (QDB1) and one FastFInsideScan (QDB2) call of the 350-nt model on targets
    of 2,300 and 8,000 nt (2nd-order Markov background, one planted homolog),
    wall time of the call (upload, launch, hit download included), median of
    --latency-reps calls; with the default chunking and with chunks of 2W
    (the rule for targets that fill the chip), hits compared between the
    two."""
    cm = ia.CM.read(args.cm)
    cm.configure(local=True)
    W = cm.W
    out = {"cm": "%s (clen %d, W %d)" % (cm.c.name.decode(), cm.clen, W), "reps": args.latency_reps, "targets": []}
    key = lambda h: (h[0], h[1])
    for L in (2300, 8000):
        bg = synth.markov2_background(L, gc=0.40, seed=900 + L)
        frags = synth.homologs(ia.CM.read(args.cm), 1, seed=901 + L)
        codes, _ = synth.plant(bg, frags, seed=902 + L)
        dsq = ia.dsq_from_codes(codes)
        rec = {"nt": int(len(codes))}
        for algo, f, q in (("cyk", ia.FastCYKScan, 1), ("finside", ia.FastFInsideScan, 2)):
            res = {}
            for mode, chunk in (("default", 0), ("chunks_2W", 2 * W)):
                ia.set_launch_params(chunk, 0)
                try:
                    f(cm, q, dsq, 1, len(codes), 0.0)                 # first use
                    ts, hits = [], None
                    for _ in range(args.latency_reps):
                        t0 = time.perf_counter()
                        r = f(cm, q, dsq, 1, len(codes), 0.0)
                        ts.append(time.perf_counter() - t0)
                        hits = sorted(r["hits"], key=key)
                finally:
                    ia.set_launch_params(0, 0)
                res[mode] = {"wall_ms": round(1e3 * float(np.median(ts)), 2), "kernel_ms": round(ia.last_call_timing()["kernel_ms"], 2),
                             "hits": hits}
            same = res["default"]["hits"] == res["chunks_2W"]["hits"]
            rec[algo] = {m: {k: v for k, v in r.items() if k != "hits"} for m, r in res.items()}
            rec[algo]["hits"] = len(res["default"]["hits"])
            rec[algo]["hits_identical"] = bool(same)
            rec[algo]["speedup"] = round(res["chunks_2W"]["wall_ms"] / max(res["default"]["wall_ms"], 1e-9), 2)
            log("latency %s %d nt: %.1f ms (2W chunks %.1f ms)" % (algo, len(codes), res["default"]["wall_ms"], res["chunks_2W"]["wall_ms"]))
        out["targets"].append(rec)
    return out


def trunc_leg(args, ia, synth, threads):
    """Truncated scanning with the 1,500-nt model: --trunc-windows windows of
    200 to --trunc-lmax nt (uniform), half an emitted homolog clipped by the
    window at a random offset, half random ACGU decoys.  RefTrCYKScan (QDB1)
    and RefITrInsideScan (QDB2) in the 5'-and-3' truncation pass (J, L, R and
    T modes filled), --trunc-batch windows per launch (0: all in one launch,
    one workgroup per window); the oracle on the host scans a sample of the
    windows (one per thread, of at most --trunc-cpu-lmax nt) with the same
    scanner, so both rates are banded cells per second."""
    path = os.path.join(ROOT, "tests", "golden", "cm", "synth_nested1500.cm")
    cm = ia.CM.read(path)
    cm.configure(local=True, trunc=True, scanmx=True, trscanmx=True)
    log("truncated leg: %d windows of 200-%d nt" % (args.trunc_windows, args.trunc_lmax))
    codes, wins, hom = synth.truncated_windows(ia.CM.read(path), args.trunc_windows, seed=33, lmax=args.trunc_lmax, pool=50)
    dsq = ia.dsq_from_codes(codes)
    pidx = ia.PLI_PASS_5P_AND_3P_FORCE
    out = {"cm": "%s (clen %d, M %d, W %d)" % (cm.c.name.decode(), cm.clen, cm.M, cm.W),
           "windows": len(wins), "window_nt": [200, args.trunc_lmax], "residues": int(len(codes)), "pass": "5P_AND_3P_FORCE",
           "note": "cells are banded (v, j, d) cells; each is computed in J, L and R mode (T too for bifurcations)"}
    order = sorted(range(len(wins)), key=lambda k: wins[k][1] - wins[k][0])
    nb = args.trunc_batch if args.trunc_batch > 0 else len(order)
    batches = [order[a:a + nb] for a in range(0, len(order), nb)]
    out["launch_windows"] = nb
    for algo, q in (("cyk", 1), ("iinside", 2)):
        ia.TrScanWindows(algo, cm, q, pidx, dsq, [wins[order[0]]], 0.0)    # first-use set-up
        wall = kern = cells = 0.0
        hitw = [False] * len(wins)
        nh = 0
        for b in batches:
            t0 = time.perf_counter()
            r = ia.TrScanWindows(algo, cm, q, pidx, dsq, [wins[k] for k in b], 0.0)
            wall += time.perf_counter() - t0
            tm = ia.last_call_timing()
            kern += tm["kernel_ms"] / 1e3
            cells += tm["cells"]
            for k, h in zip(b, r["hits"]):
                hitw[k] = len(h) > 0
                nh += len(h)
            log("truncated %s: launch of %d windows (%d-%d nt) %.2f s kernel" % (algo, len(b), wins[b[0]][1] - wins[b[0]][0] + 1,
                                                                                 wins[b[-1]][1] - wins[b[-1]][0] + 1, tm["kernel_ms"] / 1e3))
        log("truncated %s: %.2f s wall, %.2f s kernel, %d windows, %d hits" % (algo, wall, kern, len(wins), nh))
        out[algo] = {"wall_s": round(wall, 3), "kernel_s": round(kern, 3), "cells": cells,
                     "gpu_gcells_per_s": round(cells / wall / 1e9, 3),
                     "gpu_kernel_gcells_per_s": round(cells / kern / 1e9, 3),
                     "homolog_windows_hit": int(sum(h and w for h, w in zip(hitw, hom))),
                     "decoy_windows_hit": int(sum(h and not w for h, w in zip(hitw, hom)))}
    if threads:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_bind as ob
        ob.load()
        rng = np.random.default_rng(77)
        short = [k for k in range(len(wins)) if wins[k][1] - wins[k][0] + 1 <= args.trunc_cpu_lmax]
        sample = [wins[k] for k in rng.choice(short, min(threads, len(short)), replace=False)]
        out["cpu"] = {"cores": len(sample), "kind": "port",
                      "sample": "oracle, %d of the windows of at most %d nt, one per thread" % (len(sample), args.trunc_cpu_lmax)}
        for algo, oalgo, q in (("cyk", ob.ORACLE_TRCYK, 1), ("iinside", ob.ORACLE_TRIINSIDE, 2)):
            if algo == "iinside":
                cm.c.search_opts |= ia.CM_SEARCH_INSIDE
            t0 = time.perf_counter()
            with cf.ThreadPoolExecutor(len(sample)) as ex:
                list(ex.map(lambda w: ob.scan(cm, oalgo, q, dsq, int(w[0]), int(w[1]), 1e30, want_rows=False, pass_idx=pidx), sample))
            cw = time.perf_counter() - t0
            cells = sum(_cells(cm, q, int(w[0]), int(w[1]), truncated=True) for w in sample)
            log("truncated CPU %s sample: %d windows in %.1f s" % (algo, len(sample), cw))
            out["cpu"][algo + "_gcells_per_s"] = round(cells / cw / 1e9, 4)
            out["gpu_over_cpu_" + algo] = round(out[algo]["gpu_gcells_per_s"] / out["cpu"][algo + "_gcells_per_s"], 1)
    return out


def gate_leg(args, ia, synth, cm_path, codes, dsq, plain_hits, plain_s):
    """The learned gate on this workload: trained on a separate seeded
    target (10% emitted homologs), then run on the benchmark target in place
    of the plain CYK + Inside pass.  Training pairs come from the GPU CYK
    (--gate-labels cyk, the default): a residue is positive when a
    FastCYKScan hit scoring at most 5 bits below the CYK score of E = 0.01
    on the training target (the model's local-CYK tail) covers it, and the
    gate's threshold is set so the held-out CYK hits of E <= 0.01 pass;
    --gate-labels inside labels from Inside hits down to T - 5 bits instead.
    (The pipeline's CYK filter cutoff, P = 1e-4 per hit, lets through random
    hits covering a few percent of a target: as labels they leave the gate
    nothing to reject.)  Judged
    against the plain method's hits: FastFInsideScan hits with E <= 0.01
    (the model's local-Inside tail, one strand of the target).  Returns the
    JSON record."""
    from infernal_amd import gate as gt
    from infernal_amd.calibrate import read_exp_infos
    exps = read_exp_infos(cm_path)
    if "LI" not in exps:
        return {"skipped": "model has no E-value calibration"}
    n = len(codes)
    li = exps["LI"].update_for_dbsize(float(n))
    T = li.score_for_evalue(0.01)
    cm = ia.CM.read(cm_path)
    cm.configure(local=True)
    W = cm.W
    # training target and its labels
    t0 = time.perf_counter()
    ntr = int(args.gate_train_mb * 1e6)
    bg = synth.markov2_background(ntr, gc=0.40, seed=BG_SEED, sample_seed=4242)   # the benchmark's chain, another sample
    nh = max(1, int(0.10 * ntr / (0.75 * cm.clen)))
    frags = synth.homologs(ia.CM.read(cm_path), nh, seed=4243)
    tcodes, _ = synth.plant(bg, frags, seed=4244)
    tdsq = ia.dsq_from_codes(tcodes)
    if args.gate_labels == "cyk":
        if "LC" not in exps:
            return {"skipped": "model has no local-CYK calibration"}
        Tc = exps["LC"].update_for_dbsize(float(ntr)).score_for_evalue(0.01)
        r = ia.FastCYKScan(cm, 1, tdsq, 1, ntr, Tc - 5.0, want_hits=True)
        lab = gt.labels_from_hits(ntr, [(h[0], h[1]) for h in r["hits"]])
        train_hits = [(h[0], h[1], h[4]) for h in r["hits"] if h[4] >= Tc]
        label_rule = "FastCYKScan (QDB1) hits >= %.2f bits (CYK E = 0.01 on the training target: %.2f)" % (Tc - 5.0, Tc)
    else:
        r = ia.FastFInsideScan(cm, 2, tdsq, 1, ntr, T - 5.0, want_hits=True)
        lab = gt.labels_from_hits(ntr, [(h[0], h[1]) for h in r["hits"]])
        train_hits = [(h[0], h[1], h[4]) for h in r["hits"] if h[4] >= T]
        label_rule = "FastFInsideScan (QDB2) hits >= %.2f bits" % (T - 5.0)
    t_data = time.perf_counter() - t0
    # training
    net = gt.GateNet(seed=7)
    pos = float(lab.mean())
    pw = min(20.0, max(1.0, (1 - pos) / max(pos, 1e-6)) ** 0.5)
    t0 = time.perf_counter()
    ntrain = int(0.75 * ntr)          # the rest is held out to set the threshold
    losses = gt.train(net, tcodes[:ntrain], lab[:ntrain], args.gate_steps, seg=1024, nseg=32, lr=2e-3, pos_weight=pw)
    gt._call("infx_dev_sync")
    t_train = time.perf_counter() - t0
    # threshold: every held-out plain hit passes, with a margin
    tl = gt.smooth(net.score(tcodes[ntrain:]), args.gate_smooth)
    val = [(a - ntrain, b - ntrain) for a, b, _ in train_hits if min(a, b) > ntrain]
    per_hit = np.array([tl[min(a, b) - 1:max(a, b)].max() for a, b in val]) if val else np.zeros(1)
    tau = float(np.quantile(per_hit, args.gate_quantile)) - args.gate_margin
    # widening of the flagged residues: W - 1, or set on the held-out hits
    if args.gate_widen == "auto":
        widen = gt.fit_widen(tl, tau, val, W)
    else:
        widen = W - 1 if args.gate_widen == "w" else int(args.gate_widen)
    # the gated pass on the benchmark target
    ia.FastFInsideScan(cm, 2, dsq, 1, min(n, 200000), T, want_hits=True)    # warm profile
    net.score(codes[:100000])
    t0 = time.perf_counter()
    logits = gt.smooth(net.score(codes), args.gate_smooth)
    t_score = time.perf_counter() - t0
    keep = [h for h in plain_hits if h[2] >= T]
    test_max = np.array([logits[min(a, b) - 1:max(a, b)].max() for a, b, _ in keep]) if keep else np.zeros(1)
    for q in (0.0, 0.002, 0.005, 0.01, 0.02):
        tq = float(np.quantile(per_hit, q)) - args.gate_margin
        log("gate: held-out quantile %.3f -> tau %.2f: residues above %.4f, test hits reachable %.4f"
            % (q, tq, float((logits > tq).mean()), float((test_max > tq).mean())))
    wins = gt.select_windows(logits, tau, W, widen)
    t1 = time.perf_counter()
    ghits = gt.gated_inside(cm, dsq, wins, T)
    t_inside = time.perf_counter() - t1
    t_gate = t_score + (t1 - t0 - t_score) + t_inside
    rec = gt.recall(keep, ghits)
    log("gate: train %.1f s (%d steps, loss %.4f -> %.4f), tau %.2f, widen %d, %d windows (%.2f%% of residues), "
        "score %.2f s, inside %.2f s, recall %.4f of %d hits" % (t_train, args.gate_steps, losses[0], losses[-1], tau, widen,
                                                               len(wins), 100.0 * sum(j - i + 1 for i, j in wins) / n,
                                                               t_score, t_inside, rec, len(keep)))
    return {"speedup": round(plain_s / t_gate, 2), "recall": round(rec, 5), "plain_hits_e001": len(keep),
            "score_threshold_bits": round(T, 2), "gate_s": round(t_gate, 3), "plain_s": round(plain_s, 2),
            "gate_score_s": round(t_score, 3), "gated_inside_s": round(t_inside, 3), "windows": len(wins),
            "residue_fraction": round(sum(j - i + 1 for i, j in wins) / n, 5), "tau": round(tau, 3), "widen": int(widen),
            "train": {"target_mb": ntr / 1e6, "homologs": nh, "steps": args.gate_steps, "seconds": round(t_train, 1),
                      "label_seconds": round(t_data, 1), "labels": label_rule, "positive_fraction": round(pos, 5),
                      "loss_first": round(losses[0], 4), "loss_last": round(losses[-1], 4),
                      "train_hits": len(train_hits), "heldout_hits": len(val), "smooth": args.gate_smooth,
                      "margin": args.gate_margin, "quantile": args.gate_quantile}}


def gate_1500_leg(args, ia, synth):
    """The learned gate on the 1,500-nt synthetic model (the truncated
    workload's model; E-value calibrated): an inference target of
    --gate5-mb Mb (2nd-order Markov background, 10% of its residues in
    emitted homologs), the plain method over all of it (FastCYKScan QDB1 +
    FastFInsideScan QDB2, timed, target resident in device memory), then the
    gate trained on --gate5-train-mb Mb of CYK-labelled residues (one
    (window, CYK label) pair per residue) and run in place of the plain
    pass.  Judged on the plain FastFInsideScan hits with E <= 0.01."""
    import copy
    from infernal_amd.calibrate import read_exp_infos
    path = os.path.join(ROOT, "tests", "golden", "cm", "synth_nested1500.cm")
    exps = read_exp_infos(path)
    if "LI" not in exps or "LC" not in exps:
        return {"skipped": "synth_nested1500.cm has no E-value calibration"}
    cm = ia.CM.read(path)
    cm.configure(local=True)
    n = int(args.gate5_mb * 1e6)
    t0 = time.perf_counter()
    bg = synth.markov2_background(n, gc=0.40, seed=BG_SEED + 5)
    nh = max(1, int(0.10 * n / (0.75 * cm.clen)))
    frags = synth.homologs(ia.CM.read(path), nh, seed=5151)
    codes, _ = synth.plant(bg, frags, seed=5152)
    dsq = ia.dsq_from_codes(codes)
    log("gate, 1500-nt model: %.0f-Mb target with %d homologs generated in %.1f s" % (n / 1e6, nh, time.perf_counter() - t0))
    T = exps["LI"].update_for_dbsize(float(n)).score_for_evalue(0.01)
    ia.pin_target(dsq)
    try:
        ia.FastCYKScan(cm, 1, dsq, 1, min(n, 4000), 0.0)          # first-use set-up
        ia.FastFInsideScan(cm, 2, dsq, 1, min(n, 4000), 0.0)
        t0 = time.perf_counter()
        ia.FastCYKScan(cm, 1, dsq, 1, n, 0.0, want_hits=True)
        t_cyk = time.perf_counter() - t0
        log("gate, 1500-nt model: plain CYK %.1f s" % t_cyk)
        t0 = time.perf_counter()
        ri = ia.FastFInsideScan(cm, 2, dsq, 1, n, T, want_hits=True)
        t_ins = time.perf_counter() - t0
        plain_hits = [(h[0], h[1], h[4]) for h in ri["hits"]]
        log("gate, 1500-nt model: plain Inside %.1f s, %d hits with E <= 0.01 (T = %.2f bits)" % (t_ins, len(plain_hits), T))
        a5 = copy.copy(args)
        a5.gate_train_mb, a5.gate_steps = args.gate5_train_mb, args.gate5_steps
        gate = gate_leg(a5, ia, synth, path, codes, dsq, plain_hits, t_cyk + t_ins)
    finally:
        ia.unpin_target()
    gate.update({"cm": "%s (clen %d, M %d, W %d)" % (cm.c.name.decode(), cm.clen, cm.M, cm.W), "target_mb": n / 1e6,
                 "homologs": nh, "plain_cyk_s": round(t_cyk, 2), "plain_inside_s": round(t_ins, 2)})
    return gate


def _cells(cm, qdbidx, i0, j0, truncated=False):
    import ctypes as C
    import infernal_amd as ia
    lib = ia.lib()
    lib.infx_count_scan_cells.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int64, C.c_int64, C.POINTER(C.c_double)]
    out = C.c_double()
    mx = cm.c.trsmx if truncated else cm.c.smx
    lib.infx_count_scan_cells(C.cast(cm.ptr, C.c_void_p), mx, qdbidx, 1 if truncated else 0, i0, j0, C.byref(out))
    return out.value


def cpu_share():
    """Host cores this process may use: its CPU affinity, bounded by a cgroup
    CPU quota when one is set, and by 16 (the GPU host's per-GPU CPU share).
    Returns (cores, record)."""
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()[:2]
            if q != "max":
                quota = float(q) / float(per)
    except (OSError, ValueError):
        pass
    cores = aff if quota is None else max(1, min(aff, int(quota)))
    rec = dict(host_cpu(), affinity=aff, cgroup_quota_cpus=quota, per_gpu_share=16)
    return max(1, min(cores, 16)), rec


def blocks_of(n, block, W):
    """The target cut into database blocks of `block` residues, each after
    the first starting W-1 residues early (the overlap cmsearch gives its
    sequence windows, src/cmsearch.c:810), so every hit lies inside a block."""
    out, a = [], 1
    while a <= n:
        b = min(n, a + block - 1)
        out.append((max(1, a - (W - 1)), b))
        a = b + 1
    return out


def main():
    t_proc = time.perf_counter()
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--mb", type=float, default=100.0, help="target length in megabases")
    ap.add_argument("--block-mb", type=float, default=12.5, help="database block scanned by one step (megabases)")
    ap.add_argument("--budget-s", type=float, default=430.0,
                    help="wall-clock seconds (from start) after which no further timed step is begun")
    ap.add_argument("--homologs", type=int, default=2000)
    ap.add_argument("--cm", default=os.path.join(ROOT, "tests", "golden", "cm", "synth_multiloop350.cm"))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-sample", type=int, default=12000, help="residues per CPU-baseline window (16 windows: ~12 s of CPU work)")
    ap.add_argument("--only", choices=["both", "cyk", "finside"], default="both")
    ap.add_argument("--chunk", type=int, default=0)
    ap.add_argument("--gate-steps", type=int, default=6000, help="learned-gate training steps (0 = skip the gate leg)")
    ap.add_argument("--gate-budget-s", type=float, default=470.0,
                    help="run the gate leg only if it can start before this many seconds")
    ap.add_argument("--gate-train-mb", type=float, default=8.0, help="learned-gate training target in megabases")
    ap.add_argument("--gate-smooth", type=int, default=128, help="residues the gate's logits are averaged over")
    ap.add_argument("--gate-margin", type=float, default=0.5, help="logit margin below the threshold quantile")
    ap.add_argument("--gate-quantile", type=float, default=0.005, help="held-out hit quantile the threshold sits at")
    ap.add_argument("--gate-widen", default="auto",
                    help="residues each flagged residue is widened by: 'auto' (set on the held-out hits), 'w' (W-1) or a number")
    ap.add_argument("--gate-labels", choices=("cyk", "inside"), default="cyk",
                    help="training labels: GPU CYK hits at the CYK filter cutoff, or Inside hits")
    ap.add_argument("--cyk1500-mb", type=float, default=0.0,
                    help="also scan this many Mb with the 1,500-nt model (FastCYKScan QDB1), GPU and CPU; 0 = skip")
    ap.add_argument("--trunc-windows", type=int, default=0,
                    help="windows of the truncated-scan leg (1,500-nt model, 200 to --trunc-lmax nt each; "
                         "run in a child process); 0 = skip")
    ap.add_argument("--trunc-lmax", type=int, default=1000,
                    help="longest truncated-scan window")
    ap.add_argument("--trunc-batch", type=int, default=0, help="windows per truncated-scan launch (0: all in one launch)")
    ap.add_argument("--trunc-cpu-lmax", type=int, default=500, help="longest window of the truncated leg's CPU sample")
    ap.add_argument("--trunc-only", action="store_true", help="run only the truncated-scan leg and print its JSON")
    ap.add_argument("--batch-windows", type=int, default=0,
                    help="windows per model of the many-models leg (one launch for all models vs one per model); 0 = skip")
    ap.add_argument("--batch-mb", type=float, default=50.0, help="target the many-models leg samples its windows from")
    ap.add_argument("--batch-cmdir", default="", help="models of the many-models leg (.cm or .cm.gz; default tests/golden/cm)")
    ap.add_argument("--batch-tune", action="store_true", help="many-models leg: also run with autotuned launch parameters")
    ap.add_argument("--batch-only", action="store_true", help="run only the many-models leg and print its JSON")
    ap.add_argument("--latency-only", action="store_true", help="run only the single-call latency leg and print its JSON")
    ap.add_argument("--gate5-only", action="store_true", help="run only the learned gate on the 1,500-nt model and print its JSON")
    ap.add_argument("--cyk1500-only", action="store_true", help="run only the 1,500-nt CYK leg (--cyk1500-mb, default 20) and print its JSON")
    ap.add_argument("--gate5-mb", type=float, default=20.0, help="1,500-nt gate leg: inference target in megabases")
    ap.add_argument("--gate5-train-mb", type=float, default=50.0, help="1,500-nt gate leg: CYK-labelled training target in megabases")
    ap.add_argument("--gate5-steps", type=int, default=6000, help="1,500-nt gate leg: training steps")
    ap.add_argument("--latency-reps", type=int, default=5, help="calls per latency case (median reported)")
    ap.add_argument("--copy-probe", type=float, default=0.0,
                    help="only copy this many GiB device-to-device (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE) and exit")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (the single-GPU share of a sharded scan; "
              "sharding across GPUs is described but not built); --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    if args.copy_probe > 0:
        import torch
        nbytes = int(args.copy_probe * (1 << 30))
        a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": nbytes}))
        return

    import infernal_amd as ia
    from infernal_amd import synth
    if not ia.gpu_available():
        print("bench.py: no usable gfx9 GPU", file=sys.stderr)
        sys.exit(1)
    if args.chunk:
        ia.set_launch_params(args.chunk, 0)
    cores, cpu_rec = cpu_share()
    threads = args.cpu_threads or cores
    if args.trunc_only:
        print(json.dumps(trunc_leg(args, ia, synth, None if args.no_cpu else threads)))
        return
    if args.batch_only:
        print(json.dumps(batch_leg(args, ia, synth)))
        return
    if args.latency_only:
        print(json.dumps(latency_leg(args, ia, synth)))
        return
    if args.gate5_only:
        print(json.dumps(gate_1500_leg(args, ia, synth)))
        return
    if args.cyk1500_only:
        if args.cyk1500_mb <= 0:
            args.cyk1500_mb = 20.0
        print(json.dumps(cyk_1500(args, ia, synth, None if args.no_cpu else threads)))
        return

    cm = ia.CM.read(args.cm)
    cm.configure(local=True)
    n = int(args.mb * 1e6)
    t = time.perf_counter()
    bg = synth.markov2_background(n, gc=0.40, seed=BG_SEED)
    cm_g = ia.CM.read(args.cm)      # global model for emission
    frags = synth.homologs(cm_g, args.homologs, seed=7) if args.homologs else []
    codes, starts = synth.plant(bg, frags, seed=11) if frags else (bg, [])
    dsq = ia.dsq_from_codes(codes)
    log("target %.1f Mb, %d homologs, generated in %.1f s" % (n / 1e6, len(frags), time.perf_counter() - t))
    M, W, clen = cm.M, cm.W, cm.clen
    c1, _ = cm.cells_per_res(1)
    c2, _ = cm.cells_per_res(2)
    log("model %s clen %d M %d W %d; cells/res QDB1 %.0f QDB2 %.0f" % (cm.c.name.decode(), clen, M, W, c1, c2))
    blocks = blocks_of(n, int(args.block_mb * 1e6), W)
    nb = len(blocks)
    # the target stays in device memory for the whole run: steps start with
    # their residues resident (one upload here)
    ia.pin_target(dsq)
    log("%d blocks of %.1f Mb (W-1 overlap), target pinned in device memory" % (nb, args.block_mb))

    # CPU baseline first, so that every line printed below carries it
    cpu = None
    if not args.no_cpu:
        r = cpu_baseline(cm, codes, threads, args.cpu_sample, W)
        log("CPU baseline: CYK %.1f s, FInside %.1f s in %d threads" % (r["cyk"][1], r["finside"][1], threads))
        cc = r["cyk"][0] + r["finside"][0]
        ct = r["cyk"][1] + r["finside"][1]
        cpu = {"value": round(cc / ct / 1e9, 4), "unit": "Gcells/s", "cores": threads, "kind": "port", "host": cpu_rec,
               "sample": "oracle FastCYKScan (QDB1) + FastFInsideScan (QDB2), %d windows of %d nt of the target in %d threads, "
                         "%.1f s CYK + %.1f s FInside" % (threads, args.cpu_sample + W - 1, threads, r["cyk"][1], r["finside"][1]),
               "cyk_gcells_per_s": round(r["cyk"][0] / r["cyk"][1] / 1e9, 4),
               "finside_gcells_per_s": round(r["finside"][0] / r["finside"][1] / 1e9, 4),
               "one_core": {"value": round((r["cyk_1core"][0] + r["finside_1core"][0]) / (r["cyk_1core"][1] + r["finside_1core"][1]) / 1e9, 4),
                            "cyk_gcells_per_s": round(r["cyk_1core"][0] / r["cyk_1core"][1] / 1e9, 4),
                            "finside_gcells_per_s": round(r["finside_1core"][0] / r["finside_1core"][1] / 1e9, 4),
                            "sample": "the first of those windows, alone"}}

    do_cyk, do_ins = args.only in ("both", "cyk"), args.only in ("both", "finside")
    block_hits = {}

    def step(k, timed):
        a, b = blocks[k % nb]
        out = {}
        if do_cyk:
            r = ia.FastCYKScan(cm, 1, dsq, a, b, 0.0, want_hits=True)
            tm = ia.last_call_timing()
            out["cyk"] = (tm["kernel_ms"], tm["cells"], len(r["hits"]))
        if do_ins:
            r = ia.FastFInsideScan(cm, 2, dsq, a, b, 0.0, want_hits=True)
            tm = ia.last_call_timing()
            out["finside"] = (tm["kernel_ms"], tm["cells"], len(r["hits"]))
            block_hits[k % nb] = [(h[0], h[1], h[4]) for h in r["hits"]]
        return out

    for k in range(args.warmup):
        t0 = time.perf_counter()
        step(k, False)
        log("warmup step %d (block %d): %.2f s" % (k, k % nb, time.perf_counter() - t0))

    kname = {"cyk": "std_cyk_kernel", "finside": "std_inside_lds_kernel<SEMI_FINSIDE>"}
    ops_res = {"cyk": ops_per_res(cm, 1, False), "finside": ops_per_res(cm, 2, True)}
    kern = {"cyk": [], "finside": []}
    cells_by = {"cyk": 0.0, "finside": 0.0}
    res_by = 0
    hits = {}
    step_times = []
    line = None
    for s in range(args.steps):
        if step_times and time.perf_counter() - t_proc + max(step_times) > args.budget_s:
            log("wall budget: stopping after %d of %d timed steps" % (len(step_times), args.steps))
            break
        k = args.warmup + s
        a, b = blocks[k % nb]
        t0 = time.perf_counter()
        out = step(k, True)
        dt = time.perf_counter() - t0
        step_times.append(dt)
        res_by += b - a + 1
        for alg, (kms, cells, nh) in out.items():
            kern[alg].append(kms)
            cells_by[alg] += cells
            hits[alg] = hits.get(alg, 0) + nh
        log("step %d (block %d, %d-%d): %.2f s; " % (s, k % nb, a, b, dt) +
            ", ".join("%s kernel %.2f s" % (alg, v[0] / 1e3) for alg, v in out.items()))
        total_t = sum(step_times)
        ns = len(step_times)
        cells_total = cells_by["cyk"] + cells_by["finside"]
        dom = max((x for x in kern if kern[x]), key=lambda x: sum(kern[x]))
        kms_avg = sum(kern[dom]) / len(kern[dom])
        ops = ops_res[dom] * res_by
        achieved = ops / (sum(kern[dom]) / 1e3) / 1e12
        roof = {"bound": "valu", "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / FP32_PEAK_TFLOPS, 4), "traffic": None,
                "peak_nonpacked_valu": round(VALU_PEAK_TOPS, 2), "frac_nonpacked_valu": round(achieved / VALU_PEAK_TOPS, 4),
                "kernel": kname[dom], "kernel_ms_per_launch": round(kms_avg, 1),
                "ops_per_residue": round(ops_res[dom], 1),
                "gcells_per_s_kernel": {x: round(cells_by[x] / (sum(kern[x]) / 1e3) / 1e9, 2) for x in kern if kern[x]}}
        tj = os.path.join(ROOT, "profiles", "traffic_%s_%s.json" % (dom, WORKLOAD_TAG))
        if os.path.exists(tj):
            with open(tj) as f:
                tr = json.load(f)
            if tr.get("cm") == cm.c.name.decode() and tr.get("block_residues") == blocks[0][1] - blocks[0][0] + 1:
                roof["traffic"] = tr["hbm_bytes_per_launch"]
                roof["traffic_note"] = tr.get("note", "")
                roof["traffic_rate"] = traffic_rate(tr["hbm_bytes_per_launch"], kms_avg)
        line = {
            "metric": "Gcells/s QDB CYK+Inside scan on 1 GPU; learned-gate speedup at \u226599% hits kept",
            "value": round(cells_total / total_t / 1e9, 3),
            "unit": "Gcells/s",
            "n_gpus": 1,
            "steps": ns,
            "warmup": args.warmup,
            "ms_per_step": round(total_t / ns * 1e3, 1),
            "higher_is_better": True,
            "scaling": "weak",
            "vs_baseline": None,
            "dtype": "f32",
            "data": "synthetic",
            "config": {"workload": "FastCYKScan QDB1 + FastFInsideScan QDB2, one %.1f-Mb database block per step, of a %.0f-Mb "
                                   "target (2nd-order Markov, 40%% GC, %d planted homologs; %d blocks with W-1 overlap), one strand; "
                                   "residues resident in device memory" % (args.block_mb, n / 1e6, len(frags), nb),
                       "cm": "%s (clen %d, M %d, W %d)" % (cm.c.name.decode(), clen, M, W),
                       "cells_per_step": cells_total / ns, "hits": hits, "steps_requested": args.steps},
            "roofline": roof,
            "cpu_baseline": cpu,
        }
        print(json.dumps(line), flush=True)
    if line is None:
        print("bench.py: no step was timed", file=sys.stderr)
        sys.exit(1)
    ia.unpin_target()
    mean_step = sum(step_times) / len(step_times)

    extra = {}
    if args.gate_steps > 0 and args.only == "both":
        if len(block_hits) < nb:
            extra["learned_gate"] = {"skipped": "the timed and warm-up steps did not cover every block"}
        elif time.perf_counter() - t_proc > args.gate_budget_s:
            extra["learned_gate"] = {"skipped": "wall budget"}
        else:
            seen, plain_hits = set(), []
            for k in range(nb):
                for h in block_hits[k]:
                    if (h[0], h[1]) not in seen:
                        seen.add((h[0], h[1]))
                        plain_hits.append(h)
            # the plain method over the whole target: every block once, at the
            # measured mean step time
            gate = gate_leg(args, ia, synth, args.cm, codes, dsq, plain_hits, mean_step * nb)
            extra["learned_gate"] = gate
            extra["learned_gate_speedup"] = gate["speedup"] if gate.get("recall", 0) >= 0.99 else None
    if args.cyk1500_mb > 0:
        extra["cyk_1500nt"] = cyk_1500(args, ia, synth, None if args.no_cpu else threads)
    if args.trunc_windows > 0:
        # in a child process of its own: the leg's device buffers and host
        # memory are released when it ends
        import subprocess
        cmd = [sys.executable, "-u", os.path.abspath(__file__), "--trunc-only", "--trunc-windows", str(args.trunc_windows),
               "--trunc-lmax", str(args.trunc_lmax), "--trunc-batch", str(args.trunc_batch),
               "--trunc-cpu-lmax", str(args.trunc_cpu_lmax)]
        if args.no_cpu:
            cmd.append("--no-cpu")
        if args.cpu_threads:
            cmd += ["--cpu-threads", str(args.cpu_threads)]
        pr = subprocess.run(cmd, stdout=subprocess.PIPE, text=True)
        lines = [x for x in pr.stdout.splitlines() if x.startswith("{")]
        extra["truncated_1500nt"] = json.loads(lines[-1]) if (pr.returncode == 0 and lines) else \
            {"error": "truncated leg exited with status %d" % pr.returncode}
    if args.batch_windows > 0:
        extra["many_models"] = batch_leg(args, ia, synth)
    if extra:
        line.update(extra)
        print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
