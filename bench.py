#!/usr/bin/env python3
"""Benchmark of the MI355X bbFMM downward pass (one JSON line on stdout).

Workload (the BASELINE precision config): N = 1,048,576 particles i.i.d.
uniform in [0,1)^3 (shifted into the root cell [-1/2,1/2)^3), charges i.i.d.
uniform [-1,1], Laplace kernel 1/r, Chebyshev order n = 4, SVD rank r = 32,
depth 5 by the reference's level rule.  A step is one full solve (P2M, M2M,
M2L, post-M2L, L2L, L2P, P2P) on data already resident in HBM:

  chosen form : 16 charge vectors per solve on one geometry, M2L on MFMA with
                split-BF16 operands (hi + lo, 3 products) and fp32 accumulate,
                P2P in fp32 (1/r and the sum over sources);
  plain form  : the fp32 drop-in (one vector, dense-grid M2L in fp32 FMAs).

value = per-vector solves per second of the chosen form (whole job).  Errors
are relative L2 of q*phi against the FP64 direct sum (GPU) on 4,096 sampled
targets.  cpu_baseline times the reference's own CPU path (CPU8A build,
oracle/_ref) on the host on the same input: one thread per physical core
available to the process, one warm-up solve, the median of 5.
"""
from __future__ import annotations

import argparse
import hashlib
import itertools
import json
import os
import re
import subprocess
import sys
import time

import numpy as np

METRIC = "FMM s/solve + M2L Gflop/s at N=1M,n=4: chosen vs plain, rel L2 err ≤1e-4"
PEAK_BF16_TFLOPS = 2500.0
PEAK_FP32_TFLOPS = 157.3


ROOT = os.path.dirname(os.path.abspath(__file__))
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pmc_traffic_c2.txt")
CSRC = os.path.join(ROOT, "bbfmmongpu_amd", "csrc")
# source file of each kernel whose HBM traffic the bench reports
KERNEL_SOURCE = {"k_m2l_mfma_ss<4,1,true>": "m2l_grid.hip", "k_m2l_mfma_ss_pair<true>": "m2l_grid.hip",
                 "k_m2l_add_parts": "m2l_grid.hip", "k_p2p_seg16<true>": "p2p.hip", "k_p2p_seg16x<3>": "p2p.hip"}
# the launches of the split-BF16 M2L stage at C2 (automatic configuration):
# the deepest level with two 32-column tiles per wave and both class groups
# in one block, the levels above with one tile and 4 classes per wave and
# their steps in 4 parts, then the partial sums added
M2L_KERNELS = ("k_m2l_mfma_ss_pair<true>", "k_m2l_mfma_ss<4,1,true>", "k_m2l_add_parts")


def ss_issued_mfma(maxlev: int, nrhs: int, deep=(4, 2, True), upper=(4, 1)) -> int:
    """MFMAs (32x32x16) the source-stationary M2L issues: per (wave, step) it
    runs its products on, classes x tiles x 3 split products x 2 k-steps.
    A wave (32 columns = (parent, vector) pairs per tile, vector fastest,
    then pz, py, px) skips the steps whose parent shift d moves its whole
    parent box outside the level's [0, P)^3 (ghost sources only); the box is
    exact in px and conservative (the whole axis) in py and pz when the
    columns span more than one row.  deep = (classes per wave, tiles per
    wave, both class groups on the union of their step lists) of the
    deepest level, upper = (classes per wave, tiles per wave) of the others."""
    total = 0
    for lev in range(2, maxlev + 1):
        cls, tiles, union = deep if lev == maxlev else (*upper, False)
        P = 1 << (lev - 1)
        ncols = P ** 3 * nrhs
        span = 32 * tiles
        nwave = -(-ncols // (4 * span)) * 4
        col0 = np.arange(nwave) * span
        ok = col0 < ncols
        f = np.minimum(col0, ncols - 1) // nrhs
        t = np.minimum(col0 + span - 1, ncols - 1) // nrhs
        fx, fy, fz = f // (P * P), (f // P) % P, f % P
        tx, ty, tz = t // (P * P), (t // P) % P, t % P
        row = fx == tx
        line = row & (fy == ty)
        lo = np.stack([fx, np.where(row, fy, 0), np.where(line, fz, 0)], 1)
        hi = np.stack([tx, np.where(row, ty, P - 1), np.where(line, tz, P - 1)], 1)
        for g in range(8 // cls):
            group = range(8) if union else range(g * cls, (g + 1) * cls)
            ds = []
            for sc in range(8):
                for d in itertools.product((-1, 0, 1), repeat=3):
                    if d == (0, 0, 0):
                        continue
                    far = any(max(abs(2 * d[a] + ((sc >> (2 - a)) & 1) - ((c >> (2 - a)) & 1)) for a in range(3)) >= 2
                              for c in group)
                    if far:
                        ds.append(d)
            D = np.array(ds)[None, :, :]
            ghost = ((hi[:, None, :] + D < 0) | (lo[:, None, :] + D >= P)).any(axis=2)
            total += int((~ghost[ok]).sum()) * cls * tiles * 6
    return total


def source_sha(kernel: str):
    """First 12 hex digits of the SHA-256 of the source file of `kernel`."""
    name = KERNEL_SOURCE.get(kernel)
    if not name:
        return None
    with open(os.path.join(CSRC, name), "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()[:12]


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE scaled by the copy-kernel factor) and where they
    came from.  The summary records the SHA-256 of the kernel's source file
    when it was measured; if the source has changed since, the bytes are stale
    and None is returned instead."""
    src = {"file": os.path.relpath(TRAFFIC_FILE, ROOT), "src_sha": None, "current_src_sha": source_sha(kernel)}
    try:
        for line in open(TRAFFIC_FILE):
            f = dict(kv.split("=", 1) for kv in line.split() if "=" in kv)
            if f.get("kernel") == kernel and "traffic_bytes" in f:
                src["src_sha"] = f.get("src_sha")
                if src["src_sha"] and src["src_sha"] == src["current_src_sha"]:
                    return float(f["traffic_bytes"]), src
                src["stale"] = True
                return None, src
    except OSError:
        pass
    return None, src


def short_kernel_name(name: str) -> str:
    """rocprofv3 kernel name -> the short form the summaries use, e.g.
    'void bb::k_m2l_mfma_ss<4, 1, true>(bb::GridArgs)' -> 'k_m2l_mfma_ss<4,1,true>'."""
    n = name.strip()
    for pre in ("void ", "bb::"):
        if n.startswith(pre):
            n = n[len(pre):]
    n = n.replace("bb::", "")
    depth, out = 0, []
    for ch in n:  # drop the argument list (the first '(' outside template brackets)
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            break
        out.append(ch)
    return "".join(out).replace(" ", "")


def summarize_pmc(fetch_csv: str, write_csv: str, out_path: str, command: str):
    """Writes the HBM-traffic summary bench.py reads (profiles/pmc_traffic_c2.txt)
    from two rocprofv3 --pmc passes of the same run (FETCH_SIZE, WRITE_SIZE;
    KiB per dispatch).  Both counters are scaled by what they read for the
    256 MiB calibration copy (--copy-calibration: the largest
    __amd_rocclr_copyBuffer dispatch), and every kernel of KERNEL_SOURCE gets
    the SHA-256 of its source file, so that a later change marks it stale."""
    import csv
    import collections

    def load(path, counter):
        per = collections.defaultdict(list)
        copy = 0.0
        with open(path) as f:
            for row in csv.DictReader(f):
                if row["Counter_Name"] != counter:
                    continue
                v = float(row["Counter_Value"])
                name = short_kernel_name(row["Kernel_Name"])
                if name == "__amd_rocclr_copyBuffer":
                    copy = max(copy, v)
                per[name].append(v)
        return per, copy

    fetch, fcopy = load(fetch_csv, "FETCH_SIZE")
    write, wcopy = load(write_csv, "WRITE_SIZE")
    nbytes = 256 << 20
    ffac, wfac = nbytes / (fcopy * 1024), nbytes / (wcopy * 1024)
    lines = ["# rocprofv3 FETCH_SIZE and WRITE_SIZE passes (KiB per dispatch, mean over dispatches)",
             f"# command: {command}",
             f"# calibration: 256 MiB device copy reads FETCH {fcopy:.0f} KiB, WRITE {wcopy:.0f} KiB; "
             f"factors fetch {ffac:.4f} write {wfac:.4f}",
             f"copy_kernel=__amd_rocclr_copyBuffer bytes={nbytes} fetch_kib={fcopy:.0f} write_kib={wcopy:.0f}"]
    for name in sorted(set(fetch) | set(write)):
        fk = float(np.mean(fetch[name])) if fetch.get(name) else 0.0
        wk = float(np.mean(write[name])) if write.get(name) else 0.0
        traffic = (fk * ffac + wk * wfac) * 1024
        sha = source_sha(name)
        lines.append(f"kernel={name} dispatches={len(fetch.get(name, write.get(name, [])))} fetch_kib={fk:.0f} "
                     f"write_kib={wk:.0f} traffic_bytes={traffic:.0f}" + (f" src_sha={sha}" if sha else ""))
    with open(out_path, "w") as f:
        f.write("\n".join(lines) + "\n")


def host_cores():
    """Physical cores this process may run on: the (socket, core) pairs that
    lscpu lists for the CPUs of the affinity mask, capped by the cgroup CPU
    quota.  Returns (threads, detail)."""
    aff = sorted(os.sched_getaffinity(0))
    detail = {"affinity_cpus": len(aff), "lscpu_physical_cores": None, "cgroup_quota_cpus": None}
    phys = None
    try:
        out = subprocess.run(["lscpu", "-p=CPU,SOCKET,CORE"], capture_output=True, text=True, check=True).stdout
        rows = [l.split(",") for l in out.splitlines() if l and not l.startswith("#")]
        detail["lscpu_physical_cores"] = len({(r[1], r[2]) for r in rows})
        allowed = set(aff)
        phys = len({(r[1], r[2]) for r in rows if int(r[0]) in allowed})
        detail["physical_cores_in_affinity"] = phys
    except (OSError, subprocess.CalledProcessError, ValueError, IndexError):
        pass
    try:
        quota, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if quota != "max":
            detail["cgroup_quota_cpus"] = int(quota) / int(period)
    except (OSError, ValueError):
        pass
    threads = phys or len(aff)
    if detail["cgroup_quota_cpus"]:
        threads = max(1, min(threads, int(detail["cgroup_quota_cpus"])))
    return threads, detail


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def rel_l2(ref, x):
    ref = np.asarray(ref, np.float64)
    x = np.asarray(x, np.float64)
    return float(np.linalg.norm(ref - x) / np.linalg.norm(ref))


def timed_solves(plan, q, phi, steps, warmup, stream):
    import torch
    for _ in range(warmup):
        plan.solve(q, phi)
    torch.cuda.synchronize()
    plan.set(timing=steps)  # per-stage HIP events on the plan's stream, read after the region
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record(stream)
    for _ in range(steps):
        plan.solve(q, phi)
    end.record(stream)
    torch.cuda.synchronize()
    ms = start.elapsed_time(end) / steps
    stages = plan.stage_times()
    plan.set(timing=0)
    return ms, stages


def cpu_baseline(pos, q, U, K, threads, reps=5):
    """Reference CPU path (oracle/_ref, CPU8A build) on the same input: one
    warm-up solve, then the median of `reps` solves on the same tree."""
    import oracle as O
    if O.ref_available("s", fast=True):
        t0 = time.time()
        out = O.ref_run(pos, q, 4, U, K, prec="s", dump=False, fast=True, threads=threads, reps=reps + 1,
                        bind=True)
        wall = time.time() - t0
        info = out["info"]
        # the reference prints its own M2L Gflop/s and P2P rate per solve:
        # the median over the timed solves (the first line is the warm-up)
        m = [float(x) for x in re.findall(r"calc_performance: kernel = ([0-9.]+)", info["stderr"])][1:]
        p = [float(x) for x in re.findall(r"num_pairwise_interactions_per_sec = ([0-9.]+)", info["stderr"])][1:]
        return {
            "kind": "reference",
            "s_per_solve": info["t_solve"],
            "s_per_solve_all": info.get("t_solve_all"),
            "m2l_gflops": float(np.median(m)) if m else None,
            "p2p_ginteractions_per_s": float(np.median(p)) if p else None,
            "t_upward": info["t_upward"], "t_interaction": info["t_interaction"],
            "t_downwardx": info["t_downwardx"], "t_nearfield": info["t_nearfield"],
            "wall_s": wall, "phi": out["phi"],
        }
    # restated port (oracle C build) when the reference build is absent
    T = O.Tree(pos, 0, "s")
    ts = []
    for _ in range(reps + 1):
        t0 = time.time()
        res = O.fmm(T, q.astype(np.float32), 4, U, K, scheme="sibling")
        ts.append(time.time() - t0)
    t = float(np.median(ts[1:]))
    return {"kind": "port", "s_per_solve": t, "s_per_solve_all": ts, "m2l_gflops": None,
            "p2p_ginteractions_per_s": None, "wall_s": sum(ts), "phi": res["phi"]}


def run_c3(args):
    """N = 4,194,304 uniform, n = 8, r = 256, depth 6, 16 charge vectors per
    solve, M2L on FP32 MFMA (dense grid), MFMA P2P; operators precomputed in
    FP64 on the GPU.  Errors against the FP64 direct sum on sampled targets."""
    import torch
    import bbfmmongpu_amd as B

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream()
    N = 1 << 22 if args.N == 1 << 20 else args.N
    nrhs = args.nrhs
    steps = min(args.steps, 3)
    warmup = min(args.warmup, 1)
    rng = np.random.default_rng(0)
    pos = rng.random((N, 3)) - 0.5
    Q = rng.uniform(-1.0, 1.0, size=(nrhs, N))
    log(f"c3: N={N} nrhs={nrhs}: operator precompute n=8 r=256 (FP64, GPU)")
    t0 = time.time()
    U64, K64, sigma = B.operators(8, 256, prec="d", device=True)
    U, K = U64.astype(np.float32), K64.astype(np.float32)
    t_ops = time.time() - t0
    t0 = time.time()
    plan = B.Plan(pos, 8, U, K, nrhs=nrhs, prec="s", stream=stream.cuda_stream).set(
        scheme=B.SCHEME_IJ, form=B.FORM_FP32MFMA)
    t_setup = time.time() - t0
    log(f"c3: maxlev={plan.maxlev} ncell={plan.ncell} interactions={plan.m2l_interactions:.0f} "
        f"pairs={plan.p2p_pairs:.4g} ops={t_ops:.1f}s setup={t_setup:.1f}s")
    qm = torch.from_numpy(Q.astype(np.float32)).to(dev).contiguous()
    phim = torch.empty_like(qm)
    ms, st = timed_solves(plan, qm, phim, steps, warmup, stream)
    log(f"c3: {ms:.1f} ms/step ({nrhs} vectors) stages {st}")
    tg = np.sort(np.random.default_rng(1).choice(N, size=min(args.targets, N), replace=False)).astype(np.int32)
    phi = phim.cpu().numpy()
    errs = []
    for j in range(min(nrhs, 2)):
        d = B.direct(pos, Q[j], tg)
        errs.append(rel_l2(d, phi[j, tg]))
    log(f"c3: errors vs FP64 direct {errs}")
    # the fp64 plain solve of vector 0 (dense-grid M2L in fp64 FMAs), kept for
    # error comparison: its own error against the direct sum and the chosen
    # form's distance from it
    p64 = B.Plan(pos, 8, U64, K64, nrhs=1, prec="d", stream=stream.cuda_stream).set(
        scheme=B.SCHEME_IJ, form=B.FORM_PLAIN)
    q64 = torch.from_numpy(Q[0].astype(np.float64)).to(dev)
    phi64d = torch.empty_like(q64)
    ms64, st64 = timed_solves(p64, q64, phi64d, 1, 1, stream)
    phi64 = phi64d.cpu().numpy()
    p64.close()
    err64 = rel_l2(B.direct(pos, Q[0], tg), phi64[tg])
    chosen_vs_fp64 = rel_l2(phi64, phi[0])
    log(f"c3: fp64 plain solve {ms64:.1f} ms (M2L {st64['m2l']:.1f} ms), its error vs FP64 direct {err64:.3e}, "
        f"chosen vs fp64 plain {chosen_vs_fp64:.3e}")
    r = 256
    inter = plan.m2l_interactions
    flop_m2l = r * (2 * r - 1) * inter * nrhs
    roof = {"kernel": "k_m2l_fp32_r256", "bound": "mfma", "unit": "TFLOP/s",
            "achieved": flop_m2l / (st["m2l"] * 1e-3) / 1e12, "peak": PEAK_FP32_TFLOPS,
            "launch_ms": st["m2l"], "algorithmic_flop_per_launch": flop_m2l, "traffic": None}
    roof["frac"] = roof["achieved"] / roof["peak"]
    result = {
        "metric": "FMM s/solve + M2L Gflop/s at N=4M, n=8, 16 charge vectors, rel L2 err vs FP64 direct",
        "value": nrhs / (ms * 1e-3), "unit": "solve/s (per charge vector)", "n_gpus": 1,
        "steps": steps, "warmup": warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"bbFMM Laplace 1/r, N={N} uniform, n=8, r=256, depth {plan.maxlev}, "
                               f"{nrhs} charge vectors per solve",
                   "N": N, "n": 8, "r": r, "maxlev": plan.maxlev, "nrhs": nrhs,
                   "m2l_scheme": "dense grid (IJ)", "m2l_form": "FP32 MFMA"},
        "s_per_solve": ms * 1e-3 / nrhs,
        "m2l_gflops": flop_m2l / (st["m2l"] * 1e-3) / 1e9,
        "p2p_ginteractions_per_s": plan.p2p_pairs * nrhs / (st["p2p"] * 1e-3) / 1e9,
        "rel_l2_err": {"q_phi": errs, "targets": int(tg.size), "fp64_plain_q_phi": err64,
                       "chosen_vs_fp64_plain": chosen_vs_fp64},
        "fp64_plain": {"ms_per_solve": ms64, "stage_ms": st64},
        "stage_ms": st, "setup_s": {"operators": t_ops, "plan": t_setup},
        "roofline": roof,
    }
    plan.close()
    print(json.dumps(result), flush=True)


def run_c4(args):
    """64 independent instances, N log-uniform in [4096, 262144]; half Plummer
    (a = 0.1, r <= 1, mapped into the unit cube), half on a sphere of radius
    0.4; charges uniform [0, 1]; each instance on its own sparse octree with
    leaves of at most 32 points; n = 4, r = 32, fp32 data with the split-BF16
    M2L (FP32 accumulation).  One plan, M2L configuration by the autotuner,
    the whole solve replayed as one hipGraph.  Errors against the FP64 direct
    sum on sampled targets of every instance; the FP32 MFMA list kernel is
    timed on the same plan beside it."""
    import torch
    import bbfmmongpu_amd as B
    from bbfmmongpu_amd import workloads as W

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream()
    pos_list, q_list = W.uneven_instances(count=64, nmin=4096, nmax=262144, seed=0)
    Ntot = sum(p.shape[0] for p in pos_list)
    U, K, _ = B.operators(4, 32, prec="s")
    t0 = time.time()
    # the n = 4 precision form (split-BF16 M2L, FP32 accumulation; everything
    # else fp32); the tuner chooses between the basic kernel and the
    # class-grouped list MFMA kernel
    plan = B.Plan.batch(pos_list, 4, U, K, nrhs=1, prec="s", stream=stream.cuda_stream).set(
        form=B.FORM_BF16X2)
    t_setup = time.time() - t0
    depths = plan.depths
    log(f"c4: {len(pos_list)} instances, N total {Ntot}, depths {min(depths)}..{max(depths)}, "
        f"interactions {plan.m2l_interactions:.4g} pairs {plan.p2p_pairs:.4g} setup {t_setup:.1f}s")
    q = torch.from_numpy(np.concatenate(q_list).astype(np.float32)).to(dev)[None, :].contiguous()
    phi = torch.empty_like(q)
    plan.solve(q, phi)
    table = plan.tune(reps=3)
    best = min(table, key=lambda t: t[4])
    log(f"c4: tuner {[(int(a), int(b), int(u), int(c), round(d, 3)) for a, b, u, c, d in table]} -> {best}")
    ms_eager, st = timed_solves(plan, q, phi, args.steps, args.warmup, stream)
    plan.set(graph=True)
    ms_graph, _ = timed_solves(plan, q, phi, args.steps, args.warmup, stream)
    log(f"c4: eager {ms_eager:.3f} ms/solve, graph {ms_graph:.3f} ms/solve, stages {st}")
    direct = []
    for i, (p, qi) in enumerate(zip(pos_list, q_list)):
        n = p.shape[0]
        tg = np.sort(np.random.default_rng(100 + i).choice(n, min(n, 256), replace=False)).astype(np.int32)
        direct.append((tg, B.direct(p, qi, tg)))

    def errors(out):
        errs, o = [], 0
        for p, (tg, d) in zip(pos_list, direct):
            errs.append(rel_l2(d, out[o + tg]))
            o += p.shape[0]
        return errs

    errs = errors(phi.cpu().numpy()[0])
    log(f"c4: errors vs FP64 direct: max {max(errs):.3e} median {float(np.median(errs)):.3e}")
    # the same solve with the FP32 MFMA list kernel (eager), for comparison
    plan.set(graph=False).set(scheme=B.SCHEME_IJ, form=B.FORM_FP32MFMA)
    plan.solve(q, phi)
    ms_f32, st_f32 = timed_solves(plan, q, phi, args.steps, args.warmup, stream)
    errs_f32 = errors(phi.cpu().numpy()[0])
    log(f"c4: FP32 MFMA list M2L: {ms_f32:.3f} ms/solve (m2l {st_f32['m2l']:.3f} ms), "
        f"errors max {max(errs_f32):.3e} median {float(np.median(errs_f32)):.3e}")
    flop_m2l = 32 * 63 * plan.m2l_interactions
    result = {
        "metric": "FMM s/solve for 64 uneven instances in one hipGraph (n=4, r=32), rel L2 err vs FP64 direct",
        "value": len(pos_list) / (ms_graph * 1e-3), "unit": "instance solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_graph, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"64 instances, N log-uniform [4096, 262144] (total {Ntot}), half Plummer "
                               f"a=0.1, half sphere surface r=0.4, leaves <= 32 points, n=4, r=32",
                   "instances": len(pos_list), "N_total": Ntot, "depth_min": min(depths),
                   "depth_max": max(depths), "m2l_form": "split-BF16 (3 products, FP32 accumulation)",
                   "m2l_config": {"scheme": int(best[0]), "variant": int(best[1]), "variant_upper": int(best[2]),
                                  "level_switch": int(best[3])}},
        "particles_per_s": Ntot / (ms_graph * 1e-3),
        "ms_eager": ms_eager, "ms_graph": ms_graph,
        "m2l_gflops": flop_m2l / (st["m2l"] * 1e-3) / 1e9,
        "p2p_ginteractions_per_s": plan.p2p_pairs / (st["p2p"] * 1e-3) / 1e9,
        "stage_ms": st, "tuning_table": table,
        "rel_l2_err": {"max": max(errs), "median": float(np.median(errs)), "targets_per_instance": 256},
        "other_forms": {"fp32_mfma": {"ms_eager": ms_f32, "m2l_ms": st_f32["m2l"],
                                      "rel_l2_err": {"max": max(errs_f32), "median": float(np.median(errs_f32))}}},
        "setup_s": t_setup,
    }
    plan.close()
    print(json.dumps(result), flush=True)


def run_c5(args):
    """Differentiation workload: N = 262,144 uniform in [-1/2, 1/2)^3 (seed
    0), depth 4, n = 4, r = 32, fp32, 8 charge vectors uniform [-1, 1]; the
    loss 1/2 ||phi - phi_t||^2 with phi_t the FP64 direct-sum potential of a
    second charge set (seed 2).  A step is the forward solve of the 8 vectors
    plus the adjoint solve (bbfmm_amd_gradient: 16 vectors [q; g], returning
    dLoss/dq and dLoss/dpositions) on plans built once, data resident on the
    GPU, with the split-BF16 M2L; the plain fp32 forms and the
    torch.autograd.Function path (plans from its cache after the first call)
    are timed beside it."""
    import torch
    import bbfmmongpu_amd as B

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream()
    N, m = 262144, 8
    rng = np.random.default_rng(0)
    pos = rng.random((N, 3)) - 0.5
    Q = rng.uniform(-1.0, 1.0, size=(m, N))
    Q2 = np.random.default_rng(2).uniform(-1.0, 1.0, size=(m, N))
    t_np, _ = B.direct_field(pos, Q2)
    U, K, _ = B.operators(4, 32, prec="s")
    q = torch.from_numpy(Q.astype(np.float32)).to(dev).contiguous()
    t = torch.from_numpy(t_np.astype(np.float32)).to(dev).contiguous()
    tg = np.sort(np.random.default_rng(1).choice(N, size=min(args.targets, N), replace=False)).astype(np.int32)
    ex, _ = B.direct_field(pos, Q[:2], tg)

    def measure(form):
        fwd = B.Plan(pos, 4, U, K, nrhs=m, prec="s", stream=stream.cuda_stream).set(
            scheme=B.SCHEME_IJ, form=form, output="potential")
        bwd = B.Plan(pos, 4, U, K, nrhs=2 * m, prec="s", stream=stream.cuda_stream).set(
            scheme=B.SCHEME_IJ, form=form)
        phi = torch.empty_like(q)

        def step():
            fwd.solve(q, phi)
            g = (phi - t).contiguous()
            return bwd.gradient(q, g)

        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(args.steps):
            dq, dx = step()
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / args.steps
        ms_f, _ = timed_solves(fwd, q, phi, args.steps, args.warmup, stream)
        errs = [rel_l2(ex[j], phi[j].cpu().numpy()[tg]) for j in range(2)]
        fwd.close()
        bwd.close()
        return ms, ms_f, errs, dq, dx

    ms, ms_f, errs, dq, dx = measure(B.FORM_BF16X2)
    ms_p, ms_fp, errs_p, dq_p, dx_p = measure(B.FORM_PLAIN)
    d_dq = float((dq - dq_p).norm() / dq_p.norm())
    d_dx = float((dx - dx_p).norm() / dx_p.norm())
    log(f"c5: plain fp32 forms: {ms_p:.3f} ms/step (forward {ms_fp:.3f} ms), errors {errs_p}; "
        f"split-BF16 vs plain gradients: dq {d_dq:.2e} dx {d_dx:.2e}")
    # the torch.autograd.Function path: the first call builds its two plans
    # (tree, lists, uploads); later calls take them from the plan cache, and
    # positions moved inside their leaves are uploaded in place
    pos_t = torch.tensor(pos, dtype=torch.float32, device=dev, requires_grad=True)
    q_t = q.clone().requires_grad_(True)
    B.plan_cache.clear()

    def ag_step(p_t):
        p_t.grad = None
        q_t.grad = None
        ph = B.potential(p_t, q_t, 4, U, K, prec="s")
        loss = 0.5 * ((ph - t) ** 2).sum()
        loss.backward()

    torch.cuda.synchronize()
    t0 = time.time()
    ag_step(pos_t)
    torch.cuda.synchronize()
    s_ag_first = time.time() - t0
    n_ag = max(args.steps // 2, 3)
    t0 = time.time()
    for _ in range(n_ag):
        ag_step(pos_t)
    torch.cuda.synchronize()
    s_ag = (time.time() - t0) / n_ag
    agree = float((q_t.grad - dq).norm() / dq.norm())
    # positions moved by a small step (inside their leaves): update_positions
    moved = torch.tensor(pos + 1e-7 * (rng.random((N, 3)) - 0.5), dtype=torch.float32, device=dev,
                         requires_grad=True)
    builds = B.plan_cache.builds
    ag_step(moved)  # first call on the moved positions: uploads them
    torch.cuda.synchronize()
    t0 = time.time()
    for k in range(n_ag):
        with torch.no_grad():
            moved += 1e-9 * (1 if k % 2 == 0 else -1)
        ag_step(moved)
    torch.cuda.synchronize()
    s_ag_moved = (time.time() - t0) / n_ag
    rebuilt = B.plan_cache.builds - builds
    log(f"c5: forward + adjoint {ms:.3f} ms/step (forward {ms_f:.3f} ms), autograd step {s_ag * 1e3:.2f} ms on "
        f"cached plans ({s_ag_first:.3f} s for the first call with its plan builds), {s_ag_moved * 1e3:.2f} ms "
        f"with moved positions ({rebuilt} plans rebuilt), potential errors {errs}, autograd vs plan dq {agree:.2e}")
    result = {
        "metric": "forward + adjoint (dLoss/dq, dLoss/dpos) s/step at N=262K, n=4, 8 vectors",
        "value": 1e3 / ms, "unit": "steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "N=262144 uniform, depth 4, n=4, r=32, 8 charge vectors, loss 1/2||phi - phi_t||^2 "
                               "(phi_t: FP64 direct sum of a second charge set)", "N": N, "vectors": m,
                   "m2l_form": "split-BF16 (IJ), the n = 4 precision form"},
        "ms_forward": ms_f, "ms_autograd_step_cached_plans": s_ag * 1e3,
        "ms_autograd_step_moved_positions": s_ag_moved * 1e3, "autograd_plans_rebuilt_when_moved": rebuilt,
        "s_autograd_first_call_with_plan_builds": s_ag_first,
        "rel_l2_err_potential": errs, "autograd_vs_plan_dq": agree,
        "other_forms": {"plain_fp32": {"ms_per_step": ms_p, "ms_forward": ms_fp, "rel_l2_err_potential": errs_p,
                                       "gradients_vs_split_bf16": {"dq": d_dq, "dx": d_dx}}},
    }
    print(json.dumps(result), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--N", type=int, default=1 << 20)
    ap.add_argument("--nrhs", type=int, default=16)
    ap.add_argument("--targets", type=int, default=4096)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads of the CPU baseline (default: one per physical core available)")
    ap.add_argument("--cpu-reps", type=int, default=5, help="timed CPU baseline solves (median; after 1 warm-up)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one 256 MiB device copy (reference kernel for HBM byte counters)")
    ap.add_argument("--summarize-pmc", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="write the HBM-traffic summary (profiles/pmc_traffic_c2.txt) from the "
                         "pmc_counter_collection.csv files of a FETCH_SIZE and a WRITE_SIZE rocprofv3 pass "
                         "of this bench run with --copy-calibration, then exit")
    ap.add_argument("--pmc-command", default="", help="the profiling command, recorded in the summary")
    ap.add_argument("--config", choices=["c2", "c3", "c4", "c5"], default="c2",
                    help="c2: N=1M n=4 chosen vs plain (default); c3: N=4M n=8 r=256 depth 6, "
                         "16 vectors, FP32 MFMA M2L; c4: 64 uneven instances in one plan and one hipGraph; "
                         "c5: N=262K forward + adjoint (gradients) of 8 vectors")
    args = ap.parse_args()
    if args.summarize_pmc:
        summarize_pmc(args.summarize_pmc[0], args.summarize_pmc[1], TRAFFIC_FILE, args.pmc_command)
        return
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; multi-GPU sharding is a design note only "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if args.config == "c3":
        run_c3(args)
        return
    if args.config == "c4":
        run_c4(args)
        return
    if args.config == "c5":
        run_c5(args)
        return

    import torch
    import bbfmmongpu_amd as B

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.copy_calibration:
        src = torch.ones(64 << 20, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    stream = torch.cuda.current_stream()
    N, nrhs = args.N, args.nrhs
    rng = np.random.default_rng(0)
    pos = rng.random((N, 3)) - 0.5
    Q = rng.uniform(-1.0, 1.0, size=(nrhs, N))
    log(f"N={N} nrhs={nrhs}: operator precompute (FP64)")
    t0 = time.time()
    U, K, sigma = B.operators(4, 32, prec="s")
    t_ops = time.time() - t0

    t0 = time.time()
    plain = B.Plan(pos, 4, U, K, nrhs=1, prec="s", stream=stream.cuda_stream).set(scheme=B.SCHEME_IJ,
                                                                                     form=B.FORM_PLAIN)
    chosen = B.Plan(pos, 4, U, K, nrhs=nrhs, prec="s", stream=stream.cuda_stream).set(
        scheme=B.SCHEME_IJ, form=B.FORM_BF16X2, p2p_form=B.P2P_FP32)
    t_setup = time.time() - t0
    log(f"maxlev={chosen.maxlev} ncell={chosen.ncell} interactions={chosen.m2l_interactions:.0f} "
        f"pairs={chosen.p2p_pairs:.4g} setup={t_setup:.2f}s")

    q1 = torch.from_numpy(Q[0].astype(np.float32)).to(dev)
    phi1 = torch.empty_like(q1)
    qm = torch.from_numpy(Q.astype(np.float32)).to(dev).contiguous()
    phim = torch.empty_like(qm)

    log("plain form: timing")
    ms_plain, st_plain = timed_solves(plain, q1, phi1, args.steps, args.warmup, stream)
    log(f"plain {ms_plain:.3f} ms/solve stages {st_plain}")
    log("chosen form: timing")
    ms_chosen, st_chosen = timed_solves(chosen, qm, phim, args.steps, args.warmup, stream)
    log(f"chosen {ms_chosen:.3f} ms/step ({nrhs} vectors) stages {st_chosen}")

    # accuracy against the FP64 direct sum on sampled targets
    tg = np.sort(np.random.default_rng(1).choice(N, size=min(args.targets, N), replace=False)).astype(np.int32)
    phi_plain = phi1.cpu().numpy()
    phi_chosen = phim.cpu().numpy()
    errs_chosen, errs_pot = [], []
    direct_ref = {}
    nerr = min(nrhs, 4)
    for j in range(nerr):
        d = B.direct(pos, Q[j], tg)
        direct_ref[j] = d
        errs_chosen.append(rel_l2(d, phi_chosen[j, tg]))
        errs_pot.append(rel_l2(d / Q[j, tg], phi_chosen[j, tg] / Q[j, tg]))
    err_plain = rel_l2(direct_ref[0], phi_plain[tg])
    # the algorithm's own truncation error: the fp64 plain form
    plan64 = B.Plan(pos, 4, U.astype(np.float64), K.astype(np.float64), nrhs=1, prec="d").set(
        scheme=B.SCHEME_IJ, form=B.FORM_PLAIN)
    phi64 = plan64.solve_host(Q[0])
    err_fp64 = rel_l2(direct_ref[0], phi64[tg])
    err_chosen_vs_fp64 = rel_l2(phi64, phi_chosen[0])
    plan64.close()
    log(f"errors vs FP64 direct: plain {err_plain:.3e} chosen {errs_chosen} fp64-plain {err_fp64:.3e}")

    # the other arithmetic forms of the 16-vector solve on the same inputs:
    # the chosen M2L with the split-BF16 P2P sum (narrower than the fp32 P2P
    # of the stated configuration), and the FP32 MFMA and FP16 3-term split
    # M2L forms (dense grids, fp32 P2P)
    variants = {}
    for name, form, p2p in (("p2p_bf16x2", B.FORM_BF16X2, B.P2P_BF16X2),
                            ("fp32_mfma", B.FORM_FP32MFMA, B.P2P_FP32), ("fp16x3", B.FORM_FP16X3, B.P2P_FP32)):
        vp = B.Plan(pos, 4, U, K, nrhs=nrhs, prec="s", stream=stream.cuda_stream).set(scheme=B.SCHEME_IJ,
                                                                                        form=form, p2p_form=p2p)
        phiv = torch.empty_like(qm)
        ms_v, st_v = timed_solves(vp, qm, phiv, max(args.steps // 2, 1), 1, stream)
        pv = phiv.cpu().numpy()
        variants[name] = {"ms_per_step": ms_v, "m2l_ms": st_v["m2l"], "p2p_ms": st_v["p2p"],
                          "rel_l2_err_q_phi": [rel_l2(direct_ref[j], pv[j, tg]) for j in range(nerr)],
                          "vs_fp64_plain": rel_l2(phi64, pv[0])}
        vp.close()
        log(f"form {name}: {ms_v:.3f} ms/step, M2L {st_v['m2l']:.3f} ms, errors {variants[name]['rel_l2_err_q_phi']}")

    inter = chosen.m2l_interactions
    pairs = chosen.p2p_pairs
    r = 32
    flop_m2l = r * (2 * r - 1) * inter  # per vector (m2l_aux.c:967-979)
    m2l_gflops_chosen = flop_m2l * nrhs / (st_chosen["m2l"] * 1e-3) / 1e9
    m2l_gflops_plain = flop_m2l / (st_plain["m2l"] * 1e-3) / 1e9
    p2p_gips_chosen = pairs * nrhs / (st_chosen["p2p"] * 1e-3) / 1e9
    p2p_gips_plain = pairs / (st_plain["p2p"] * 1e-3) / 1e9
    s_per_solve_chosen = ms_chosen * 1e-3 / nrhs
    s_per_solve_plain = ms_plain * 1e-3

    # roofline of the dominant kernel of the chosen form
    p2p_flop = pairs * (18 + 2 * nrhs)  # geometry (20-flop convention) shared by the vectors
    p2p_valu_flop = pairs * 18          # the 1/r part alone (the sum over sources runs on MFMA)
    # issued MFMA flops of the source-stationary kernel (deepest level: 4
    # field classes and two tiles per wave, both class groups on the union of
    # their steps; above: 4 classes, one tile): per wave and step it runs,
    # classes x tiles x 6 products of 32x32x16 (zero operator included, ghost
    # steps skipped)
    issued = ss_issued_mfma(chosen.maxlev, nrhs) * (2 * 32 * 32 * 16)
    # the M2L stage is three launches; its roofline takes their sum (time
    # between the stage's events, HBM bytes of the three)
    m2l_kernel = " + ".join(M2L_KERNELS)
    tr_parts = [measured_traffic(k) for k in M2L_KERNELS]
    tr_m2l = sum(t for t, _ in tr_parts) if all(t is not None for t, _ in tr_parts) else None
    tr_m2l_src = {k: src for k, (_, src) in zip(M2L_KERNELS, tr_parts)}
    roof_m2l = {"kernel": m2l_kernel, "bound": "mfma", "unit": "TFLOP/s",
                "achieved": flop_m2l * nrhs / (st_chosen["m2l"] * 1e-3) / 1e12, "peak": PEAK_BF16_TFLOPS,
                "launch_ms": st_chosen["m2l"], "algorithmic_flop_per_launch": flop_m2l * nrhs,
                "issued_mfma_flop_per_launch": issued,
                "issued_mfma_frac": issued / (st_chosen["m2l"] * 1e-3) / (PEAK_BF16_TFLOPS * 1e12),
                "traffic": tr_m2l, "traffic_source": tr_m2l_src}
    roof_m2l["frac"] = roof_m2l["achieved"] / roof_m2l["peak"]
    p2p_kernel = "k_p2p_seg16<true>"
    tr_p2p, tr_p2p_src = measured_traffic(p2p_kernel)
    roof_p2p = {"kernel": p2p_kernel, "bound": "valu", "unit": "TFLOP/s",
                "achieved": p2p_flop / (st_chosen["p2p"] * 1e-3) / 1e12, "peak": PEAK_FP32_TFLOPS,
                "launch_ms": st_chosen["p2p"], "algorithmic_flop_per_launch": p2p_flop,
                "flop_convention": "(18 + 2 * vectors) per pair: 1/r once, one FMA per vector",
                "valu_only_achieved": p2p_valu_flop / (st_chosen["p2p"] * 1e-3) / 1e12,
                "traffic": tr_p2p, "traffic_source": tr_p2p_src}
    roof_p2p["frac"] = roof_p2p["achieved"] / roof_p2p["peak"]
    roof_p2p["valu_only_frac"] = roof_p2p["valu_only_achieved"] / roof_p2p["peak"]
    dominant = roof_m2l if st_chosen["m2l"] >= st_chosen["p2p"] else roof_p2p

    result = {
        "metric": METRIC,
        "value": nrhs / (ms_chosen * 1e-3),
        "unit": "solve/s (per charge vector)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_chosen,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"bbFMM Laplace 1/r, N={N} uniform, n=4, r=32, depth {chosen.maxlev}, "
                               f"{nrhs} charge vectors per solve",
                   "N": N, "n": 4, "r": 32, "maxlev": chosen.maxlev, "nrhs": nrhs,
                   "m2l_scheme": "dense grid (IJ)", "m2l_form": "split-BF16 MFMA (hi+lo, 3 products), fp32 accumulate",
                   "p2p_form": "fp32 (1/r on the vector ALUs, the sum over sources on FP32 MFMA)",
                   "plain_form": "fp32 drop-in, dense-grid M2L in fp32 FMAs, one vector"},
        "s_per_solve": s_per_solve_chosen,
        "m2l_gflops": m2l_gflops_chosen,
        "p2p_ginteractions_per_s": p2p_gips_chosen,
        "speedup_per_vector_vs_plain": s_per_solve_plain / s_per_solve_chosen,
        "rel_l2_err": {"chosen_q_phi": max(errs_chosen), "chosen_potential": max(errs_pot),
                       "plain_q_phi": err_plain, "fp64_plain_q_phi": err_fp64,
                       "chosen_vs_fp64_plain": err_chosen_vs_fp64, "targets": int(tg.size),
                       "vectors_checked": nerr},
        # the metric's bar (rel L2 <= 1e-4 against the FP64 direct sum) and
        # the floor the algorithm itself sets at n = 4, r = 32: the fp64
        # plain FMM's own error
        "err_bar": 1e-4,
        "meets_err_bar": bool(max(errs_pot) <= 1e-4),
        "err_floor_fp64_plain": err_fp64,
        "err_bar_reachable_at_n4": bool(err_fp64 <= 1e-4),
        "plain": {"s_per_solve": s_per_solve_plain, "m2l_gflops": m2l_gflops_plain,
                  "p2p_ginteractions_per_s": p2p_gips_plain, "stage_ms": st_plain},
        "chosen_stage_ms": st_chosen,
        "other_forms": variants,
        "setup_s": {"operators": t_ops, "plans": t_setup},
        "roofline": dominant,
        "roofline_m2l": roof_m2l,
        "roofline_p2p": roof_p2p,
    }

    if not args.no_cpu_baseline:
        threads, cores = host_cores()
        if args.cpu_threads:
            threads = args.cpu_threads
        log(f"cpu baseline: reference CPU path, {threads} threads ({cores}), full N={N} solves of one vector")
        cb = cpu_baseline(pos, Q[0], U, K, threads, reps=args.cpu_reps)
        phi_cpu = cb.pop("phi")
        cb_err = rel_l2(direct_ref[0], phi_cpu[tg])
        result["cpu_baseline"] = {
            "value": 1.0 / cb["s_per_solve"], "unit": "solve/s (per charge vector)", "cores": threads,
            "kind": cb["kind"],
            "sample": f"full solves of vector 0 at N={N} (upward + interaction + downward + near field; "
                      f"tree build excluded) on one tree: 1 warm-up, median of {args.cpu_reps}; "
                      f"OMP_NUM_THREADS={threads} (one per physical core), OMP_PROC_BIND=close, OMP_PLACES=cores",
            "s_per_solve": cb["s_per_solve"], "s_per_solve_all": cb["s_per_solve_all"],
            "m2l_gflops": cb["m2l_gflops"],
            "p2p_ginteractions_per_s": cb["p2p_ginteractions_per_s"], "rel_l2_err_q_phi": cb_err,
            "host": cores,
            "detail": {k: v for k, v in cb.items() if k.startswith("t_") or k == "wall_s"},
        }
    plain.close()
    chosen.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
