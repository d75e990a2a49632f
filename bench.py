"""Throughput of the fictitious-wave-domain FDTD forward step on one MI355X.

Workload (default): 256^3 interior + 16-cell CPML (288^3 = 23,887,872 cells),
FP64, seeded layered + box conductivity, x-dipole Gaussian-derivative source,
256 receivers on a 16x16 surface grid, full-volume running DFT of E at 8
log-spaced frequencies in 0.1-4 Hz, Nt = 8000 samples.  A "step" is one time
step of that shot: E + CPML-E + source, H + CPML-H, receiver gather, and its
share of the volume-DFT flush.  Inputs (model, fields, accumulators) are
resident in HBM before the timed region.

Prints one JSON line:
  value = Gcell-updates/s (cells incl. CPML x steps / wall time of the K steps),
  roofline = dominant kernel's algorithmic bytes / its mean launch time (HIP
             events on the library's stream) against 8 TB/s,
  cpu_baseline = the CPU oracle (reference loop structure) on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gcell-updates/s + % HBM peak (256³ fp64); served shots/s, p50/p99 latency"
HBM_PEAK_GBS = 8000.0
# the launches that make up each timed half-step: the LDS family from 2.5 M
# cells up (kLdsMinCells in fdtd_api.hip), the per-cell family alone below
LDS_MIN_CELLS = 2_500_000
HALF_STEP_KERNELS = {
    "efield": "E half-step: k_efield_f (CPML-shell tiles) + k_efield_l (interior, LDS z-march)",
    "hfield": "H half-step: k_hfield_f (CPML-shell tiles) + k_hfield_l (interior, LDS z-march)",
    "flush": "k_flush (volume DFT, per ring)",
}
HALF_STEP_KERNELS_PER_CELL = {
    "efield": "E half-step: k_efield_f (per-cell tiles over the whole grid)",
    "hfield": "H half-step: k_hfield_f (per-cell tiles over the whole grid)",
    "flush": "k_flush (volume DFT, per ring)",
}
# the fused schedule: shell E, one fused E+H launch over the CPML-free
# interior, shell H
FUSED_KERNELS = {
    "efield": "E shell: k_efield_f (CPML-shell tiles)",
    "ehfield": "k_ehfield_l (fused E+H z-march over the CPML-free interior)",
    "hfield": "H shell: k_hfield_f (CPML-shell tiles)",
    "flush": "k_flush (volume DFT, per ring)",
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def measured_traffic(workload, ring, kernel):
    """HBM bytes per launch of `kernel` from the committed PMC summary
    (FETCH_SIZE + WRITE_SIZE, calibrated against a plain copy kernel), or None
    when there is none for this workload and ring or the kernel sources have
    changed since it was measured."""
    import hashlib
    path = os.path.join(ROOT, "profiles", f"{workload}_pmc_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f)
    src = b"".join(open(os.path.join(ROOT, "mittet_main_amd", "csrc", n), "rb").read()
                   for n in ("fdtd_kernels.hip", "fdtd_dev.h"))
    if rec.get("ring") != ring or rec.get("sources_sha256_16") != hashlib.sha256(src).hexdigest()[:16]:
        return None
    k = rec["kernels"].get(kernel)
    return None if k is None else int(round(k["traffic_bytes"]))


def slab_cells(g):
    nc2 = 2 * g.ncpml
    return g.iz * g.iy * nc2, g.iz * nc2 * g.ix, nc2 * g.iy * g.ix


def region_cells(g):
    """Cells of the CPML-free interior the LDS kernels march: x from ncpml
    rounded up to 16, y and z from ncpml, each up to n - ncpml."""
    xlo = (g.ncpml + 15) // 16 * 16
    return (g.ix - g.ncpml - xlo) * (g.iy - 2 * g.ncpml) * (g.iz - 2 * g.ncpml)


def kernel_bytes(g, nf, ring, fused=False):
    """Algorithmic HBM bytes per launch of each kernel (FP64 = 8 B).  With the
    fused schedule the E and H launches cover the cells outside the interior
    region and the fused launch the region."""
    N = g.ncells
    sx, sy, sz = slab_cells(g)
    shell = sx + sy + sz
    nr = region_cells(g) if fused else 0
    return {
        # read H(3) + E(3) + eps', write E(3); psi-E: 2 arrays read+write per slab cell
        "efield": (N - nr) * (3 * 8 + 3 * 8 + 8 + 3 * 8) + shell * 2 * 16,
        # read E(3), read+write H(3); eps' in the shells; psi-H 2 arrays read+write
        "hfield": (N - nr) * (3 * 8 + 3 * 16) + shell * (8 + 2 * 16),
        # read E(3), H(3), dt/eps', write E(3) and H(3)
        "ehfield": nr * (4 * 3 * 8 + 8),
        # ring slots of E(3) read once + complex accumulators read+write
        "flush": N * (ring * 3 * 8 + 3 * nf * 32),
    }


def step_bytes(g, nf, ring):
    """Algorithmic bytes of one whole step: the per-step byte count of the
    benchmark definition (6 fields read and written, one material word, the
    psi slabs), with the volume DFT charged as the ring moves it: each E slot
    read once (24 B per cell) and the accumulators read and written once per
    ring (3 nf 32 B per cell over `ring` steps)."""
    N = g.ncells
    shell = sum(slab_cells(g))
    return N * (6 * 16 + 8) + shell * 64 + N * 24 + N * 3 * nf * 32 // ring


def host_info():
    """CPU model, sockets, physical cores and logical CPUs of this host
    (from /proc/cpuinfo, as lscpu reports them), and the CPUs this process
    may use."""
    model, phys, logical = "unknown", set(), 0
    try:
        cur = {}
        with open("/proc/cpuinfo") as f:
            for line in f:
                if ":" not in line:
                    if cur:
                        phys.add((cur.get("physical id", "0"), cur.get("core id", str(logical))))
                        logical += 1
                        cur = {}
                    continue
                k, v = (t.strip() for t in line.split(":", 1))
                cur[k] = v
                if k == "model name":
                    model = v
        if cur:
            phys.add((cur.get("physical id", "0"), cur.get("core id", str(logical))))
            logical += 1
    except OSError:
        pass
    sockets = len({p for p, _ in phys}) or 1
    return {"cpu_model": model, "sockets": sockets, "physical_cores": len(phys) or None,
            "logical_cpus": logical or os.cpu_count(),
            "usable_cpus": len(os.sched_getaffinity(0))}


def cpu_baseline(wl, seconds):
    """The oracle (the reference's loop structure, OpenMP C) on the same
    workload, on this host, in about `seconds` of wall time:
      * STREAM triad GB/s with the same threads;
      * three untimed warm-up steps, then three consecutive windows of W steps
        with the running volume DFT and three without, all threads (the
        OpenMP default, OMP_NUM_THREADS on the GPU host), medians reported;
      * one window with the DFT and one without on 1 thread.
    W fills about seconds/8 per window at the rate of warm-up steps 1-2 (at
    least 2 steps), capped so that all windows fit in the workload."""
    import statistics
    import oracle
    g = wl.grid
    host = host_info()
    threads = oracle.num_threads()
    triad = oracle.stream_triad(40_000_000, 5)
    S = oracle.OracleSim(g.ix, g.iy, g.iz, g.ncpml, g.dx, g.dy, g.dz, g.dt, g.omega0, g.cmax)
    S.set_sigma(wl.sigma)
    S.init_cpml()
    S.media_coeff_3d()
    w = oracle.dft_weights(wl.it, g.dt, g.omega0, wl.omegas)
    acc = np.zeros((3, len(wl.omegas), g.ncells), dtype=np.complex128)
    acc.fill(0.0)                          # fault the pages in before any timing
    traces = np.zeros((3, len(wl.receivers), wl.it))
    src1 = tuple(c + 1 for c in wl.src)
    step = [0]

    def window(n, dft):
        t0 = time.perf_counter()
        S.forward(0, wl.signal, src1, n, rec1=wl.receivers + 1, it=wl.it, step0=step[0],
                  weights=w if dft else None, acc=acc if dft else None, traces=traces)
        step[0] += n
        return time.perf_counter() - t0

    left = wl.it - 4                      # steps 3 .. it-2 follow the warm-up steps 0-2
    if left < 8:
        raise ValueError(f"{wl.name}: {wl.it} samples are too few for 8 timed windows")
    window(1, True)
    t_step = window(2, True) / 2.0        # rate probe, part of the warm-up
    W = max(2, int(seconds / 8.0 / max(t_step, 1e-6)))
    one = max(1, int(seconds / 8.0 / max(t_step * threads, 1e-6)))
    one = min(one, W, max(1, left // 8))
    W = max(1, min(W, (left - 2 * one) // 6))
    rate = lambda n, t: g.ncells * n / t / 1e9
    with_dft = [rate(W, window(W, True)) for _ in range(3)]
    plain = [rate(W, window(W, False)) for _ in range(3)]
    oracle.set_num_threads(1)
    try:
        one_dft = rate(one, window(one, True))
        one_plain = rate(one, window(one, False))
    finally:
        oracle.set_num_threads(threads)
    value = statistics.median(with_dft)
    return {"value": round(value, 5), "unit": "Gcell-updates/s", "cores": threads, "kind": "port",
            "sample": f"steps 3-{step[0] - 2 * one - 1} of the {wl.name} workload ({g.ix}^3 cells incl. "
                      f"CPML, {len(wl.receivers)} receivers) in the reference's loop structure: "
                      f"median of 3 windows of {W} steps with the {len(wl.omegas)}-frequency running "
                      f"volume DFT on {threads} threads, after 3 warm-up steps",
            "plain_step_value": round(statistics.median(plain), 5),
            "windows_with_dft": [round(v, 5) for v in with_dft],
            "windows_plain": [round(v, 5) for v in plain],
            "one_thread": {"steps": one, "with_dft": round(one_dft, 5),
                           "plain_step": round(one_plain, 5)},
            "last_step": step[0] - 1,
            "stream_triad_GBps": round(triad, 1), **host}


def serve_mode(args):
    """Served forward modelling: Poisson stream (mean 20 req/s), grids
    {96,128,192,256}^3 + 12-cell CPML, 1-16 shots, Nt = 30 x n, receiver spectra
    at 4 frequencies, 200 GB HBM budget.  Compared with the sequential plain
    path (one shot at a time in arrival order, graph-captured fused loop) on the
    whole stream, and with the reference-named per-step operators on the
    smallest requests."""
    from mittet_main_amd import serving
    reqs = serving.request_stream(args.seed, args.requests, max_shots=args.max_shots,
                                  steps_per_cell=args.steps_per_cell)
    log(f"[serve] {len(reqs)} requests, {sum(r.nshots for r in reqs)} shots, "
        f"(n_int, shots): {[(r.n_int, r.nshots) for r in reqs]}, Nt = {args.steps_per_cell} x n_int")
    done = lambda r: log(f"[serve]   request {r.rid} done, latency {r.latency:.2f} s")
    runs = {}
    for launch in args.launch.split(","):
        srv = serving.Server(budget_bytes=200e9, chunk=32, launch=launch)
        runs[launch] = srv.run(reqs, realtime=True, on_done=done)
        srv.close()
        r = runs[launch]
        log(f"[serve] served ({launch}): {r['shots_per_s']:.3f} shots/s, p50 {r['p50_s']:.2f} s, "
            f"p99 {r['p99_s']:.2f} s")
    served = runs[args.launch.split(",")[0]]
    seq = serving.run_sequential(reqs, mode="fused", realtime=True, on_done=done)
    log(f"[serve] sequential fused: {seq['shots_per_s']:.3f} shots/s, p50 {seq['p50_s']:.2f} s")
    strip = lambda d: {k: (round(v, 4) if isinstance(v, float) else v)
                       for k, v in d.items() if k != "results"}
    out = {"mode": "serve", "metric": "served shots/s, p50/p95/p99 latency",
           "requests": len(reqs), "shots": served["shots"], "seed": args.seed,
           "steps_per_cell": args.steps_per_cell, "launch": args.launch.split(",")[0],
           "served": strip(served), "sequential_fused": strip(seq),
           "throughput_vs_sequential": round(served["shots_per_s"] / seq["shots_per_s"], 3),
           "p50_vs_sequential": round(served["p50_s"] / seq["p50_s"], 3)}
    for launch, r in runs.items():
        out[f"served_{launch}"] = strip(r)
        out[f"throughput_vs_sequential_{launch}"] = round(r["shots_per_s"] / seq["shots_per_s"], 3)
    if args.ops_requests:
        sub = sorted(reqs, key=serving.request_work)[:args.ops_requests]
        ops = serving.run_sequential(sub, mode="ops", realtime=False)
        fus = serving.run_sequential(sub, mode="fused", realtime=False)
        out["per_operator_vs_fused_subset"] = {
            "requests": [(r.n_int, r.nshots) for r in sub],
            "ops_shots_per_s": round(ops["shots_per_s"], 4),
            "fused_shots_per_s": round(fus["shots_per_s"], 4)}
        log(f"[serve] subset: ops {ops['shots_per_s']:.3f} vs fused {fus['shots_per_s']:.3f} shots/s")
    same = all(np.array_equal(a, b) for sv in runs.values() for rid, r in seq["results"].items()
               for a, b in zip(r.spectra, sv["results"][rid].spectra))
    out["served_equals_alone_bitwise"] = bool(same)
    print(json.dumps(out), flush=True)


def merge_mode(args):
    """Merged multi-shot launches against the same shots back to back: k shots
    of one grid (96^3 interior + 12-cell CPML, seed 7, 1024 steps, one receiver,
    no volume DFT) advanced by one batch (every kernel launch covers all k)
    and run one after another (fwi3d_run), for k = 1, 2, 3, 4, 8.  Each
    configuration runs twice; the second run is timed."""
    from mittet_main_amd import Batch, Model, configs
    from mittet_main_amd.fwi3d import Grid, firstderiv_gauss
    n_int, nc, it = 96, 12, 1024
    sig, _ = configs.layered_box_sigma(n_int, nc, 7)
    n = n_int + 2 * nc
    dt, cmax = configs.dt_for(50.0, 50.0, 50.0)
    g = Grid(n, n, n, nc, 50.0, 50.0, 50.0, dt, configs.OMEGA0, cmax)
    signal = firstderiv_gauss(it, dt, configs.FMAX_W)
    rec = np.array([[nc + 5, nc + 5, nc + n_int // 3]], dtype=np.int32)

    def rate(k, merged):
        m = Model(g)
        m.media_coeff_3d(sig)
        src = np.random.default_rng(3).integers(nc + 2, nc + n_int - 3, size=(k, 3))
        shots = [m.shot(it, signal, tuple(int(v) for v in s), 0, rec) for s in src]
        t0 = time.perf_counter()
        if merged:
            b = Batch(k + 1, chunk=32)
            for sh in shots:
                b.add(sh)
            b.step(it)
            b.sync()
            for sh in shots:
                b.remove(sh)
            b.close()
        else:
            for sh in shots:
                sh.run()
                sh.sync()
        t = time.perf_counter() - t0
        m.close()
        return g.ncells * (it - 1) * k / t / 1e9

    rows = []
    for k in (1, 2, 3, 4, 8):
        alone = [rate(k, False) for _ in range(2)][-1]
        merged = [rate(k, True) for _ in range(2)][-1]
        rows.append({"shots": k, "fields_mb": round(k * g.ncells * 48 / 1e6), "alone_gcell_s": round(alone, 3),
                     "merged_gcell_s": round(merged, 3), "merged_over_alone": round(merged / alone, 3)})
        log(f"[merge] {k} shots of {n}^3 ({rows[-1]['fields_mb']} MB of E and H): alone {alone:.2f}, "
            f"merged {merged:.2f} Gcell-updates/s")
    print(json.dumps({"mode": "merge", "grid": n, "steps": it - 1, "rows": rows}), flush=True)


def serve_fwi_mode(args):
    """Served FWI gradients with checkpointing, bounded: 384^3 interior +
    16-cell CPML (416^3 cells), a seeded Poisson stream of
    forward (70 %) and gradient (30 %) requests with 1-2 x-dipole shots each,
    8 frequencies, 200 GB HBM budget; Nt = --fwi-it (default 1000) instead of
    a full-length run, so the stream finishes in minutes.  At this size the
    sqrt(Nt) checkpoint plan (about 66 states of 4 GB) does not fit the
    budget, so each gradient recomputes recursively under it.  Observed data of
    the gradient requests are computed before the stream starts.  Compared
    with the same requests run one after another."""
    from mittet_main_amd import serving
    reqs = serving.fwi_stream(args.seed, args.requests, n_int=args.fwi_interior,
                              it=args.fwi_it, max_shots=args.max_shots)
    log(f"[fwi] {len(reqs)} requests on {args.fwi_interior + 32}^3 cells, it={args.fwi_it}: "
        f"{[(r.kind, r.nshots) for r in reqs]}")
    t0 = time.perf_counter()
    for r in reqs:
        if r.kind == "gradient":
            r.d_obs = serving.observed_data(r)
    log(f"[fwi] observed data for the gradient requests: {time.perf_counter() - t0:.1f} s")
    done = lambda r: log(f"[fwi]   request {r.rid} done, latency {r.latency:.2f} s"
                         + (f", levels {r.levels}" if r.grad is not None else ""))
    srv = serving.Server(budget_bytes=200e9, chunk=32)
    served = srv.run(reqs, realtime=True, on_done=done)
    srv.close()
    strip = lambda d: {k: (round(v, 4) if isinstance(v, float) else v)
                       for k, v in d.items() if k != "results"}
    gshots = sum(r.nshots for r in reqs if r.kind == "gradient")
    out = {"mode": "serve-fwi", "metric": "gradients/s, p99 latency, peak HBM (416^3 served FWI)",
           "workload": f"fwi_stream_{args.fwi_interior}cube_fp64", "seed": args.seed,
           "cells": (args.fwi_interior + 32) ** 3, "nt": args.fwi_it, "budget_gb": 200,
           "requests": len(reqs), "gradient_shots": gshots,
           "gradients_per_s": round(gshots / served["makespan_s"], 5),
           "levels": {rid: r.levels for rid, r in served["results"].items() if r.grad is not None},
           "served": strip(served)}
    log(f"[fwi] served: makespan {served['makespan_s']:.1f} s, peak {served['peak_hbm_gb']:.1f} GB")
    if not args.no_sequential:
        seq = serving.run_sequential(reqs, mode="fused", realtime=True, on_done=done)
        out["sequential"] = strip(seq)
        out["makespan_vs_sequential"] = round(served["makespan_s"] / seq["makespan_s"], 3)
        same = all((np.array_equal(r.grad, served["results"][rid].grad) if r.grad is not None else
                    all(np.array_equal(a, b) for a, b in zip(r.spectra, served["results"][rid].spectra)))
                   for rid, r in seq["results"].items())
        out["served_equals_alone_bitwise"] = bool(same)
    print(json.dumps(out, default=float), flush=True)


def gradient_mode(args):
    """Adjoint gradient: 128^3 interior + 12-cell CPML, 8 x-dipole shots, 64
    receivers, L2 misfit on complex receiver E at 4 frequencies, Nt = 30 x 128;
    observed data from the true model, gradient at the smoothed start model."""
    from mittet_main_amd import configs
    from mittet_main_amd.adjoint import GradientProblem, ShotSpec
    wl = configs.gradient_128(it=args.grad_nt)
    g = wl.grid
    shots = [ShotSpec(tuple(s), wl.signal) for s in wl.sources]
    P = GradientProblem(g, shots, wl.receivers, wl.it, wl.omegas, segment=args.segment)
    log(f"[grad] {wl.name}: {g.ix}^3 cells, it={wl.it}, {len(shots)} shots, "
        f"{len(wl.receivers)} receivers, {len(wl.omegas)} frequencies")
    t0 = time.perf_counter()
    P.d_obs = P.forward_data(wl.sigma_true)
    t_fwd = time.perf_counter() - t0
    log(f"[grad] observed data: {t_fwd:.2f} s")
    logs0 = np.log(wl.sigma_start)
    P.value_and_grad(logs0)                       # warm-up (allocations, graphs)
    t0 = time.perf_counter()
    J, grad = P.value_and_grad(logs0)
    t_grad = time.perf_counter() - t0
    log(f"[grad] gradient: {t_grad:.2f} s")
    cell_steps = g.ncells * (wl.it - 1) * len(shots)
    out = {"mode": "gradient", "workload": wl.name, "cells": g.ncells, "nt": wl.it,
           "shots": len(shots), "nrec": len(wl.receivers), "nf": len(wl.omegas),
           "forward_all_shots_s": round(t_fwd, 4), "gradient_s": round(t_grad, 4),
           "gradients_per_s": round(1.0 / t_grad, 5),
           "gradient_over_forward": round(t_grad / t_fwd, 3),
           "forward_gcell_updates_per_s": round(cell_steps / t_fwd / 1e9, 3),
           "misfit": J, "grad_norm": float(np.linalg.norm(grad)),
           "checkpoint_storage_gb": round(P.storage_bytes() / 1e9, 3), "segment": args.segment,
           "dtype": "f64", "data": "synthetic"}
    P.close()
    print(json.dumps(out), flush=True)


def forward_step_counts(args, workload_steps):
    """Fills the forward mode's unset step counts (0) with the workload's
    defaults: 2048 timed steps and a 128-slot ring at 288^3, 1024 and 64 on the
    84^3 workload of 1999 steps.  Returns an error message when warm-up, timed
    and profiled steps together exceed the workload, else None."""
    small = args.workload == "plain_64"
    if args.steps <= 0:
        args.steps = 1024 if small else 2048
    if args.ring <= 0:
        args.ring = 64 if small else 128
    if args.profile_steps <= 0:
        args.profile_steps = args.ring
    need = args.warmup + args.steps + args.profile_steps
    if need > workload_steps:
        return (f"--warmup + --steps + --profile-steps = {need} exceeds the {workload_steps} "
                "steps of the workload")
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mode", default="forward",
                    choices=["forward", "serve", "gradient", "serve-fwi", "merge"])
    ap.add_argument("--seed", type=int, default=11)
    ap.add_argument("--requests", type=int, default=4)
    ap.add_argument("--max-shots", type=int, default=16)
    ap.add_argument("--ops-requests", type=int, default=1)
    ap.add_argument("--steps-per-cell", type=int, default=30,
                    help="serve: Nt = this x interior size (30 as the served workload defines it)")
    ap.add_argument("--launch", default="cache",
                    help="serve: server launch modes to run, comma-separated (cache, interleave, merged)")
    ap.add_argument("--grad-nt", type=int, default=0, help="time steps (default 30 x 128)")
    ap.add_argument("--segment", type=int, default=0)
    ap.add_argument("--fwi-it", type=int, default=1000, help="serve-fwi: time samples per shot")
    ap.add_argument("--fwi-interior", type=int, default=384)
    ap.add_argument("--no-sequential", action="store_true")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0, help="timed steps (default 2048; 1024 for plain_64)")
    ap.add_argument("--warmup", type=int, default=128)
    ap.add_argument("--workload", default="roofline_256", choices=["roofline_256", "plain_64"])
    ap.add_argument("--ring", type=int, default=0, help="DFT ring slots (default 128; 64 for plain_64)")
    ap.add_argument("--zchunk", type=int, default=0)
    ap.add_argument("--volume-dft-h", action="store_true",
                    help="forward: also sum H over the volume (H ring, fwi3d_enable_volume_dft_h)")
    ap.add_argument("--profile-steps", type=int, default=0,
                    help="steps timed per kernel after the timed region (default: one ring)")
    ap.add_argument("--cpu-seconds", type=float, default=24.0,
                    help="wall time of the CPU-oracle sample (8 windows of whole steps)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run the plain copy kernel over MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        import ctypes as C
        from mittet_main_amd import _lib
        ms = C.c_double(0.0)
        nbytes = args.copy_calibration << 20
        _lib.check(_lib.load().fwi3d_diag_copy(nbytes, 5, C.byref(ms)))
        print(json.dumps({"copy_bytes": nbytes, "reps": 5, "ms": ms.value,
                          "GBps": 2 * nbytes / (ms.value * 1e-3) / 1e9}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(shots shard across GPUs by construction)", file=sys.stderr)
        sys.exit(2)
    if args.mode == "serve":
        return serve_mode(args)
    if args.mode == "serve-fwi":
        return serve_fwi_mode(args)
    if args.mode == "merge":
        return merge_mode(args)
    if args.mode == "gradient":
        return gradient_mode(args)

    from mittet_main_amd import Model, configs

    wl = configs.roofline_256() if args.workload == "roofline_256" else configs.plain_64()
    g = wl.grid
    nf = len(wl.omegas)
    err = forward_step_counts(args, wl.steps)
    if err:
        sys.exit(err)
    log(f"[bench] workload {wl.name}: {g.ix}x{g.iy}x{g.iz} cells, it={wl.it}, nf={nf}, "
        f"nrec={len(wl.receivers)}, ring={args.ring}")
    model = Model(g)
    model.media_coeff_3d(wl.sigma)
    shot = model.shot(wl.it, wl.signal, wl.src, wl.src_comp, wl.receivers, wl.omegas,
                      ring=args.ring, zchunk=args.zchunk)
    if args.volume_dft_h:
        shot.enable_volume_dft_h()
    shot.step_n(args.warmup)
    shot.sync()
    log(f"[bench] warmup {args.warmup} steps done")
    t0 = time.perf_counter()
    shot.step_n(args.steps)
    shot.sync()
    t1 = time.perf_counter()
    wall = t1 - t0
    log(f"[bench] timed {args.steps} steps: {wall:.3f} s")
    prof = shot.profile_steps(args.profile_steps)
    log(f"[bench] per-kernel ms: {prof}")

    # the profiled flush covers min(ring, profile-steps) slots
    flushed = min(args.ring, args.profile_steps)
    fused = prof["fused"] == 1.0
    kb = kernel_bytes(g, nf, flushed, fused)
    per_step = {"efield": prof["efield"], "hfield": prof["hfield"],
                "flush": prof["flush"] / flushed}
    if fused:
        per_step["ehfield"] = prof["ehfield"]
    dom = max(per_step, key=per_step.get)
    dom_ms = prof[dom]
    achieved = kb[dom] / (dom_ms * 1e-3) / 1e9
    ms_step = wall / args.steps * 1e3
    value = g.ncells * args.steps / wall / 1e9
    sb = step_bytes(g, nf, args.ring)
    result = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": wl.name, "interior": g.ix - 2 * g.ncpml, "ncpml": g.ncpml,
                   "cells": g.ncells, "nt": wl.it, "nf_vol": nf, "nrec": len(wl.receivers),
                   "ring": args.ring, "zchunk": args.zchunk, "volume_dft_h": args.volume_dft_h,
                   "sigma_sha256_16": configs.sigma_digest(wl.sigma)},
        "roofline": {"bound": "hbm",
                     "kernel": (FUSED_KERNELS if fused else
                                HALF_STEP_KERNELS if g.ncells >= LDS_MIN_CELLS
                                else HALF_STEP_KERNELS_PER_CELL).get(dom, dom),
                     "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(wl.name, args.ring, dom),
                     "bytes_per_launch": kb[dom], "ms_per_launch": round(dom_ms, 5)},
        # the whole timed step against the same peak, with the step_bytes model
        "step_roofline": {"bytes_per_step": sb, "achieved": round(sb / (ms_step * 1e-3) / 1e9, 1),
                          "frac": round(sb / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
        "kernel_ms": {k: round(v, 5) for k, v in prof.items()},
        "interior_gcells": round((g.ix - 2 * g.ncpml) ** 3 * args.steps / wall / 1e9, 4),
    }
    shot.close()
    model.close()
    if not args.no_cpu_baseline:
        log(f"[bench] cpu baseline: about {args.cpu_seconds:.0f} s of oracle steps")
        result["cpu_baseline"] = cpu_baseline(wl, args.cpu_seconds)
        result["gpu_over_cpu"] = round(value / result["cpu_baseline"]["value"], 1)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
