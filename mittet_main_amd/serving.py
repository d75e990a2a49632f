"""Serving: continuous batching of forward-modelling shots, and gradient
requests served beside them.

A request carries a conductivity model (seeded layered + box generator), a
grid size, 1-16 shots (x-dipoles at random positions) and asks for receiver
spectra at a few frequencies.  The server

* admits a request when the HBM its model and shots need fits under the budget
  (``fwi3d_shot_bytes`` + model bytes), otherwise it waits in the queue;
* advances every admitted shot `chunk` steps per round, so requests that
  arrive while others run join as soon as memory allows and overlap with them.
  Launch modes: "merged" puts all live shots in ONE persistent multi-shot
  batch (``fwi3d_batch_*``) whose every kernel launch advances all members,
  whatever their grid size; "interleave" launches each live shot's own
  graph-captured chunk in turn, one shot at a time on the device; "cache"
  (default) merges shots into groups whose E and H fit together in
  `cache_bytes` and launches the groups in turn.  On MI355X a shot of up to
  ~150^3 cells keeps its fields in the 256 MB Infinity Cache across the
  steps of its chunk: a merged launch pays while its members fit there too
  (two 120^3 shots, 166 MB: 1.07x the same shots back to back) and loses as
  soon as they do not (three and more: 0.80x; 152^3 pairs: 0.89x);
* retires each shot when its own time loop ends, frees its memory, and
  completes the request when all its shots are done.

Each shot's result is bit-identical to the same shot run alone (per-shot
descriptors, no shared accumulators, no data-dependent reduction order).

A gradient request (kind "gradient") brings observed receiver spectra for
its shots and asks for the misfit and dJ/dlog(sigma) at its starting model
(the smoothed background of its seed).  Gradient requests run one at a time in
a worker thread, on their own model's HIP stream, while the main loop keeps
advancing the forward batch, so their kernels overlap on the device.  A
gradient request is admitted when its model, one shot and the checkpoint
storage fit in the budget left; the storage plan is the fewest recursion
levels (fwi3d_grad_create_levels) that fit, so a large grid recomputes more
instead of waiting.  The gradient does not depend on the plan (bitwise).

Latency = completion time - arrival time (wall clock, the request stream is
replayed in real time); throughput = shots / makespan.
"""
from __future__ import annotations

import math
import os
import threading
import time
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional

import numpy as np

from . import _lib, configs
from .adjoint import dft_weights, receiver_spectra, shot_value_and_grad
from .fwi3d import Batch, Grid, Model, Shot, firstderiv_gauss, shot_bytes

MODEL_BYTES_PER_CELL = 8 * 3      # sigma, eps', dt/eps'


@dataclass
class Request:
    rid: int
    arrival: float                  # s after stream start
    n_int: int                      # interior cells per axis
    ncpml: int
    it: int
    seed: int
    sources: np.ndarray             # (nshots, 3) 0-based cells
    receivers: np.ndarray           # (nrec, 3)
    freqs_hz: List[float]
    kind: str = "forward"           # "forward" or "gradient"
    d_obs: Optional[List[np.ndarray]] = None   # gradient: per shot (3, nrec, nf) complex

    @property
    def nshots(self) -> int:
        return len(self.sources)


@dataclass
class Result:
    rid: int
    spectra: List[np.ndarray]       # forward: per shot (6, nrec, nf) complex
    arrival: float
    done: float
    misfit: Optional[float] = None  # gradient requests
    grad: Optional[np.ndarray] = None   # dJ/dlog(sigma), [iz, iy, ix]
    levels: int = 0                 # checkpoint recursion levels used

    @property
    def latency(self) -> float:
        return self.done - self.arrival


def request_stream(seed: int, n_requests: int, rate: float = 20.0,
                   sizes=(96, 128, 192, 256), ncpml: int = 12, max_shots: int = 16,
                   nf: int = 4, nrec_side: int = 8, steps_per_cell: int = 30) -> List[Request]:
    """Poisson arrivals (mean `rate`/s); grid size uniform over `sizes`;
    1..max_shots shots at random interior positions; Nt = steps_per_cell x n_int."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    freqs = list(np.geomspace(0.25, 2.0, nf))
    for rid in range(n_requests):
        t += rng.exponential(1.0 / rate)
        n_int = int(rng.choice(sizes))
        n = n_int + 2 * ncpml
        ns = int(rng.integers(1, max_shots + 1))
        lo, hi = ncpml + 2, ncpml + n_int - 3
        src = rng.integers(lo, hi + 1, size=(ns, 3)).astype(np.int32)
        g = np.linspace(lo, hi, nrec_side).round().astype(int)
        X, Y = np.meshgrid(g, g, indexing="xy")
        k_top = ncpml + n_int // 3
        rec = np.stack([X.ravel(), Y.ravel(), np.full(X.size, k_top)], axis=1).astype(np.int32)
        out.append(Request(rid, t, n_int, ncpml, steps_per_cell * n_int, int(rng.integers(1 << 30)),
                           src, rec, freqs))
    return out


def fwi_stream(seed: int, n_requests: int, n_int: int = 384, ncpml: int = 16,
               rate: float = 20.0, grad_frac: float = 0.3, max_shots: int = 2, nf: int = 8,
               it: int = 1000, nrec_side: int = 8) -> List[Request]:
    """Mixed stream of forward and gradient requests on one grid size:
    Poisson arrivals (mean `rate`/s), each request a gradient with probability
    `grad_frac`, 1..max_shots x-dipole shots at random interior positions,
    receivers on an nrec_side^2 grid on the top interface, `nf` frequencies
    log-spaced in 0.1-4 Hz.  Gradient requests still need observed data
    (observed_data) before they are served."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    freqs = list(np.geomspace(0.1, 4.0, nf))
    lo, hi = ncpml + 2, ncpml + n_int - 3
    g = np.linspace(lo, hi, nrec_side).round().astype(int)
    X, Y = np.meshgrid(g, g, indexing="xy")
    k_top = ncpml + n_int // 3
    rec = np.stack([X.ravel(), Y.ravel(), np.full(X.size, k_top)], axis=1).astype(np.int32)
    for rid in range(n_requests):
        t += rng.exponential(1.0 / rate)
        kind = "gradient" if rng.random() < grad_frac else "forward"
        ns = int(rng.integers(1, max_shots + 1))
        src = rng.integers(lo, hi + 1, size=(ns, 3)).astype(np.int32)
        out.append(Request(rid, t, n_int, ncpml, it, int(rng.integers(1 << 30)), src, rec.copy(),
                           freqs, kind=kind))
    return out


def build_model(req: Request):
    sig, _ = configs.layered_box_sigma(req.n_int, req.ncpml, req.seed)
    n = req.n_int + 2 * req.ncpml
    dt, cmax = configs.dt_for(50.0, 50.0, 50.0)
    grid = Grid(n, n, n, req.ncpml, 50.0, 50.0, 50.0, dt, configs.OMEGA0, cmax)
    return grid, sig


def start_model(req: Request) -> np.ndarray:
    """Starting model of a gradient request: the smoothed three-layer
    background of its seed."""
    return configs.smoothed_background(req.n_int, req.ncpml, req.seed)


def observed_data(req: Request) -> List[np.ndarray]:
    """The receiver E spectra a gradient request brings: its shots run on the
    true model of its seed (layered + boxes)."""
    grid, sig = build_model(req)
    model = Model(grid)
    try:
        model.media_coeff_3d(sig)
        w = dft_weights(req.it, grid.dt, grid.omega0, 2 * np.pi * np.asarray(req.freqs_hz))
        signal = firstderiv_gauss(req.it, grid.dt, configs.FMAX_W)
        out = []
        for src in req.sources:
            sh = model.shot(req.it, signal, tuple(int(v) for v in src), 0, req.receivers)
            sh.run()
            out.append(receiver_spectra(sh.traces()[:3], w))
            sh.close()
        return out
    finally:
        model.close()


def gradient_bytes(req: Request, budget: float):
    """(levels, bytes) of the cheapest checkpoint plan whose model + shot +
    storage fits `budget`, or None."""
    grid, _ = build_model(req)
    fixed = grid.ncells * MODEL_BYTES_PER_CELL + shot_bytes(grid, req.it, nrec=len(req.receivers))
    nt = req.it - 1
    for lv in range(1, 9):
        need = fixed + plan_bytes(grid, nt, len(req.receivers), lv)
        if need <= budget:
            return lv, need
    return None


def _adjoint_tile_doubles(grid: Grid) -> int:
    """Doubles held by the adjoint's per-cell tile list: 16-byte entries for
    the 32 x 8 tiles (16 x 16 over the x edges) around its LDS region
    [ncpml + 1, n - ncpml - 1) (x start rounded up to 8); none without that
    region (kappa != 1, ncpml < 2, LDS kernels switched off)."""
    nc = grid.ncpml
    n = (grid.ix, grid.iy, grid.iz)
    if nc < 2 or grid.kappa_max != 1.0 or int(os.environ.get("MITTET_FDTD_NOLDS", "0") or 0):
        return 0
    lo = [(nc + 1 + 7) // 8 * 8, nc + 1, nc + 1]
    hi = [v - nc - 1 for v in n]
    if any(a >= b for a, b in zip(lo, hi)):
        return 0
    ftx, fty = -(-grid.ix // 32), -(-grid.iy // 8)
    ya, yb = -(-lo[1] // 8), hi[1] // 8
    inside = max(0, yb - ya)
    edge = -(-lo[0] // 16) + -(-(grid.ix - hi[0]) // 16)
    per_in = (fty - inside) * ftx + (-(-inside * 8 // 16) * edge if inside else 0)
    count = (hi[2] - lo[2]) * per_in + (grid.iz - (hi[2] - lo[2])) * fty * ftx
    return 2 * (-(-count // 8) * 8)


def plan_bytes(grid: Grid, nt: int, nrec: int, levels: int) -> int:
    """Device bytes of a gradient's checkpoint storage (host-side mirror of
    fwi3d_grad_plan_bytes: states of 6 fields + shell psi, each padded to
    256 bytes, scratch, gradient, the adjoint's tile list)."""
    import ctypes as C
    sl = (C.c_int * levels)()
    ns, st = C.c_int(0), C.c_longlong(0)
    _lib.check(_lib.load().fwi3d_grad_plan(nt, levels, sl, C.byref(ns), C.byref(st)))
    n, nc = grid.ncells, grid.ncpml
    psi = 4 * 2 * nc * (grid.iy * grid.iz + grid.ix * grid.iz + grid.ix * grid.iy)
    state = 6 * n + psi
    other = ((state - 3 * n) + n + 1 + max(1, 3 * nrec * (nt + 1)) +
             2 * (grid.ix + grid.iy + grid.iz) + _adjoint_tile_doubles(grid))
    return 8 * (st.value * ((state + 31) // 32 * 32) + other)


def run_gradient(req: Request, budget: float) -> Result:
    """A gradient request, one shot at a time on its own model: misfit and
    dJ/dlog(sigma) at the starting model against req.d_obs."""
    fit = gradient_bytes(req, budget)
    if fit is None:
        raise ValueError(f"request {req.rid}: no checkpoint plan fits {budget / 1e9:.1f} GB")
    levels = fit[0]
    grid, _ = build_model(req)
    sig0 = start_model(req)
    model = Model(grid)
    try:
        model.media_coeff_3d(sig0)
        w = dft_weights(req.it, grid.dt, grid.omega0, 2 * np.pi * np.asarray(req.freqs_hz))
        signal = firstderiv_gauss(req.it, grid.dt, configs.FMAX_W)
        total, terms = np.zeros(grid.ncells), []
        for src, obs in zip(req.sources, req.d_obs):
            sh = model.shot(req.it, signal, tuple(int(v) for v in src), 0, req.receivers)
            try:
                J, dj, _ = shot_value_and_grad(sh, obs, w, levels)
            finally:
                sh.close()
            terms.append(J)
            total += dj
    finally:
        model.close()
    g = (total * sig0.reshape(-1)).reshape(sig0.shape)
    return Result(req.rid, [], req.arrival, 0.0, misfit=math.fsum(terms), grad=g, levels=levels)


def request_bytes(req: Request, grid: Grid) -> int:
    per_shot = shot_bytes(grid, req.it, nrec=len(req.receivers))
    return grid.ncells * MODEL_BYTES_PER_CELL + req.nshots * per_shot


def shot_result(shot: Shot, req: Request) -> np.ndarray:
    om = 2 * np.pi * np.asarray(req.freqs_hz)
    return shot.receiver_dft(om, h_factor=True)


def request_work(req: Request) -> int:
    """Cell updates the request needs: cells (incl. CPML) x steps x shots
    (a gradient shot counts as the forward plus its recomputation and adjoint,
    about 6 forward runs)."""
    w = (req.n_int + 2 * req.ncpml) ** 3 * (req.it - 1) * req.nshots
    return 6 * w if req.kind == "gradient" else w


class Server:
    """Continuous-batching server: shots in flight advance a chunk per round.

    Scheduling: requests that have arrived wait in a shortest-work-first queue
    (cell updates, ties in arrival order).  Shots are admitted one by one while
    the cells in flight stay under `target_cells` (enough to fill every CU;
    beyond that, more shots in flight only share the GPU and raise every
    request's latency) and the HBM in use stays under `budget_bytes` (model
    bytes while any of its shots is unfinished, plus each admitted shot's
    bytes).  A request started earlier keeps priority for its remaining shots,
    so requests complete one after another instead of all at the end.  The batch
    advances `chunk` steps per round (one graph launch per half-step for every
    member); finished shots leave, their memory is released, and waiting shots
    join before the next round."""

    def __init__(self, budget_bytes: float = 200e9, chunk: int = 32, max_shots: int = 512,
                 target_cells: float = 24e6, policy: str = "swf", launch: str = "cache",
                 cache_bytes: float = 200e6):
        if policy not in ("swf", "fifo"):
            raise ValueError("policy must be 'swf' (shortest work first) or 'fifo'")
        if launch not in ("cache", "interleave", "merged"):
            raise ValueError("launch must be 'cache', 'interleave' or 'merged'")
        self.budget = float(budget_bytes)
        self.chunk = chunk
        self.target_cells = float(target_cells)
        self.policy = policy
        self.launch = launch
        self.max_shots = max_shots
        # group cap in E and H bytes: unbounded when merged, 0 (one shot per
        # group) when interleaved
        self.cache_bytes = {"merged": math.inf, "interleave": 0.0}.get(launch, float(cache_bytes))
        self.groups: List[dict] = []          # {"batch", "bytes", "shots"}, launched in turn
        self.live: List[Shot] = []            # shots in flight, in admission order
        self.in_use = 0
        self.peak = 0
        self.admissions = []                  # (rid, time its first shot joined)

    def close(self):
        for g in self.groups:
            g["batch"].close()
        self.groups = []

    def _add(self, sh: Shot):
        fb = 48 * sh.model.grid.ncells           # E and H of the shot
        g = next((g for g in self.groups if g["bytes"] + fb <= self.cache_bytes), None)
        if g is None:
            g = {"batch": Batch(self.max_shots, chunk=self.chunk), "bytes": 0, "shots": []}
            self.groups.append(g)
        g["batch"].add(sh)
        g["bytes"] += fb
        g["shots"].append(sh)
        self.live.append(sh)

    def _remove(self, sh: Shot):
        g = next(g for g in self.groups if sh in g["shots"])
        g["batch"].remove(sh)
        g["shots"].remove(sh)
        g["bytes"] -= 48 * sh.model.grid.ncells
        if not g["shots"]:
            g["batch"].close()
            self.groups.remove(g)
        self.live.remove(sh)

    def _advance(self):
        """Every live shot `chunk` steps further: each group's multi-shot
        launch sequence in turn (one group when merged, one shot per group
        when interleaved)."""
        for g in self.groups:
            g["batch"].step(self.chunk)
            g["batch"].sync()

    def _start(self, req: Request, now: float) -> dict:
        grid, sig = build_model(req)
        model = Model(grid)
        model.media_coeff_3d(sig)
        self.in_use += grid.ncells * MODEL_BYTES_PER_CELL
        self.admissions.append((req.rid, now))
        return {"req": req, "grid": grid, "model": model,
                "signal": firstderiv_gauss(req.it, grid.dt, configs.FMAX_W),
                "shot_bytes": shot_bytes(grid, req.it, nrec=len(req.receivers)),
                "next": 0, "live": {}, "spectra": [None] * req.nshots, "left": req.nshots}

    def _start_gradient(self, req: Request, now: float) -> Optional[dict]:
        fit = gradient_bytes(req, self.budget - self.in_use)
        if fit is None:
            return None
        need = fit[1]
        self.in_use += need
        self.peak = max(self.peak, self.in_use)
        self.admissions.append((req.rid, now))
        job = {"req": req, "need": need, "result": None, "error": None,
               "budget": self.budget - self.in_use + need}

        def work():
            try:
                job["result"] = run_gradient(req, job["budget"])
            except BaseException as e:          # re-raised on the main thread
                job["error"] = e

        job["thread"] = threading.Thread(target=work, name=f"gradient-{req.rid}", daemon=True)
        job["thread"].start()
        return job

    def run(self, requests: List[Request], realtime: bool = True,
            on_done: Optional[Callable[[Result], None]] = None) -> Dict:
        for req in requests:
            grid, _ = build_model(req)
            if req.kind == "gradient":
                if req.d_obs is None or len(req.d_obs) != req.nshots:
                    raise ValueError(f"gradient request {req.rid} needs observed data per shot")
                if gradient_bytes(req, self.budget) is None:
                    raise ValueError(f"gradient request {req.rid}: no checkpoint plan fits the budget")
                continue
            one = grid.ncells * MODEL_BYTES_PER_CELL + shot_bytes(grid, req.it, nrec=len(req.receivers))
            if one > self.budget:
                raise ValueError(f"request {req.rid} needs {one / 1e9:.1f} GB for one shot > budget")
        incoming = sorted(requests, key=lambda r: (r.arrival, r.rid))
        ready: List[Request] = []
        gready: List[Request] = []            # gradient requests waiting
        job: Optional[dict] = None            # the gradient request in flight
        started: List[dict] = []              # in start order
        results: List[Result] = []
        cells = 0
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0

        def finish(res: Result):
            results.append(res)
            if on_done:
                on_done(res)

        while incoming or ready or started or gready or job:
            while incoming and (not realtime or incoming[0].arrival <= now()):
                r = incoming.pop(0)
                (gready if r.kind == "gradient" else ready).append(r)
            if self.policy == "swf":
                ready.sort(key=lambda r: (request_work(r), r.arrival, r.rid))
                gready.sort(key=lambda r: (request_work(r), r.arrival, r.rid))
            # a finished gradient request leaves; the next one starts if it fits
            if job and not job["thread"].is_alive():
                job["thread"].join()
                self.in_use -= job["need"]
                if job["error"] is not None:
                    raise job["error"]
                res = job["result"]
                res.done = now()
                finish(res)
                job = None
            if job is None and gready:
                job = self._start_gradient(gready[0], now())
                if job is not None:
                    gready.pop(0)
            # admit shots: unfinished started requests first, then the queue head
            while True:
                nb = len(self.live)
                a = next((a for a in started if a["next"] < a["req"].nshots), None)
                if a is None:
                    if not ready:
                        break
                    req = ready[0]
                    grid, _ = build_model(req)
                    need = grid.ncells * MODEL_BYTES_PER_CELL + \
                        shot_bytes(grid, req.it, nrec=len(req.receivers))
                    if nb and (cells >= self.target_cells or self.in_use + need > self.budget):
                        break
                    if self.in_use + need > self.budget:
                        break                 # a gradient holds the memory; wait for it
                    ready.pop(0)
                    a = self._start(req, now())
                    started.append(a)
                gcells = a["grid"].ncells
                if nb and (cells + gcells > self.target_cells or
                           self.in_use + a["shot_bytes"] > self.budget):
                    break
                if self.in_use + a["shot_bytes"] > self.budget:
                    break                     # only the model fits now; wait for memory
                q = a["next"]
                sh = a["model"].shot(a["req"].it, a["signal"], tuple(int(v) for v in a["req"].sources[q]),
                                     0, a["req"].receivers)
                self._add(sh)
                a["live"][q] = sh
                a["next"] += 1
                cells += gcells
                self.in_use += a["shot_bytes"]
                self.peak = max(self.peak, self.in_use)
            if not self.live:
                if job is not None or incoming:
                    # wait for the gradient in flight or the next arrival
                    if job is not None:
                        job["thread"].join(timeout=0.005)
                    else:
                        time.sleep(max(0.0, min(incoming[0].arrival - now(), 0.01)))
                    continue
                if not ready and not started and not gready:
                    break
                raise RuntimeError("no request fits the HBM budget while nothing runs")
            self._advance()
            for a in list(started):
                for q, sh in list(a["live"].items()):
                    if sh.step >= sh.it - 1:
                        self._remove(sh)
                        a["spectra"][q] = shot_result(sh, a["req"])
                        sh.close()
                        del a["live"][q]
                        a["left"] -= 1
                        cells -= a["grid"].ncells
                        self.in_use -= a["shot_bytes"]
                if a["left"] == 0:
                    a["model"].close()
                    self.in_use -= a["grid"].ncells * MODEL_BYTES_PER_CELL
                    started.remove(a)
                    finish(Result(a["req"].rid, a["spectra"], a["req"].arrival, now()))
        makespan = now()
        return summarize(results, requests, makespan, peak=self.peak)


def _percentiles(lat) -> Dict:
    lat = np.asarray(lat) if len(lat) else np.zeros(1)
    return {"mean_s": float(lat.mean()), "p50_s": float(np.percentile(lat, 50)),
            "p95_s": float(np.percentile(lat, 95)), "p99_s": float(np.percentile(lat, 99))}


def summarize(results: List[Result], requests: List[Request], makespan: float, peak=0) -> Dict:
    """Throughput and latency over all requests; with gradient requests in the
    stream also per kind (forward shots, gradient shots)."""
    kind = {r.rid: r.kind for r in requests}
    shots = sum(r.nshots for r in requests)
    cells = sum((r.n_int + 2 * r.ncpml) ** 3 * (r.it - 1) * r.nshots for r in requests)
    out = {"requests": len(requests), "shots": shots, "makespan_s": makespan,
           "shots_per_s": shots / makespan, "gcell_updates_per_s": cells / makespan / 1e9,
           **_percentiles([r.latency for r in results]), "peak_hbm_gb": peak / 1e9,
           "results": {r.rid: r for r in results}}
    if any(k == "gradient" for k in kind.values()):
        for k in ("forward", "gradient"):
            sel = [r for r in results if kind[r.rid] == k]
            n = sum(q.nshots for q in requests if q.kind == k)
            out[k] = {"requests": len(sel), "shots": n, "shots_per_s": n / makespan,
                      **_percentiles([r.latency for r in sel])}
    return out


def run_sequential(requests: List[Request], mode: str = "fused", realtime: bool = True,
                   on_done: Optional[Callable[[Result], None]] = None,
                   budget_bytes: float = 200e9) -> Dict:
    """The plain path: one shot at a time, in arrival order.  mode 'fused' runs
    each shot's graph-captured time loop; mode 'ops' calls the reference-named
    per-step operators (six launches per step) as the reference loop does.
    Gradient requests run alone with the whole budget (run_gradient)."""
    results = []
    peak = 0
    t0 = time.perf_counter()
    now = lambda: time.perf_counter() - t0
    for req in sorted(requests, key=lambda r: r.arrival):
        if realtime and req.arrival > now():
            time.sleep(req.arrival - now())
        if req.kind == "gradient":
            peak = max(peak, gradient_bytes(req, budget_bytes)[1])
            res = run_gradient(req, budget_bytes)
            res.done = now()
            results.append(res)
            if on_done:
                on_done(res)
            continue
        grid, sig = build_model(req)
        model = Model(grid)
        model.media_coeff_3d(sig)
        signal = firstderiv_gauss(req.it, grid.dt, configs.FMAX_W)
        peak = max(peak, grid.ncells * MODEL_BYTES_PER_CELL +
                   shot_bytes(grid, req.it, nrec=len(req.receivers)))
        spectra = []
        for src in req.sources:
            sh = model.shot(req.it, signal, tuple(int(v) for v in src), 0, req.receivers)
            if mode == "fused":
                sh.run()
            else:
                for step in range(req.it - 1):
                    sh.e_field_cpml42()
                    sh.epml_abcs4()
                    sh.read_source_3d(step)
                    sh.h_field_cpml42()
                    sh.hpml_abcs4()
                    sh.output(step)
            spectra.append(shot_result(sh, req))
            sh.close()
        model.close()
        results.append(Result(req.rid, spectra, req.arrival, now()))
        if on_done:
            on_done(results[-1])
    return summarize(results, requests, now(), peak=peak)
