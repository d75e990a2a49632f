"""Serving: continuous batching of forward-modelling shots.

A request carries a conductivity model (seeded layered + box generator), a
grid size, 1-16 shots (x-dipoles at random positions) and asks for receiver
spectra at a few frequencies.  The server

* admits a request when the HBM its model and shots need fits under the budget
  (``fwi3d_shot_bytes`` + model bytes), otherwise it waits in the queue;
* adds the admitted shots to ONE persistent multi-shot batch
  (``fwi3d_batch_*``): every member, whatever its grid size, advances in the
  same kernel launches; membership changes between chunks, so requests that
  arrive while others run join as soon as memory allows and overlap with them;
* retires each shot when its own time loop ends, frees its memory, and
  completes the request when all its shots are done.

Each shot's result is bit-identical to the same shot run alone (per-shot
descriptors, no shared accumulators, no data-dependent reduction order).

Latency = completion time - arrival time (wall clock, the request stream is
replayed in real time); throughput = shots / makespan.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from . import configs
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

    @property
    def nshots(self) -> int:
        return len(self.sources)


@dataclass
class Result:
    rid: int
    spectra: List[np.ndarray]       # per shot (6, nrec, nf) complex
    arrival: float
    done: float

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


def build_model(req: Request):
    sig, _ = configs.layered_box_sigma(req.n_int, req.ncpml, req.seed)
    n = req.n_int + 2 * req.ncpml
    dt, cmax = configs.dt_for(50.0, 50.0, 50.0)
    grid = Grid(n, n, n, req.ncpml, 50.0, 50.0, 50.0, dt, configs.OMEGA0, cmax)
    return grid, sig


def request_bytes(req: Request, grid: Grid) -> int:
    per_shot = shot_bytes(grid, req.it, nrec=len(req.receivers))
    return grid.ncells * MODEL_BYTES_PER_CELL + req.nshots * per_shot


def shot_result(shot: Shot, req: Request) -> np.ndarray:
    om = 2 * np.pi * np.asarray(req.freqs_hz)
    return shot.receiver_dft(om, h_factor=True)


class Server:
    """Continuous-batching server over one persistent multi-shot batch."""

    def __init__(self, budget_bytes: float = 200e9, chunk: int = 32, max_shots: int = 512):
        self.budget = float(budget_bytes)
        self.chunk = chunk
        self.batch = Batch(max_shots, chunk=chunk)
        self.in_use = 0
        self.peak = 0
        self.admissions = []                  # (rid, admit time)

    def close(self):
        self.batch.close()

    def run(self, requests: List[Request], realtime: bool = True) -> Dict:
        pending = sorted(requests, key=lambda r: r.arrival)
        active: Dict[int, dict] = {}
        results: List[Result] = []
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0
        while pending or active:
            # admit everything that has arrived and fits under the budget (FIFO)
            while pending and (not realtime or pending[0].arrival <= now()):
                req = pending[0]
                grid, sig = build_model(req)
                need = request_bytes(req, grid)
                if need > self.budget:
                    raise ValueError(f"request {req.rid} needs {need / 1e9:.1f} GB > budget")
                if self.in_use + need > self.budget:
                    break                     # wait for memory to free up
                pending.pop(0)
                model = Model(grid)
                model.media_coeff_3d(sig)
                signal = firstderiv_gauss(req.it, grid.dt, configs.FMAX_W)
                shots = []
                for src in req.sources:
                    sh = model.shot(req.it, signal, tuple(int(v) for v in src), 0, req.receivers)
                    self.batch.add(sh)
                    shots.append(sh)
                self.in_use += need
                self.peak = max(self.peak, self.in_use)
                self.admissions.append((req.rid, now()))
                active[req.rid] = {"req": req, "model": model, "shots": shots, "left": len(shots),
                                   "spectra": [None] * len(shots), "bytes": need}
            if not active:
                if pending:                   # idle until the next arrival
                    time.sleep(max(0.0, min(pending[0].arrival - now(), 0.01)))
                continue
            self.batch.step(self.chunk)
            self.batch.sync()
            for rid in list(active):
                a = active[rid]
                for q, sh in enumerate(a["shots"]):
                    if sh is not None and sh.step >= sh.it - 1:
                        self.batch.remove(sh)
                        a["spectra"][q] = shot_result(sh, a["req"])
                        sh.close()
                        a["shots"][q] = None
                        a["left"] -= 1
                if a["left"] == 0:
                    a["model"].close()
                    self.in_use -= a["bytes"]
                    results.append(Result(rid, a["spectra"], a["req"].arrival, now()))
                    del active[rid]
        makespan = now()
        return summarize(results, requests, makespan, peak=self.peak)


def summarize(results: List[Result], requests: List[Request], makespan: float, peak=0) -> Dict:
    lat = np.array([r.latency for r in results]) if results else np.zeros(1)
    shots = sum(r.nshots for r in requests)
    cells = sum((r.n_int + 2 * r.ncpml) ** 3 * (r.it - 1) * r.nshots for r in requests)
    return {"requests": len(requests), "shots": shots, "makespan_s": makespan,
            "shots_per_s": shots / makespan, "gcell_updates_per_s": cells / makespan / 1e9,
            "p50_s": float(np.percentile(lat, 50)), "p95_s": float(np.percentile(lat, 95)),
            "p99_s": float(np.percentile(lat, 99)), "peak_hbm_gb": peak / 1e9,
            "results": {r.rid: r for r in results}}


def run_sequential(requests: List[Request], mode: str = "fused", realtime: bool = True) -> Dict:
    """The plain path: one shot at a time, in arrival order.  mode 'fused' runs
    each shot's graph-captured time loop; mode 'ops' calls the reference-named
    per-step operators (six launches per step) as the reference loop does."""
    results = []
    t0 = time.perf_counter()
    now = lambda: time.perf_counter() - t0
    for req in sorted(requests, key=lambda r: r.arrival):
        if realtime and req.arrival > now():
            time.sleep(req.arrival - now())
        grid, sig = build_model(req)
        model = Model(grid)
        model.media_coeff_3d(sig)
        signal = firstderiv_gauss(req.it, grid.dt, configs.FMAX_W)
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
    return summarize(results, requests, now())
