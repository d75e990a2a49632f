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
from typing import Callable, Dict, List, Optional

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


def request_work(req: Request) -> int:
    """Cell updates the request needs: cells (incl. CPML) x steps x shots."""
    return (req.n_int + 2 * req.ncpml) ** 3 * (req.it - 1) * req.nshots


class Server:
    """Continuous-batching server over one persistent multi-shot batch.

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
                 target_cells: float = 24e6, policy: str = "swf"):
        if policy not in ("swf", "fifo"):
            raise ValueError("policy must be 'swf' (shortest work first) or 'fifo'")
        self.budget = float(budget_bytes)
        self.chunk = chunk
        self.target_cells = float(target_cells)
        self.policy = policy
        self.batch = Batch(max_shots, chunk=chunk)
        self.in_use = 0
        self.peak = 0
        self.admissions = []                  # (rid, time its first shot joined)

    def close(self):
        self.batch.close()

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

    def run(self, requests: List[Request], realtime: bool = True,
            on_done: Optional[Callable[[Result], None]] = None) -> Dict:
        for req in requests:
            grid, _ = build_model(req)
            one = grid.ncells * MODEL_BYTES_PER_CELL + shot_bytes(grid, req.it, nrec=len(req.receivers))
            if one > self.budget:
                raise ValueError(f"request {req.rid} needs {one / 1e9:.1f} GB for one shot > budget")
        incoming = sorted(requests, key=lambda r: (r.arrival, r.rid))
        ready: List[Request] = []
        started: List[dict] = []              # in start order
        results: List[Result] = []
        cells = 0
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0
        while incoming or ready or started:
            while incoming and (not realtime or incoming[0].arrival <= now()):
                ready.append(incoming.pop(0))
            if self.policy == "swf":
                ready.sort(key=lambda r: (request_work(r), r.arrival, r.rid))
            # admit shots: unfinished started requests first, then the queue head
            while True:
                nb = len(self.batch.members)
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
                self.batch.add(sh)
                a["live"][q] = sh
                a["next"] += 1
                cells += gcells
                self.in_use += a["shot_bytes"]
                self.peak = max(self.peak, self.in_use)
            if not self.batch.members:
                if incoming:                  # idle until the next arrival
                    time.sleep(max(0.0, min(incoming[0].arrival - now(), 0.01)))
                    continue
                if not ready and not started:
                    break
                raise RuntimeError("no shot fits the HBM budget while the batch is empty")
            self.batch.step(self.chunk)
            self.batch.sync()
            for a in list(started):
                for q, sh in list(a["live"].items()):
                    if sh.step >= sh.it - 1:
                        self.batch.remove(sh)
                        a["spectra"][q] = shot_result(sh, a["req"])
                        sh.close()
                        del a["live"][q]
                        a["left"] -= 1
                        cells -= a["grid"].ncells
                        self.in_use -= a["shot_bytes"]
                if a["left"] == 0:
                    a["model"].close()
                    self.in_use -= a["grid"].ncells * MODEL_BYTES_PER_CELL
                    results.append(Result(a["req"].rid, a["spectra"], a["req"].arrival, now()))
                    started.remove(a)
                    if on_done:
                        on_done(results[-1])
        makespan = now()
        return summarize(results, requests, makespan, peak=self.peak)


def summarize(results: List[Result], requests: List[Request], makespan: float, peak=0) -> Dict:
    lat = np.array([r.latency for r in results]) if results else np.zeros(1)
    shots = sum(r.nshots for r in requests)
    cells = sum((r.n_int + 2 * r.ncpml) ** 3 * (r.it - 1) * r.nshots for r in requests)
    return {"requests": len(requests), "shots": shots, "makespan_s": makespan,
            "shots_per_s": shots / makespan, "gcell_updates_per_s": cells / makespan / 1e9,
            "mean_s": float(lat.mean()), "p50_s": float(np.percentile(lat, 50)), "p95_s": float(np.percentile(lat, 95)),
            "p99_s": float(np.percentile(lat, 99)), "peak_hbm_gb": peak / 1e9,
            "results": {r.rid: r for r in results}}


def run_sequential(requests: List[Request], mode: str = "fused", realtime: bool = True,
                   on_done: Optional[Callable[[Result], None]] = None) -> Dict:
    """The plain path: one shot at a time, in arrival order.  mode 'fused' runs
    each shot's graph-captured time loop; mode 'ops' calls the reference-named
    per-step operators (six launches per step) as the reference loop does."""
    results = []
    peak = 0
    t0 = time.perf_counter()
    now = lambda: time.perf_counter() - t0
    for req in sorted(requests, key=lambda r: r.arrival):
        if realtime and req.arrival > now():
            time.sleep(req.arrival - now())
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
