"""Serving: a continuous-batching request loop over device-resident MHD solves.

Requests are independent solves (2D field-loop advection, 2D or 3D blast with a given plasma
beta) of different sizes, each advanced a fixed number of steps with its own CFL dt.  They
arrive as a seeded Poisson stream.  The loop

  * prepares each arrival on the host (mesh and initial state) in a small thread pool,
  * admits prepared requests into a batch (one per dimension) while their HBM footprint fits
    the budget -- state and scratch stay on the device until the request finishes,
  * advances every admitted request together, one batch step at a time in ticks no longer than
    the shortest remaining request, so that requests join and leave between ticks; the 2D and
    the 3D batch tick on streams of their own, independently, so their work overlaps,
  * reads finished results back to the host and frees their slots.

Latency is arrival -> result on the host.  A request's result does not depend on what else
was in the batch (tests/test_gpu_serving.py).  The reference has no serving layer; this is the
batching layer of this package (batch.BatchGodunovMHD) driven as a server.
"""
from __future__ import annotations

import math
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from . import problems
from .batch import BatchGodunovMHD
from .implem import IMPL_FUSED, hydro_settings
from .mesh import uniform_mesh


@dataclass
class Request:
    rid: int
    arrival: float          # seconds after the stream start
    kind: str               # "loop2d", "blast2d", "blast3d"
    n: int                  # cells per axis (2D: along y; the field loop is 2n x n)
    b: int                  # block edge
    beta: float = 1.0
    nsteps: int = 200
    seed: int = 0
    # filled by the server
    admitted: float = math.nan
    done: float = math.nan
    result: tuple | None = field(default=None, repr=False)

    @property
    def latency(self) -> float:
        return self.done - self.arrival


def request_stream(nreq: int, rate: float = 50.0, seed: int = 0, nsteps: int = 200,
                   sizes2d=(64, 128, 256, 512), sizes3d=(32, 64, 128), blocks=(8, 16)) -> list[Request]:
    """Seeded Poisson arrivals at `rate` per second; kind, size, block edge and beta (log-uniform
    in [0.1, 10]) drawn per request."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(nreq):
        t += rng.exponential(1.0 / rate)
        kind = ("loop2d", "blast2d", "blast3d")[rng.integers(0, 3)]
        n = int(rng.choice(sizes3d if kind == "blast3d" else sizes2d))
        b = int(rng.choice(blocks))
        beta = float(10 ** rng.uniform(-1.0, 1.0))
        out.append(Request(i, t, kind, n, b, beta, nsteps, seed=int(rng.integers(0, 2 ** 31))))
    return out


def prepare(req: Request):
    """Mesh, initial state and settings of a request (host side)."""
    b = req.b
    if req.kind == "loop2d":
        # [-1,1] x [-0.5,0.5] periodic: a 2 x 1 brick of trees, n cells along y
        level = int(round(math.log2(req.n // b)))
        mesh = uniform_mesh(2, b, level, nbrick=(2, 1, 1))
        U, B = problems.field_loop_2d(mesh, scaling=1.0)
        st = hydro_settings(gamma0=5.0 / 3.0, cfl=0.8, slope_type=1.0)
    else:
        dim = 3 if req.kind == "blast3d" else 2
        level = int(round(math.log2(req.n // b)))
        mesh = uniform_mesh(dim, b, level)
        p_out = 0.1
        bmag = math.sqrt(2.0 * p_out / req.beta)
        bdir = (1 / math.sqrt(3),) * 3 if dim == 3 else (1 / math.sqrt(2), 1 / math.sqrt(2), 0.0)
        centre = (0.5, 0.5, 0.5) if dim == 3 else (0.5, 0.5, 0.0)
        U, B = problems.blast(mesh, radius=0.1, center=centre, p_in=10.0, p_out=p_out,
                              bfield=tuple(bmag * c for c in bdir))
        st = hydro_settings(gamma0=5.0 / 3.0, cfl=0.8, slope_type=1.0)
    return mesh, U, B, st


class Server:
    """Continuous-batching loop.  `hbm_budget` bounds the summed footprint of admitted
    requests (state + scratch); `max_tick` bounds the steps between admission points.

    The 2D and the 3D batch run on streams of their own, so their ticks overlap on the device;
    the loop polls each batch's tick by an event, retires and admits requests of a batch while
    it is idle, and starts its next tick at once, without waiting for the other batch."""

    def __init__(self, hbm_budget: int, impl: int = IMPL_FUSED, max_tick: int = 20, use_graph: bool = False,
                 prep_workers: int = 4):
        self.budget = int(hbm_budget)
        self.impl = impl
        self.max_tick = int(max_tick)
        self.use_graph = use_graph
        self.prep_workers = prep_workers
        self.batches: dict[int, BatchGodunovMHD] = {}
        self.streams: dict = {}

    def _batch(self, dim):
        if dim not in self.batches:
            import torch

            bt = BatchGodunovMHD(dim, self.impl)
            bt.set_graph(self.use_graph)
            self.batches[dim] = bt
            self.streams[dim] = torch.cuda.Stream()
        return self.batches[dim]

    def run(self, requests: list[Request], keep_results: bool = False, log=None) -> dict:
        import torch

        reqs = sorted(requests, key=lambda r: r.arrival)
        pool = ThreadPoolExecutor(max_workers=self.prep_workers)
        futures = {}
        pending = list(reqs)         # not arrived yet
        ready = []                   # arrived; (req, future)
        active = {}                  # (dim, slot) -> [req, remaining, footprint, cells]
        inflight = {}                # dim -> (event, tick) of the batch's running tick
        used = 0
        done = []
        cell_updates = 0
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0  # noqa: E731
        while pending or ready or active:
            progressed = False
            # arrivals: start host preparation
            while pending and pending[0].arrival <= now():
                r = pending.pop(0)
                futures[r.rid] = pool.submit(prepare, r)
                ready.append(r)
                progressed = True
            # finished ticks: every request of that batch advanced; retire the finished ones
            for dim in [d for d, (ev, _) in inflight.items() if ev.query()]:
                _, tick = inflight.pop(dim)
                progressed = True
                bt = self.batches[dim]
                for key, a in active.items():
                    if key[0] == dim:
                        a[1] -= tick
                        cell_updates += a[3] * tick
                for key in [k for k, a in active.items() if k[0] == dim and a[1] == 0]:
                    r, _, need, _ = active.pop(key)
                    sid = key[1]
                    if keep_results:
                        r.result = bt.state(sid)
                    else:
                        # the state after r.nsteps steps is in buffer r.nsteps % 2: one value to
                        # the host stands for the result's read-back
                        U, _ = bt.slots[sid].bufs[r.nsteps & 1]
                        _ = float(U[0].item())
                    bt.remove(sid)
                    used -= need
                    r.done = now()
                    done.append(r)
                    if log:
                        log(f"[serve] request {r.rid} {r.kind} n={r.n} b={r.b} latency {r.latency * 1e3:.1f} ms")
            # admissions, in arrival order, into batches that are not running a tick, while the
            # footprint fits
            still = []
            for r in ready:
                f = futures[r.rid]
                if not f.done():
                    still.append(r)
                    continue
                mesh, U, B, st = f.result()
                need = BatchGodunovMHD.footprint_bytes(mesh, self.impl)
                if mesh.dim in inflight or (active and used + need > self.budget):
                    still.append(r)
                    continue
                bt = self._batch(mesh.dim)
                sid = bt.add(mesh, U, B, st)
                r.admitted = now()
                active[(mesh.dim, sid)] = [r, r.nsteps, need, mesh.num_leaves * mesh.cells_per_block]
                used += need
                del futures[r.rid]
                progressed = True
            ready = still
            # a tick on every idle batch with requests: as many steps as its nearest finish
            for dim, bt in self.batches.items():
                if dim in inflight:
                    continue
                rem = [a[1] for k, a in active.items() if k[0] == dim]
                if not rem:
                    continue
                tick = min(self.max_tick, min(rem))
                s = self.streams[dim]
                bt.step(tick, s)
                ev = torch.cuda.Event()
                ev.record(s)
                inflight[dim] = (ev, tick)
                progressed = True
            if not progressed:
                time.sleep(0.0002)
        wall = now()
        pool.shutdown()
        torch.cuda.synchronize()
        for bt in self.batches.values():
            bt.check()
        lat = np.array([r.latency for r in done]) * 1e3
        return {
            "requests": len(done),
            "wall_s": wall,
            "cell_updates_per_s": cell_updates / wall if wall > 0 else 0.0,
            "latency_ms_p50": float(np.percentile(lat, 50)) if len(lat) else None,
            "latency_ms_p99": float(np.percentile(lat, 99)) if len(lat) else None,
            "latency_ms_mean": float(lat.mean()) if len(lat) else None,
            "done": done,
        }
