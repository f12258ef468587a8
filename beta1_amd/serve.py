"""Continuous-batching solve server over one device context.

Requests (a case description and a number of steps) arrive over time.  A
request is admitted into the running context as soon as the HBM budget has
room for its state (and for its geometry unless a case of the same grid shape
is already resident); from then on every step of the context advances it
together with all other resident cases (each with its own dt).  When a case
has taken its steps its conserved fields are copied to the host and it is
retired, freeing memory for the queue.  Host-side case set-up (grid,
geometry, weight tables, op lists) runs on worker threads so it overlaps the
device work of the cases already running.
"""
import time
from collections import deque
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from .capi import MhdError
from .solver import HostCase, MhdContext

CONSERVED = ("U_H", "B_xi", "B_eta", "B_zeta")


@dataclass
class Request:
    rid: int
    desc: object
    steps: int
    arrival: float          # seconds after the start of the run
    kind: str = "lunar"
    cid: int = -1
    admitted: float = -1.0
    done: float = -1.0
    failed: bool = False    # its dt met the abort condition: retired without a result
    result: dict = field(default_factory=dict)

    @property
    def cells(self):
        return self.desc.cells


class ServingEngine:
    def __init__(self, precision="mixed", hbm_budget_gb=200.0, chunk=10, workers=4, graph=True, keep_results=False):
        self.ctx = MhdContext(precision=precision, hbm_budget_gb=hbm_budget_gb, autotune=False)
        self.ctx.finalize()
        self.chunk = chunk
        self.graph = graph
        self.keep = keep_results
        self.pool = ThreadPoolExecutor(max_workers=workers)

    def close(self):
        self.pool.shutdown(wait=True)
        self.ctx.close()

    def run(self, requests, log=None):
        """Serve `requests` (sorted by arrival); returns per-request records."""
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        building = deque()  # (request, future HostCase) in arrival order
        active = {}
        finished = []
        t0 = time.perf_counter()
        cell_steps = 0
        while pending or building or active:
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                r = pending.popleft()
                building.append((r, self.pool.submit(HostCase, r.desc)))
            # admit in arrival order while memory allows (head-of-line)
            while building and building[0][1].done():
                r, fut = building[0]
                cid = self.ctx.adopt(fut.result())
                if cid is None:
                    if not active:
                        raise MhdError(f"request {r.rid} ({r.cells} cells) does not fit the HBM budget")
                    break
                building.popleft()
                r.cid = cid
                r.admitted = time.perf_counter() - t0
                active[cid] = r
            if not active:
                wait = (pending[0].arrival - now) if pending and not building else 0.002
                time.sleep(max(0.0, min(wait, 0.05)))
                continue
            left = min(r.steps - self.ctx.case_steps(cid) for cid, r in active.items())
            n = max(1, min(self.chunk, left))
            try:
                self.ctx.step(n, graph=self.graph and n >= 2)
            except MhdError as e:
                if e.code != -3:
                    raise
                # a case met the dt abort condition: it leaves the batch
                # without a result; every other case stepped normally
                for cid in [c for c in active if self.ctx.case_error(c)]:
                    r = active.pop(cid)
                    r.failed = True
                    r.done = time.perf_counter() - t0
                    self.ctx.retire(cid)
                    finished.append(r)
                    if log:
                        log(f"[serve] request {r.rid} ({r.kind}) failed: dt abort after "
                            f"{self.ctx.case_steps(cid)} steps")
            cell_steps += n * sum(r.cells for r in active.values())
            for cid in [c for c, r in active.items() if self.ctx.case_steps(c) >= r.steps]:
                r = active.pop(cid)
                nb = self.ctx.info(cid)[0]
                res = {(f, b): self.ctx.field(cid, f, b) for f in CONSERVED for b in range(nb)}
                r.done = time.perf_counter() - t0
                if self.keep:
                    r.result = res
                self.ctx.retire(cid)
                finished.append(r)
                if log:
                    log(f"[serve] request {r.rid} ({r.kind}, {r.cells} cells) done: "
                        f"latency {r.done - r.arrival:.2f}s, {len(active)} active, {len(building)} queued")
        wall = time.perf_counter() - t0
        lat = np.array([r.done - r.arrival for r in finished if not r.failed])
        return {
            "requests": len(finished),
            "failed": sum(r.failed for r in finished),
            "wall_s": wall,
            "p50_s": float(np.percentile(lat, 50)) if len(lat) else None,
            "p99_s": float(np.percentile(lat, 99)) if len(lat) else None,
            "mcell_updates_per_s": cell_steps / wall / 1e6,
            "records": finished,
        }


def poisson_requests(n, rate_hz, seed=20261203, steps=1000, sizes=None, mercury_frac=0.3, arrival_offset=0.0,
                     mercury_r_in=0.85):
    """Poisson arrivals at `rate_hz`; 70/30 Lunar/Mercury solar-wind cases with
    grids from 6x32^2x64 to 6x128^2x256.  The Mercury shells reach inside the
    planet to `mercury_r_in` (< 1: a conducting body, the reference's interior
    resistivity inver_Rem yita0(r) acting on 0.84 < r < 1.04 with its explicit
    subcycle, Z5_Mercury/src/physics/MercurySolver_MagneticRHS.cpp:79-141)."""
    from .solver import WindDesc, MercuryDesc

    rng = np.random.default_rng(seed)
    sizes = sizes or [(32, 64), (48, 96), (64, 128), (96, 192), (128, 256)]
    t = arrival_offset
    out = []
    for rid in range(n):
        t += rng.exponential(1.0 / rate_hz)
        nang, nrad = sizes[rng.integers(len(sizes))]
        if rng.random() < mercury_frac:
            desc = MercuryDesc(nang=nang, nrad=nrad, r_in=mercury_r_in, v_kms=float(rng.uniform(300, 500)),
                               n_cc=float(rng.uniform(20, 60)))
            kind = "mercury"
        else:
            z = rng.uniform(-1, 1)
            phi = rng.uniform(0, 2 * np.pi)
            s = np.sqrt(1 - z * z)
            bmag = float(rng.uniform(2, 10))
            desc = WindDesc(nang=nang, nrad=nrad, v_kms=float(rng.uniform(300, 700)), n_cc=float(rng.uniform(1, 20)),
                            B_nT=(bmag * s * np.cos(phi), bmag * s * np.sin(phi), bmag * z))
            kind = "lunar"
        out.append(Request(rid, desc, steps, t, kind))
    return out
