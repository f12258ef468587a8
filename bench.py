"""Benchmark: explicit MHD step throughput on one MI355X.

Workload: Lunar wake, plain FP64 -- 6-block cubed-sphere shell, r in [1, 8],
6 x 64 x 64 x 128 cells, solar wind n=5 cm^-3, v=400 km/s along -x, T=10 eV,
IMF 5 nT at 45 deg in the x-y plane, absorbing insulating body.  A "step" is
one full explicit step (dt reduction, Rusanov fluxes, J x B, face and UCT
edge EMF, CT curl, Euler update, halos, derived fields) over all 6 blocks.

Prints one JSON line (value = Mcell-updates/s, whole job on one GPU) with a
roofline object for the dominant kernel (HIP-event timing on the library's
stream) and a CPU baseline (the FP64 oracle on the same workload, bounded
number of steps, host cores of this machine).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Algorithmic HBM bytes per interior cell and launch for each step kernel:
# every array the kernel must touch counted once at its stored width, as
# (per-case state, metric and node coordinates (always FP64), Hodge weights
# and weight tables (FP64, or float in mixed precision: half)).
KERNEL_PARTS = {
    "k_dt": (40 + 24, 8 + 72, 0),                  # U, B_cell | Jac, JD
    "k_ucontra": (40 + 24, 72 + 8, 0),             # U -> UC | JD, Jac
    "k_face_emf": (24 + 24 + 24 + 72, 0, 0),       # B, Badd, UC -> Eface
    "k_edge_emf": (72 + 24 + 24 + 24, 0, 0),       # Eface, UC, B -> E
    "k_emf": (24 + 24 + 24 + 24, 0, 0),            # B, Badd, UC -> E (face and edge EMF in one pass)
    "k_fluid": (40 + 24 + 24 + 40, 72 + 8, 0),     # U, B_cell, J_cell -> U' | JD, Jac
    "k_curl_update": (24 + 24 + 24, 0, 0),         # E, B -> B
    # U, B, Badd -> UC, B_cell (+ next step's dt) | JD, Jac | B_cell least squares Mi (6), mu (6)
    "k_derived": (40 + 24 + 24 + 24 + 24, 72 + 8, 96),
    # B -> J_cell (J edges formed in the pass) | nodes | beta, alpha, Mi (6), active-edge mask
    "k_jcell": (24 + 24, 24, 24 + 24 + 48 + 8),
    # two-fluid step: U_H, U_Na -> U_H', U_Na', B_cell, J_cell, p_e, photoionization, drag | JD, Jac
    "k_fluid2": (80 + 80 + 24 + 24 + 8 + 8 + 8, 72 + 8, 0),
    # U_H, U_Na, B, Badd -> UC, p_e, B_cell (+ dt) | JD, Jac | Mi, mu
    "k_derived2": (80 + 24 + 24 + 24 + 8 + 24, 72 + 8, 96),
    # E R/W, p_e, rho_H, rho_Na, B, B_cell, J_cell, Hall taper, eta | JD, nodes (dr) | beta, alpha
    "k_emf_extra": (48 + 8 + 16 + 24 + 24 + 24 + 24 + 24, 72 + 24, 24 + 24),
}


def kernel_bytes(prec="fp64"):
    return {k: s + m + (w if prec == "fp64" else w // 2) for k, (s, m, w) in KERNEL_PARTS.items()}


KERNEL_BYTES = kernel_bytes("fp64")
HBM_PEAK_GBS = 8000.0
# Whole-step algorithmic bytes per cell-update, each array once per step:
# U R/W 80 + B_face R/W 48 + Badd R 24 (per case) + geometry: metric Jac 8,
# JD 72, nodes 24 (FP64 always) and the weights beta 24, alpha 24, the B_cell
# least squares Mi 48 + mu 48, the J_cell Mi 48 + mask 8 (FP64, float in
# mixed precision), counted
# once per case (K = 1: the geometry-shared tile order lets cases of one grid
# meet their geometry in L2, which is a cache effect, not a guaranteed single
# read).
STATE_BYTES = 152
GEO_BYTES = {"fp64": 104 + 200, "mixed": 104 + 100}
STEP_BYTES = STATE_BYTES + GEO_BYTES["fp64"]
SEED = 20261203


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _record(metric, value, unit, steps, warmup, ms, dtype, config, **extra):
    out = {"metric": metric, "value": round(value, 3), "unit": unit, "n_gpus": 1, "steps": steps,
           "warmup": warmup, "ms_per_step": round(ms, 4) if ms is not None else None, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": dtype, "data": "synthetic", "config": config}
    out.update(extra)
    return out


def batch_descs(n, seed=SEED):
    """Seeded shells: v~U[300,700] km/s, n~U[1,20] cm^-3, |B|~U[2,10] nT with
    the IMF direction uniform on the sphere; angular size from {32,48,64},
    radial from {64,96,128}, independently."""
    import numpy as np
    from beta1_amd import WindDesc

    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        nang = int(rng.choice([32, 48, 64]))
        nrad = int(rng.choice([64, 96, 128]))
        v = float(rng.uniform(300, 700))
        ncc = float(rng.uniform(1, 20))
        bmag = float(rng.uniform(2, 10))
        z = float(rng.uniform(-1, 1))
        phi = float(rng.uniform(0, 2 * np.pi))
        s = (1 - z * z) ** 0.5
        out.append(WindDesc(nang=nang, nrad=nrad, v_kms=v, n_cc=ncc,
                            B_nT=(bmag * s * np.cos(phi), bmag * s * np.sin(phi), bmag * z)))
    return out


def bench_batch(args):
    """64 seeded shells of mixed sizes advanced together in one context
    (mixed precision, graph-captured steps, tuned workgroups)."""
    from concurrent.futures import ThreadPoolExecutor
    from beta1_amd import HostCase, MhdContext

    descs = batch_descs(args.cases)
    cells = sum(d.cells for d in descs)
    prec = args.precision
    t0 = time.time()
    ctx = MhdContext(precision=prec, autotune=True, xcd_order=True)
    # the largest case first: the context is finalized (and its workgroup
    # shapes tuned) on it, the others are admitted live as their host set-up
    # completes, so host arrays are released case by case
    order = sorted(range(len(descs)), key=lambda i: -descs[i].cells)
    with ThreadPoolExecutor(max_workers=4) as pool:
        futs = [pool.submit(HostCase, descs[i]) for i in order]
        for n, f in enumerate(futs):
            hc = f.result()
            if ctx.adopt(hc) is None:
                raise RuntimeError("batch does not fit the device")
            futs[n] = None
            del hc
            if n == 0:
                ctx.finalize()
            if n % 8 == 7:
                log(f"batch set-up: {n + 1} cases placed after {time.time() - t0:.1f}s")
    log(f"batch set-up {time.time() - t0:.1f}s: {len(descs)} cases, {cells} cells, "
        f"{ctx.memory_bytes() / 1e9:.1f} GB on device")
    # the tuner ran on the first case alone: tune again on the whole batch
    # (this is where the geometry-shared tile order can pay)
    ctx.autotune()
    log(f"batch launch shapes: {ctx.tuning()}")
    ctx.step(args.warmup, graph=True)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.step(args.steps, graph=True)
    ctx.sync()
    el = time.perf_counter() - t0
    value = cells * args.steps / el / 1e6
    log(f"batch: {args.steps} steps in {el:.2f}s, {value:.1f} Mcell-updates/s")
    # cases per geometry set (one set per distinct grid); with member-looping
    # tiles the B_cell least-squares data (read by k_derived only) is read
    # once per set, the other geometry arrays once per case
    K = len(descs) / len({(d.nang, d.nrad) for d in descs})
    loop = ctx.tuning()["k_derived"][2] == "geo-loop"
    bcell_w = 96 if prec == "fp64" else 48
    step_bytes = STATE_BYTES + GEO_BYTES[prec] - (bcell_w * (1.0 - 1.0 / K) if loop else 0.0)
    hbm = value * 1e6 * step_bytes / (HBM_PEAK_GBS * 1e9)
    ctx.profile(True)
    ctx.step(4)
    prof = ctx.profile_read()
    ctx.profile(False)
    kb = kernel_bytes(prec)
    per_kernel = {k: {"ms": round(ms / n, 4), "alg_GBs": round(kb[k] * cells / (ms / n * 1e-3) / 1e9, 1)
                      if k in kb else None} for k, (ms, n) in prof.items() if n}
    log(f"batch per-kernel avg ms and algorithmic GB/s: {per_kernel}")
    return _record("Mcell-updates/s (64-case batch)", value, "Mcell-updates/s", args.steps, args.warmup,
                   1e3 * el / args.steps, "f32/f64" if prec == "mixed" else "f64",
                   {"workload": f"batch_{len(descs)}_shells_{prec}", "cases": len(descs), "cells": cells,
                    "seed": SEED, "graph": True, "autotune": ctx.tuning()},
                   step_hbm={"bytes_per_cell_update": round(step_bytes, 1), "frac": round(hbm, 4),
                             "geometry_sharing_K": round(K, 2) if loop else 1,
                             "geometry_sharing": {k: v[2] for k, v in ctx.tuning().items()
                                                  if k in ("k_derived", "k_fluid", "k_jcell")}},
                   per_kernel=per_kernel)


def bench_serve(args):
    """Poisson request stream (70% Lunar, 30% Mercury shells of 6x32^2x64 ..
    6x128^2x256 cells, `--req-steps` steps each) served by continuous batching
    under an HBM budget; latency from arrival to result on the host."""
    import numpy as np
    from beta1_amd.serve import ServingEngine, poisson_requests

    prec = args.precision
    # the request list at unit rate, rescaled below to the offered load
    reqs = poisson_requests(args.requests, 1.0, seed=SEED, steps=args.req_steps, mercury_r_in=args.mercury_r_in)
    mean_work = float(np.mean([r.cells for r in reqs])) * args.req_steps
    eng = ServingEngine(precision=prec, hbm_budget_gb=args.budget_gb, chunk=20, workers=2)
    from beta1_amd import HostCase, MercuryDesc, WindDesc
    # capacity of each model on a 6x64^2x128 shell alone: a request's service
    # time is its cell-updates at its model's rate
    cap = {}
    for name, d in (("lunar", WindDesc(nang=64, nrad=128)),
                    ("mercury", MercuryDesc(nang=64, nrad=128, r_in=args.mercury_r_in))):
        cal = eng.ctx.adopt(HostCase(d))
        eng.ctx.step(4, graph=True)
        eng.ctx.sync()
        t0 = time.perf_counter()
        eng.ctx.step(20, graph=True)
        eng.ctx.sync()
        cap[name] = d.cells * 20 / (time.perf_counter() - t0)  # cell-updates/s
        eng.ctx.retire(cal)
    service = [r.cells * args.req_steps / cap["mercury" if isinstance(r.desc, MercuryDesc) else "lunar"]
               for r in reqs]
    rate = args.load / float(np.mean(service))
    for r in reqs:  # exponential gaps of mean 1 -> mean 1 / rate
        r.arrival /= rate
    log(f"serve: capacity ~{cap['lunar'] / 1e6:.0f} (Lunar) / {cap['mercury'] / 1e6:.0f} (Mercury) "
        f"Mcell-updates/s, mean request {mean_work / 1e9:.2f} G cell-updates, rate {rate:.3f}/s "
        f"(offered load {args.load})")
    res = eng.run(reqs, log=log)
    eng.close()
    offered = rate * float(np.mean(service))
    return _record("p50/p99 solve latency + Mcell-updates/s (serving)", res["mcell_updates_per_s"],
                   "Mcell-updates/s", args.req_steps, 0, None, "f32/f64" if prec == "mixed" else "f64",
                   {"workload": f"serve_poisson_{args.requests}req_{prec}", "requests": args.requests,
                    "rate_hz": round(rate, 4), "offered_load": round(offered, 3), "mix": "70% Lunar / 30% Mercury",
                    "steps_per_request": args.req_steps, "hbm_budget_gb": args.budget_gb, "seed": SEED,
                    "mercury_body": f"conducting: shell from r = {args.mercury_r_in} R_M, interior resistivity "
                                    f"(0.84 < r < 1.04), 1 explicit resistive subcycle"
                                    if args.mercury_r_in < 1.0 else "absorbing insulating wall at r = 1"},
                   latency={"p50_s": round(res["p50_s"], 3), "p99_s": round(res["p99_s"], 3),
                            "wall_s": round(res["wall_s"], 2), "requests_failed": res["failed"]})


def bench_ot(args):
    """Orszag-Tang vortex in a periodic 64^3 box (the Cartesian case: uniform
    metric, Badd = 0), plain FP64, first-order and minmod."""
    from beta1_amd import MhdContext, OTDesc

    desc = OTDesc(n=64)
    cells = desc.n ** 3
    res = {}
    for recon in (0, 1):
        ctx = MhdContext(precision="fp64", recon=recon, autotune=True)
        cid = ctx.add(desc)
        ctx.finalize()
        ctx.step(args.warmup, graph=True)
        ctx.sync()
        t0 = time.perf_counter()
        ctx.step(args.steps, graph=True)
        ctx.sync()
        el = time.perf_counter() - t0
        res[recon] = (cells * args.steps / el / 1e6, 1e3 * el / args.steps, ctx.divb_max(cid))
        log(f"orszag-tang recon={recon}: {res[recon][0]:.1f} Mcell-updates/s, {res[recon][1]:.3f} ms/step, "
            f"div B {res[recon][2]:.2e}")
        ctx.profile(True)
        ctx.step(4)
        prof = ctx.profile_read()
        ctx.profile(False)
        log(f"orszag-tang recon={recon} per-kernel avg ms: "
            f"{ {k: round(ms / n, 4) for k, (ms, n) in prof.items() if n} }")
        ctx.close()
    return _record("Mcell-updates/s (Orszag-Tang 64^3)", res[0][0], "Mcell-updates/s", args.steps, args.warmup,
                   res[0][1], "f64", {"workload": "orszag_tang_fp64_64^3", "cells": cells, "recon": "first-order"},
                   minmod={"value": round(res[1][0], 3), "ms_per_step": round(res[1][1], 4)},
                   divb_max=float(max(res[0][2], res[1][2])),
                   fp64=fp64_record("orszag_tang_fp64_64cubed", res[0][0]),
                   step_hbm={"bytes_per_cell_update": STEP_BYTES,
                             "frac": round(res[0][0] * 1e6 * STEP_BYTES / (HBM_PEAK_GBS * 1e9), 4),
                             "note": "the box keeps stored metric and weight arrays like the shell (no closed-form "
                                     f"Cartesian specialisation), so it reads the shell's {STEP_BYTES} B per "
                                     "cell-update"})


def bench_precision(args):
    """FP64 and mixed precision on the same shell: throughput of each and the
    per-variable relative L-inf / L2 of mixed against FP64 (the FP64 path is
    bit-identical to the oracle), plus max |div B| of both."""
    import numpy as np
    from beta1_amd import MhdContext, WindDesc

    nang, nrad, steps = args.nang, args.nrad, args.steps
    desc = WindDesc(nang=nang, nrad=nrad)
    res, rate, divb = {}, {}, {}
    for prec in ("fp64", "mixed"):
        ctx = MhdContext(precision=prec, autotune=True)
        cid = ctx.add(desc)
        ctx.finalize()
        ctx.step(2, graph=True)  # graph capture outside the timed region; 'steps' in total
        ctx.sync()
        t0 = time.perf_counter()
        ctx.step(steps - 2, graph=True)
        ctx.sync()
        el = time.perf_counter() - t0
        rate[prec] = desc.cells * (steps - 2) / el / 1e6
        divb[prec] = ctx.divb_max(cid)
        res[prec] = {(f, b): ctx.field(cid, f, b) for f in ("U_H", "B_xi", "B_eta", "B_zeta") for b in range(6)}
        log(f"precision {prec}: {steps} steps in {el:.2f}s, {rate[prec]:.1f} Mcell-updates/s, div B {divb[prec]:.2e}")
        ctx.close()
    from beta1_amd.fields import interior
    names = {"U_H": ["rho", "rho_u", "rho_v", "rho_w", "E"], "B_xi": ["B_xi"], "B_eta": ["B_eta"],
             "B_zeta": ["B_zeta"]}
    err = {}
    n = (nang, nang, nrad)
    for f, labels in names.items():
        for m, lab in enumerate(labels):
            a = np.concatenate([res["mixed"][(f, b)][interior(f, n, 2)][..., m].ravel() for b in range(6)])
            r = np.concatenate([res["fp64"][(f, b)][interior(f, n, 2)][..., m].ravel() for b in range(6)])
            d = a - r
            err[lab] = {"rel_linf": float(np.abs(d).max() / max(np.abs(r).max(), 1e-300)),
                        "rel_l2": float(np.sqrt(np.sum(d * d) / max(np.sum(r * r), 1e-300)))}
    return _record("Mcell-updates/s FP64 vs mixed", rate["mixed"], "Mcell-updates/s", steps, 0, None, "f32/f64",
                   {"workload": f"precision_lunar_6x{nang}x{nang}x{nrad}", "cells": desc.cells},
                   fp64_value=round(rate["fp64"], 3), errors_mixed_vs_fp64=err,
                   divb_max={k: float(v) for k, v in divb.items()})


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(desc, cells, args):
    """The FP64 oracle on the same workload on the host cores this process may
    use (its CPU affinity; OMP_NUM_THREADS when the machine sets a smaller
    share), OpenMP over (block, i-plane), and on one thread: each timed over
    whole steps after one untimed step, about 10 s of work each."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from beta1_amd import HostCase
    from oracle_bind import Oracle

    share = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS")
    threads = max(1, min(share, int(env))) if env and env.isdigit() else max(1, share)
    hc = HostCase(desc, ghost=2)
    rates = {}
    for nt in (1, threads):
        orc = Oracle(hc, threads=nt)
        t0 = time.perf_counter()
        orc.step(1)                      # untimed: first touch of the arrays
        per = time.perf_counter() - t0
        nsteps = max(args.cpu_steps, int(10.0 / max(per, 1e-3)))
        t0 = time.perf_counter()
        orc.step(nsteps)
        cs = time.perf_counter() - t0
        rates[nt] = (cells * nsteps / cs / 1e6, nsteps, cs)
        log(f"cpu oracle, {nt} thread(s): {cs / nsteps:.3f} s/step over {nsteps} steps")
        del orc
    del hc
    v, nsteps, cs = rates[threads]
    return {"value": round(v, 4), "unit": "Mcell-updates/s", "cores": threads, "kind": "port",
            "sample": f"{nsteps} steps (after 1 untimed) of the same {desc.cells}-cell shell, FP64 oracle with OpenMP "
                      f"over (block, i-plane) on {threads} threads, {cs:.1f}s",
            "s_per_MCell_step": round(1.0 / v, 5),
            "cpu_model": cpu_model(), "host_cpus": os.cpu_count(), "process_cpus": share,
            "single_thread": {"value": round(rates[1][0], 4), "steps": rates[1][1], "seconds": round(rates[1][2], 2),
                              "s_per_MCell_step": round(1.0 / rates[1][0], 4)}}


def extra_configs(args):
    """The other benchmark configurations, each at its own size and step
    count, for the default JSON line: C1 (Orszag-Tang 64^3, first order and
    minmod), C3 (6x128x128x256, 500 steps, FP64 and mixed with the error of
    mixed against FP64), C4 (the 64-case mixed batch)."""
    import copy
    out = {}
    a = copy.copy(args)
    a.steps, a.warmup = 100, 5
    r = bench_ot(a)
    out["c1_orszag_tang_64"] = {"first_order": {"value": r["value"], "ms_per_step": r["ms_per_step"]},
                                "minmod": r["minmod"], "divb_max": r["divb_max"], "steps": 100,
                                "dtype": "f64", "step_hbm": r["step_hbm"], "fp64": r.get("fp64")}
    a = copy.copy(args)
    a.nang, a.nrad, a.steps = 128, 256, 500
    r = bench_precision(a)
    out["c3_precision_6x128x128x256"] = {"mixed_value": r["value"], "fp64_value": r["fp64_value"], "steps": 500,
                                         "errors_mixed_vs_fp64": r["errors_mixed_vs_fp64"],
                                         "divb_max": r["divb_max"],
                                         "note": "the FP64 step equals the FP64 oracle bit for bit at every size the "
                                                 "tests run, so mixed vs FP64 is mixed vs the oracle"}
    a = copy.copy(args)
    a.cases, a.steps, a.warmup, a.precision = 64, 40, 4, "mixed"
    r = bench_batch(a)
    out["c4_batch_64_mixed"] = {"value": r["value"], "ms_per_step": r["ms_per_step"], "steps": 40,
                                "cases": r["config"]["cases"], "cells": r["config"]["cells"],
                                "step_hbm": r["step_hbm"], "dtype": r["dtype"]}
    return out


def measured_traffic(workload, kernel):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3 counter
    summary of this workload (profiles/traffic_<workload>.md), or None."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.md")
    if not os.path.exists(path):
        return None
    import re
    col = None
    for line in open(path):
        cols = [c.strip() for c in line.strip().strip("|").split("|")]
        if col is None and "calibrated GB" in cols:
            col = cols.index("calibrated GB")
            continue
        m = re.search(r"\b(k_\w+)", cols[0]) if cols else None
        if col is not None and len(cols) > col and m and m.group(1) == kernel:
            try:
                return float(cols[col]) * 1e9
            except ValueError:
                return None
    return None


FP64_PEAK_TFLOPS = 78.6


def measured_flops(workload):
    """FP64 flops per cell-update of a step from the committed VALU counter
    summary of this workload (profiles/valu_<workload>.md), or None."""
    path = os.path.join(ROOT, "profiles", f"valu_{workload}.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("FP64 flops per cell-update:"):
            return float(line.split(":")[1])
    return None


def fp64_record(workload, mcells):
    """The step's FP64 rate against MI355X's 78.6 TFLOP/s vector FP64 peak,
    from the measured flops per cell-update."""
    f = measured_flops(workload)
    if f is None:
        return None
    tf = mcells * 1e6 * f / 1e12
    return {"flops_per_cell_update": f, "tflops": round(tf, 2), "peak_tflops": FP64_PEAK_TFLOPS,
            "frac": round(tf / FP64_PEAK_TFLOPS, 4), "source": f"profiles/valu_{workload}.md"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default 200; batch 100; precision 500)")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--precision", default=None, choices=["fp64", "mixed"],
                    help="default fp64 for the single shell, mixed for batch and serve")
    ap.add_argument("--nang", type=int, default=None, help="default 64 (precision: 128)")
    ap.add_argument("--nrad", type=int, default=None, help="default 128 (precision: 256)")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="default workload only: skip the C1 / C3 / C4 figures added to the JSON line")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device copy five times (counter calibration)")
    ap.add_argument("--workload", default="lunar", choices=["lunar", "mercury", "batch", "serve", "precision", "ot"],
                    help="lunar: one 6-block shell (default); mercury: the two-fluid Mercury shell (H+, Na+, "
                         "Hall, ambipolar, resistive terms); batch: 64 seeded shells of mixed sizes in one "
                         "context; serve: Poisson request stream with continuous batching; precision: FP64 vs "
                         "mixed on one shell; ot: Orszag-Tang 64^3 periodic box")
    ap.add_argument("--cases", type=int, default=64)
    ap.add_argument("--requests", type=int, default=100)
    ap.add_argument("--req-steps", type=int, default=1000)
    ap.add_argument("--load", type=float, default=0.8, help="offered load of the request stream")
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--mercury-r-in", type=float, default=0.85,
                    help="inner radius of the served Mercury shells (< 1: conducting interior)")
    args = ap.parse_args()
    big = args.workload == "precision"
    args.nang = args.nang or (128 if big else 64)
    args.nrad = args.nrad or (256 if big else 128)
    args.steps = args.steps or {"lunar": 200, "mercury": 200, "batch": 100, "serve": 0, "precision": 500,
                                "ot": 100}[args.workload]
    args.precision = args.precision or ("fp64" if args.workload in ("lunar", "mercury") else "mixed")
    if args.calibrate_copy:
        import torch
        a = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            torch.mul(a, 1.0, out=b)  # one plain elementwise read + write kernel
        torch.cuda.synchronize()
        print("streamed 5 x (1073741824 read + 1073741824 written) bytes", flush=True)
        return
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (the step shards by case or by block across GPUs, "
              "which is not built here); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.workload not in ("lunar", "mercury"):
        out = {"batch": bench_batch, "serve": bench_serve, "precision": bench_precision,
               "ot": bench_ot}[args.workload](args)
        print(json.dumps(out), flush=True)
        return

    from beta1_amd import MercuryDesc, MhdContext, WindDesc

    mercury = args.workload == "mercury"
    desc = (MercuryDesc if mercury else WindDesc)(nang=args.nang, nrad=args.nrad)
    kbytes = kernel_bytes(args.precision)
    # whole-step algorithmic bytes: the second species adds U_Na R/W 80
    step_bytes = STATE_BYTES + GEO_BYTES[args.precision] + (80 if mercury else 0)
    cells = desc.cells
    t0 = time.time()
    ctx = MhdContext(precision=args.precision, recon=0, ghost=2, autotune=True, xcd_order=True)
    cid = ctx.add(desc)
    ctx.finalize()
    setup_s = time.time() - t0
    print(f"[bench] setup {setup_s:.1f}s, {cells} cells, device memory {ctx.memory_bytes() / 1e9:.2f} GB",
          file=sys.stderr, flush=True)

    ctx.step(args.warmup, graph=True)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.step(args.steps, graph=True)
    ctx.sync()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    value = cells * args.steps / elapsed / 1e6
    divb = ctx.divb_max(cid)
    print(f"[bench] {args.steps} steps in {elapsed:.3f}s, {ms_per_step:.3f} ms/step, {value:.1f} Mcell-updates/s, "
          f"div B {divb:.2e}", file=sys.stderr, flush=True)

    # per-kernel timing (HIP events on the library stream, eager launches)
    prof_steps = min(20, max(4, args.steps // 10))
    ctx.profile(True)
    ctx.step(prof_steps)
    prof = ctx.profile_read()
    ctx.profile(False)
    per_kernel = {k: (ms / max(n, 1), n) for k, (ms, n) in prof.items() if n}
    dom = max((k for k in per_kernel if k in kbytes), key=lambda k: per_kernel[k][0] * per_kernel[k][1])
    avg_ms = per_kernel[dom][0]
    alg_bytes = kbytes[dom] * cells
    achieved = alg_bytes / (avg_ms * 1e-3) / 1e9
    kernels_txt = {k: round(v[0], 4) for k, v in per_kernel.items()}
    print(f"[bench] per-kernel avg ms: {kernels_txt}", file=sys.stderr, flush=True)
    print(f"[bench] tuned workgroup rows: {ctx.tuning()}", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu and args.cpu_steps > 0:
        cpu = cpu_baseline(desc, cells, args)

    out = {
        "metric": "Mcell-updates/s + %HBM roofline (FP64 vs mixed/batched); p50/p99 solve latency",
        "value": round(value, 3),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32/f64",
        "data": "synthetic",
        "config": {"workload": f"{'mercury_two_fluid' if mercury else 'lunar_wake'}_{args.precision}_6x{args.nang}x"
                               f"{args.nang}x{args.nrad}",
                   "cells": cells, "blocks": 6, "recon": "first-order", "cfl": desc.cfl},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(f"{args.workload}_{args.precision}_6x{args.nang}x{args.nang}x"
                                                 f"{args.nrad}", dom),
                     "alg_bytes_per_cell": kbytes[dom], "avg_ms": round(avg_ms, 4)},
        "cpu_baseline": cpu,
        "step_hbm": {"bytes_per_cell_update": step_bytes,
                     "frac": round(value * 1e6 * step_bytes / (HBM_PEAK_GBS * 1e9), 4)},
        "fp64": fp64_record(f"{args.workload}_{args.precision}_6x{args.nang}x{args.nang}x{args.nrad}", value)
        if args.precision == "fp64" else None,
    }
    if mercury:
        out["metric"] = "Mcell-updates/s (two-fluid Mercury shell)"
    elif not args.no_extra:
        del ctx
        out["configs"] = extra_configs(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
