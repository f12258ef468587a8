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
# every array the kernel must touch counted once at its stored width.
KERNEL_BYTES = {
    "k_dt": 40 + 24 + 8 + 72,                 # U, B_cell, Jac, JD
    "k_ucontra": 40 + 72 + 8 + 24,            # U, JD, Jac -> UC
    "k_face_emf": 24 + 24 + 24 + 72,          # B, Badd, UC -> Eface
    "k_edge_emf": 72 + 24 + 24 + 24,          # Eface, UC, B -> E
    "k_fluid": 40 + 72 + 8 + 24 + 24 + 40,    # U, JD, Jac, B_cell, J_cell -> U'
    "k_curl_update": 24 + 24 + 24,            # E, B -> B
    # B, Badd, beta, alpha, Wb, Wj, U, JD, Jac -> B_cell, J_cell, UC (+ next step's dt)
    "k_derived": 24 + 24 + 24 + 24 + 144 + 288 + 40 + 72 + 8 + 24 + 24 + 24,
}
HBM_PEAK_GBS = 8000.0


def measured_traffic(workload, kernel):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3 counter
    summary of this workload (profiles/traffic_<workload>.md), or None."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.md")
    if not os.path.exists(path):
        return None
    import re
    for line in open(path):
        cols = [c.strip() for c in line.strip().strip("|").split("|")]
        m = re.search(r"\b(k_\w+)", cols[0]) if cols else None
        if len(cols) >= 5 and m and m.group(1) == kernel:
            try:
                return float(cols[4]) * 1e9
            except ValueError:
                return None
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--nang", type=int, default=64)
    ap.add_argument("--nrad", type=int, default=128)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device copy five times (counter calibration)")
    args = ap.parse_args()
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

    from beta1_amd import MhdContext, WindDesc

    desc = WindDesc(nang=args.nang, nrad=args.nrad)
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
    dom = max((k for k in per_kernel if k in KERNEL_BYTES), key=lambda k: per_kernel[k][0] * per_kernel[k][1])
    avg_ms = per_kernel[dom][0]
    alg_bytes = KERNEL_BYTES[dom] * cells
    achieved = alg_bytes / (avg_ms * 1e-3) / 1e9
    kernels_txt = {k: round(v[0], 4) for k, v in per_kernel.items()}
    print(f"[bench] per-kernel avg ms: {kernels_txt}", file=sys.stderr, flush=True)
    print(f"[bench] tuned workgroup rows: {ctx.tuning()}", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu and args.cpu_steps > 0:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from beta1_amd import HostCase
        from oracle_bind import Oracle

        hc = HostCase(desc, ghost=2)
        orc = Oracle(hc)
        del hc
        orc.step(1)
        t0 = time.perf_counter()
        orc.step(args.cpu_steps)
        cs = time.perf_counter() - t0
        cpu = {"value": round(cells * args.cpu_steps / cs / 1e6, 4), "unit": "Mcell-updates/s", "cores": 1,
               "kind": "port",
               "sample": f"{args.cpu_steps} steps (after 1 untimed) of the same 6x{args.nang}x{args.nang}x{args.nrad} "
                         f"shell, single-threaded FP64 oracle, {cs:.1f}s"}
        print(f"[bench] cpu oracle {cs / args.cpu_steps:.2f} s/step", file=sys.stderr, flush=True)

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
        "config": {"workload": f"lunar_wake_{args.precision}_6x{args.nang}x{args.nang}x{args.nrad}",
                   "cells": cells, "blocks": 6, "recon": "first-order", "cfl": desc.cfl},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(f"lunar_{args.precision}_6x{args.nang}x{args.nang}x{args.nrad}", dom),
                     "alg_bytes_per_cell": KERNEL_BYTES[dom], "avg_ms": round(avg_ms, 4)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
