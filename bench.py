#!/usr/bin/env python3
"""Throughput of the boundary-element hot path on one MI355X.

Workload: floating hemisphere R = 1 m, 64 x 64 latitude-longitude mesh
(2,048 immersed panels), 64 frequencies log-spaced over 0.3-3.0 rad/s,
Rankine + image + Wu Green function, direct formulation, heave radiation and
one diffraction heading (beta = 0) per frequency, g = 9.81, rho = 1025.
One step = one batched sweep over the 64 frequencies: assembly of S and D,
one LU per frequency, both right-hand sides, forces.  Inputs (mesh, dof
fields) are resident in HBM before the timed region; outputs are copied back
inside it.

Prints one JSON line: value = freq-solves/s (frequencies x steps / time).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "marinehydro.jl_amd"))

METRIC = "freq-solves/s incl. adjoint grads at 2k panels, mixed vs FP64 at ≤1e-6 rel err"
PEAK = {"fp64": 78.6, "fp32": 157.3}  # MI355X dense matrix peaks, TFLOP/s
HBM_PEAK = 8000.0  # GB/s


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--ntheta", type=int, default=64)
    ap.add_argument("--nphi", type=int, default=64)
    ap.add_argument("--nomega", type=int, default=64)
    ap.add_argument("--cpu-sample", type=int, default=2,
                    help="frequencies timed on the host oracle (0 = skip)")
    ap.add_argument("--profile-only", action="store_true",
                    help="one sweep, no timing loop (for rocprofv3 runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the frequency sweep is "
              "independent per (body, frequency) and would shard by frequency, but multi-GPU runs "
              "are not part of this build", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import marinehydro_amd as mh

    if not torch.cuda.is_available():
        log("no GPU visible")
        sys.exit(1)
    rho, g = 1025.0, 9.81
    body = mh.hemisphere(1.0, ntheta=args.ntheta, nphi=args.nphi)
    mesh = body.mesh
    n = mesh.nfaces
    omegas = 0.3 * 10.0 ** (np.arange(args.nomega) / max(args.nomega - 1, 1))
    dofs = [np.array([0.0, 0.0, 1.0])]
    betas = [0.0]
    dm = mh.DeviceMesh(mesh)
    log(f"bench: N={n} panels, {args.nomega} frequencies, precision={args.precision}")

    def step():
        return mh.freq_sweep(dm, omegas, dofs, betas, precision=args.precision, g=g, rho=rho)

    if args.profile_only:
        step()
        torch.cuda.synchronize()
        log("bench: profile-only sweep done")
        return

    for w in range(args.warmup):
        step()
        log(f"bench: warmup {w + 1}/{args.warmup} done")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        A, B, F = step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    value = args.nomega / dt
    log(f"bench: {dt * 1e3:.1f} ms per sweep, {value:.1f} freq-solves/s")

    # kernel timer over one more sweep: dominant kernel and its roofline
    mh._lib.profile_enable(True)
    step()
    stats = mh._lib.profile_report()
    mh._lib.profile_enable(False)
    dom_name, dom = max(stats.items(), key=lambda kv: kv[1]["ms"])
    avg_ms = dom["ms"] / dom["launches"]
    if dom["flops"] > 0:
        achieved = dom["flops"] / (dom["ms"] * 1e-3) / 1e12
        peak = PEAK["fp64" if args.precision == "fp64" else "fp32"]
        roof = {"bound": "mfma", "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
                "frac": round(achieved / peak, 4), "traffic": None}
    else:
        achieved = dom["bytes"] / (dom["ms"] * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK, 4), "traffic": None}
    roof["kernel"] = dom_name
    roof["avg_launch_ms"] = round(avg_ms, 4)
    roof["launches_per_step"] = dom["launches"]
    kernel_ms = {k: round(v["ms"], 3) for k, v in sorted(stats.items(), key=lambda kv: -kv[1]["ms"])}
    log(f"bench: kernel ms per sweep {kernel_ms}")

    cpu = None
    max_err = None
    if args.cpu_sample > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O

        threads = min(16, os.cpu_count() or 1)
        idx = np.linspace(0, args.nomega - 1, args.cpu_sample).round().astype(int)
        errs = []
        t0 = time.perf_counter()
        for q in idx:
            Ao, Bo, Fo = O.freq_solve(O.RANKINE | O.RANKINE_REFLECTED | O.WU, mesh, omegas[q], dofs,
                                      betas, direct=True, g=g, rho=rho, nthreads=threads)
            errs += [abs(A[q, 0, 0] - Ao[0, 0]) / abs(Ao[0, 0]),
                     abs(B[q, 0, 0] - Bo[0, 0]) / abs(Bo[0, 0]),
                     abs(F[q, 0, 0] - Fo[0, 0]) / abs(Fo[0, 0])]
        tc = time.perf_counter() - t0
        max_err = float(max(errs))
        cpu = {"value": round(len(idx) / tc, 4), "unit": "freq-solves/s", "cores": threads,
               "kind": "port",
               "sample": f"{len(idx)} of the {args.nomega} frequencies (indices {idx.tolist()}), "
                         f"same mesh, FP64 C oracle with OpenMP ({tc:.1f} s)"}
        log(f"bench: cpu oracle {cpu['value']} freq-solves/s on {threads} threads; "
            f"max rel err GPU vs oracle {max_err:.2e}")

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "freq-solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"hemisphere R=1, {args.ntheta}x{args.nphi} lat-long mesh "
                               f"({n} panels), {args.nomega} frequencies 0.3-3.0 rad/s, "
                               "Wu GF, direct, heave + 1 heading",
                   "panels": n, "frequencies": args.nomega, "precision": args.precision,
                   "adjoint_gradients": False},
        "roofline": roof,
        "cpu_baseline": cpu,
        "max_rel_err_vs_oracle": max_err,
        "kernel_ms_per_step": kernel_ms,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
