#!/usr/bin/env python3
"""Throughput of the boundary-element hot path on one MI355X.

Workload: floating hemisphere R = 1 m, 64 x 64 latitude-longitude mesh
(2,048 immersed panels), 64 frequencies log-spaced over 0.3-3.0 rad/s,
Rankine + image + Wu Green function, direct formulation, heave radiation and
one diffraction heading (beta = 0) per frequency, g = 9.81, rho = 1025.
One step = one batched sweep over the 64 frequencies: assembly of S and D,
one LU per frequency, both right-hand sides, forces, and the adjoint
gradients of A, B, F with respect to theta = (radius, omega).  Default
precision: mixed (FP32 LU + FP64 iterative refinement); the same sweep in
FP64 is run once outside the timed region and the largest relative
difference is reported (the metric's bound is 1e-6).  Inputs (mesh, dof
fields, tangents) are resident in HBM before the timed region; outputs are
copied back inside it.

Prints one JSON line: value = freq-solves/s (frequencies x steps / time).
"""
import argparse
import json
import os
import sys
import time

# One hardware queue per HIP stream: the multi-body scheduler runs up to 6
# worker contexts of 3 streams each, and HIP multiplexes streams onto
# GPU_MAX_HW_QUEUES queues (4 by default), which serialises streams that
# share a queue.  Set before HIP initialises (the import of torch), over any
# value in the environment (HIP's default of 4 may be exported there).
os.environ["GPU_MAX_HW_QUEUES"] = "24"

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "marinehydro.jl_amd"))

METRIC = "freq-solves/s incl. adjoint grads at 2k panels, mixed vs FP64 at ≤1e-6 rel err"
PEAK = {"fp64": 78.6, "fp32": 157.3, "fp16": 2500.0}  # MI355X dense matrix peaks, TFLOP/s
# matrix-core kernels of the sweep: timer name -> (peak key, executed products
# per algorithmic complex-GEMM flop).  The mixed LU's trailing updates run on
# fp16 matrix cores as three split products per real product.
MFMA_KERNELS = {"lu_gemm_big": ("fp64", 1), "lu_gemm": ("fp64", 1), "lu_gemm_f32": ("fp32", 1),
                "lu_gemm_f16x3": ("fp16", 3)}
HBM_PEAK = 8000.0  # GB/s


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_threads():
    """Host cores this process may use (the affinity mask, capped by
    OMP_NUM_THREADS when it is set)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit():
        n = min(n, int(env))
    return max(n, 1)


def cpu_solve(O, gfm, mesh, omega, dofs, betas, g, rho, threads):
    """One frequency on the host: the oracle's FP64 assembly (OpenMP) and a
    LAPACK solve (numpy's zgesv, OpenBLAS threads) for every right-hand side
    at once, the arithmetic the reference's `\\` uses; boundary conditions
    and forces as in src/waves.jl (radiation_bc, AiryBC, integrate_pressure)."""
    import numpy as np

    k = omega * omega / g
    S, D = O.assemble(gfm, mesh, k, True, nthreads=threads)
    N, C, ar = mesh.normals, mesh.centers, mesh.areas
    cols = [-1j * omega * (N @ np.asarray(d)) for d in dofs]
    for b in betas:
        cb, sb = np.cos(b), np.sin(b)
        pre = g * k / omega * np.exp(1j * k * (C[:, 0] * cb + C[:, 1] * sb)) * np.exp(k * C[:, 2])
        cols.append(-(pre * cb * N[:, 0] + pre * sb * N[:, 1] - 1j * pre * N[:, 2]))
    phi = np.linalg.solve(D, S @ np.stack(cols, axis=1))
    amp = np.stack([-(N @ np.asarray(d)) * ar for d in dofs], axis=0)  # influenced dof x panel
    Fm = (1j * rho * omega) * (amp @ phi)  # [influenced, rhs]
    nd = len(dofs)
    return Fm[:, :nd].real / omega ** 2, Fm[:, :nd].imag / omega, Fm[:, nd:]


def cpu_baseline(args, mh, mesh, omegas, dofs, betas, g, rho, res, grad):
    """Host FP64 baseline on a bounded sample of the workload, on all host
    cores and on one core, plus the GPU's error against it."""
    import numpy as np
    from threadpoolctl import threadpool_limits

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    gfm = O.RANKINE | O.RANKINE_REFLECTED | O.WU
    threads = host_threads()
    A, B, F = res[:3]
    idx = np.linspace(0, args.nomega - 1, args.cpu_sample).round().astype(int)
    # the host path agrees with the oracle's own end-to-end solve
    Ao, Bo, Fo = O.freq_solve(gfm, mesh, omegas[idx[0]], dofs, betas, direct=True, g=g, rho=rho,
                              nthreads=threads)
    Ac, Bc, Fc = cpu_solve(O, gfm, mesh, omegas[idx[0]], dofs, betas, g, rho, threads)
    assert max(abs(Ac[0, 0] - Ao[0, 0]) / abs(Ao[0, 0]), abs(Fc[0, 0] - Fo[0, 0]) / abs(Fo[0, 0])) < 1e-9
    errs = {"A": 0.0, "B": 0.0, "F": 0.0}
    gerrs = {"dA": 0.0, "dB": 0.0, "dF": 0.0} if grad else None
    h = 1e-6
    with threadpool_limits(limits=threads):
        t0 = time.perf_counter()
        for q in idx:
            Ao, Bo, Fo = cpu_solve(O, gfm, mesh, omegas[q], dofs, betas, g, rho, threads)
            errs["A"] = max(errs["A"], abs(A[q, 0, 0] - Ao[0, 0]) / abs(Ao[0, 0]))
            errs["B"] = max(errs["B"], abs(B[q, 0, 0] - Bo[0, 0]) / abs(Bo[0, 0]))
            errs["F"] = max(errs["F"], abs(F[q, 0, 0] - Fo[0, 0]) / abs(Fo[0, 0]))
            if grad:  # central differences, h = 1e-6, for theta = (radius, omega)
                for m in range(2):
                    vals = []
                    for sgn in (1, -1):
                        if m == 0:
                            mm = mh.hemisphere(1.0 + sgn * h, ntheta=args.ntheta, nphi=args.nphi).mesh
                            om = omegas[q]
                        else:
                            mm, om = mesh, omegas[q] + sgn * h
                        vals.append(cpu_solve(O, gfm, mm, om, dofs, betas, g, rho, threads))
                    fd = [(vals[0][c] - vals[1][c]) / (2 * h) for c in range(3)]
                    for c, key in enumerate(("dA", "dB", "dF")):
                        got = res[3 + c][q, m, 0, 0]
                        gerrs[key] = max(gerrs[key], abs(got - fd[c][0, 0]) / abs(fd[c][0, 0]))
        tc = time.perf_counter() - t0
    solves = len(idx) * (5 if grad else 1)
    # one core: one forward solve, scaled to the same solves per frequency
    with threadpool_limits(limits=1):
        t0 = time.perf_counter()
        cpu_solve(O, gfm, mesh, omegas[idx[0]], dofs, betas, g, rho, 1)
        t1 = time.perf_counter() - t0
    per_freq_1 = t1 * (5 if grad else 1)
    what = "forward + central-FD gradients of (radius, omega), 5 solves per frequency" if grad else "forward"
    cpu = {"value": round(len(idx) / tc, 4), "unit": "freq-solves/s", "cores": threads, "kind": "port",
           "sample": f"{len(idx)} of the {args.nomega} frequencies (indices {idx.tolist()}), same mesh; "
                     f"FP64 oracle assembly (C, OpenMP) + LAPACK zgesv (numpy/OpenBLAS) for all "
                     f"right-hand sides; {what}; {solves} solves in {tc:.1f} s",
           "single_core": {"value": round(1.0 / per_freq_1, 5), "cores": 1,
                           "sample": f"one forward solve ({t1:.1f} s) x {5 if grad else 1} per frequency"}}
    log(f"bench: cpu {cpu['value']} freq-solves/s on {threads} cores, {cpu['single_core']['value']} on 1; "
        f"GPU vs host {errs}, gradients vs host FD {gerrs}")
    return cpu, {k: float(v) for k, v in errs.items()}, \
        ({k: float(v) for k, v in gerrs.items()} if gerrs else None)


def uneven_batch(mh, torch, precision):
    """The uneven multi-body batch: 256 seeded spheres, vertical cylinders
    and boxes of 200-3,000 panels, 16 random frequencies each, heave and
    surge, heading 0, gradients w.r.t. each body's dimensions, in one
    scheduled call (mh_multi_sweep; worker streams, longest body first).
    The autotuner (worker count, LU leaf width) first, on every 8th body in
    order of size (32 bodies spanning the panel counts, the largest
    included, which also sizes the worker contexts' workspaces for the
    largest body); then one timed call on all 256, inputs on the device."""
    items = mh.random_workload(5, 256)
    prep = mh.batches.prepare(items)
    by_size = sorted(items, key=lambda it: -it.mesh.nfaces)
    sample = by_size[::8]
    workers, leaf, tune_ms = mh.multi_autotune(sample, precision)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    mh.multi_sweep(items, precision, prepared=prep)
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    nfs = sum(len(it.omegas) for it in items)
    panels = [it.mesh.nfaces for it in items]
    log(f"bench: uneven batch, {len(items)} bodies ({min(panels)}-{max(panels)} panels), {nfs} freq-solves "
        f"with gradients in {sec:.2f} s ({nfs / sec:.0f}/s)")
    return {"workload": f"{len(items)} seeded bodies (spheres, vertical cylinders, boxes) of {min(panels)}-"
                        f"{max(panels)} panels, 16 frequencies each, heave + surge, heading 0, gradients "
                        "w.r.t. each body's dimensions",
            "freq_solves": nfs, "seconds": round(sec, 3), "freq_solves_per_s": round(nfs / sec, 1),
            "precision": precision, "workers": workers, "leaf_width_cap": leaf}


def shape_modes(U, count=16):
    """Radial shape modes of a sphere's vertices: unit directions U times
    low-order monomials of U (1, x, y, z, x^2, ...), `count` of them."""
    import numpy as np

    x, y, z = U[:, 0], U[:, 1], U[:, 2]
    f = [np.ones_like(x), x, y, z, x * x, y * y, z * z, x * y, x * z, y * z, x ** 3, y ** 3, z ** 3,
         x * y * z, x * x * z, y * y * z]
    return [U * fm[:, None] for fm in f[:count]]


def many_parameters(mh, torch, args, omegas, dofs, betas, g, rho, nparam=16):
    """The same sweep with gradients w.r.t. nparam shape parameters (radial
    modes of the hemisphere's vertices) against the FP64 forward + central
    finite-difference sweep that the adjoint replaces (2 nparam + 1 forward
    sweeps on meshes already on the device), and the adjoint gradients'
    difference to those finite differences (h = 1e-6)."""
    import numpy as np

    body = mh.hemisphere(1.0, ntheta=args.ntheta, nphi=args.nphi)
    mesh = body.mesh
    modes = shape_modes(body.dvertices["radius"], nparam)
    T = mh.shape_tangents(mesh, modes)
    dm = mh.DeviceMesh(mesh)
    dofs_dev = mh.device_dof_fields(mesh, dofs)
    T_dev = mh.device_tangents(T, mesh, len(dofs))

    def adj():
        return mh.freq_sweep_grad(dm, omegas, dofs_dev, betas, T_dev, precision=args.precision, g=g, rho=rho)

    adj()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(2):
        res = adj()
    torch.cuda.synchronize()
    adj_ms = (time.perf_counter() - t0) / 2 * 1e3
    h = 1e-6
    V0, F = mesh.vertices, mesh.faces
    pert = [mh.DeviceMesh(mh.mesh_from_vertices(V0 + sg * h * dv, F)) for dv in modes for sg in (1, -1)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    mh.freq_sweep(dm, omegas, dofs_dev, betas, precision="fp64", g=g, rho=rho)
    outs = [mh.freq_sweep(pm, omegas, dofs_dev, betas, precision="fp64", g=g, rho=rho) for pm in pert]
    torch.cuda.synchronize()
    fd_ms = (time.perf_counter() - t0) * 1e3
    err = {}
    for c, key in enumerate(("dA", "dB", "dF")):
        fd = np.stack([(outs[2 * m][c] - outs[2 * m + 1][c]) / (2 * h) for m in range(nparam)], axis=1)
        got = res[3 + c][:, :nparam]
        err[key] = float(np.max(np.abs(got - fd)) / max(float(np.max(np.abs(fd))), 1e-300))
    log(f"bench: {nparam} shape parameters: adjoint sweep {adj_ms:.1f} ms, FP64 forward + central-FD sweep "
        f"({2 * nparam + 1} forward sweeps) {fd_ms:.1f} ms ({fd_ms / adj_ms:.1f}x); adjoint vs FD " +
        str({k: f"{v:.1e}" for k, v in err.items()}))
    return {"parameters": f"{nparam} radial shape modes of the hemisphere's vertices (unit direction x "
                          "monomials of degree <= 3)",
            "adjoint_sweep_ms": round(adj_ms, 3), "fp64_forward_fd_sweep_ms": round(fd_ms, 3),
            "forward_sweeps_in_fd": 2 * nparam + 1, "speedup": round(fd_ms / adj_ms, 3),
            "rel_err_adjoint_vs_fd": err}


def delhommeau_assembly(mh, torch, np):
    """S and D of the 4,000-panel cylinder at 32 wavenumbers (omega 0.2-2.5
    rad/s), exact and tabulated Delhommeau term, direct formulation: wall
    time of each assembly (after one warm-up call that also builds the
    table) and the largest difference relative to the largest entry."""
    c = mh.vertical_cylinder(5.0, 10.0, ntheta=80, nz=30, nr=20).mesh
    dm = mh.DeviceMesh(c)
    ks = np.linspace(0.2, 2.5, 32) ** 2 / 9.81
    out = {}
    for name, wave in (("exact", mh.ExactGuevelDelhommeau()), ("table", mh.TabulatedDelhommeau())):
        gfs = (mh.Rankine(), mh.RankineReflected(), wave)
        mh.assemble_matrices_batched(gfs, dm, ks[:2])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        S, D = mh.assemble_matrices_batched(gfs, dm, ks)
        torch.cuda.synchronize()
        out[name] = (time.perf_counter() - t0, S, D)
        log(f"bench: Delhommeau ({name}) assembly, 4000 panels x 32 wavenumbers: {out[name][0] * 1e3:.1f} ms")
    (te, Se, De), (tt, St, Dt) = out["exact"], out["table"]
    eS = (St - Se).abs().max().item() / Se.abs().max().item()
    eD = (Dt - De).abs().max().item() / De.abs().max().item()
    log(f"bench: table vs exact: {te / tt:.1f}x faster; max difference S {eS:.2e}, D {eD:.2e} of the largest entry")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precision", default="mixed", choices=["fp64", "mixed"])
    ap.add_argument("--ntheta", type=int, default=64)
    ap.add_argument("--nphi", type=int, default=64)
    ap.add_argument("--nomega", type=int, default=64)
    ap.add_argument("--cpu-sample", type=int, default=4,
                    help="frequencies timed on the host oracle (0 = skip)")
    ap.add_argument("--no-grad", action="store_true",
                    help="forward sweep only (no adjoint gradients)")
    ap.add_argument("--profile-only", action="store_true",
                    help="one sweep with the kernel timer on (serialised schedule), no timing loop; "
                         "for rocprofv3 runs")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only a 1 GiB device-to-device copy (counter calibration for rocprofv3 --pmc)")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the uneven multi-body batch (256 bodies of 200-3,000 panels)")
    ap.add_argument("--no-many", action="store_true",
                    help="skip the 16-parameter adjoint vs finite-difference comparison")
    ap.add_argument("--batch-only", action="store_true",
                    help="only the uneven multi-body batch (autotuner, then one timed call); for rocprofv3 runs")
    ap.add_argument("--delhommeau-assembly", action="store_true",
                    help="only the assembly of the 4,000-panel cylinder (R 5 m, draft 10 m) at 32 wavenumbers "
                         "with the exact and the tabulated Delhommeau term (timings and their difference)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the frequency sweep is "
              "independent per (body, frequency) and would shard by frequency, but multi-GPU runs "
              "are not implemented", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import marinehydro_amd as mh

    if not torch.cuda.is_available():
        log("no GPU visible")
        sys.exit(1)
    if args.copy_calibration:
        nbytes = 1 << 30
        src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"bench: copied {nbytes} bytes 3 times")
        return
    if args.delhommeau_assembly:
        delhommeau_assembly(mh, torch, np)
        return
    if args.batch_only:
        log("bench: " + json.dumps(uneven_batch(mh, torch, args.precision)))
        return
    rho, g = 1025.0, 9.81
    body = mh.hemisphere(1.0, ntheta=args.ntheta, nphi=args.nphi)
    mesh = body.mesh
    n = mesh.nfaces
    omegas = 0.3 * 10.0 ** (np.arange(args.nomega) / max(args.nomega - 1, 1))
    dofs = [np.array([0.0, 0.0, 1.0])]
    betas = [0.0]
    dm = mh.DeviceMesh(mesh)
    grad = not args.no_grad
    # gradient parameters: theta = (radius, omega)
    T = mh.shape_tangents(mesh, [body.dvertices["radius"], None], nomega=args.nomega, omega_param=1)
    # dof fields and tangents resident on the device before the timed region
    dofs_dev = mh.device_dof_fields(mesh, dofs)
    T_dev = mh.device_tangents(T, mesh, len(dofs))
    log(f"bench: N={n} panels, {args.nomega} frequencies, precision={args.precision}, "
        f"adjoint gradients={'on, theta=(radius, omega)' if grad else 'off'}")

    def forward(prec=None):
        return mh.freq_sweep(dm, omegas, dofs_dev, betas, precision=prec or args.precision, g=g, rho=rho)

    def step():
        if grad:
            return mh.freq_sweep_grad(dm, omegas, dofs_dev, betas, T_dev, precision=args.precision, g=g,
                                      rho=rho)
        return forward()

    if args.profile_only:
        mh._lib.profile_enable(True)
        step()
        torch.cuda.synchronize()
        stats = mh._lib.profile_report()
        mh._lib.profile_enable(False)
        log("bench: profile-only sweep done; kernel timer (ms, launches, avg ms): " +
            str({k: (round(v["ms"], 3), v["launches"], round(v["ms"] / v["launches"], 4))
                 for k, v in sorted(stats.items(), key=lambda kv: -kv[1]["ms"])}))
        return

    for w in range(args.warmup):
        step()
        log(f"bench: warmup {w + 1}/{args.warmup} done")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        res = step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    value = args.nomega / dt
    A, B, F = res[:3]
    log(f"bench: {dt * 1e3:.1f} ms per sweep, {value:.1f} freq-solves/s")
    # the FP64 forward + central finite-difference sweep (2 |theta| + 1
    # forward sweeps) that the adjoint path replaces
    fd_ms = None
    if grad:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(2 * T.nparam + 1):
            forward("fp64")
        torch.cuda.synchronize()
        fd_ms = (time.perf_counter() - t0) * 1e3
        log(f"bench: FP64 forward + central-FD sweep {fd_ms:.1f} ms "
            f"({fd_ms / (dt * 1e3):.2f}x the timed step)")

    # the same sweep in FP64 (outside the timed region): mixed vs FP64 error
    rel_fp64 = None
    if args.precision != "fp64":
        ref = mh.freq_sweep_grad(dm, omegas, dofs_dev, betas, T_dev, precision="fp64", g=g, rho=rho) if grad \
            else forward("fp64")
        # per quantity: max |mixed - fp64| / max |fp64| over all frequencies
        rel_fp64 = {}
        for name, a, b in zip(("A", "B", "F", "dA", "dB", "dF"), res, ref):
            rel_fp64[name] = float(np.max(np.abs(np.asarray(a) - np.asarray(b))) /
                                   max(float(np.max(np.abs(np.asarray(b)))), 1e-300))
        log(f"bench: rel difference {args.precision} vs FP64 per quantity " +
            str({k: f"{v:.1e}" for k, v in rel_fp64.items()}))

    # the same sweep with the Rankine and image far field (pairs beyond 7
    # panel radii) assembled in FP32 (mh_set_far_field_fp32): its time and
    # its difference to the FP64 sweep per quantity
    far32 = None
    if rel_fp64 is not None:
        try:
            mh._lib.set_far_field_fp32(True)
            step()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            res32 = step()
            torch.cuda.synchronize()
            far_ms = (time.perf_counter() - t0) * 1e3
        finally:
            mh._lib.set_far_field_fp32(False)
        far32 = {"ms_per_step": round(far_ms, 3),
                 "rel_diff_vs_fp64": {name: float(np.max(np.abs(np.asarray(a) - np.asarray(b))) /
                                                  max(float(np.max(np.abs(np.asarray(b)))), 1e-300))
                                      for name, a, b in zip(("A", "B", "F", "dA", "dB", "dF"), res32, ref)}}
        log(f"bench: FP32 far-field assembly: {far_ms:.1f} ms per sweep, rel difference vs FP64 " +
            str({k: f"{v:.1e}" for k, v in far32["rel_diff_vs_fp64"].items()}))

    # kernel timer over one more sweep (the library serialises its two streams
    # while the timer is on, so each duration is the kernel's own): dominant
    # kernel and its roofline
    mh._lib.profile_enable(True)
    step()
    stats = mh._lib.profile_report()
    mh._lib.profile_enable(False)
    # The roofline is the dominant matrix-core kernel's (the LU trailing
    # update); the pair kernels are FP64 VALU-bound with no algorithmic flop
    # count (Bessel, log, exp), so the longest-running kernel is named and,
    # when it is one of them, its VALU issue utilisation from rocprofv3
    # counters is attached (profiles/valu.json).
    top_name, top = max(stats.items(), key=lambda kv: kv[1]["ms"])
    mfma = {k: v for k, v in stats.items() if k in MFMA_KERNELS and v["flops"] > 0}
    dom_name, dom = max(mfma.items(), key=lambda kv: kv[1]["ms"])
    avg_ms = dom["ms"] / dom["launches"]
    pkey, mult = MFMA_KERNELS[dom_name]
    # The roofline model: the kernel's arithmetic intensity (algorithmic
    # complex-GEMM flops, 8 per complex multiply-add, over its algorithmic HBM
    # bytes: C read and written once, the operands once) against the ridge
    # point of the unit that executes it (its peak in algorithmic flops over
    # 8 TB/s).  Below the ridge the kernel is bound by HBM and its roofline is
    # algorithmic bytes per second against 8 TB/s; above it, algorithmic
    # flops per second against the unit's peak.  Both views are reported.
    # For the split-fp16 kernel the unit is the fp16 matrix core, whose peak
    # in algorithmic flops is a third of 2.5 PFLOP/s (three fp16 products per
    # real product).
    sec = dom["ms"] * 1e-3
    alg = dom["flops"] / sec / 1e12  # TFLOP/s
    gbs = dom["bytes"] / sec / 1e9  # GB/s
    intensity = dom["flops"] / dom["bytes"]
    ridge = PEAK[pkey] / mult * 1e12 / (HBM_PEAK * 1e9)
    mfma_view = {"achieved_tflops": round(alg, 3), "peak_tflops": PEAK[pkey], "frac": round(alg / PEAK[pkey], 4),
                 "frac_of_algorithmic_peak": round(alg / (PEAK[pkey] / mult), 4)}
    hbm_view = {"achieved_gbs": round(gbs, 1), "peak_gbs": HBM_PEAK, "frac": round(gbs / HBM_PEAK, 4)}
    if intensity < ridge:
        roof = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK, "unit": "GB/s",
                "frac": round(gbs / HBM_PEAK, 4), "traffic": None}
    else:
        roof = {"bound": "mfma", "achieved": round(alg, 3), "peak": PEAK[pkey], "unit": "TFLOP/s",
                "frac": round(alg / PEAK[pkey], 4), "traffic": None}
    roof["kernel"] = dom_name
    roof["avg_launch_ms"] = round(avg_ms, 4)
    roof["launches_per_step"] = dom["launches"]
    roof["algorithmic_flops_per_launch"] = round(dom["flops"] / dom["launches"], 1)
    roof["algorithmic_bytes_per_launch"] = round(dom["bytes"] / dom["launches"], 1)
    roof["arithmetic_intensity_flop_per_byte"] = round(intensity, 2)
    roof["ridge_flop_per_byte"] = round(ridge, 2)
    roof["mfma_view"] = mfma_view
    roof["hbm_view"] = hbm_view
    if mult != 1:
        roof["frac_of_fp32_matrix_peak"] = round(alg / PEAK["fp32"], 4)
        roof["issued_fp16_tflops"] = round(mult * alg, 3)
        roof["note"] = (f"FP32 operands split into fp16 halves, {mult} fp16 products per real product; "
                        "achieved flops count algorithmic flops only")
    roof["longest_kernel"] = {"name": top_name, "ms_per_step": round(top["ms"], 3)}
    vpath = os.path.join(ROOT, "profiles", "valu.json")
    if top_name != dom_name and os.path.exists(vpath):
        with open(vpath) as f:
            vj = json.load(f)
        ent = vj.get(f"{top_name}|{args.precision}|{n}x{args.nomega}")
        if ent:
            roof["longest_kernel"]["valu_issue_utilisation"] = ent["valu_issue_utilisation"]
            for key in ("valu_active_per_wave", "waves_per_simd"):
                if key in ent:
                    roof["longest_kernel"][key] = ent[key]
            roof["longest_kernel"]["source"] = ent["summary"]
    # HBM bytes per launch from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of
    # the same workload, scaled by the counters' reading for a copy of known
    # size (profiles/traffic.json, written from those runs)
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tr = json.load(f)
        ent = tr.get(f"{dom_name}|{args.precision}|{n}x{args.nomega}")
        if ent:
            roof["traffic"] = ent["bytes_per_launch"]
    # matrix writes of the assembly, live: S and D in FP64 (+ D in FP32 on
    # the mixed path) written once per frequency, over the kernel's time
    asm_write = None
    if "assemble" in stats:
        wbytes = (32.0 + (8.0 if args.precision == "mixed" else 0.0)) * n * n * args.nomega
        asm_write = {"bytes": wbytes, "ms": round(stats["assemble"]["ms"], 3),
                     "tb_per_s": round(wbytes / (stats["assemble"]["ms"] * 1e-3) / 1e12, 3),
                     "frac_of_hbm": round(wbytes / (stats["assemble"]["ms"] * 1e-3) / (HBM_PEAK * 1e9), 3)}
    kernel_ms = {k: round(v["ms"], 3) for k, v in sorted(stats.items(), key=lambda kv: -kv[1]["ms"])}
    log(f"bench: kernel ms per sweep {kernel_ms}")

    many = many_parameters(mh, torch, args, omegas, dofs, betas, g, rho) if grad and not args.no_many else None

    batch = None if args.no_batch else uneven_batch(mh, torch, args.precision)

    cpu = None
    oracle_err = None
    grad_err = None
    if args.cpu_sample > 0:
        cpu, oracle_err, grad_err = cpu_baseline(args, mh, mesh, omegas, dofs, betas, g, rho, res, grad)
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
                   "adjoint_gradients": "theta=(radius, omega)" if grad else None},
        "roofline": roof,
        "cpu_baseline": cpu,
        "max_rel_err_vs_oracle": max(oracle_err.values()) if oracle_err else None,
        "rel_err_vs_oracle": oracle_err,
        "max_rel_diff_vs_fp64": max(rel_fp64.values()) if rel_fp64 else None,
        "rel_diff_vs_fp64": rel_fp64,
        "max_rel_err_grad_vs_oracle_fd": max(grad_err.values()) if grad_err else None,
        "rel_err_grad_vs_oracle_fd": grad_err,
        "fp64_forward_fd_sweep_ms": round(fd_ms, 3) if fd_ms else None,
        "speedup_vs_fp64_fd_sweep": round(fd_ms / (dt * 1e3), 3) if fd_ms else None,
        "many_parameters": many,
        "far_field_fp32": far32,
        "assembly_matrix_writes": asm_write,
        "kernel_ms_per_step": kernel_ms,
        "uneven_batch": batch,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
