"""Uneven multi-body batch (config 5): 256 seeded spheres, vertical
cylinders and boxes of 200-3,000 panels, 16 random frequencies each, heave
and surge, heading 0, gradients w.r.t. each body's dimensions, in one
scheduled call (mh_multi_sweep); mixed precision against FP64 for every
body, the oracle for the smallest bodies (forward and central-difference
gradients), and the single-body sweep for two bodies."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GF_WU = 1 | 2 | 8


def relmax(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def workload(mh, gpu):
    items = mh.random_workload(5, 256)
    prep = mh.batches.prepare(items)
    return items, prep


def test_config5_mixed_against_fp64_all_bodies(mh, gpu, workload):
    items, prep = workload
    t0 = time.time()
    mixed = mh.multi_sweep(items, "mixed", prepared=prep)
    t1 = time.time()
    fp64 = mh.multi_sweep(items, "fp64", prepared=prep)
    t2 = time.time()
    nf = sum(len(it.omegas) for it in items)
    print(f"config5: {len(items)} bodies, {nf} freq-solves with gradients: mixed {t1 - t0:.2f} s "
          f"({nf / (t1 - t0):.0f}/s), fp64 {t2 - t1:.2f} s", flush=True)
    worst = {k: 0.0 for k in ("A", "B", "F", "dA", "dB", "dF")}
    for m_, f_ in zip(mixed, fp64):
        for k, a, b in zip(worst, m_, f_):
            worst[k] = max(worst[k], relmax(a, b))
    print("config5 mixed vs fp64 per quantity:", {k: f"{v:.1e}" for k, v in worst.items()}, flush=True)
    assert max(worst.values()) < 1e-6


def test_config5_against_oracle_and_single_sweeps(mh, oracle, gpu, workload):
    items, prep = workload
    res = mh.multi_sweep(items, "fp64", prepared=prep)
    order = np.argsort([it.mesh.nfaces for it in items])
    for q in order[:3]:  # the three smallest bodies, every frequency, forward
        it = items[q]
        A, B, F = res[q][:3]
        for w in range(0, len(it.omegas), 5):
            Ao, Bo, Fo = oracle.freq_solve(GF_WU, it.mesh, it.omegas[w], it.dof_fields, list(it.betas))
            assert relmax(A[w], Ao) < 1e-9 and relmax(B[w], Bo) < 1e-9 and relmax(F[w, 0], Fo[:, 0]) < 1e-9
    # gradients of the smallest body against central differences (h = 1e-6)
    it = items[order[0]]
    params, rebuild = it.tag["params"], it.tag["rebuild"]
    dA, dB, dF = res[order[0]][3:]
    h = 1e-6
    for m, k in enumerate(params):
        for w in (0, len(it.omegas) - 1):
            outs = []
            for s in (+1, -1):
                p = dict(params)
                p[k] += s * h
                outs.append(oracle.freq_solve(GF_WU, rebuild(p).mesh, it.omegas[w], it.dof_fields, list(it.betas)))
            fA, fB, fF = [(outs[0][c] - outs[1][c]) / (2 * h) for c in range(3)]
            print(f"config5 body {order[0]} ({it.tag['kind']}, {it.mesh.nfaces} panels) d/d{k} omega "
                  f"{it.omegas[w]:.2f}: dA {relmax(dA[w, m], fA):.1e} dB {relmax(dB[w, m], fB):.1e} "
                  f"dF {relmax(dF[w, m, 0], fF[:, 0]):.1e}", flush=True)
            assert relmax(dA[w, m], fA) < 1e-6 and relmax(dB[w, m], fB) < 1e-6
            assert relmax(dF[w, m, 0], fF[:, 0]) < 1e-6
    # the scheduler gives what one sweep per body gives
    for q in (order[0], order[-1]):
        it = items[q]
        one = mh.freq_sweep_grad(it.mesh, it.omegas, it.dof_fields, it.betas, it.tangents, precision="fp64")
        for a, b in zip(res[q], one):
            assert relmax(a, b) < 1e-12


def test_config5_autotune_on_a_sample(mh, gpu, workload):
    # on a stream of its own, so the tuned configuration stays with that
    # stream's library context and the other tests keep the defaults
    torch = gpu
    items, prep = workload
    sample = list(range(0, len(items), 16))  # 16 bodies across the size range
    with torch.cuda.stream(torch.cuda.Stream()):
        w, leaf, ms = mh.multi_autotune([items[i] for i in sample], "mixed",
                                        prepared=[prep[i] for i in sample])
        print(f"config5 autotune: {w} workers, leaf cap {leaf}, {ms:.1f} ms for the sample", flush=True)
        assert w in (1, 2, 3, 4, 6) and leaf in (0, 8)
        t0 = time.time()
        tuned = mh.multi_sweep(items, "mixed", prepared=prep)  # nworkers = 0: the tuned configuration
        dt = time.time() - t0
    nf = sum(len(it.omegas) for it in items)
    print(f"config5 tuned: {dt:.2f} s, {nf / dt:.0f} freq-solves/s with gradients", flush=True)
    ref = mh.multi_sweep(items, "fp64", prepared=prep)
    assert max(relmax(a, b) for t, r in zip(tuned, ref) for a, b in zip(t, r)) < 1e-6
