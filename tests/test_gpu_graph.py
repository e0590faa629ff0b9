"""Captured FP64 sweeps (mh_sweep_graph_*): every replay against the eager
sweep of the same inputs (the same kernels in the same order), with eager
sweeps on the shared context between the replays; a body reshaped in place
between launches; a bad mesh index written in place, reported after the
launch (the packing kernel clamps the index and raises a flag: no fault);
and, printed with -s, the time of a replay against an eager sweep."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def close(got, ref, name):
    scale = max(float(np.abs(ref).max()), 1e-300)
    err = float(np.abs(np.asarray(got) - np.asarray(ref)).max()) / scale
    assert err < 1e-12, (name, err)


@pytest.mark.parametrize("direct", [True, False])
def test_graph_replays_match_eager_sweeps(mh, gpu, direct):
    body = mh.hemisphere(1.0, ntheta=20, nphi=20)
    assert body.mesh.nfaces == 200
    dm = mh.DeviceMesh(body.mesh)
    dofs = [[0, 0, 1.0], [1.0, 0, 0]]
    graph = mh.SweepGraph(dm, 6, dofs, nbeta=2, direct=direct)
    rng = np.random.default_rng(7)
    try:
        for rep in range(4):
            om = np.sort(rng.uniform(0.3, 3.0, 6))
            be = rng.uniform(0.0, np.pi, 2)
            got = graph.launch(om, be)
            ref = mh.freq_sweep(dm, om, dofs, be, direct=direct, precision="fp64")
            for g_, r_, name in zip(got, ref, "ABF"):
                close(g_, r_, f"{name} replay {rep}")
    finally:
        graph.close()


def test_graph_follows_a_body_reshaped_in_place(mh, gpu):
    small = mh.hemisphere(1.0, ntheta=16, nphi=16).mesh
    large = mh.hemisphere(1.3, ntheta=16, nphi=16).mesh
    dm = mh.DeviceMesh(small)
    om = [0.5, 1.0, 1.5, 2.0]
    graph = mh.SweepGraph(dm, len(om), [[0, 0, 1.0]], nbeta=1)
    try:
        A1, B1, F1 = graph.launch(om, [0.0])
        ref1 = mh.freq_sweep(small, om, [[0, 0, 1.0]], [0.0])
        for g_, r_, name in zip((A1, B1, F1), ref1, "ABF"):
            close(g_, r_, name)
        other = mh.DeviceMesh(large)
        for f in ("vertices", "centers", "normals", "areas", "radii"):
            getattr(dm, f).copy_(getattr(other, f))
        A2, B2, F2 = graph.launch(om, [0.0])
        ref2 = mh.freq_sweep(large, om, [[0, 0, 1.0]], [0.0])
        for g_, r_, name in zip((A2, B2, F2), ref2, "ABF"):
            close(g_, r_, name)
        assert np.abs(A2 - A1).max() > 0.1 * np.abs(A1).max()  # the larger body, not a stale result
    finally:
        graph.close()


def test_graph_reports_a_bad_mesh_index_after_the_launch(mh, gpu):
    mesh = mh.hemisphere(1.0, ntheta=10, nphi=10).mesh
    dm = mh.DeviceMesh(mesh)
    graph = mh.SweepGraph(dm, 2, [[0, 0, 1.0]])
    try:
        good = dm.faces[0, 3].item()
        dm.faces[0, 3] = mesh.nvertices + 5
        with pytest.raises(ValueError, match="vertex index"):
            graph.launch([0.7, 1.4])
        dm.faces[0, 3] = good
        A, B, _ = graph.launch([0.7, 1.4])
        ref = mh.freq_sweep(mesh, [0.7, 1.4], [[0, 0, 1.0]])
        close(A, ref[0], "A")
        close(B, ref[1], "B")
    finally:
        graph.close()


@pytest.mark.parametrize("ntheta,nomega", [(20, 16), (40, 16)])
def test_graph_replay_time(mh, gpu, ntheta, nomega):
    import torch

    body = mh.hemisphere(1.0, ntheta=ntheta, nphi=ntheta)
    dm = mh.DeviceMesh(body.mesh)
    om = np.linspace(0.3, 3.0, nomega)
    dofs = [[0, 0, 1.0]]
    graph = mh.SweepGraph(dm, nomega, dofs, nbeta=1)
    try:
        for _ in range(2):
            graph.launch(om, [0.0])
            mh.freq_sweep(dm, om, dofs, [0.0])
        torch.cuda.synchronize()
        reps = 10
        t0 = time.perf_counter()
        for _ in range(reps):
            mh.freq_sweep(dm, om, dofs, [0.0])
        t1 = time.perf_counter()
        for _ in range(reps):
            graph.launch(om, [0.0])
        t2 = time.perf_counter()
        eager, replay = (t1 - t0) / reps * 1e3, (t2 - t1) / reps * 1e3
        print(f"\n{body.mesh.nfaces} panels x {nomega} frequencies, FP64 forward sweep: eager {eager:.3f} ms, "
              f"graph replay {replay:.3f} ms ({eager / replay:.2f}x)")
    finally:
        graph.close()


# ---------------------------------------------------------------------------
# captured batches (mh_batch_graph_*): many bodies in one graph
# ---------------------------------------------------------------------------
GRAPH_SKIP = ("captured mixed-precision and gradient sweeps are switched off: a one-body mixed "
              "sweep with gradients, captured without error, crashed the host process inside its first "
              "launch (twice, the second time with kernel copies in place of its memcpy nodes); the "
              "cause was not found and mh_batch_graph_create refuses such sweeps")

def _grad_item(mh, ntheta, omegas, radius=1.0):
    body = mh.hemisphere(radius, ntheta=ntheta, nphi=ntheta)
    T = mh.shape_tangents(body.mesh, [body.dvertices["radius"], None], nomega=len(omegas), omega_param=1)
    return mh.SweepItem(body.mesh, np.asarray(omegas, dtype=float), [[0, 0, 1.0], [1.0, 0, 0]], (0.0, 0.7), T)


@pytest.mark.skip(reason=GRAPH_SKIP)
@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_batch_graph_one_body_with_gradients(mh, gpu, precision):
    """One body, gradients w.r.t. (radius, omega): every replay (new
    frequencies and headings, eager sweeps on the shared context between
    them) against the eager sweep of the same inputs; no item runs again."""
    rng = np.random.default_rng(11)
    om0 = np.sort(rng.uniform(0.4, 2.5, 8))
    it = _grad_item(mh, 20, om0)
    print(f"\nbatch graph, one body: {it.mesh.nfaces} panels, {len(om0)} frequencies, {precision}", flush=True)
    graph = mh.BatchGraph([it], precision, nworkers=1)
    try:
        nodes, reruns, workers = graph.info()
        assert nodes > 0 and reruns == 0 and workers == 1
        for rep in range(3):
            om = np.sort(rng.uniform(0.4, 2.5, 8))
            be = rng.uniform(0.0, np.pi, 2)
            got = graph.launch([om], [be])[0]
            ref = mh.freq_sweep_grad(it.mesh, om, it.dof_fields, be, it.tangents, precision=precision)
            for g_, r_, name in zip(got, ref, ("A", "B", "F", "dA", "dB", "dF")):
                close(g_, r_, f"{name} replay {rep}")
        assert graph.info()[1] == 0
    finally:
        graph.close()


@pytest.mark.skip(reason=GRAPH_SKIP)
def test_batch_graph_uneven_bodies(mh, gpu):
    """Five seeded bodies of 150-500 panels (spheres, cylinders, boxes),
    four frequencies each, heave + surge, gradients w.r.t. their dimensions,
    mixed precision, on three workers: replays with new frequencies against
    the eager scheduled batch (mh_multi_sweep)."""
    items = mh.random_workload(21, 5, nomega=4, n_min=150, n_max=500)
    print(f"\nbatch graph, uneven: panels {[it.mesh.nfaces for it in items]}", flush=True)
    prep = mh.batches.prepare(items)
    graph = mh.BatchGraph(items, "mixed", nworkers=3, prepared=prep)
    rng = np.random.default_rng(3)
    try:
        print(f"graph: nodes, reruns, workers = {graph.info()}", flush=True)
        for rep in range(2):
            oms = [np.sort(rng.uniform(0.3, 3.0, 4)) for _ in items]
            got = graph.launch(oms)
            for it, om in zip(items, oms):
                it.omegas = om
            ref = mh.multi_sweep(items, "mixed", nworkers=2)
            for q, (g_item, r_item) in enumerate(zip(got, ref)):
                for g_, r_, name in zip(g_item, r_item, ("A", "B", "F", "dA", "dB", "dF")):
                    close(g_, r_, f"item {q} {name} replay {rep}")
        assert graph.info()[1] == 0
    finally:
        graph.close()


@pytest.mark.skip(reason=GRAPH_SKIP)
def test_batch_graph_reruns_an_item_whose_checks_fail(mh, gpu):
    """A NaN written into one body's dof field between launches: that item's
    refinement verdict fails on the device, the item runs again without the
    graph and its error (non-finite residual) names it; after the field is
    repaired the next launch is clean again."""
    items = [_grad_item(mh, 10, [0.8, 1.6]), _grad_item(mh, 12, [0.6, 1.2], radius=1.3)]
    prep = mh.batches.prepare(items)
    graph = mh.BatchGraph(items, "mixed", nworkers=2, prepared=prep)
    try:
        ref = graph.launch()
        good = prep[1].dofs[0, 2, 5].item()
        prep[1].dofs[0, 2, 5] = float("nan")
        with pytest.raises(ArithmeticError, match="item 1.*non-finite"):
            graph.launch()
        assert graph.info()[1] == 1
        prep[1].dofs[0, 2, 5] = good
        got = graph.launch()
        assert graph.info()[1] == 1
        for q in range(2):
            for g_, r_, name in zip(got[q], ref[q], ("A", "B", "F", "dA", "dB", "dF")):
                close(g_, r_, f"item {q} {name}")
    finally:
        graph.close()


def test_batch_graph_refuses_mixed_and_gradient_sweeps(mh, gpu):
    items = [_grad_item(mh, 8, [1.0, 2.0])]
    prep = mh.batches.prepare(items)
    with pytest.raises(NotImplementedError, match="only FP64"):
        mh.BatchGraph(items, "mixed", prepared=prep)
    with pytest.raises(NotImplementedError, match="gradients are not captured"):
        mh.BatchGraph(items, "fp64", prepared=prep)
    two = _fwd_items(mh, 9, 2, 1, 60, 120)
    with pytest.raises(NotImplementedError, match="one item on one worker"):
        mh.BatchGraph(two, "fp64", nworkers=1)


def _fwd_items(mh, seed, nbodies, nomega, n_min, n_max):
    items = mh.random_workload(seed, nbodies, nomega=nomega, n_min=n_min, n_max=n_max, gradients=False)
    for it in items:
        it.tangents = None
    return items


MULTI_SKIP = ("captures of several bodies are switched off: six FP64 bodies on three workers crashed the host "
              "process inside mh_batch_graph_create; the cause was not found and the entry point refuses "
              "several items or workers")


@pytest.mark.skip(reason=MULTI_SKIP)
def test_batch_graph_uneven_bodies_fp64(mh, gpu):
    """Six seeded bodies of 150-600 panels (spheres, cylinders, boxes), four
    frequencies each, heave + surge, heading 0, FP64 forward sweeps on three
    workers in one graph: replays with new frequencies (eager sweeps on the
    shared context between them) against the eager scheduled batch."""
    items = _fwd_items(mh, 21, 6, 4, 150, 600)
    print(f"\nbatch graph, uneven FP64: panels {[it.mesh.nfaces for it in items]}", flush=True)
    prep = mh.batches.prepare(items)
    graph = mh.BatchGraph(items, "fp64", nworkers=3, prepared=prep)
    rng = np.random.default_rng(3)
    try:
        print(f"graph: nodes, reruns, workers = {graph.info()}", flush=True)
        for rep in range(3):
            oms = [np.sort(rng.uniform(0.3, 3.0, 4)) for _ in items]
            got = graph.launch(oms)
            for it, om in zip(items, oms):
                it.omegas = om
            ref = mh.multi_sweep(items, "fp64", nworkers=2)
            for q, (g_item, r_item) in enumerate(zip(got, ref)):
                for g_, r_, name in zip(g_item, r_item, "ABF"):
                    close(g_, r_, f"item {q} {name} replay {rep}")
        assert graph.info()[1] == 0
    finally:
        graph.close()


@pytest.mark.skip(reason=MULTI_SKIP)
def test_batch_graph_reports_the_item_with_a_bad_mesh_index(mh, gpu):
    """A face index out of range written into the second body's mesh between
    launches: the packing kernel clamps it and raises the item's flag, the
    item runs again without the graph and its error names it; after the
    index is repaired the next launch matches the first."""
    items = _fwd_items(mh, 4, 2, 2, 100, 200)
    prep = mh.batches.prepare(items)
    graph = mh.BatchGraph(items, "fp64", nworkers=2, prepared=prep)
    try:
        ref = graph.launch()
        dm = prep[1].dm
        good = dm.faces[0, 2].item()
        dm.faces[0, 2] = dm.mesh.nvertices + 3
        with pytest.raises(ValueError, match="item 1.*vertex index"):
            graph.launch()
        assert graph.info()[1] == 1
        dm.faces[0, 2] = good
        got = graph.launch()
        assert graph.info()[1] == 1
        for q in range(2):
            for g_, r_, name in zip(got[q], ref[q], "ABF"):
                close(g_, r_, f"item {q} {name}")
    finally:
        graph.close()


def test_batch_graph_rejects_other_arrays(mh, gpu):
    items = _fwd_items(mh, 9, 1, 1, 60, 120)
    prep = mh.batches.prepare(items)
    graph = mh.BatchGraph(items, "fp64", prepared=prep)
    try:
        other = mh.batches.prepare(items)
        arr = mh.batches._array(other)
        rc = mh._lib.lib().mh_batch_graph_launch(mh._lib.context(0), graph.handle, 1, arr)
        assert rc == 1 and b"differs from the recorded" in mh._lib.lib().mh_last_error()
    finally:
        graph.close()
