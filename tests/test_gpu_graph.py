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
