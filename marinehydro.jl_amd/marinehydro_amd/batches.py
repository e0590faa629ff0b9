"""Uneven multi-body batches over libmhip's work-list scheduler
(mh_multi_sweep, mh_multi_autotune), and the seeded random-body workload
used to exercise them.

The reference solves a list of problems one at a time
(src/solve.jl:17-21); here every body is one batched sweep over its
frequencies and the bodies run longest-first on a pool of worker streams.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from .api import (SETTINGS, Rankine, RankineReflected, GFWu, Tangents, _dev_tangent, _dmesh, _Outputs,
                  _torch, gf_mask, shape_tangents)
from .meshes import box, floating_sphere, vertical_cylinder


@dataclass
class SweepItem:
    """One body: its mesh, frequencies, dof fields, headings and (optional)
    tangents of its parameters (as for freq_sweep_grad)."""
    mesh: object
    omegas: np.ndarray
    dof_fields: list
    betas: tuple = (0.0,)
    tangents: Tangents = None
    green_functions: tuple = None
    direct: bool = True
    tag: dict = field(default_factory=dict)


class _Prepared:
    def __init__(self, torch, it: SweepItem, device):
        self.dm = _dmesh(it.mesh, device)
        dev = f"cuda:{self.dm.device}"
        n, nv = self.dm.nfaces, self.dm.mesh.nvertices
        self.om = np.ascontiguousarray(np.asarray(it.omegas, dtype=np.float64).reshape(-1))
        self.be = np.ascontiguousarray(np.asarray(it.betas, dtype=np.float64).reshape(-1))
        fields = [np.broadcast_to(np.asarray(f, dtype=np.float64), (n, 3)) for f in it.dof_fields]
        self.ndof, self.nbeta, self.nw = len(fields), self.be.shape[0], self.om.shape[0]
        self.dofs = torch.from_numpy(np.ascontiguousarray(np.stack([f.T for f in fields]))).to(dev)
        T = it.tangents
        self.npar = T.nparam if T is not None else 0
        mp, nb, nd = max(self.npar, 1), max(self.nbeta, 1), self.ndof
        self.out = _Outputs(torch, dev, [("A", (self.nw, nd, nd), False), ("B", (self.nw, nd, nd), False),
                                         ("F", (self.nw, nb, nd), True), ("gA", (self.nw, mp, nd, nd), False),
                                         ("gB", (self.nw, mp, nd, nd), False),
                                         ("gF", (self.nw, mp, nb, nd), True)])
        self.keep = []
        s = L.SweepItemT()
        s.mesh = C.cast(C.pointer(self.dm.struct), C.c_void_p)
        gfs = it.green_functions if it.green_functions is not None else (Rankine(), RankineReflected(), GFWu())
        s.gf_mask = gf_mask(gfs)
        s.direct = int(bool(it.direct))
        s.nomega, s.omegas = self.nw, self.om.ctypes.data
        s.ndof, s.dof_fields = self.ndof, self.dofs.data_ptr()
        s.nbeta, s.betas = self.nbeta, (self.be.ctypes.data if self.nbeta else None)
        s.nparam = self.npar
        if self.npar:
            ptr = lambda t: t.data_ptr() if t is not None else None
            parts = [_dev_tangent(torch, T.dvertices, (self.npar, nv, 3), dev),
                     _dev_tangent(torch, T.dcenters, (self.npar, n, 3), dev),
                     _dev_tangent(torch, T.dnormals, (self.npar, n, 3), dev),
                     _dev_tangent(torch, T.dareas, (self.npar, n), dev),
                     _dev_tangent(torch, T.ddof_fields, (self.npar, self.ndof, n, 3), dev)]
            self.keep += parts
            s.dvertices, s.dcenters, s.dnormals, s.dareas, s.ddof_fields = [ptr(t) for t in parts]
            if T.domegas is not None:
                self.dom = np.ascontiguousarray(np.asarray(T.domegas, dtype=np.float64).reshape(self.npar, self.nw))
                s.domegas = self.dom.ctypes.data
            s.dA, s.dB, s.dF = self.out.ptr("gA"), self.out.ptr("gB"), self.out.ptr("gF")
        s.A, s.B, s.F = self.out.ptr("A"), self.out.ptr("B"), self.out.ptr("F")
        self.struct = s

    def results(self):
        o = self.out
        o.host = None
        A = np.ascontiguousarray(o.get("A").transpose(0, 2, 1))
        B = np.ascontiguousarray(o.get("B").transpose(0, 2, 1))
        F = np.ascontiguousarray(o.get("F")[:, :self.nbeta, :])
        if not self.npar:
            return A, B, F
        gA = np.ascontiguousarray(o.get("gA").transpose(0, 1, 3, 2)[:, :self.npar])
        gB = np.ascontiguousarray(o.get("gB").transpose(0, 1, 3, 2)[:, :self.npar])
        gF = np.ascontiguousarray(o.get("gF")[:, :self.npar, :self.nbeta, :])
        return A, B, F, gA, gB, gF


def prepare(items, device=0):
    """Device copies of the items' inputs and outputs (reusable across calls)."""
    torch = _torch()
    return [_Prepared(torch, it, device) for it in items]


def _array(prepared):
    arr = (L.SweepItemT * len(prepared))()
    for i, p in enumerate(prepared):
        arr[i] = p.struct
    return arr


def multi_sweep(items, precision="mixed", nworkers=0, device=0, g=None, rho=None, prepared=None):
    """All items in one scheduled call; returns per item (A, B, F) or, with
    tangents, (A, B, F, dA, dB, dF), laid out as freq_sweep(_grad)."""
    prep = prepared if prepared is not None else prepare(items, device)
    arr = _array(prep)
    prec = {"fp64": L.FP64, "mixed": L.MIXED}[precision]
    L.check(L.lib().mh_multi_sweep(L.context(device), len(prep), arr, prec, float(g or SETTINGS.g),
                                   float(rho or SETTINGS.rho), int(nworkers)), "multi_sweep")
    return [p.results() for p in prep]


def multi_autotune(items, precision="mixed", device=0, g=None, rho=None, prepared=None):
    """Times the items under each candidate configuration; the context keeps
    the fastest.  Returns (workers, leaf width cap (0 = default), ms)."""
    prep = prepared if prepared is not None else prepare(items, device)
    arr = _array(prep)
    prec = {"fp64": L.FP64, "mixed": L.MIXED}[precision]
    w, lf, ms = C.c_int(0), C.c_int(0), C.c_double(0.0)
    L.check(L.lib().mh_multi_autotune(L.context(device), len(prep), arr, prec, float(g or SETTINGS.g),
                                      float(rho or SETTINGS.rho), C.byref(w), C.byref(lf), C.byref(ms)),
            "multi_autotune")
    return w.value, lf.value, ms.value


class BatchGraph:
    """The items of multi_sweep recorded once as one HIP graph and replayed
    by launch() (mh_batch_graph_*): FP64 or mixed precision, with or
    without gradients, one body or many.  The items' device arrays (meshes,
    dof fields, tangents, outputs) are recorded: their contents may change
    between launches, not their sizes.  launch() takes new host frequencies,
    headings and frequency tangents per item (None keeps the current ones)
    and returns per item what multi_sweep returns.  Creation runs every
    item once, then records; an item whose device checks fail at a launch
    (mesh index, pivots, refinement verdict) runs again without the graph."""

    def __init__(self, items, precision="mixed", nworkers=0, device=0, g=None, rho=None, prepared=None):
        self.prep = prepared if prepared is not None else prepare(items, device)
        self.arr = _array(self.prep)
        self.device = device
        prec = {"fp64": L.FP64, "mixed": L.MIXED}[precision]
        self.handle = C.c_void_p()
        L.check(L.lib().mh_batch_graph_create(L.context(device), len(self.prep), self.arr, prec,
                                              float(g or SETTINGS.g), float(rho or SETTINGS.rho), int(nworkers),
                                              C.byref(self.handle)), "batch_graph_create")

    def launch(self, omegas=None, betas=None, domegas=None):
        for i, p in enumerate(self.prep):
            if omegas is not None and omegas[i] is not None:
                p.om[:] = np.asarray(omegas[i], dtype=np.float64).reshape(p.om.shape)
            if betas is not None and betas[i] is not None and p.nbeta:
                p.be[:] = np.asarray(betas[i], dtype=np.float64).reshape(p.be.shape)
            if domegas is not None and domegas[i] is not None:
                if not hasattr(p, "dom"):
                    raise ValueError(f"BatchGraph.launch: item {i} was recorded without frequency tangents")
                p.dom[:] = np.asarray(domegas[i], dtype=np.float64).reshape(p.dom.shape)
        L.check(L.lib().mh_batch_graph_launch(L.context(self.device), self.handle, len(self.prep), self.arr),
                "batch_graph_launch")
        return [p.results() for p in self.prep]

    def info(self):
        """(recorded nodes, items run again without the graph so far, workers)"""
        nodes, reruns, workers = C.c_int64(0), C.c_int64(0), C.c_int(0)
        L.check(L.lib().mh_batch_graph_info(self.handle, C.byref(nodes), C.byref(reruns), C.byref(workers)),
                "batch_graph_info")
        return nodes.value, reruns.value, workers.value

    def close(self):
        if self.handle:
            L.lib().mh_batch_graph_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# seeded uneven workload: spheres, vertical cylinders and boxes
# ---------------------------------------------------------------------------
def _sphere_res(n_target):
    nphi = max(8, int(round(np.sqrt(2.0 * n_target))))
    return max(4, int(round(n_target / nphi))), nphi


def random_body(rng, n_min=200, n_max=3000):
    """One body: kind uniform over {sphere, cylinder, box}; dimensions
    U[0.5, 5] m; draft U[0.3, 1] x height (spheres U[0.3, 0.95] x diameter,
    since a sphere touching the free surface with its top pole has no
    waterline); a resolution aiming at a panel count drawn U[n_min, n_max].
    Returns (Body, kind, names of its gradient parameters, rebuild), where
    rebuild(params) makes the same body at other parameter values and the
    same resolution."""
    kind = ("sphere", "cylinder", "box")[rng.integers(3)]
    nt = float(rng.uniform(n_min, n_max))
    for _ in range(20):
        if kind == "sphere":
            D = rng.uniform(0.5, 5.0)
            draft = rng.uniform(0.3, 0.95) * D
            ntheta, nphi = _sphere_res(nt)
            b = floating_sphere(D / 2, draft, ntheta=ntheta, nphi=nphi)
            names = ["radius", "draft"]
            rebuild = (lambda nth, nph: lambda p: floating_sphere(p["radius"], p["draft"], ntheta=nth,
                                                                  nphi=nph))(ntheta, nphi)
        elif kind == "cylinder":
            D, H = rng.uniform(0.5, 5.0, 2)
            R, draft = D / 2, rng.uniform(0.3, 1.0) * H
            ntheta = max(8, int(round(np.sqrt(nt) * 1.3)))
            rest = max(4.0, nt / ntheta)
            nz = max(2, int(round(rest * draft / (draft + R))))
            nr = max(2, int(round(rest - nz)))
            b = vertical_cylinder(R, draft, ntheta=ntheta, nz=nz, nr=nr)
            names = ["radius", "draft"]
            rebuild = (lambda a, b_, c: lambda p: vertical_cylinder(p["radius"], p["draft"], ntheta=a, nz=b_,
                                                                    nr=c))(ntheta, nz, nr)
        else:
            Lx, Ly, H = rng.uniform(0.5, 5.0, 3)
            draft = rng.uniform(0.3, 1.0) * H
            area = Lx * Ly + 2 * (Lx + Ly) * draft
            h = np.sqrt(area / nt)
            nx, ny, nz = (max(2, int(round(v / h))) for v in (Lx, Ly, draft))
            b = box(Lx, Ly, draft, nx=nx, ny=ny, nz=nz)
            names = ["length", "width", "draft"]
            rebuild = (lambda a, b_, c: lambda p: box(p["length"], p["width"], p["draft"], nx=a, ny=b_,
                                                      nz=c))(nx, ny, nz)
        if n_min <= b.mesh.nfaces <= n_max:
            return b, kind, names, rebuild
        nt = float(np.clip(nt * nt / b.mesh.nfaces, n_min, n_max))
    return b, kind, names, rebuild


def random_workload(seed, nbodies=256, nomega=16, n_min=200, n_max=3000, gradients=True):
    """The uneven batch: nbodies seeded bodies, each with nomega frequencies
    U[0.3, 3] rad/s, heave and surge, heading 0, and (gradients) tangents of
    its dimensions."""
    rng = np.random.default_rng(seed)
    items = []
    for q in range(nbodies):
        b, kind, names, rebuild = random_body(rng, n_min, n_max)
        om = np.sort(rng.uniform(0.3, 3.0, nomega))
        dofs = [[0.0, 0.0, 1.0], [1.0, 0.0, 0.0]]
        T = shape_tangents(b.mesh, [b.dvertices[k] for k in names]) if gradients else None
        items.append(SweepItem(b.mesh, om, dofs, (0.0,), T,
                               tag={"kind": kind, "params": {k: b.params[k] for k in names}, "body": b,
                                    "rebuild": rebuild}))
    return items
