"""Problems, results and hydrodynamic coefficients over the batched GPU sweep.

Mirrors the reference's orchestration layer (same names, argument meaning
and output layout):

  FloatingBody, combine_floatingbodies      src/bodies.jl:4-164
  DiffractionProblem, RadiationProblem      src/problems_and_results.jl:6-37
  DiffractionResult, RadiationResult,
  make_result                               src/problems_and_results.jl:41-66
  problems_from_data                        src/problems_and_results.jl:70-100
  assemble_hydrodynamic_coefficients        src/problems_and_results.jl:105-179
  create_DimStack                           src/problems_and_results.jl:184-226
  compute_hydrodynamic_coefficients,
  compute_and_label_hydrodynamic_coefficients
                                            src/problems_and_results.jl:231-243
  solve_problem, solve_all_problems         src/solve.jl:4-21

The reference solves the problems one by one, assembling and factorising
the influence matrices again for every problem (src/solve.jl:4-21).  Here
the problems of one body are grouped by frequency: each (body, frequency)
is assembled and factorised once, with every radiating dof and every
heading as a right-hand side of that one factorisation, and all
frequencies of a body run in one batched sweep (libmhip's mh_freq_sweep).
The forces are the same quantities; only the order of the floating-point
work differs.
"""
from __future__ import annotations

from collections import OrderedDict
from dataclasses import dataclass, field

import numpy as np

from .meshes import Mesh, combine_meshes, rotational_dofs, translational_dofs

_TRANSLATIONS = ("Surge", "Sway", "Heave")
_ROTATIONS = ("Roll", "Pitch", "Yaw")


# ---------------------------------------------------------------------------
# bodies (src/bodies.jl)
# ---------------------------------------------------------------------------
class FloatingBody:
    """A mesh with named dof fields (each nfaces x 3), src/bodies.jl:4-15.

    FloatingBody(mesh, dofs, body_name) with dofs a mapping name -> field, or
    FloatingBody(mesh, ["Surge", "Heave", "Pitch"], body_name,
    rotation_center=...) for rigid-body dofs (src/bodies.jl:19-44, :83-92).
    """

    def __init__(self, mesh: Mesh, dofs, body_name: str = "body", rotation_center=None):
        self.mesh = mesh
        self.body_name = body_name
        if isinstance(dofs, (list, tuple)):
            rc = (0.0, 0.0, 0.0) if rotation_center is None else rotation_center
            fields = OrderedDict()
            for name in dofs:
                name = str(name)
                if name in _TRANSLATIONS:
                    fields[name] = translational_dofs(mesh, name)
                elif name in _ROTATIONS:
                    fields[name] = rotational_dofs(mesh, name, rc)
                else:
                    raise ValueError(f"unknown rigid dof {name!r}")
            dofs = fields
        self.dofs = OrderedDict()
        for name, f in dofs.items():
            f = np.asarray(f, dtype=np.float64)
            if f.shape != (mesh.nfaces, 3):
                raise ValueError(f"dof {name!r}: expected a ({mesh.nfaces}, 3) field, got {f.shape}")
            self.dofs[str(name)] = f

    def __add__(self, other):
        others = other if isinstance(other, (list, tuple)) else [other]
        return combine_floatingbodies([self, *others])

    def __repr__(self):
        return f"FloatingBody({self.body_name!r}, {self.mesh.nfaces} faces, dofs={list(self.dofs)})"


def _unique_names(names):
    """src/bodies.jl:94-119: repeated names get _2, _3, ... suffixes."""
    seen, out = {}, []
    for s in names:
        if s not in seen:
            out.append(s)
            seen[s] = 2
        else:
            c = seen[s]
            new = f"{s}_{c}"
            while new in seen or new in names:
                c += 1
                new = f"{s}_{c}"
            out.append(new)
            seen[s] = c + 1
            seen[new] = 1
    return out


def combine_floatingbodies(bodies, new_body_name=None) -> FloatingBody:
    """src/bodies.jl:122-156: one body whose mesh is the concatenation and
    whose dofs are every body's dofs, named body__dof, zero off their body."""
    names = _unique_names([b.body_name.replace(" ", "_") for b in bodies])
    if new_body_name is None:
        new_body_name = "+".join(names)
    mesh = combine_meshes([b.mesh for b in bodies])
    dofs = OrderedDict()
    off = 0
    for name, b in zip(names, bodies):
        for dn, f in b.dofs.items():
            g = np.zeros((mesh.nfaces, 3))
            g[off:off + b.mesh.nfaces] = f
            dofs[f"{name}__{dn}"] = g
        off += b.mesh.nfaces
    return FloatingBody(mesh, dofs, new_body_name)


# ---------------------------------------------------------------------------
# problems and results (src/problems_and_results.jl:6-66)
# ---------------------------------------------------------------------------
class LinearPotentialFlowProblem:
    pass


@dataclass(eq=False)
class DiffractionProblem(LinearPotentialFlowProblem):
    floatingbody: FloatingBody
    omega: float
    beta: float
    influenced_dofs: list

    def __post_init__(self):
        assert set(self.influenced_dofs) <= set(self.floatingbody.dofs), \
            "the influenced_dofs must be keys of floatingbody.dofs"


@dataclass(eq=False)
class RadiationProblem(LinearPotentialFlowProblem):
    floatingbody: FloatingBody
    omega: float
    radiating_dof: str
    influenced_dofs: list

    def __post_init__(self):
        assert self.radiating_dof in self.floatingbody.dofs, \
            "the radiating_dof must be a key of floatingbody.dofs"
        assert set(self.influenced_dofs) <= set(self.floatingbody.dofs), \
            "the influenced_dofs must be keys of floatingbody.dofs"


class LinearPotentialFlowResult:
    pass


@dataclass(eq=False)
class DiffractionResult(LinearPotentialFlowResult):
    problem: DiffractionProblem
    forces: dict  # influenced dof -> complex force


@dataclass(eq=False)
class RadiationResult(LinearPotentialFlowResult):
    problem: RadiationProblem
    forces: dict


def make_result(problem, forces):
    if isinstance(problem, RadiationProblem):
        return RadiationResult(problem, forces)
    return DiffractionResult(problem, forces)


def _inf_dofs(parameters, body):
    return list(parameters.get("influenced_dofs", list(body.dofs)))


def problems_from_data(parameters: dict, floatingbody: FloatingBody):
    """src/problems_and_results.jl:70-100: diffraction problems (headings
    fastest, then frequencies) followed by radiation problems (dofs fastest,
    then frequencies)."""
    inf = _inf_dofs(parameters, floatingbody)
    probs = []
    omegas = list(parameters["wave_frequencies"])
    if "wave_directions" in parameters:
        probs += [DiffractionProblem(floatingbody, w, b, inf) for w in omegas
                  for b in parameters["wave_directions"]]
    if "radiating_dofs" in parameters:
        probs += [RadiationProblem(floatingbody, w, d, inf) for w in omegas
                  for d in parameters["radiating_dofs"]]
    return probs


# ---------------------------------------------------------------------------
# solving (src/solve.jl:4-21), batched per body
# ---------------------------------------------------------------------------
def solve_all_problems(problems, green_functions=None, direct=True, precision="fp64", device=0):
    """Results in the order of `problems`.  Problems sharing a body are
    solved together: one sweep over their distinct frequencies, one
    factorisation per frequency for all radiating dofs and headings."""
    from .api import freq_sweep

    results = [None] * len(problems)
    groups = OrderedDict()
    for idx, p in enumerate(problems):
        groups.setdefault(id(p.floatingbody), []).append(idx)
    for idxs in groups.values():
        body = problems[idxs[0]].floatingbody
        omegas = sorted({float(problems[i].omega) for i in idxs})
        betas = sorted({float(problems[i].beta) for i in idxs
                        if isinstance(problems[i], DiffractionProblem)})
        used = []
        for i in idxs:
            p = problems[i]
            names = list(p.influenced_dofs)
            if isinstance(p, RadiationProblem):
                names.append(p.radiating_dof)
            for n in names:
                if n not in used:
                    used.append(n)
        if len(used) + len(betas) > 16:
            raise ValueError("at most 16 dofs + headings per body in one sweep")
        A, B, F = freq_sweep(body.mesh, omegas, [body.dofs[n] for n in used], betas=betas,
                             green_functions=green_functions, direct=direct,
                             precision=precision, device=device)
        wi = {w: k for k, w in enumerate(omegas)}
        bi = {b: k for k, b in enumerate(betas)}
        di = {n: k for k, n in enumerate(used)}
        for i in idxs:
            p = problems[i]
            w = wi[float(p.omega)]
            if isinstance(p, RadiationProblem):
                j = di[p.radiating_dof]
                # integrate_pressure's complex force: A omega^2 + i B omega
                forces = OrderedDict((n, complex(A[w, di[n], j] * p.omega ** 2,
                                                 B[w, di[n], j] * p.omega))
                                     for n in p.influenced_dofs)
            else:
                b = bi[float(p.beta)]
                forces = OrderedDict((n, complex(F[w, b, di[n]])) for n in p.influenced_dofs)
            results[i] = make_result(p, forces)
    return results


def solve_problem(problem, green_functions=None, direct=True, precision="fp64", device=0):
    """src/solve.jl:4-14 (Wu, direct, by default)."""
    return solve_all_problems([problem], green_functions, direct, precision, device)[0]


# ---------------------------------------------------------------------------
# coefficients (src/problems_and_results.jl:105-243)
# ---------------------------------------------------------------------------
def assemble_hydrodynamic_coefficients(parameters: dict, floatingbody: FloatingBody, results,
                                       device=0):
    """src/problems_and_results.jl:105-179.  Arrays:
    added_mass, radiation_damping: [influenced dof, radiating dof, frequency];
    diffraction_force, Froude_Krylov_force, excitation_force:
    [influenced dof, frequency, heading] (complex).  Missing kinds are []."""
    from .api import froude_krylov

    omegas = list(parameters["wave_frequencies"])
    inf = _inf_dofs(parameters, floatingbody)
    data = {"added_mass": [], "radiation_damping": [], "diffraction_force": [],
            "Froude_Krylov_force": [], "excitation_force": []}
    if "wave_directions" in parameters:
        betas = list(parameters["wave_directions"])
        look = {(float(r.problem.omega), float(r.problem.beta)): r.forces for r in results
                if isinstance(r, DiffractionResult)}
        diff = np.array([[[look[(float(w), float(b))][n] for b in betas] for w in omegas]
                         for n in inf], dtype=np.complex128)
        FK = froude_krylov(floatingbody.mesh, omegas, [floatingbody.dofs[n] for n in inf],
                           betas=betas, device=device)  # [w, b, i]
        fk = np.ascontiguousarray(FK.transpose(2, 0, 1))
        data["diffraction_force"] = diff
        data["Froude_Krylov_force"] = fk
        data["excitation_force"] = diff + fk
    if "radiating_dofs" in parameters:
        rad = list(parameters["radiating_dofs"])
        look = {(r.problem.radiating_dof, float(r.problem.omega)): r.forces for r in results
                if isinstance(r, RadiationResult)}
        data["added_mass"] = np.array(
            [[[look[(d, float(w))][n].real / w ** 2 for w in omegas] for d in rad] for n in inf])
        data["radiation_damping"] = np.array(
            [[[look[(d, float(w))][n].imag / w for w in omegas] for d in rad] for n in inf])
    return data


def compute_hydrodynamic_coefficients(parameters: dict, floatingbody: FloatingBody,
                                      green_functions=None, direct=True, precision="fp64",
                                      device=0):
    """src/problems_and_results.jl:231-236."""
    problems = problems_from_data(parameters, floatingbody)
    results = solve_all_problems(problems, green_functions, direct, precision, device)
    return assemble_hydrodynamic_coefficients(parameters, floatingbody, results, device)


@dataclass
class DimArray:
    """A labelled array: values with named dimensions and coordinates (the
    DimensionalData.DimArray the reference returns)."""
    values: np.ndarray
    dims: tuple
    coords: dict = field(default_factory=dict)

    def sel(self, **kw):
        idx = []
        for d in self.dims:
            if d in kw:
                idx.append(list(self.coords[d]).index(kw[d]))
            else:
                idx.append(slice(None))
        return self.values[tuple(idx)]


def create_DimStack(data: dict, parameters: dict, floatingbody: FloatingBody):
    """src/problems_and_results.jl:184-226: dict of DimArray."""
    inf = _inf_dofs(parameters, floatingbody)
    om = list(parameters["wave_frequencies"])
    be = list(parameters.get("wave_directions", []))
    rad = list(parameters.get("radiating_dofs", []))
    rdims = ("influenced_dofs", "radiating_dofs", "wave_frequencies")
    ddims = ("influenced_dofs", "wave_frequencies", "wave_directions")
    rc = {"influenced_dofs": inf, "radiating_dofs": rad, "wave_frequencies": om}
    dc = {"influenced_dofs": inf, "wave_frequencies": om, "wave_directions": be}
    out = OrderedDict()
    for key in ("added_mass", "radiation_damping"):
        out[key] = DimArray(np.asarray(data[key]), rdims, rc)
    for key in ("excitation_force", "diffraction_force", "Froude_Krylov_force"):
        out[key] = DimArray(np.asarray(data[key]), ddims, dc)
    return out


def compute_and_label_hydrodynamic_coefficients(parameters: dict, floatingbody: FloatingBody,
                                                **kw):
    """src/problems_and_results.jl:239-243."""
    data = compute_hydrodynamic_coefficients(parameters, floatingbody, **kw)
    return create_DimStack(data, parameters, floatingbody)
