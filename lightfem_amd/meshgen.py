"""Synthetic structured Q4 cantilever workloads and `.fem` text output.

Numbering follows the structured grid of the reference tests
(tests/test_solver.cpp:371-396 and tools/gen_mesh.py:26-63): node id
``j*(nx+1)+i+1`` at ``(i*lx/nx, j*ly/ny)``, element ``e = j*nx+i+1`` with the
counter-clockwise connectivity ``(a, a+1, a+nx+2, a+nx+1)``, ``a = j*(nx+1)+i+1``.
The left edge is fully fixed. The right edge carries a tributary nodal load
(interior nodes ``F/ny``, the two corners ``F/(2ny)``) whose total is ``F``.

Random fields use splitmix64 (Steele, Lea and Flood, OOPSLA 2014) with
``u = (x >> 11) * 2**-53`` in [0, 1), drawn in element-id order from a recorded
seed, so every configuration is reproducible bit for bit.
"""
from __future__ import annotations

import dataclasses
from typing import Optional, Sequence

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, count: int) -> np.ndarray:
    """The first ``count`` outputs of splitmix64 seeded with ``seed``."""
    with np.errstate(over="ignore"):
        k = np.arange(1, count + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + k * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform01(seed: int, count: int) -> np.ndarray:
    return (splitmix64(seed, count) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


@dataclasses.dataclass
class Q4Model:
    """A Q4 model in `.fem` terms (ids are the user's ids, not indices)."""

    node_ids: np.ndarray          # (N,) int64
    xy: np.ndarray                # (N, 2) float64
    elem_ids: np.ndarray          # (E,) int64
    enodes: np.ndarray            # (E, 4) int64 node ids, counter-clockwise
    elem_mat: np.ndarray          # (E,) int64 material ids
    mat_ids: np.ndarray           # (M,) int64
    mat_props: np.ndarray         # (M, 3) float64: E, nu, thickness
    bc_nodes: np.ndarray          # (B,) int64 node ids
    bc_values: np.ndarray         # (B, 2) float64, NaN = free
    load_nodes: np.ndarray        # (L,) int64
    load_values: np.ndarray       # (L, 2) float64
    settings: dict = dataclasses.field(default_factory=dict)

    @property
    def n_nodes(self) -> int:
        return int(self.node_ids.shape[0])

    @property
    def n_elems(self) -> int:
        return int(self.elem_ids.shape[0])

    def element_young(self) -> np.ndarray:
        """Per-element E, nu, thickness (E, 3)."""
        order = np.argsort(self.mat_ids, kind="stable")
        pos = np.searchsorted(self.mat_ids[order], self.elem_mat)
        return self.mat_props[order][pos]

    def to_fem(self, path: str) -> None:
        """Write `.fem` text with round-trip ``%.17g`` numbers."""
        lines = ["[settings]"]
        for key, value in self.settings.items():
            lines.append(f"{key} {value}")
        lines.append("[materials]")
        lines.extend(f"{i} {p[0]:.17g} {p[1]:.17g} {p[2]:.17g}" for i, p in zip(self.mat_ids.tolist(), self.mat_props.tolist()))
        lines.append("[nodes]")
        lines.extend(f"{i} {p[0]:.17g} {p[1]:.17g}" for i, p in zip(self.node_ids.tolist(), self.xy.tolist()))
        lines.append("[elements]")
        lines.extend(f"{i} {n[0]} {n[1]} {n[2]} {n[3]} {m}"
                     for i, n, m in zip(self.elem_ids.tolist(), self.enodes.tolist(), self.elem_mat.tolist()))
        lines.append("[boundary]")

        def fmt(v: float) -> str:
            return "free" if np.isnan(v) else f"{v:.17g}"

        lines.extend(f"{i} {fmt(v[0])} {fmt(v[1])}" for i, v in zip(self.bc_nodes.tolist(), self.bc_values.tolist()))
        lines.append("[loads]")
        lines.extend(f"{i} {v[0]:.17g} {v[1]:.17g}" for i, v in zip(self.load_nodes.tolist(), self.load_values.tolist()))
        with open(path, "w") as handle:
            handle.write("\n".join(lines) + "\n")


def structured_cantilever(nx: int, ny: int, lx: float = 1.0, ly: float = 1.0,
                          young: Optional[np.ndarray] = None, e_uniform: float = 1.0,
                          nu: float = 0.3, thickness: float = 1.0,
                          plane_mode: str = "plane_stress",
                          total_load: Sequence[float] = (0.0, -1.0),
                          settings: Optional[dict] = None) -> Q4Model:
    """Structured nx x ny cantilever on an lx x ly rectangle.

    ``young`` gives one modulus per element (one material row per element), else
    a single material with ``e_uniform`` is used.
    """
    n_nodes = (nx + 1) * (ny + 1)
    jj, ii = np.meshgrid(np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    node_ids = (jj * (nx + 1) + ii + 1).reshape(-1).astype(np.int64)
    xy = np.empty((n_nodes, 2), dtype=np.float64)
    xy[:, 0] = (ii.reshape(-1).astype(np.float64) * lx) / nx
    xy[:, 1] = (jj.reshape(-1).astype(np.float64) * ly) / ny
    ej, ei = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    a = (ej * (nx + 1) + ei + 1).reshape(-1).astype(np.int64)
    enodes = np.stack([a, a + 1, a + nx + 2, a + nx + 1], axis=1)
    n_elems = nx * ny
    elem_ids = np.arange(1, n_elems + 1, dtype=np.int64)
    if young is None:
        mat_ids = np.array([1], dtype=np.int64)
        mat_props = np.array([[e_uniform, nu, thickness]], dtype=np.float64)
        elem_mat = np.ones(n_elems, dtype=np.int64)
    else:
        young = np.asarray(young, dtype=np.float64)
        assert young.shape == (n_elems,)
        mat_ids = elem_ids.copy()
        mat_props = np.empty((n_elems, 3), dtype=np.float64)
        mat_props[:, 0] = young
        mat_props[:, 1] = nu
        mat_props[:, 2] = thickness
        elem_mat = elem_ids.copy()
    left = (np.arange(ny + 1) * (nx + 1) + 1).astype(np.int64)
    right = (np.arange(ny + 1) * (nx + 1) + nx + 1).astype(np.int64)
    share = np.full(ny + 1, 1.0 / ny)
    share[0] = share[-1] = 0.5 / ny
    loads = np.stack([share * total_load[0], share * total_load[1]], axis=1)
    base = {"plane_mode": plane_mode, "solver": "pcg", "tolerance": "1e-10", "max_iterations": "100000",
            "reordering": "none", "assembly": "serial", "output": "result.vtk"}
    if settings:
        base.update(settings)
    return Q4Model(node_ids=node_ids, xy=xy, elem_ids=elem_ids, enodes=enodes, elem_mat=elem_mat,
                   mat_ids=mat_ids, mat_props=mat_props, bc_nodes=left,
                   bc_values=np.zeros((ny + 1, 2)), load_nodes=right, load_values=loads, settings=base)


# Workload configurations (seeds recorded here are the ones the fixtures and the bench use).
SEED_LOGE = 20260301     # 1000x1000 log-uniform E
SEED_DENSITY = 20260401  # 400x400 SIMP density
SEED_BATCH = 20260501    # batch of uneven problems


def cantilever_50() -> Q4Model:
    """50x50 plane-stress cantilever, E=1, nu=0.3, serial, no RCM, Jacobi."""
    return structured_cantilever(50, 50, settings={"assembly": "serial", "reordering": "none"})


def cantilever_200_rcm(assembly: str = "coloring", preconditioner: str = "jacobi") -> Q4Model:
    """200x200 cantilever, RCM on."""
    return structured_cantilever(200, 200, settings={"assembly": assembly, "reordering": "rcm",
                                                     "preconditioner": preconditioner})


def loguniform_young(n_elems: int, seed: int, decades: float = 3.0) -> np.ndarray:
    """E_e = 10**(-decades*u_e), u_e in [0,1) from splitmix64."""
    return np.power(10.0, -decades * uniform01(seed, n_elems))


def loguniform_cantilever(n: int = 1000, seed: int = SEED_LOGE) -> Q4Model:
    """n x n plane-strain cantilever with log-uniform E in [1e-3, 1]."""
    return structured_cantilever(n, n, young=loguniform_young(n * n, seed), plane_mode="plane_strain")


def simp_young(density: np.ndarray, emin: float = 1e-9, penal: float = 3.0) -> np.ndarray:
    return emin + density ** penal * (1.0 - emin)


def simp_density(n: int = 400, seed: int = SEED_DENSITY) -> np.ndarray:
    return 0.3 + 0.7 * uniform01(seed, n * n)


def simp_cantilever(n: int = 400, seed: int = SEED_DENSITY) -> Q4Model:
    return structured_cantilever(n, n, young=simp_young(simp_density(n, seed)))


def batch_problems(n_problems: int = 256, seed: int = SEED_BATCH, n_min: int = 32, n_max: int = 256,
                n_cases: int = 4) -> list:
    """Uneven batch: nx, ny ~ U{n_min..n_max} on an nx x ny-unit rectangle, nu ~ U[0.2, 0.45],
    plane stress or strain 50/50, per-element E log-uniform in [1e-2, 1], left edge fixed,
    n_cases right-edge tractions of unit total force at angles theta ~ U[0, 2 pi).

    Draws: 8 per problem from splitmix64(seed) (nx, ny, nu, mode, 4 angles), and the
    element moduli from splitmix64(seed + 1 + p)."""
    span = n_max - n_min + 1
    draws = uniform01(seed, 8 * n_problems).reshape(n_problems, 8)
    out = []
    for p in range(n_problems):
        d = draws[p]
        nx = n_min + min(int(d[0] * span), span - 1)
        ny = n_min + min(int(d[1] * span), span - 1)
        out.append({"nx": nx, "ny": ny, "nu": 0.2 + 0.25 * d[2],
                    "plane_mode": "plane_strain" if d[3] < 0.5 else "plane_stress",
                    "theta": 2.0 * np.pi * d[4:4 + n_cases],
                    "young": loguniform_young(nx * ny, seed + 1 + p, decades=2.0)})
    return out


def batch_problem_model(spec: dict) -> Q4Model:
    """One batch problem (first load case as the model's load)."""
    return structured_cantilever(spec["nx"], spec["ny"], lx=float(spec["nx"]), ly=float(spec["ny"]),
                                 young=spec["young"], nu=spec["nu"], plane_mode=spec["plane_mode"],
                                 total_load=(float(np.cos(spec["theta"][0])), float(np.sin(spec["theta"][0]))))


def batch_case_loads(model: Q4Model, spec: dict) -> np.ndarray:
    """(cases, loaded nodes, 2) tractions: the model's right-edge shares times (cos, sin) theta."""
    share = np.hypot(model.load_values[:, 0], model.load_values[:, 1])
    return np.stack([np.stack([share * np.cos(t), share * np.sin(t)], axis=1) for t in spec["theta"]])
