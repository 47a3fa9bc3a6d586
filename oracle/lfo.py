"""Test infrastructure only: numpy wrappers over the C oracle (``lfo.c``).

Each function calls the restatement of one reference stage; arrays are int64 /
float64. See ``lfo.h`` for what pins each function.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "liblfo.so")
_lib = None

PLANE_STRESS, PLANE_STRAIN = 0, 1
JACOBI, BLOCK1, IC0, BLOCK = 0, 1, 2, 3
STOP_REFERENCE, STOP_RELATIVE = 0, 1
STORAGE_CSR, STORAGE_BSR = 0, 1

_P = ctypes.c_void_p
_I = ctypes.c_int64


class PcgReport(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int64), ("converged", ctypes.c_int32), ("status", ctypes.c_int32),
                ("initial_residual", ctypes.c_double), ("final_residual", ctypes.c_double)]


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle not built: {_LIB_PATH} (run `make -C oracle port`)")
        _lib = ctypes.CDLL(_LIB_PATH)
        _lib.lfo_pattern_nnz.restype = ctypes.c_int64
        _lib.lfo_colors.restype = ctypes.c_int64
        _lib.lfo_ic0_factor_size.restype = ctypes.c_int64
        _lib.lfo_ic0_factor.restype = ctypes.c_int64
        _lib.lfo_block1_factor.restype = ctypes.c_int64
        _lib.lfo_sensitivity.restype = ctypes.c_int64
        _lib.lfo_bsr_blocks.restype = ctypes.c_int64
        _lib.lfo_pattern_nnz_var.restype = ctypes.c_int64
        _lib.lfo_colors_var.restype = ctypes.c_int64
        _lib.lfo_pin_zero_pivots.restype = ctypes.c_int64
        _lib.lfo_split_chunks.restype = ctypes.c_int64
        _lib.lfo_block_create.restype = ctypes.c_void_p
        _lib.lfo_block_apply.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        _lib.lfo_block_stats.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        _lib.lfo_block_free.argtypes = [ctypes.c_void_p]
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(_P)


def _i64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int64)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def q4_stiffness(xy, young: float, nu: float, thickness: float, plane_mode: int) -> np.ndarray:
    xy = _f64(xy).reshape(8)
    ke = np.empty(64)
    if lib().lfo_q4_stiffness(_p(xy), ctypes.c_double(young), ctypes.c_double(nu), ctypes.c_double(thickness),
                              ctypes.c_int(plane_mode), _p(ke)):
        raise RuntimeError("Q4 stiffness evaluation received a non-positive Jacobian")
    return ke.reshape(8, 8)


def q4_fields(xy, young: float, nu: float, thickness: float, plane_mode: int, u):
    xy = _f64(xy).reshape(8)
    u = _f64(u).reshape(8)
    centre = np.empty(7)
    gp = np.empty(36)
    if lib().lfo_q4_fields(_p(xy), ctypes.c_double(young), ctypes.c_double(nu), ctypes.c_double(thickness),
                           ctypes.c_int(plane_mode), _p(u), _p(centre), _p(gp)):
        raise RuntimeError("Q4 field evaluation received a non-positive Jacobian")
    return centre, gp.reshape(4, 9)


def element_stiffness(xy_nodes: np.ndarray, enode_idx: np.ndarray, props: np.ndarray, plane_mode: int) -> np.ndarray:
    """Ke for every element (E, 64); props is (E, 3) E/nu/t, enode_idx node indices."""
    n = enode_idx.shape[0]
    out = np.empty((n, 64))
    for e in range(n):
        out[e] = q4_stiffness(xy_nodes[enode_idx[e]].reshape(8), props[e, 0], props[e, 1], props[e, 2],
                              plane_mode).reshape(64)
    return out


def pattern(ndofs: int, edofs: np.ndarray):
    edofs = _i64(edofs)
    nelem = edofs.shape[0]
    nnz = lib().lfo_pattern_nnz(_I(ndofs), _I(nelem), _p(edofs))
    row_ptr = np.empty(ndofs + 1, dtype=np.int64)
    col = np.empty(nnz, dtype=np.int64)
    emap = np.empty((nelem, 64), dtype=np.int64)
    if lib().lfo_pattern_build(_I(ndofs), _I(nelem), _p(edofs), _p(row_ptr), _p(col), _p(emap)):
        raise RuntimeError("requested entry is absent from preallocated CSR pattern")
    return row_ptr, col, emap


def rcm(node_ids, enode_ids) -> np.ndarray:
    node_ids = _i64(node_ids)
    enode_ids = _i64(enode_ids)
    out = np.empty(node_ids.shape[0], dtype=np.int64)
    if lib().lfo_rcm(_I(node_ids.shape[0]), _p(node_ids), _I(enode_ids.shape[0]), _p(enode_ids), _p(out)):
        raise RuntimeError("RCM: element references an unknown node")
    return out


def colors(enode_ids):
    enode_ids = _i64(enode_ids)
    out = np.empty(enode_ids.shape[0], dtype=np.int64)
    count = lib().lfo_colors(_I(enode_ids.shape[0]), _p(enode_ids), _p(out))
    return out, int(count)


def csr_profile(row_ptr, col):
    row_ptr = _i64(row_ptr)
    col = _i64(col)
    out = np.empty(3, dtype=np.int64)
    lib().lfo_csr_profile(_I(row_ptr.shape[0] - 1), _p(row_ptr), _p(col), _p(out))
    return {"nonzeros": int(out[0]), "bandwidth": int(out[1]), "profile": int(out[2])}


def assemble_serial(ke, emap, nnz: int) -> np.ndarray:
    ke = _f64(ke)
    emap = _i64(emap)
    values = np.empty(nnz)
    lib().lfo_assemble_serial(_I(emap.shape[0]), _p(ke), _p(emap), _I(nnz), _p(values))
    return values


def dirichlet(row_ptr, col, values, rhs, cdofs, cvals):
    row_ptr, col = _i64(row_ptr), _i64(col)
    values = np.array(values, dtype=np.float64, copy=True)
    rhs = np.array(rhs, dtype=np.float64, copy=True)
    cdofs, cvals = _i64(cdofs), _f64(cvals)
    lib().lfo_dirichlet(_I(row_ptr.shape[0] - 1), _p(row_ptr), _p(col), _p(values), _p(rhs), _I(cdofs.shape[0]),
                        _p(cdofs), _p(cvals))
    return values, rhs


def spmv(row_ptr, col, values, x) -> np.ndarray:
    row_ptr, col, values, x = _i64(row_ptr), _i64(col), _f64(values), _f64(x)
    y = np.empty(row_ptr.shape[0] - 1)
    lib().lfo_spmv(_I(y.shape[0]), _p(row_ptr), _p(col), _p(values), _p(x), _p(y))
    return y


@dataclass
class Ic0Factor:
    lrow: np.ndarray
    lcol: np.ndarray
    lval: np.ndarray
    inv_diag: np.ndarray
    breakdowns: int


def ic0_factor(row_ptr, col, values) -> Ic0Factor:
    row_ptr, col, values = _i64(row_ptr), _i64(col), _f64(values)
    n = row_ptr.shape[0] - 1
    size = lib().lfo_ic0_factor_size(_I(n), _p(row_ptr), _p(col))
    f = Ic0Factor(np.empty(n + 1, dtype=np.int64), np.empty(size, dtype=np.int64), np.empty(size), np.empty(n), 0)
    f.breakdowns = int(lib().lfo_ic0_factor(_I(n), _p(row_ptr), _p(col), _p(values), _p(f.lrow), _p(f.lcol),
                                            _p(f.lval), _p(f.inv_diag)))
    return f


def ic0_apply(f: Ic0Factor, vec) -> np.ndarray:
    vec = _f64(vec)
    out = np.empty_like(vec)
    lib().lfo_ic0_apply(_I(vec.shape[0]), _p(f.lrow), _p(f.lcol), _p(f.lval), _p(f.inv_diag), _p(vec), _p(out))
    return out


def block1_factor(row_ptr, col, values):
    row_ptr, col, values = _i64(row_ptr), _i64(col), _f64(values)
    n = row_ptr.shape[0] - 1
    factor = np.empty(3 * (n // 2))
    degenerate = lib().lfo_block1_factor(_I(n), _p(row_ptr), _p(col), _p(values), _p(factor))
    return factor, int(degenerate)


def block1_apply(factor, vec) -> np.ndarray:
    factor, vec = _f64(factor), _f64(vec)
    out = np.empty_like(vec)
    lib().lfo_block1_apply(_I(vec.shape[0]), _p(factor), _p(vec), _p(out))
    return out


def bsr_spmv(row_ptr, col, values, x) -> np.ndarray:
    """Block storage product (sparse.cpp:128-198)."""
    row_ptr, col, values, x = _i64(row_ptr), _i64(col), _f64(values), _f64(x)
    y = np.empty(row_ptr.shape[0] - 1)
    lib().lfo_bsr_spmv(_I(y.shape[0]), _p(row_ptr), _p(col), _p(values), _p(x), _p(y))
    return y


def bsr_blocks(row_ptr, col) -> int:
    row_ptr, col = _i64(row_ptr), _i64(col)
    return int(lib().lfo_bsr_blocks(_I(row_ptr.shape[0] - 1), _p(row_ptr), _p(col)))


def split_chunks(n_nodes: int, nodes_per_subdomain: int):
    """Chunk subdomains (precond.cpp:19-28): (node -> subdomain, count)."""
    out = np.empty(n_nodes, dtype=np.int64)
    count = lib().lfo_split_chunks(_I(n_nodes), _I(nodes_per_subdomain), _p(out))
    return out, int(count)


class BlockJacobi:
    """SubdomainBlockJacobi (precond.cpp:121-281) on a dof -> subdomain map."""

    def __init__(self, row_ptr, col, values, dof_subdomain, n_sub: int):
        self._keep = (_i64(row_ptr), _i64(col), _f64(values), _i64(dof_subdomain))
        rp, cl, vals, dsub = self._keep
        self.n = rp.shape[0] - 1
        self._h = lib().lfo_block_create(_I(self.n), _p(rp), _p(cl), _p(vals), _p(dsub), _I(n_sub))
        st = np.empty(3)
        lib().lfo_block_stats(self._h, _p(st))
        self.degenerate, self.factor_entries, self.capture_ratio = int(st[0]), int(st[1]), float(st[2])

    def apply(self, vec) -> np.ndarray:
        vec = _f64(vec)
        out = np.empty_like(vec)
        lib().lfo_block_apply(self._h, _p(vec), _p(out))
        return out

    def __del__(self):
        if getattr(self, "_h", None):
            lib().lfo_block_free(self._h)
            self._h = None


def pcg(row_ptr, col, values, rhs, precond: int = JACOBI, stop_rule: int = STOP_REFERENCE,
        tolerance: float = 1e-10, max_iterations: int = 100000, storage: int = 0, dof_subdomain=None,
        n_sub: int = 0):
    """pcg_solve (solver.cpp:532-627); precond BLOCK takes dof_subdomain / n_sub and
    storage BSR multiplies with the block storage product."""
    row_ptr, col, values, rhs = _i64(row_ptr), _i64(col), _f64(values), _f64(rhs)
    dsub = _i64(dof_subdomain if dof_subdomain is not None else np.zeros(1, dtype=np.int64))
    x = np.empty(rhs.shape[0])
    rep = PcgReport()
    lib().lfo_pcg_ex(_I(rhs.shape[0]), _p(row_ptr), _p(col), _p(values), _p(rhs), ctypes.c_int(precond),
                     ctypes.c_int(storage), _p(dsub), _I(n_sub), ctypes.c_int(stop_rule),
                     ctypes.c_double(tolerance), _I(max_iterations), _p(x), ctypes.byref(rep))
    return x, {"iterations": rep.iterations, "converged": bool(rep.converged), "status": rep.status,
               "initial_residual": rep.initial_residual, "final_residual": rep.final_residual}


# ---------------------------------------------------------------- bars

def pattern_var(ndofs: int, element_dofs: list):
    """(row_ptr, col, emap_flat) for element dof lists of any length (Q4 then bars)."""
    eptr = np.concatenate([[0], np.cumsum([len(d) for d in element_dofs])]).astype(np.int64)
    edofs = _i64(np.concatenate([np.asarray(d, dtype=np.int64) for d in element_dofs]))
    nelem = len(element_dofs)
    nnz = lib().lfo_pattern_nnz_var(_I(ndofs), _I(nelem), _p(eptr), _p(edofs))
    row_ptr = np.empty(ndofs + 1, dtype=np.int64)
    col = np.empty(nnz, dtype=np.int64)
    emap = np.empty(int(sum(len(d) ** 2 for d in element_dofs)), dtype=np.int64)
    if lib().lfo_pattern_build_var(_I(ndofs), _I(nelem), _p(eptr), _p(edofs), _p(row_ptr), _p(col), _p(emap)):
        raise RuntimeError("requested entry is absent from preallocated CSR pattern")
    return row_ptr, col, emap


def colors_var(element_node_ids: list):
    nptr = np.concatenate([[0], np.cumsum([len(n) for n in element_node_ids])]).astype(np.int64)
    ids = _i64(np.concatenate([np.asarray(n, dtype=np.int64) for n in element_node_ids]))
    out = np.empty(len(element_node_ids), dtype=np.int64)
    count = lib().lfo_colors_var(_I(len(element_node_ids)), _p(nptr), _p(ids), _p(out))
    return out, int(count)


def assemble_flat(ke_flat, emap_flat, nnz: int) -> np.ndarray:
    ke_flat, emap_flat = _f64(ke_flat), _i64(emap_flat)
    values = np.empty(nnz)
    lib().lfo_assemble_flat(_I(ke_flat.shape[0]), _p(ke_flat), _p(emap_flat), _I(nnz), _p(values))
    return values


def truss_stiffness(xy4, young: float, area: float) -> np.ndarray:
    xy4 = _f64(xy4)
    ke = np.empty(16)
    if lib().lfo_truss_stiffness(_p(xy4), ctypes.c_double(young), ctypes.c_double(area), _p(ke)):
        raise RuntimeError("truss stiffness evaluation received a zero-length member")
    return ke


def truss_fields(xy4, young: float, u4) -> np.ndarray:
    xy4, u4 = _f64(xy4), _f64(u4)
    out = np.empty(7)
    if lib().lfo_truss_fields(_p(xy4), ctypes.c_double(young), _p(u4), _p(out)):
        raise RuntimeError("truss strain evaluation received a zero-length member")
    return out


def pin_zero_pivots(row_ptr, col, values, rhs):
    """The zero-pivot sweep after the Dirichlet elimination (solver.cpp:916-940):
    (values, rhs, pinned count, first loaded zero-pivot dof or -1)."""
    row_ptr, col = _i64(row_ptr), _i64(col)
    values = np.array(values, dtype=np.float64, copy=True)
    rhs = np.array(rhs, dtype=np.float64, copy=True)
    bad = ctypes.c_int64(-1)
    pinned = lib().lfo_pin_zero_pivots(_I(row_ptr.shape[0] - 1), _p(row_ptr), _p(col), _p(values), _p(rhs),
                                       ctypes.byref(bad))
    return values, rhs, int(max(pinned, 0)), int(bad.value) if pinned < 0 else -1


def smooth_nodal_ex(nnodes: int, enode_idx, gp, tnode_idx, tstress) -> np.ndarray:
    enode_idx, gp = _i64(enode_idx).reshape(-1, 4), _f64(gp)
    tnode_idx, tstress = _i64(tnode_idx).reshape(-1, 2), _f64(tstress).reshape(-1, 3)
    out = np.empty((nnodes, 4))
    lib().lfo_smooth_nodal_ex(_I(nnodes), _I(enode_idx.shape[0]), _p(enode_idx), _p(gp), _I(tnode_idx.shape[0]),
                              _p(tnode_idx), _p(tstress), _p(out))
    return out


def smooth_nodal(nnodes: int, enode_idx, gp) -> np.ndarray:
    enode_idx, gp = _i64(enode_idx), _f64(gp)
    out = np.empty((nnodes, 4))
    lib().lfo_smooth_nodal(_I(nnodes), _I(enode_idx.shape[0]), _p(enode_idx), _p(gp), _p(out))
    return out


def element_fields(xy_nodes, enode_idx, props, plane_mode, u):
    """Centre (E,7) and Gauss-point (E,4,9) fields for all elements."""
    n = enode_idx.shape[0]
    centre = np.empty((n, 7))
    gp = np.empty((n, 4, 9))
    for e in range(n):
        dofs = np.stack([2 * enode_idx[e], 2 * enode_idx[e] + 1], axis=1).reshape(8)
        c, g = q4_fields(xy_nodes[enode_idx[e]].reshape(8), props[e, 0], props[e, 1], props[e, 2], plane_mode,
                         u[dofs])
        centre[e] = c
        gp[e] = g
    return centre, gp


def solve_model(xy_nodes, enode_idx, props, plane_mode, f, cdofs, cvals, precond=JACOBI,
                stop_rule=STOP_REFERENCE, tolerance=1e-10, max_iterations=100000):
    """Whole reference pipeline on an already-numbered model (node i -> dofs 2i, 2i+1).

    Returns a dict with pattern, K, Kc, fc, u, R, fields and the PCG report.
    """
    enode_idx = _i64(enode_idx)
    n = 2 * xy_nodes.shape[0]
    edofs = np.stack([2 * enode_idx, 2 * enode_idx + 1], axis=2).reshape(-1, 8)
    row_ptr, col, emap = pattern(n, edofs)
    ke = element_stiffness(xy_nodes, enode_idx, props, plane_mode)
    K = assemble_serial(ke, emap, col.shape[0])
    Kc, fc = dirichlet(row_ptr, col, K, f, cdofs, cvals)
    u, rep = pcg(row_ptr, col, Kc, fc, precond, stop_rule, tolerance, max_iterations)
    R = spmv(row_ptr, col, K, u) - f
    centre, gp = element_fields(xy_nodes, enode_idx, props, plane_mode, u)
    nodal = smooth_nodal(xy_nodes.shape[0], enode_idx, gp.reshape(-1, 36))
    return {"row_ptr": row_ptr, "col": col, "emap": emap, "edofs": edofs, "ke": ke, "K": K, "Kc": Kc, "fc": fc,
            "u": u, "R": R, "centre": centre, "gp": gp, "nodal": nodal, "report": rep}


def sensitivity(xy_nodes, enode_idx, props, plane_mode, u, lam, dE=None) -> np.ndarray:
    """-lambda_e^T Ke(E=1) u_e (* dE[e]) per element (lfo_sensitivity)."""
    xy_nodes, enode_idx, props = _f64(xy_nodes), _i64(enode_idx), _f64(props)
    u, lam = _f64(u), _f64(lam)
    dE = None if dE is None else _f64(dE)
    grad = np.empty(enode_idx.shape[0])
    bad = lib().lfo_sensitivity(_I(enode_idx.shape[0]), _p(xy_nodes), _p(enode_idx), _p(props),
                                ctypes.c_int(plane_mode), _p(u), _p(lam), None if dE is None else _p(dE), _p(grad))
    if bad:
        raise RuntimeError("Q4 stiffness evaluation received a non-positive Jacobian")
    return grad


class DirectSystem:
    """FP64 direct solves of the constrained system (scipy SuperLU on the oracle's
    serial K and Dirichlet elimination), with iterative refinement. The CPU
    reference for the adjoint gradient; the reference itself has none."""

    def __init__(self, xy_nodes, enode_idx, props, plane_mode, f, cdofs, cvals, factor: bool = True):
        import scipy.sparse
        import scipy.sparse.linalg
        enode_idx = _i64(enode_idx)
        self.n = 2 * xy_nodes.shape[0]
        edofs = np.stack([2 * enode_idx, 2 * enode_idx + 1], axis=2).reshape(-1, 8)
        self.row_ptr, self.col, self.emap = pattern(self.n, edofs)
        self.ke = element_stiffness(xy_nodes, enode_idx, props, plane_mode)
        self.K = assemble_serial(self.ke, self.emap, self.col.shape[0])
        self.Kc, self.fc = dirichlet(self.row_ptr, self.col, self.K, f, cdofs, cvals)
        self.cdofs = _i64(cdofs)
        self.A = scipy.sparse.csr_matrix((self.Kc, self.col, self.row_ptr), shape=(self.n, self.n))
        self.lu = scipy.sparse.linalg.splu(self.A.tocsc()) if factor else None

    def solve(self, rhs, refine: int = 3) -> np.ndarray:
        x = self.lu.solve(rhs)
        for _ in range(refine):
            x = x + self.lu.solve(rhs - self.A @ x)
        return x

    def adjoint_rhs(self, dJdu) -> np.ndarray:
        g = np.array(dJdu, dtype=np.float64, copy=True)
        g[self.cdofs] = 0.0
        return g
