"""ctypes binding of the two C ABIs: ``include/lightfem_gpu.h`` (the gfx950
backend, liblightfem_gpu.so) and ``include/lightfem_host.h`` (host stages,
liblightfem.so). Both libraries are built in-tree by ``lightfem_amd.build``;
loading fails loudly when they are missing - there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_PKG, "lib")
GPU_LIB_PATH = os.path.join(LIB_DIR, "liblightfem_gpu.so")
HOST_LIB_PATH = os.path.join(LIB_DIR, "liblightfem.so")

OK, ERR_ARG, ERR_JACOBIAN, ERR_NOT_SPD, ERR_NONFINITE, ERR_NOT_CONVERGED, ERR_HIP, ERR_STATE, ERR_UNSUPPORTED, \
    ERR_MECHANISM = range(10)
PLANE_STRESS, PLANE_STRAIN = 0, 1
ASSEMBLY = {"serial": 0, "atomic": 1, "coloring": 2, "thread_local": 3}
PRECOND = {"jacobi": 0, "block": 1, "block2": 1, "ic0": 2, "subdomain": 3}
STORAGE = {"csr": 0, "bsr": 1}
SPMV_BLOCK_ORDER = 256
SPLIT = {"rcb": 0, "chunk": 1}
PRECISION = {"fp64": 0, "mixed": 1}
STOP = {"reference": 0, "relative": 1}

GPU_SYMBOLS = ["lfg_create", "lfg_destroy", "lfg_last_error", "lfg_status_string", "lfg_set_mesh", "lfg_set_pattern",
               "lfg_set_colors", "lfg_element_stiffness", "lfg_assemble", "lfg_get_stiffness", "lfg_set_stiffness",
               "lfg_set_load", "lfg_dirichlet", "lfg_get_system", "lfg_pcg", "lfg_reactions", "lfg_postprocess",
               "lfg_precond_apply", "lfg_spmv", "lfg_bench_spmv", "lfg_bench_copy", "lfg_bench_stream", "lfg_set_materials",
               "lfg_adjoint_solve", "lfg_sensitivity", "lfg_set_material_modes", "lfg_batch_set_problems",
               "lfg_batch_set_loads", "lfg_batch_pcg", "lfg_layout_info_get", "lfg_set_subdomains", "lfg_set_bars",
               "lfg_pin_zero_pivots", "lfg_spmv_noalloc"]
HOST_SYMBOLS = ["lfh_csr_pattern_nnz", "lfh_csr_pattern", "lfh_rcm", "lfh_element_colors", "lfh_csr_profile",
                "lfh_subdomain_split", "lfh_csr_pattern_nnz_bars", "lfh_csr_pattern_bars", "lfh_rcm_bars",
                "lfh_element_colors_bars", "lfh_ic0_node_schedule"]


class LightFemError(RuntimeError):
    """A failed stage; ``code`` is the lfg_status, the message the reference's text."""

    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


class PcgOptions(ctypes.Structure):
    _fields_ = [("preconditioner", ctypes.c_int32), ("precision", ctypes.c_int32), ("stop_rule", ctypes.c_int32),
                ("check_every", ctypes.c_int32), ("tolerance", ctypes.c_double), ("max_iterations", ctypes.c_int64),
                ("replace_drop", ctypes.c_double), ("mixed_switch", ctypes.c_double),
                ("direct_launch", ctypes.c_int32),
                ("spmv_variant", ctypes.c_int32), ("storage", ctypes.c_int32), ("refine_passes", ctypes.c_int32),
                ("refine_reduction", ctypes.c_double)]


class PcgReport(ctypes.Structure):
    _fields_ = [("converged", ctypes.c_int32), ("status", ctypes.c_int32), ("iterations", ctypes.c_int64),
                ("residual_replacements", ctypes.c_int64), ("initial_residual", ctypes.c_double),
                ("final_residual", ctypes.c_double), ("true_residual", ctypes.c_double),
                ("solve_ms", ctypes.c_double), ("setup_ms", ctypes.c_double), ("breakdowns", ctypes.c_int64),
                ("degenerate_blocks", ctypes.c_int64), ("levels", ctypes.c_int64), ("refine_passes", ctypes.c_int64)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class AssemblyReport(ctypes.Structure):
    _fields_ = [("element_stiffness_ms", ctypes.c_double), ("assembly_ms", ctypes.c_double),
                ("coloring_count", ctypes.c_int32), ("jacobian_failures", ctypes.c_int32)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class LayoutInfo(ctypes.Structure):
    _fields_ = [("node_blocked", ctypes.c_int32), ("n_slices", ctypes.c_int32),
                ("stencil_exact_planes", ctypes.c_int32), ("stencil_sym_planes", ctypes.c_int32),
                ("fused_window_pairs", ctypes.c_int32), ("stencil_offsets", ctypes.c_int32 * 17),
                ("grid_row_len", ctypes.c_int32), ("sym_slices", ctypes.c_int32)]

    def as_dict(self) -> dict:
        out = {name: getattr(self, name) for name, _ in self._fields_[:5]}
        out["stencil_offsets"] = list(self.stencil_offsets[:self.stencil_exact_planes])
        out["grid_row_len"] = self.grid_row_len
        out["sym_slices"] = self.sym_slices
        return out


class PivotReport(ctypes.Structure):
    _fields_ = [("pinned", ctypes.c_int64), ("dof", ctypes.c_int64), ("load", ctypes.c_double),
                ("diag", ctypes.c_double)]


class BatchOptions(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int32), ("stop_rule", ctypes.c_int32), ("check_every", ctypes.c_int32),
                ("tune", ctypes.c_int32), ("tolerance", ctypes.c_double), ("max_iterations", ctypes.c_int64),
                ("mixed_switch", ctypes.c_double), ("block_threads", ctypes.c_int32),
                ("slices_per_wave", ctypes.c_int32)]


class BatchSystemReport(ctypes.Structure):
    _fields_ = [("converged", ctypes.c_int32), ("status", ctypes.c_int32), ("iterations", ctypes.c_int64),
                ("switch_iteration", ctypes.c_int64), ("initial_residual", ctypes.c_double),
                ("final_residual", ctypes.c_double),
                ("true_residual", ctypes.c_double)]


class BatchSummary(ctypes.Structure):
    _fields_ = [("solve_ms", ctypes.c_double), ("tune_ms", ctypes.c_double), ("graph_launches", ctypes.c_int64),
                ("problems_switched", ctypes.c_int32), ("block_threads", ctypes.c_int32),
                ("slices_per_wave", ctypes.c_int32),
                ("systems_converged", ctypes.c_int32), ("max_iterations", ctypes.c_int64),
                ("multiply_ms", ctypes.c_double), ("sym_slices", ctypes.c_int32)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


_gpu: Optional[ctypes.CDLL] = None
_host: Optional[ctypes.CDLL] = None
_P = ctypes.c_void_p
_I64 = ctypes.c_int64


def _load(path: str) -> ctypes.CDLL:
    if not os.path.exists(path):
        raise RuntimeError(f"LightFEM native library is not built: {path} (run python -m lightfem_amd.build)")
    return ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)


def gpu_lib() -> ctypes.CDLL:
    global _gpu
    if _gpu is None:
        lib = _load(GPU_LIB_PATH)
        lib.lfg_last_error.restype = ctypes.c_char_p
        lib.lfg_last_error.argtypes = [_P]
        lib.lfg_status_string.restype = ctypes.c_char_p
        lib.lfg_create.argtypes = [ctypes.c_int32, ctypes.POINTER(_P)]
        lib.lfg_destroy.argtypes = [_P]
        lib.lfg_destroy.restype = None
        lib.lfg_set_mesh.argtypes = [_P, _I64, _P, _I64, _P, _P, _I64, _P, ctypes.c_int32]
        lib.lfg_set_pattern.argtypes = [_P, _I64, _I64, _P, _P, _P]
        lib.lfg_set_colors.argtypes = [_P, ctypes.c_int32, _P, _P]
        lib.lfg_element_stiffness.argtypes = [_P, _P]
        lib.lfg_assemble.argtypes = [_P, ctypes.c_int32, ctypes.POINTER(AssemblyReport)]
        lib.lfg_get_stiffness.argtypes = [_P, _P]
        lib.lfg_set_stiffness.argtypes = [_P, _P]
        lib.lfg_set_load.argtypes = [_P, _P]
        lib.lfg_dirichlet.argtypes = [_P, _I64, _P, _P]
        lib.lfg_get_system.argtypes = [_P, _P, _P]
        lib.lfg_pcg.argtypes = [_P, ctypes.POINTER(PcgOptions), _P, ctypes.POINTER(PcgReport)]
        lib.lfg_reactions.argtypes = [_P, _P]
        lib.lfg_postprocess.argtypes = [_P, _P, _P, _P]
        lib.lfg_precond_apply.argtypes = [_P, ctypes.c_int32, _P, _P]
        lib.lfg_spmv.argtypes = [_P, ctypes.c_int32, ctypes.c_int32, _P, _P]
        lib.lfg_spmv_noalloc.argtypes = [_P, ctypes.c_int32, ctypes.c_int32, _P, _P, _P]
        lib.lfg_bench_spmv.argtypes = [_P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                       ctypes.POINTER(ctypes.c_double)]
        lib.lfg_bench_copy.argtypes = [_P, _I64, ctypes.c_int32, ctypes.POINTER(ctypes.c_double)]
        lib.lfg_bench_stream.argtypes = [_P, _I64, ctypes.c_int32, ctypes.POINTER(ctypes.c_double)]
        lib.lfg_set_materials.argtypes = [_P, _I64, _P]
        lib.lfg_adjoint_solve.argtypes = [_P, ctypes.POINTER(PcgOptions), _P, _P, ctypes.POINTER(PcgReport)]
        lib.lfg_sensitivity.argtypes = [_P, _P, _P, _P]
        lib.lfg_set_material_modes.argtypes = [_P, _P]
        lib.lfg_batch_set_problems.argtypes = [_P, ctypes.c_int32, _P]
        lib.lfg_batch_set_loads.argtypes = [_P, ctypes.c_int32, _P, _I64, _P, _P]
        lib.lfg_batch_pcg.argtypes = [_P, ctypes.POINTER(BatchOptions), _P, _P, ctypes.POINTER(BatchSummary)]
        lib.lfg_layout_info_get.argtypes = [_P, ctypes.POINTER(LayoutInfo)]
        lib.lfg_set_subdomains.argtypes = [_P, _I64, _P]
        lib.lfg_set_bars.argtypes = [_P, _I64, _P, _P, _P, _P]
        lib.lfg_pin_zero_pivots.argtypes = [_P, ctypes.POINTER(PivotReport)]
        _gpu = lib
    return _gpu


def host_lib() -> ctypes.CDLL:
    global _host
    if _host is None:
        gpu_lib()  # liblightfem.so links against it
        lib = _load(HOST_LIB_PATH)
        lib.lfh_csr_pattern_nnz.restype = ctypes.c_int64
        lib.lfh_csr_pattern_nnz.argtypes = [_I64, _I64, _P]
        lib.lfh_csr_pattern.argtypes = [_I64, _I64, _P, _P, _P, _P]
        lib.lfh_rcm.argtypes = [_I64, _P, _I64, _P, _P]
        lib.lfh_element_colors.argtypes = [_I64, _I64, _P, _P, _P, _P, _P]
        lib.lfh_csr_profile.argtypes = [_I64, _P, _P, _P]
        lib.lfh_subdomain_split.argtypes = [_I64, _P, ctypes.c_int32, _I64, _P, _P]
        lib.lfh_csr_pattern_nnz_bars.restype = ctypes.c_int64
        lib.lfh_csr_pattern_nnz_bars.argtypes = [_I64, _I64, _P, _I64, _P]
        lib.lfh_csr_pattern_bars.argtypes = [_I64, _I64, _P, _I64, _P, _P, _P, _P, _P]
        lib.lfh_rcm_bars.argtypes = [_I64, _P, _I64, _P, _I64, _P, _P]
        lib.lfh_ic0_node_schedule.argtypes = [_I64, _P, _P, ctypes.c_int32, _P, _P, _P, _I64, _P, _P, _P]
        lib.lfh_element_colors_bars.argtypes = [_I64, _I64, _P, _I64, _P, _P, _P, _P, _P]
        _host = lib
    return _host


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(_P)


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


# ------------------------------------------------------------- host stages

def csr_pattern(n_nodes: int, enodes: np.ndarray, bnodes: Optional[np.ndarray] = None):
    """(row_ptr, col, emap) of build_csr_assembly_pattern for Q4 node-index lists; with
    bars (node-index pairs) also bmap: (row_ptr, col, emap, bmap)."""
    enodes = _i32(enodes).reshape(-1, 4)
    bars = _i32(bnodes if bnodes is not None else np.zeros((0, 2))).reshape(-1, 2)
    lib = host_lib()
    nnz = lib.lfh_csr_pattern_nnz_bars(n_nodes, enodes.shape[0], _ptr(enodes), bars.shape[0], _ptr(bars))
    if nnz < 0:
        raise LightFemError(ERR_ARG, "element DOF is outside CSR matrix dimension")
    row_ptr = np.empty(2 * n_nodes + 1, dtype=np.int32)
    col = np.empty(nnz, dtype=np.int32)
    emap = np.empty((enodes.shape[0], 64), dtype=np.int32)
    bmap = np.empty((bars.shape[0], 16), dtype=np.int32)
    rc = lib.lfh_csr_pattern_bars(n_nodes, enodes.shape[0], _ptr(enodes), bars.shape[0], _ptr(bars),
                                  _ptr(row_ptr), _ptr(col), _ptr(emap), _ptr(bmap))
    if rc:
        raise LightFemError(ERR_ARG, "CSR pattern exceeds the int32 index range")
    return (row_ptr, col, emap) if bnodes is None else (row_ptr, col, emap, bmap)


def rcm(node_ids: np.ndarray, enode_ids: np.ndarray, bnode_ids: Optional[np.ndarray] = None) -> np.ndarray:
    node_ids = _i32(node_ids)
    enode_ids = _i32(enode_ids).reshape(-1, 4)
    bars = _i32(bnode_ids if bnode_ids is not None else np.zeros((0, 2))).reshape(-1, 2)
    out = np.empty(node_ids.shape[0], dtype=np.int32)
    if host_lib().lfh_rcm_bars(node_ids.shape[0], _ptr(node_ids), enode_ids.shape[0], _ptr(enode_ids),
                               bars.shape[0], _ptr(bars), _ptr(out)):
        raise LightFemError(ERR_ARG, "RCM: element references an unknown node")
    return out


def element_colors(n_nodes: int, enodes: np.ndarray, bnodes: Optional[np.ndarray] = None):
    """(color per element, color_ptr, color_elems); bars follow the Q4 elements."""
    enodes = _i32(enodes).reshape(-1, 4)
    bars = _i32(bnodes if bnodes is not None else np.zeros((0, 2))).reshape(-1, 2)
    ne = enodes.shape[0] + bars.shape[0]
    colors = np.empty(ne, dtype=np.int32)
    count = ctypes.c_int32(0)
    ptr = np.empty(ne + 2, dtype=np.int32)
    elems = np.empty(ne, dtype=np.int32)
    if host_lib().lfh_element_colors_bars(n_nodes, enodes.shape[0], _ptr(enodes), bars.shape[0], _ptr(bars),
                                          _ptr(colors), ctypes.byref(count), _ptr(ptr), _ptr(elems)):
        raise LightFemError(ERR_ARG, "colouring: element references an unknown node")
    return colors, ptr[:count.value + 1].copy(), elems


def csr_profile(row_ptr: np.ndarray, col: np.ndarray) -> dict:
    row_ptr, col = _i32(row_ptr), _i32(col)
    out = np.empty(3, dtype=np.int64)
    host_lib().lfh_csr_profile(row_ptr.shape[0] - 1, _ptr(row_ptr), _ptr(col), _ptr(out))
    return {"nonzeros": int(out[0]), "bandwidth": int(out[1]), "profile": int(out[2])}


def ic0_node_schedule(row_ptr: np.ndarray, col: np.ndarray, backward: bool = False):
    """The level schedule of the device IC(0) sweeps for the factor of this CSR pattern
    (lfh_ic0_node_schedule): dict with ord, qptr, key, parts, reach, or None when the
    factor's rows 2i, 2i+1 do not share their dependency nodes."""
    row_ptr, col = _i32(row_ptr), _i32(col)
    n = row_ptr.shape[0] - 1
    nn = n // 2
    ord_ = np.empty(nn, dtype=np.int32)
    qptr = np.empty(nn + 1, dtype=np.int32)
    key = np.empty(max(1, col.shape[0]), dtype=np.int32)
    parts = np.empty(nn + 2, dtype=np.int32)
    n_parts = ctypes.c_int64(0)
    reach = ctypes.c_int32(0)
    rc = host_lib().lfh_ic0_node_schedule(n, _ptr(row_ptr), _ptr(col), int(backward), _ptr(ord_), _ptr(qptr),
                                          _ptr(key), key.shape[0], _ptr(parts), ctypes.byref(n_parts),
                                          ctypes.byref(reach))
    if rc == 2:
        return None
    if rc != 0:
        raise LightFemError(rc, "lfh_ic0_node_schedule failed")
    return {"ord": ord_, "qptr": qptr, "key": key[:qptr[-1]].copy(), "parts": parts[:n_parts.value].copy(),
            "reach": int(reach.value)}


def subdomain_split(xy: np.ndarray, split: str = "rcb", nodes_per_subdomain: int = 64):
    """build_dof_subdomain's node split (solver.cpp:631-654): (node -> subdomain, count)
    for node positions xy (n_nodes x 2, node-index order)."""
    xy = _f64(xy).reshape(-1, 2)
    out = np.empty(xy.shape[0], dtype=np.int32)
    count = ctypes.c_int64(0)
    if host_lib().lfh_subdomain_split(xy.shape[0], _ptr(xy), SPLIT[split], nodes_per_subdomain, _ptr(out),
                                      ctypes.byref(count)):
        raise LightFemError(ERR_ARG, "invalid subdomain split")
    return out, int(count.value)


# ------------------------------------------------------------- device stages

class Device:
    """One lfg context: device buffers and one HIP stream on GPU `device`."""

    def __init__(self, device: int = 0):
        self._lib = gpu_lib()
        handle = _P()
        rc = self._lib.lfg_create(device, ctypes.byref(handle))
        if rc != OK:
            raise LightFemError(rc, f"cannot create a HIP context on device {device}")
        self._ctx = handle
        self._keep: dict = {}
        self.n = 0
        self.nnz = 0
        self.n_nodes = 0
        self.n_elems = 0
        self.lightfem_owner = None  # (problem, solve token) of the state held on the device

    def close(self) -> None:
        if getattr(self, "_ctx", None):
            self._lib.lfg_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc: int) -> int:
        if rc not in (OK,):
            message = self._lib.lfg_last_error(self._ctx).decode()
            raise LightFemError(rc, message)
        return rc

    def set_mesh(self, xy, enodes, elem_material, materials, plane_mode: int) -> None:
        xy = _f64(xy).reshape(-1, 2)
        enodes = _i32(enodes).reshape(-1, 4)
        elem_material = _i32(elem_material)
        materials = _f64(materials).reshape(-1, 3)
        self._check(self._lib.lfg_set_mesh(self._ctx, xy.shape[0], _ptr(xy), enodes.shape[0], _ptr(enodes),
                                           _ptr(elem_material), materials.shape[0], _ptr(materials), plane_mode))
        self.n_nodes = xy.shape[0]
        self.n_elems = enodes.shape[0]
        self.n_bars = 0

    def set_materials(self, materials) -> None:
        materials = _f64(materials).reshape(-1, 3)
        self._check(self._lib.lfg_set_materials(self._ctx, materials.shape[0], _ptr(materials)))

    def set_material_modes(self, modes) -> None:
        modes = None if modes is None else _i32(modes)
        self._check(self._lib.lfg_set_material_modes(self._ctx, _ptr(modes)))

    def batch_set_problems(self, node_ptr) -> None:
        node_ptr = _i32(node_ptr)
        self._check(self._lib.lfg_batch_set_problems(self._ctx, node_ptr.shape[0] - 1, _ptr(node_ptr)))
        self.batch_problems = node_ptr.shape[0] - 1

    def batch_set_loads(self, loads, dofs, values) -> None:
        loads = _f64(loads).reshape(-1, self.n)
        dofs, values = _i32(dofs), _f64(values)
        self._check(self._lib.lfg_batch_set_loads(self._ctx, loads.shape[0], _ptr(loads), dofs.shape[0], _ptr(dofs),
                                                  _ptr(values)))
        self.batch_cases = loads.shape[0]

    def batch_pcg(self, precision: str = "fp64", stop_rule: str = "relative", tolerance: float = 1e-10,
                  max_iterations: int = 1_000_000, check_every: int = 0, mixed_switch: float = 0.0,
                  tune: bool = True, block_threads: int = 0, slices_per_wave: int = 0, raise_on_error: bool = True):
        """All (problem, case) systems; returns u (cases x dofs), per-system reports, summary.
        mixed_switch (mixed only): relative residual at which a problem moves to FP32
        values; 0 = 1e-3, negative = FP32 from the first iteration."""
        opt = BatchOptions(PRECISION[precision], STOP[stop_rule], check_every, int(tune), tolerance, max_iterations,
                           mixed_switch, block_threads, slices_per_wave)
        n_sys = self.batch_problems * self.batch_cases
        reports = (BatchSystemReport * n_sys)()
        summary = BatchSummary()
        u = np.empty((self.batch_cases, self.n))
        rc = self._lib.lfg_batch_pcg(self._ctx, ctypes.byref(opt), _ptr(u), ctypes.cast(reports, _P),
                                     ctypes.byref(summary))
        if rc != OK and raise_on_error:
            self._check(rc)
        rep = {name: np.array([getattr(r, name) for r in reports]).reshape(self.batch_problems, self.batch_cases)
               for name, _ in BatchSystemReport._fields_}
        return u, rep, summary.as_dict()

    def set_pattern(self, row_ptr, col, emap) -> None:
        row_ptr, col, emap = _i32(row_ptr), _i32(col), _i32(emap)
        self._check(self._lib.lfg_set_pattern(self._ctx, row_ptr.shape[0] - 1, col.shape[0], _ptr(row_ptr),
                                              _ptr(col), _ptr(emap)))
        self.n = row_ptr.shape[0] - 1
        self.nnz = col.shape[0]
        self.n_bars = 0  # the library drops bars with a new pattern

    def set_bars(self, bar_nodes, bar_material, area, bar_map) -> None:
        """Bars after the Q4 elements (after set_pattern, before set_colors)."""
        bar_nodes = _i32(bar_nodes).reshape(-1, 2)
        bar_material, area, bar_map = _i32(bar_material), _f64(area), _i32(bar_map).reshape(-1, 16)
        self._check(self._lib.lfg_set_bars(self._ctx, bar_nodes.shape[0], _ptr(bar_nodes), _ptr(bar_material),
                                           _ptr(area), _ptr(bar_map)))
        self.n_bars = bar_nodes.shape[0]

    def pin_zero_pivots(self, raise_on_error: bool = True) -> dict:
        """The zero-pivot sweep after dirichlet(): pinned count, first loaded zero pivot."""
        rep = PivotReport()
        rc = self._lib.lfg_pin_zero_pivots(self._ctx, ctypes.byref(rep))
        if rc != OK and raise_on_error:
            self._check(rc)
        return {name: getattr(rep, name) for name, _ in PivotReport._fields_}

    def set_colors(self, color_ptr, color_elems) -> None:
        color_ptr, color_elems = _i32(color_ptr), _i32(color_elems)
        self._check(self._lib.lfg_set_colors(self._ctx, color_ptr.shape[0] - 1, _ptr(color_ptr), _ptr(color_elems)))

    def element_stiffness(self) -> np.ndarray:
        out = np.empty((self.n_elems, 64))
        self._check(self._lib.lfg_element_stiffness(self._ctx, _ptr(out)))
        return out

    def assemble(self, strategy: str = "serial") -> dict:
        rep = AssemblyReport()
        self._check(self._lib.lfg_assemble(self._ctx, ASSEMBLY[strategy], ctypes.byref(rep)))
        return rep.as_dict()

    def stiffness(self) -> np.ndarray:
        out = np.empty(self.nnz)
        self._check(self._lib.lfg_get_stiffness(self._ctx, _ptr(out)))
        return out

    def set_stiffness(self, values) -> None:
        values = _f64(values)
        self._check(self._lib.lfg_set_stiffness(self._ctx, _ptr(values)))

    def set_load(self, f) -> None:
        f = _f64(f)
        self._check(self._lib.lfg_set_load(self._ctx, _ptr(f)))

    def dirichlet(self, dofs, values) -> None:
        dofs, values = _i32(dofs), _f64(values)
        self._check(self._lib.lfg_dirichlet(self._ctx, dofs.shape[0], _ptr(dofs), _ptr(values)))

    def system(self):
        kc = np.empty(self.nnz)
        fc = np.empty(self.n)
        self._check(self._lib.lfg_get_system(self._ctx, _ptr(kc), _ptr(fc)))
        return kc, fc

    def pcg(self, preconditioner: str = "jacobi", precision: str = "fp64", stop_rule: str = "reference",
            tolerance: float = 1e-10, max_iterations: int = 100000, check_every: int = 0,
            replace_drop: float = 0.0, direct_launch: bool = False, spmv_variant: int = 0,
            raise_on_error: bool = True, mixed_switch: float = 0.0, storage: str = "csr",
            refine_passes: int = 0, refine_reduction: float = 0.0):
        opt = PcgOptions(PRECOND[preconditioner], PRECISION[precision], STOP[stop_rule], check_every, tolerance,
                         max_iterations, replace_drop, mixed_switch, int(direct_launch), spmv_variant,
                         STORAGE[storage], refine_passes, refine_reduction)
        rep = PcgReport()
        u = np.empty(self.n)
        rc = self._lib.lfg_pcg(self._ctx, ctypes.byref(opt), _ptr(u), ctypes.byref(rep))
        if rc != OK and raise_on_error:
            self._check(rc)
        return u, rep.as_dict()

    def adjoint_solve(self, dJdu, preconditioner: str = "jacobi", precision: str = "fp64",
                      stop_rule: str = "reference", tolerance: float = 1e-10, max_iterations: int = 100000,
                      check_every: int = 0, spmv_variant: int = 0, storage: str = "csr"):
        """K lambda = dJ/du on the constrained system of the last forward solve."""
        dJdu = _f64(dJdu)
        opt = PcgOptions(PRECOND[preconditioner], PRECISION[precision], STOP[stop_rule], check_every, tolerance,
                         max_iterations, 0.0, 0.0, 0, spmv_variant, STORAGE[storage], 0)
        rep = PcgReport()
        lam = np.empty(self.n)
        self._check(self._lib.lfg_adjoint_solve(self._ctx, ctypes.byref(opt), _ptr(dJdu), _ptr(lam),
                                                ctypes.byref(rep)))
        return lam, rep.as_dict()

    def sensitivity(self, lam=None, dE_drho=None) -> np.ndarray:
        """-lambda_e^T (dKe/dE) u_e (times dE_drho[e]) per element; lam None = last adjoint."""
        lam = None if lam is None else _f64(lam)
        dE = None if dE_drho is None else _f64(dE_drho)
        grad = np.empty(self.n_elems)
        self._check(self._lib.lfg_sensitivity(self._ctx, _ptr(lam), _ptr(dE), _ptr(grad)))
        return grad

    def reactions(self) -> np.ndarray:
        out = np.empty(self.n)
        self._check(self._lib.lfg_reactions(self._ctx, _ptr(out)))
        return out

    def postprocess(self):
        """Element fields (Q4 elements, then bars) and smoothed nodal stresses."""
        e = self.n_elems + getattr(self, "n_bars", 0)
        centre = np.empty((e, 7))
        gp = np.empty((e, 4, 9))
        nodal = np.empty((self.n_nodes, 4))
        self._check(self._lib.lfg_postprocess(self._ctx, _ptr(centre), _ptr(gp), _ptr(nodal)))
        return centre, gp, nodal

    def precond_apply(self, preconditioner: str, vec) -> np.ndarray:
        vec = _f64(vec)
        out = np.empty_like(vec)
        self._check(self._lib.lfg_precond_apply(self._ctx, PRECOND[preconditioner], _ptr(vec), _ptr(out)))
        return out

    def layout_info(self) -> dict:
        """SpMV layouts the current pattern admits (node slices, node-stencil forms)."""
        info = LayoutInfo()
        self._check(self._lib.lfg_layout_info_get(self._ctx, ctypes.byref(info)))
        return info.as_dict()

    def spmv(self, x, precision: str = "fp64", variant: int = 0, block_order: bool = False) -> np.ndarray:
        """y = Kc x with the given kernel; block_order: the block storage summation order."""
        x = _f64(x)
        y = np.empty_like(x)
        v = variant | (SPMV_BLOCK_ORDER if block_order else 0)
        self._check(self._lib.lfg_spmv(self._ctx, PRECISION[precision], v, _ptr(x), _ptr(y)))
        return y

    def spmv_noalloc(self, x_dev: int, y_dev: int, xy_dev: int = 0, precision: str = "fp64", variant: int = 0,
                     raise_on_error: bool = True) -> int:
        """y_dev = Kc x_dev (and x.y into xy_dev) on caller-owned device pointers, without any
        device allocation (lfg_spmv_noalloc); returns the status."""
        rc = self._lib.lfg_spmv_noalloc(self._ctx, PRECISION[precision], variant, ctypes.c_void_p(x_dev),
                                        ctypes.c_void_p(y_dev), ctypes.c_void_p(xy_dev or None))
        if rc != OK and raise_on_error:
            self._check(rc)
        return rc

    def set_subdomains(self, node_subdomain, n_subdomains: int) -> None:
        """Subdomains of the block preconditioner (LFG_PRECOND_SUBDOMAIN), per node index."""
        ns = _i32(node_subdomain)
        self._check(self._lib.lfg_set_subdomains(self._ctx, n_subdomains, _ptr(ns)))

    def bench_copy(self, nbytes: int, reps: int = 20) -> float:
        ms = ctypes.c_double(0.0)
        self._check(self._lib.lfg_bench_copy(self._ctx, nbytes, reps, ctypes.byref(ms)))
        return ms.value

    def bench_stream(self, nbytes: int, reps: int = 20) -> float:
        """Average ms of the streaming calibration over five buffers of nbytes (4 read, 1 written)."""
        ms = ctypes.c_double(0.0)
        self._check(self._lib.lfg_bench_stream(self._ctx, nbytes, reps, ctypes.byref(ms)))
        return ms.value

    def bench_spmv(self, precision: str = "fp64", variant: int = 0, reps: int = 20) -> float:
        ms = ctypes.c_double(0.0)
        self._check(self._lib.lfg_bench_spmv(self._ctx, PRECISION[precision], variant, reps, ctypes.byref(ms)))
        return ms.value
