"""Python mirror of ``LightElasticSolver::solve`` (src/solver.cpp:875-960) over the
two C ABIs, for Q4 models held as :class:`lightfem_amd.meshgen.Q4Model`.

Stage order is the reference's: sort by id and number dofs (model.cpp:24-92),
optional RCM (reordering.cpp:11-62, model.cpp:41-72), CSR pattern and map
(sparse.cpp:15-58), element stiffness + assembly strategy on the GPU
(solver.cpp:219-301), nodal loads (:303-308), Dirichlet elimination on the GPU
(:340-413), PCG on the GPU (:532-747), reactions (:953-954) and element / nodal
fields (:955-956) on the GPU.
"""
from __future__ import annotations

import dataclasses
import time
from typing import Optional

import numpy as np

from . import native
from .meshgen import Q4Model


@dataclasses.dataclass
class IndexedModel:
    node_ids: np.ndarray    # node ids in dof order (node i owns dofs 2i, 2i+1)
    xy: np.ndarray          # coordinates in dof order
    enodes: np.ndarray      # (E, 4) node indices, elements in ascending id order
    elem_ids: np.ndarray
    elem_mat: np.ndarray    # material index (into materials) per element
    materials: np.ndarray   # (M, 3) E, nu, t in ascending material id order
    plane_mode: int
    load: np.ndarray        # assembled nodal load vector
    cdofs: np.ndarray       # constrained dofs, ascending
    cvals: np.ndarray
    rcm_order: Optional[np.ndarray] = None

    @property
    def n_nodes(self) -> int:
        return int(self.node_ids.shape[0])


def index_model(model: Q4Model, reordering: str = "none") -> IndexedModel:
    """sort_and_index (model.cpp:24-39) and, with reordering="rcm", index_with_node_order."""
    node_sort = np.argsort(model.node_ids, kind="stable")
    sorted_ids = model.node_ids[node_sort]
    if np.any(np.diff(sorted_ids) == 0):
        raise native.LightFemError(native.ERR_ARG, "duplicate node ID")
    elem_sort = np.argsort(model.elem_ids, kind="stable")
    enode_ids = model.enodes[elem_sort]
    order_ids = sorted_ids
    rcm_order = None
    if reordering == "rcm":
        rcm_order = native.rcm(sorted_ids, enode_ids).astype(np.int64)
        order_ids = rcm_order
    elif reordering != "none":
        raise native.LightFemError(native.ERR_ARG, "reordering must be none or rcm")
    # id -> position in the (possibly reordered) node list
    rank = np.searchsorted(sorted_ids, order_ids)
    position = np.empty(sorted_ids.shape[0], dtype=np.int64)
    position[rank] = np.arange(sorted_ids.shape[0])

    def pos_of(ids):
        idx = np.searchsorted(sorted_ids, ids)
        if np.any(idx >= sorted_ids.shape[0]) or np.any(sorted_ids[np.minimum(idx, sorted_ids.shape[0] - 1)] != ids):
            raise native.LightFemError(native.ERR_ARG, "validated element references an unknown node")
        return position[idx]

    xy = model.xy[node_sort][rank]
    enodes = pos_of(enode_ids)
    mat_sort = np.argsort(model.mat_ids, kind="stable")
    mat_ids = model.mat_ids[mat_sort]
    elem_mat = np.searchsorted(mat_ids, model.elem_mat[elem_sort])
    n_nodes = sorted_ids.shape[0]
    load = np.zeros(2 * n_nodes)
    lpos = pos_of(model.load_nodes)
    np.add.at(load, 2 * lpos, model.load_values[:, 0])
    np.add.at(load, 2 * lpos + 1, model.load_values[:, 1])
    bpos = pos_of(model.bc_nodes)
    dofs, vals = [], []
    for p, (ux, uy) in zip(bpos.tolist(), model.bc_values.tolist()):
        if not np.isnan(ux):
            dofs.append(2 * p)
            vals.append(ux)
        if not np.isnan(uy):
            dofs.append(2 * p + 1)
            vals.append(uy)
    dofs_a = np.array(dofs, dtype=np.int64)
    vals_a = np.array(vals, dtype=np.float64)
    order = np.argsort(dofs_a, kind="stable")
    plane = native.PLANE_STRAIN if model.settings.get("plane_mode") == "plane_strain" else native.PLANE_STRESS
    return IndexedModel(node_ids=order_ids.astype(np.int64), xy=xy, enodes=enodes.astype(np.int64),
                        elem_ids=model.elem_ids[elem_sort], elem_mat=elem_mat, materials=model.mat_props[mat_sort],
                        plane_mode=plane, load=load, cdofs=dofs_a[order], cvals=vals_a[order], rcm_order=rcm_order)


@dataclasses.dataclass
class DeviceSystem:
    """A model uploaded, assembled and constrained on one device."""

    device: native.Device
    indexed: IndexedModel
    row_ptr: np.ndarray
    col: np.ndarray
    emap: np.ndarray
    colors: np.ndarray
    color_ptr: np.ndarray
    assembly_report: dict
    host_ms: dict


def setup(model: Q4Model, device: Optional[native.Device] = None, assembly: str = "serial",
          reordering: str = "none") -> DeviceSystem:
    host_ms = {}
    t0 = time.perf_counter()
    ix = index_model(model, reordering)
    host_ms["index_ms"] = 1e3 * (time.perf_counter() - t0)
    t0 = time.perf_counter()
    row_ptr, col, emap = native.csr_pattern(ix.n_nodes, ix.enodes)
    host_ms["pattern_ms"] = 1e3 * (time.perf_counter() - t0)
    t0 = time.perf_counter()
    colors, color_ptr, color_elems = native.element_colors(ix.n_nodes, ix.enodes)
    host_ms["coloring_ms"] = 1e3 * (time.perf_counter() - t0)
    dev = device or native.Device(0)
    dev.set_mesh(ix.xy, ix.enodes, ix.elem_mat, ix.materials, ix.plane_mode)
    dev.set_pattern(row_ptr, col, emap)
    dev.set_colors(color_ptr, color_elems)
    rep = dev.assemble(assembly)
    dev.set_load(ix.load)
    dev.dirichlet(ix.cdofs, ix.cvals)
    return DeviceSystem(dev, ix, row_ptr, col, emap, colors, color_ptr, rep, host_ms)


def solve(model: Q4Model, assembly: str = "serial", reordering: str = "none", preconditioner: str = "jacobi",
          precision: str = "fp64", stop_rule: str = "reference", tolerance: float = 1e-10,
          max_iterations: int = 100000, fields: bool = True, device: Optional[native.Device] = None) -> dict:
    sys_ = setup(model, device, assembly, reordering)
    dev = sys_.device
    u, rep = dev.pcg(preconditioner, precision, stop_rule, tolerance, max_iterations)
    out = {"u": u, "pcg": rep, "assembly": sys_.assembly_report, "host_ms": sys_.host_ms, "system": sys_,
           "reaction": dev.reactions()}
    if fields:
        out["centre"], out["gauss_points"], out["nodal"] = dev.postprocess()
    return out
