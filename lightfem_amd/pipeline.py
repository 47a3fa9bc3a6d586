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
    load_pos: Optional[np.ndarray] = None  # node index of each [loads] row

    def load_vector(self, values: np.ndarray) -> np.ndarray:
        """Nodal load vector for per-row (fx, fy) values of the model's [loads] rows."""
        load = np.zeros(2 * self.n_nodes)
        np.add.at(load, 2 * self.load_pos, values[:, 0])
        np.add.at(load, 2 * self.load_pos + 1, values[:, 1])
        return load

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
                        plane_mode=plane, load=load, cdofs=dofs_a[order], cvals=vals_a[order], rcm_order=rcm_order,
                        load_pos=lpos)


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


@dataclasses.dataclass
class BatchSystem:
    """Independent problems merged into one disjoint mesh on one device."""

    device: native.Device
    indexed: list            # IndexedModel per problem
    node_ptr: np.ndarray     # (P+1,) node ranges, multiples of 64 (padding nodes unused)
    loads: np.ndarray        # (cases, dofs) merged load vectors
    cdofs: np.ndarray
    cvals: np.ndarray
    row_ptr: np.ndarray
    col: np.ndarray
    host_ms: dict

    def split(self, u: np.ndarray) -> list:
        """Per-problem (cases, problem dofs) views of a (cases, dofs) batch solution."""
        return [u[:, 2 * self.node_ptr[p]:2 * (self.node_ptr[p] + ix.n_nodes)] for p, ix in enumerate(self.indexed)]


def batch_setup(models: list, case_loads: list, device: Optional[native.Device] = None,
                assembly: str = "coloring") -> BatchSystem:
    """Numbers every model (model.cpp:24-92), pads each to whole 64-node slices, merges them
    into one disjoint mesh with one material table (plane mode per material), builds the
    pattern and colours of the union on the host, assembles on the device and applies each
    load case's Dirichlet elimination. case_loads[p] is (cases, rows of the model's [loads], 2)."""
    host_ms = {}
    t0 = time.perf_counter()
    ixs = [index_model(m) for m in models]
    n_cases = int(np.asarray(case_loads[0]).shape[0])
    sizes = np.array([ix.n_nodes for ix in ixs], dtype=np.int64)
    padded = (sizes + 63) // 64 * 64
    node_ptr = np.concatenate([[0], np.cumsum(padded)]).astype(np.int64)
    n_nodes = int(node_ptr[-1])
    xy = np.zeros((n_nodes, 2))
    enodes, elem_mat, mats, modes, cdofs, cvals = [], [], [], [], [], []
    loads = np.zeros((n_cases, 2 * n_nodes))
    mat_off = 0
    for p, ix in enumerate(ixs):
        off = int(node_ptr[p])
        xy[off:off + ix.n_nodes] = ix.xy
        enodes.append(ix.enodes + off)
        elem_mat.append(ix.elem_mat + mat_off)
        mats.append(ix.materials)
        modes.append(np.full(ix.materials.shape[0], ix.plane_mode, dtype=np.int32))
        mat_off += ix.materials.shape[0]
        cdofs.append(ix.cdofs + 2 * off)
        cvals.append(ix.cvals)
        for c in range(n_cases):
            loads[c, 2 * off:2 * (off + ix.n_nodes)] = ix.load_vector(np.asarray(case_loads[p][c]))
    enodes = np.concatenate(enodes)
    elem_mat = np.concatenate(elem_mat)
    host_ms["index_ms"] = 1e3 * (time.perf_counter() - t0)
    t0 = time.perf_counter()
    row_ptr, col, emap = native.csr_pattern(n_nodes, enodes)
    host_ms["pattern_ms"] = 1e3 * (time.perf_counter() - t0)
    t0 = time.perf_counter()
    _, color_ptr, color_elems = native.element_colors(n_nodes, enodes)
    host_ms["coloring_ms"] = 1e3 * (time.perf_counter() - t0)
    dev = device or native.Device(0)
    dev.set_mesh(xy, enodes, elem_mat, np.concatenate(mats), native.PLANE_STRESS)
    dev.set_material_modes(np.concatenate(modes))
    dev.set_pattern(row_ptr, col, emap)
    dev.set_colors(color_ptr, color_elems)
    dev.assemble(assembly)
    dev.batch_set_problems(node_ptr)
    cd, cv = np.concatenate(cdofs), np.concatenate(cvals)
    dev.batch_set_loads(loads, cd, cv)
    return BatchSystem(dev, ixs, node_ptr, loads, cd, cv, row_ptr, col, host_ms)
