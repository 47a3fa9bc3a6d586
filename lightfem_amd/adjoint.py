"""Adjoint gradients of the Q4 forward solve on the gfx950 backend.

The forward map E -> u assembles K(E) on the GPU, eliminates the Dirichlet dofs and
solves K u = f by PCG (src/solver.cpp:243-413, 532-747). For a functional J(u) the
gradient with respect to the per-element moduli is

    dJ/dE_e = -lambda_e^T (dKe/dE_e) u_e,   K lambda = dJ/du,  lambda = 0 on constrained dofs,

one adjoint PCG solve on the same constrained operator (reusing the forward's
preconditioner) and one element-sensitivity kernel launch. The reference has no
gradient code; this is the derivative of its forward path. ``lightfem_amd.autograd``
wraps it as a ``torch.autograd.Function``.
"""
from __future__ import annotations

import itertools
from typing import Optional

import numpy as np

from . import native, pipeline
from .meshgen import Q4Model


class ElasticProblem:
    """A Q4 model on one device whose per-element moduli change between solves.

    Every element gets its own material row (E, nu, t), so ``solve(young)`` only
    replaces the moduli, re-assembles, re-applies the supports and solves; mesh,
    pattern and colours stay on the device. Solutions and gradients are in the
    model's dof order (nodes sorted by id, node i -> dofs 2i, 2i+1) and elements in
    ascending id order.
    """

    _tokens = itertools.count(1)

    def __init__(self, model: Q4Model, device: Optional[native.Device] = None, assembly: str = "coloring",
                 preconditioner: str = "jacobi", tolerance: float = 1e-12, stop_rule: str = "relative",
                 max_iterations: int = 1_000_000):
        self.ix = pipeline.index_model(model)
        self.device = device or native.Device(0)
        self.assembly = assembly
        self.solver = dict(preconditioner=preconditioner, tolerance=tolerance, stop_rule=stop_rule,
                           max_iterations=max_iterations)
        ix = self.ix
        self.props = np.ascontiguousarray(ix.materials[ix.elem_mat], dtype=np.float64)
        self.n = 2 * ix.n_nodes
        self.n_elems = ix.enodes.shape[0]
        self.row_ptr, self.col, self._emap = native.csr_pattern(ix.n_nodes, ix.enodes)
        _, self._color_ptr, self._color_elems = native.element_colors(ix.n_nodes, ix.enodes)
        self.token = 0
        self.last_report: dict = {}
        self._bind()

    def _bind(self) -> None:
        """Loads this problem's mesh, pattern and colours (a device holds one problem at a time)."""
        dev, ix = self.device, self.ix
        dev.set_mesh(ix.xy, ix.enodes, np.arange(self.n_elems), self.props, ix.plane_mode)
        dev.set_pattern(self.row_ptr, self.col, self._emap)
        dev.set_colors(self._color_ptr, self._color_elems)
        dev.lightfem_owner = (self, 0)
        self.token = 0

    def _current(self, token: int) -> bool:
        return token != 0 and getattr(self.device, "lightfem_owner", None) == (self, token)

    @property
    def young(self) -> np.ndarray:
        return self.props[:, 0].copy()

    def solve(self, young) -> np.ndarray:
        """u(E): assembly, Dirichlet elimination and PCG on the device."""
        young = np.ascontiguousarray(young, dtype=np.float64)
        if young.shape != (self.n_elems,):
            raise ValueError(f"expected {self.n_elems} element moduli, got shape {young.shape}")
        self.props[:, 0] = young
        dev = self.device
        owner = getattr(dev, "lightfem_owner", None)
        if owner is None or owner[0] is not self:
            self._bind()
        self.token = 0
        dev.lightfem_owner = (self, 0)
        dev.set_materials(self.props)
        dev.assemble(self.assembly)
        dev.set_load(self.ix.load)
        dev.dirichlet(self.ix.cdofs, self.ix.cvals)
        u, rep = dev.pcg(precision="fp64", **self.solver)
        self.last_report = rep
        self.token = next(self._tokens)
        dev.lightfem_owner = (self, self.token)
        return u

    def gradient(self, dJdu, dE_drho=None, lam: Optional[np.ndarray] = None) -> np.ndarray:
        """dJ/dE_e (or dJ/drho_e with dE_drho) for the last solve; lam skips the adjoint solve."""
        if not self._current(self.token):
            raise RuntimeError("solve() must precede gradient() (the device holds another solve)")
        dev = self.device
        if lam is None:
            lam, rep = dev.adjoint_solve(dJdu, **self.solver)
            self.last_report = rep
        return dev.sensitivity(lam, dE_drho)

    def vjp(self, token: int, young, grad_u) -> np.ndarray:
        """Vector-Jacobian product of solve() at `young` (re-solved when another solve came after)."""
        if not self._current(token):
            self.solve(young)
        return self.gradient(grad_u)

    def compliance_gradient(self, u: np.ndarray, dE_drho=None) -> np.ndarray:
        """Self-adjoint case J = f^T u: lambda = u, no adjoint solve."""
        return self.gradient(None, dE_drho, lam=u)
