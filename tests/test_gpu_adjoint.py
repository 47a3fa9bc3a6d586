"""GPU adjoint: the adjoint PCG solve and the element-sensitivity kernel against the
CPU FP64 gradient (oracle sensitivities on direct solves), central finite
differences through the GPU forward path, Euler's identity, and the
torch.autograd harness.

Bars: the sensitivity kernel is bit-identical to the oracle on the same u and
lambda; GPU gradients within 1e-6 relative (norm-wise) of the CPU FP64 gradient;
per element within 1e-6 relative of central differences with h = 1e-6 in rho.
"""
import numpy as np
import pytest
import scipy.sparse

from lightfem_amd import meshgen, native
from lightfem_amd.adjoint import ElasticProblem
from oracle import lfo

pytestmark = pytest.mark.gpu

EMIN = 1e-9


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.int64)


def simp(rho):
    return EMIN + rho ** 3 * (1.0 - EMIN)


def dsimp(rho):
    return 3.0 * rho ** 2 * (1.0 - EMIN)


def tip_selector(problem):
    """dJ/du of J = vertical displacement of the bottom-right node."""
    xy = problem.ix.xy
    corner = int(np.argmax(xy[:, 0] - xy[:, 1]))
    e_d = np.zeros(problem.n)
    e_d[2 * corner + 1] = 1.0
    return e_d, 2 * corner + 1


@pytest.fixture(scope="module")
def dev():
    d = native.Device(0)
    yield d
    d.close()


def test_adjoint_small_matches_oracle(dev):
    print("case adjoint 24x16 SIMP", flush=True)
    n = (24, 16)
    rho = 0.3 + 0.7 * meshgen.uniform01(99, n[0] * n[1])
    model = meshgen.structured_cantilever(*n, young=simp(rho))
    prob = ElasticProblem(model, device=dev)
    u = prob.solve(simp(rho))
    e_d, d = tip_selector(prob)
    lam, rep = dev.adjoint_solve(e_d, tolerance=1e-12, stop_rule="relative")
    assert rep["converged"] == 1
    ix = prob.ix
    ds = lfo.DirectSystem(ix.xy, ix.enodes, prob.props, ix.plane_mode, ix.load, ix.cdofs, ix.cvals)
    u_ref, lam_ref = ds.solve(ds.fc), ds.solve(ds.adjoint_rhs(e_d))
    assert np.linalg.norm(u - u_ref) <= 1e-9 * np.linalg.norm(u_ref)
    assert np.linalg.norm(lam - lam_ref) <= 1e-9 * np.linalg.norm(lam_ref)
    assert np.all(lam[ix.cdofs] == 0.0)
    # the kernel on the device-resident u and lambda equals the oracle on the same vectors
    dE = dsimp(rho)
    g = dev.sensitivity(None, dE)
    assert np.array_equal(bits(g), bits(lfo.sensitivity(ix.xy, ix.enodes, prob.props, ix.plane_mode, u, lam, dE)))
    g_ref = lfo.sensitivity(ix.xy, ix.enodes, prob.props, ix.plane_mode, u_ref, lam_ref, dE)
    assert np.linalg.norm(g - g_ref) <= 1e-8 * np.linalg.norm(g_ref)
    # compliance is self-adjoint: lambda = u
    gc = prob.compliance_gradient(u, dE)
    assert np.array_equal(bits(gc), bits(lfo.sensitivity(ix.xy, ix.enodes, prob.props, ix.plane_mode, u, u, dE)))
    assert np.all(gc <= 0.0)


def without_rigid_motion(ix, u, e):
    """u minus the least-squares rigid motion (translation + small rotation) of element e's nodes."""
    nodes = ix.enodes[e]
    xy = ix.xy[nodes]
    c = xy.mean(axis=0)
    a = np.zeros((8, 3))
    a[0::2, 0] = 1.0
    a[1::2, 1] = 1.0
    a[0::2, 2] = -(xy[:, 1] - c[1])
    a[1::2, 2] = xy[:, 0] - c[0]
    dofs = np.stack([2 * nodes, 2 * nodes + 1], axis=1).reshape(8)
    t = np.linalg.lstsq(a, u[dofs], rcond=None)[0]
    w = u.copy()
    w[dofs] -= a @ t
    return w


def correction_fd(prob, k_base, base_u, young, rho, e, h, selectors):
    """Central differences of c^T u(rho_e +- h) through the GPU forward path (assembly
    at the perturbed modulus, Dirichlet elimination, PCG), each solved for its
    correction d = u(rho_e +- h) - u0 from K(rho_e +- h) d = -(K(rho_e +- h) - K) u0
    with the device-assembled values:
    * solving for d keeps the O(h) difference from cancelling against u0;
    * the base residual f - K u0 stays out of the right-hand side, so its
      K^-1-amplified FP64 floor does not swamp d;
    * K(rho_e +- h) - K is nonzero only on element e and annihilates its rigid
      motions, so u0 enters with element e's rigid motion removed: near the free end
      that motion is ~1e6 times the element's deformation, and the rounding of the
      difference times it would otherwise dominate at h = 1e-6."""
    dev = prob.device
    ix = prob.ix
    w = without_rigid_motion(ix, base_u, e)
    deltas = []
    for step in (h, -h):
        props = prob.props.copy()
        props[:, 0] = young
        props[e, 0] = simp(rho[e] + step)
        dev.set_materials(props)
        dev.assemble(prob.assembly)
        dK = scipy.sparse.csr_matrix((dev.stiffness() - k_base, prob.col, prob.row_ptr), shape=(prob.n, prob.n))
        rhs = -(dK @ w)
        rhs[ix.cdofs] = 0.0
        dev.set_load(rhs)
        dev.dirichlet(ix.cdofs, np.zeros(ix.cdofs.shape[0]))
        delta, rep = dev.pcg("jacobi", "fp64", "relative", 1e-12, 1_000_000)
        assert rep["converged"] == 1
        deltas.append(delta)
    dev.lightfem_owner = None  # the device no longer holds prob's last solve
    diff = deltas[0] - deltas[1]
    return [c @ diff / (2 * h) for c in selectors]


def test_c4_gradients_400(dev):
    """400x400 SIMP cantilever: compliance and tip-displacement gradients."""
    print("case adjoint 400x400 SIMP (compliance, tip uy)", flush=True)
    rho = meshgen.simp_density(400)
    young = simp(rho)
    dE = dsimp(rho)
    prob = ElasticProblem(meshgen.simp_cantilever(400), device=dev)
    u = prob.solve(young)
    k_base = dev.stiffness()
    e_d, d = tip_selector(prob)
    g_c = prob.compliance_gradient(u, dE)
    g_d = prob.gradient(e_d, dE)
    J_c, J_d = prob.ix.load @ u, u[d]
    # CPU FP64 gradient on direct solves
    print("case adjoint 400x400: CPU FP64 direct gradient", flush=True)
    ix = prob.ix
    ds = lfo.DirectSystem(ix.xy, ix.enodes, prob.props, ix.plane_mode, ix.load, ix.cdofs, ix.cvals)
    u_ref = ds.solve(ds.fc)
    lam_ref = ds.solve(ds.adjoint_rhs(e_d))
    gc_ref = lfo.sensitivity(ix.xy, ix.enodes, prob.props, ix.plane_mode, u_ref, u_ref, dE)
    gd_ref = lfo.sensitivity(ix.xy, ix.enodes, prob.props, ix.plane_mode, u_ref, lam_ref, dE)
    err_c = np.linalg.norm(g_c - gc_ref) / np.linalg.norm(gc_ref)
    err_d = np.linalg.norm(g_d - gd_ref) / np.linalg.norm(gd_ref)
    print(f"gradient vs CPU FP64: compliance {err_c:.3e}, tip {err_d:.3e}", flush=True)
    assert err_c <= 1e-6 and err_d <= 1e-6
    # Euler: sum_e E_e dJ/dE_e = -J (supports are homogeneous)
    for g, J in ((g_c, J_c), (g_d, J_d)):
        assert abs(young @ (g / dE) + J) <= 1e-7 * abs(J)
    # central differences on 20 seeded elements
    print("case adjoint 400x400: finite differences on 20 elements", flush=True)
    elems = np.random.default_rng(meshgen.SEED_DENSITY).choice(prob.n_elems, 20, replace=False)
    fd = np.array([correction_fd(prob, k_base, u, young, rho, int(e), 1e-6, (ix.load, e_d)) for e in elems])
    for k, (name, g) in enumerate((("compliance", g_c), ("tip uy", g_d))):
        for e, a, b in zip(elems, g[elems], fd[:, k]):
            print(f"{name} element {e}: adjoint {a:.15e} fd {b:.15e} rel {abs(a - b) / abs(b):.2e}", flush=True)
        err = np.linalg.norm(g[elems] - fd[:, k]) / np.linalg.norm(fd[:, k])
        worst = np.max(np.abs(g[elems] - fd[:, k]) / np.abs(fd[:, k]))
        print(f"{name}: FD relative error over the 20 elements {err:.3e}, worst element {worst:.3e}", flush=True)
        assert err <= 1e-6 and worst <= 1e-6, name


def test_autograd_function(dev):
    import torch
    from lightfem_amd.autograd import displacement
    print("case autograd 8x5", flush=True)
    n = (8, 5)
    rho0 = 0.3 + 0.7 * meshgen.uniform01(5, n[0] * n[1])
    prob = ElasticProblem(meshgen.structured_cantilever(*n, young=simp(rho0)), device=dev)
    e_d, d = tip_selector(prob)
    load = torch.from_numpy(prob.ix.load)
    rho = torch.tensor(rho0, dtype=torch.float64, requires_grad=True)
    u = displacement(EMIN + rho ** 3 * (1.0 - EMIN), prob)
    J = u[d] + load @ u
    J.backward()
    ix = prob.ix
    ds = lfo.DirectSystem(ix.xy, ix.enodes, prob.props, ix.plane_mode, ix.load, ix.cdofs, ix.cvals)
    u_ref = ds.solve(ds.fc)
    lam_ref = ds.solve(ds.adjoint_rhs(e_d + ix.load))
    g_ref = lfo.sensitivity(ix.xy, ix.enodes, prob.props, ix.plane_mode, u_ref, lam_ref, dsimp(rho0))
    assert np.linalg.norm(rho.grad.numpy() - g_ref) <= 1e-8 * np.linalg.norm(g_ref)
    # a later forward at other moduli does not change an earlier graph's gradient
    rho2 = torch.tensor(rho0, dtype=torch.float64, requires_grad=True)
    u2 = displacement(EMIN + rho2 ** 3 * (1.0 - EMIN), prob)
    _ = displacement(torch.from_numpy(simp(rho0[::-1].copy())), prob)
    (u2[d] + load @ u2).backward()
    assert np.linalg.norm(rho2.grad.numpy() - g_ref) <= 1e-8 * np.linalg.norm(g_ref)
    # torch's own finite-difference check of the harness (solves tight enough that
    # the differences of u are not solver noise)
    tight = ElasticProblem(meshgen.structured_cantilever(*n, young=simp(rho0)), device=dev, tolerance=1e-14)
    small = torch.tensor(rho0, dtype=torch.float64, requires_grad=True)
    assert torch.autograd.gradcheck(lambda r: displacement(EMIN + r ** 3 * (1.0 - EMIN), tight)[[d]], (small,),
                                    eps=1e-6, atol=1e-7, rtol=1e-4)
