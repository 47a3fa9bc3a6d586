"""GPU bars (truss members): element matrices and every assembly strategy, the
zero-pivot sweep, the solve, reactions, bar fields and nodal smoothing on the
reference's own bar models, against its dumps and the oracle on the same inputs.

Bars: K (serial and coloured), Kc, the pinned system, reactions and fields bit for bit
(given u); atomic K within 1e-13; displacements within 1e-9 of the reference's with
iteration counts within 1% + 2."""
import numpy as np
import pytest

from conftest import load_golden
from lightfem_amd import native
from oracle import lfo
from test_oracle_trusses import TRUSS_CASES, decode

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.int64)


@pytest.fixture(scope="module")
def dev():
    d = native.Device(0)
    yield d
    d.close()


def upload_bars(dev, g):
    """Mesh, pattern, bars and colours of a golden bar model; returns the oracle inputs."""
    q4, bars = decode(g)
    xy = g["xy"]
    n_nodes, n_q4, n_bar = xy.shape[0], q4.shape[0], bars.shape[0]
    enode, bnode = q4[:, 0::2] // 2, bars[:, 0::2] // 2
    # one material row per element: Q4 (E, nu, t), bars (E, 0, 1) (a bar uses E only)
    bar_rows = np.column_stack([g["truss_props"][:, 0], np.zeros(n_bar), np.ones(n_bar)])
    materials = np.concatenate([g["elem_material"].reshape(-1, 3), bar_rows])
    dev.set_mesh(xy, enode, np.arange(n_q4), materials, native.PLANE_STRESS)
    row_ptr, col, emap, bmap = native.csr_pattern(n_nodes, enode, bnode)
    dev.set_pattern(row_ptr, col, emap)
    dev.set_bars(bnode, n_q4 + np.arange(n_bar), g["truss_props"][:, 1], bmap)
    colors, ptr, elems = native.element_colors(n_nodes, enode, bnode)
    dev.set_colors(ptr, elems)
    return q4, bars, row_ptr, col, colors


@pytest.mark.parametrize("name", TRUSS_CASES)
def test_bar_models_match_reference(dev, golden_index, name):
    g = load_golden(name)
    meta = golden_index[name]["meta"]
    print(f"case bars {name}: dofs={g['f'].shape[0]} nnz={g['col'].shape[0]}", flush=True)
    q4, bars, row_ptr, col, colors = upload_bars(dev, g)
    nnz = col.shape[0]
    n_q4, n_bar = q4.shape[0], bars.shape[0]
    ke_flat = np.concatenate([g["ke"].ravel(), g["ke_truss"].ravel()])
    emap = g["emap_flat"]
    serial = lfo.assemble_flat(ke_flat, emap, nnz)
    # coloured: colour by colour (Q4 and bars of a colour touch disjoint values)
    split = np.cumsum([64] * n_q4 + [16] * n_bar)[:-1]
    kes, maps = np.split(ke_flat, split), np.split(emap, split)
    colored = np.zeros(nnz)
    for c in range(int(colors.max()) + 1):
        for e in np.flatnonzero(colors == c):
            colored[maps[e]] += kes[e]
    dev.assemble("coloring")
    assert np.array_equal(bits(dev.stiffness()), bits(colored))
    dev.assemble("atomic")
    assert np.max(np.abs(dev.stiffness() - serial)) <= 1e-13 * np.max(np.abs(serial))
    dev.assemble("serial")
    assert np.array_equal(bits(dev.stiffness()), bits(serial))
    K = g["K"]
    assert np.array_equal(bits(serial if meta["assembly"] == "serial" else colored), bits(K))
    dev.set_stiffness(K)
    dev.set_load(g["f"])
    dev.dirichlet(g["cdofs"], g["cvals"])
    kc, fc = dev.system()
    assert np.array_equal(bits(kc), bits(g["Kc"])) and np.array_equal(bits(fc), bits(g["fc"]))
    piv = dev.pin_zero_pivots()
    kp, fp, pinned, bad = lfo.pin_zero_pivots(row_ptr, col, g["Kc"], g["fc"])
    assert bad == -1 and piv["dof"] == -1 and piv["pinned"] == pinned
    kc, fc = dev.system()
    assert np.array_equal(bits(kc), bits(kp)) and np.array_equal(bits(fc), bits(fp))
    if dev.layout_info()["sym_slices"]:  # bar blocks mirrored like the Q4 ones (variant 15)
        x = np.random.default_rng(3).standard_normal(row_ptr.shape[0] - 1)
        scale = lfo.spmv(row_ptr, col, np.abs(kp), np.abs(x))
        assert np.all(np.abs(dev.spmv(x, "fp64", 15) - lfo.spmv(row_ptr, col, kp, x)) <= 1e-14 * scale)
    u, rep = dev.pcg("jacobi", "fp64", "reference", meta["tolerance"])
    assert rep["converged"] == 1 and abs(rep["iterations"] - meta["iterations"]) <= 2 + 0.01 * meta["iterations"]
    assert np.linalg.norm(u - g["u"]) <= 1e-9 * np.linalg.norm(g["u"])
    R = dev.reactions()
    assert np.array_equal(bits(R), bits(lfo.spmv(row_ptr, col, K, u) - g["f"]))
    assert np.max(np.abs(R - g["R"])) <= 1e-9 * np.max(np.abs(g["R"]))
    centre, gp, nodal = dev.postprocess()
    xy = g["xy"]
    enode = q4[:, 0::2] // 2
    c_q4, gp_q4 = lfo.element_fields(xy, enode, g["elem_material"], lfo.PLANE_STRESS, u) if n_q4 else (
        np.zeros((0, 7)), np.zeros((0, 4, 9)))
    f_bar = np.array([lfo.truss_fields(xy[d[0::2] // 2].ravel(), g["truss_props"][t][0], u[d])
                      for t, d in enumerate(bars)]).reshape(n_bar, 7)
    gp_bar = np.concatenate([np.zeros((n_bar, 4, 2)), np.repeat(f_bar[:, None, :], 4, axis=1)], axis=2)
    assert np.array_equal(bits(centre), bits(np.concatenate([c_q4, f_bar])))
    assert np.array_equal(bits(gp), bits(np.concatenate([gp_q4.reshape(-1, 4, 9), gp_bar])))
    o_nodal = lfo.smooth_nodal_ex(xy.shape[0], enode, gp_q4.reshape(-1, 36), bars[:, 0::2] // 2, f_bar[:, 3:6])
    assert np.array_equal(bits(nodal), bits(o_nodal))
    assert np.max(np.abs(centre - g["elem_centre"])) <= 1e-9 * np.max(np.abs(g["elem_centre"]))


def test_loaded_zero_pivot_is_a_mechanism(dev):
    """A bar along x with its free end loaded across the bar: the sweep reports the dof
    (solver.cpp:922-934) instead of pinning it; unloaded it is pinned."""
    print("case bars mechanism", flush=True)
    xy = np.array([[0.0, 0.0], [1.0, 0.0]])
    materials = np.array([[2.1e11, 0.0, 1.0]])
    dev.set_mesh(xy, np.zeros((0, 4), dtype=np.int32), np.zeros(0, dtype=np.int32), materials, native.PLANE_STRESS)
    row_ptr, col, emap, bmap = native.csr_pattern(2, np.zeros((0, 4)), np.array([[0, 1]]))
    dev.set_pattern(row_ptr, col, emap)
    dev.set_bars(np.array([[0, 1]]), np.array([0]), np.array([1e-3]), bmap)
    dev.assemble("serial")
    for load, expect_bad in ((1000.0, True), (0.0, False)):
        dev.set_load(np.array([0.0, 0.0, 5.0, load]))
        dev.dirichlet(np.array([0, 1]), np.array([0.0, 0.0]))
        rep = dev.pin_zero_pivots(raise_on_error=False)
        if expect_bad:
            assert rep["dof"] == 3 and rep["load"] == 1000.0 and rep["diag"] == 0.0
            with pytest.raises(native.LightFemError) as err:
                dev.pin_zero_pivots()
            assert err.value.code == native.ERR_MECHANISM
        else:
            assert rep["dof"] == -1 and rep["pinned"] == 1
            u, r = dev.pcg("jacobi", "fp64", "reference", 1e-13)
            assert abs(u[2] - 5.0 / 2.1e8) <= 1e-18 and u[3] == 0.0
