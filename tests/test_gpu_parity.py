"""GPU parity: every device stage against the reference's golden vectors and the
CPU oracle on the same inputs, through the C ABI.

Bars: Ke, serial and coloured K, Dirichlet, preconditioner applications,
reactions and fields are bit-identical (same operation order, no contraction);
atomic K within 1e-13 of serial relative to max|K| (summation order is not
fixed); PCG displacements within 1e-9 relative (norm-wise) of the reference,
with iteration counts within 1% + 2 (dot products sum in a different order).
"""
import hashlib

import numpy as np
import pytest

from conftest import FULL_CASES, case_precond, case_storage, load_golden
from lightfem_amd import meshgen, native, pipeline
from oracle import lfo

pytestmark = pytest.mark.gpu

PRECOND_NAME = {0: "jacobi", 1: "block", 2: "ic0", 3: "subdomain"}


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.int64)


def digest(a):
    dt = "<i8" if a.dtype.kind in "iu" else "<f8"
    return hashlib.sha256(np.ascontiguousarray(a, dtype=dt).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def dev():
    d = native.Device(0)
    yield d
    d.close()


def upload(dev, g, meta):
    enode = g["edofs"][:, 0::2] // 2
    n_nodes = g["xy"].shape[0]
    plane = native.PLANE_STRAIN if meta["plane_mode"] == "plane_strain" else native.PLANE_STRESS
    dev.set_mesh(g["xy"], enode, np.arange(enode.shape[0]), g["elem_material"], plane)
    row_ptr, col, emap = native.csr_pattern(n_nodes, enode)
    colors, ptr, elems = native.element_colors(n_nodes, enode)
    dev.set_pattern(row_ptr, col, emap)
    dev.set_colors(ptr, elems)
    return enode, plane, (colors, ptr, elems)


def colored_sum(ke, emap, ptr, elems, nnz):
    values = np.zeros(nnz)
    for c in range(ptr.shape[0] - 1):
        group = elems[ptr[c]:ptr[c + 1]]
        values[emap[group].ravel()] += ke[group].ravel()
    return values


def fan_mesh(sectors: int = 8):
    """`sectors` quads around node 0 (node 0, ring 2i, ring 2i+1, ring 2i+2 of a 2*sectors
    node ring), plus a quad outside the ring: node 0's rows hold 2 * (2 * sectors + 1)
    values."""
    ring = 2 * sectors
    ang = 2.0 * np.pi * np.arange(ring) / ring
    xy = np.vstack([[0.0, 0.0], np.column_stack([np.cos(ang), np.sin(ang)]),
                    [[1.6, -0.2], [1.6, 0.6]]])
    quads = [[0, 1 + 2 * i, 1 + (2 * i + 1) % ring, 1 + (2 * i + 2) % ring] for i in range(sectors)]
    quads.append([1, ring + 1, ring + 2, 2])
    return xy, np.array(quads, dtype=np.int32)


def test_assembly_orders_on_long_rows(dev):
    """Serial and coloured assembly keep the reference's sums where a row holds more values
    than the row kernel's LDS copy (34 at the fan's centre: the store + gather and coloured
    scatter kernels), and where it does not (the same fan with 4 sectors: the row kernel)."""
    for sectors in (8, 4):
        print(f"case fan mesh, {sectors} sectors", flush=True)
        xy, enode = fan_mesh(sectors)
        n_nodes = xy.shape[0]
        young = 1.0 + 0.1 * np.arange(enode.shape[0])
        materials = np.column_stack([young, np.full_like(young, 0.3), np.ones_like(young)])
        dev.set_mesh(xy, enode, np.arange(enode.shape[0]), materials, native.PLANE_STRESS)
        row_ptr, col, emap = native.csr_pattern(n_nodes, enode)
        colors, ptr, elems = native.element_colors(n_nodes, enode)
        dev.set_pattern(row_ptr, col, emap)
        dev.set_colors(ptr, elems)
        assert np.max(np.diff(row_ptr)) == 2 * (2 * sectors + 1)
        ke = dev.element_stiffness()
        nnz = col.shape[0]
        emap2 = np.asarray(emap).reshape(-1, 64)
        dev.assemble("serial")
        assert np.array_equal(bits(dev.stiffness()), bits(lfo.assemble_serial(ke, emap2, nnz))), sectors
        dev.assemble("coloring")
        assert np.array_equal(bits(dev.stiffness()), bits(colored_sum(ke, emap2, ptr, elems, nnz))), sectors


def grid_mesh(nx: int, ny: int, copies: int = 1):
    """nx x ny unit quads numbered row by row, each listed `copies` times in a row."""
    x, y = np.meshgrid(np.arange(nx + 1, dtype=float), np.arange(ny + 1, dtype=float))
    xy = np.column_stack([x.ravel() + 0.05 * np.sin(y.ravel()), y.ravel()])
    quads = []
    for j in range(ny):
        for i in range(nx):
            n0 = j * (nx + 1) + i
            quads += [[n0, n0 + 1, n0 + nx + 2, n0 + nx + 1]] * copies
    return xy, np.array(quads, dtype=np.int32)


@pytest.mark.parametrize("case", ["fan7", "grid_doubled", "grid_tiles"])
def test_block_row_assembly_paths(dev, case):
    """The block-row assembly's general path (a node with 7 elements: more slots than the
    branch-free sum takes) and the per-node-row kernel it falls back to (every element
    listed twice: 64-node blocks with more (element, corner) slots than its LDS holds) keep
    the reference's serial and coloured sums bit for bit; on a 13 x 11 element grid (tiles
    of 8 x 8 elements cut at both edges) the tiled atomic assembly is within 1e-13 of the
    serial sums, as is the element-wise atomic kernel on the doubled grid and the fan."""
    print(f"case {case}", flush=True)
    xy, enode = {"fan7": lambda: fan_mesh(7), "grid_doubled": lambda: grid_mesh(12, 9, copies=2),
                 "grid_tiles": lambda: grid_mesh(13, 11)}[case]()
    n_nodes = xy.shape[0]
    young = 1.0 + 0.1 * np.arange(enode.shape[0])
    materials = np.column_stack([young, np.full_like(young, 0.3), np.ones_like(young)])
    dev.set_mesh(xy, enode, np.arange(enode.shape[0]), materials, native.PLANE_STRAIN)
    row_ptr, col, emap = native.csr_pattern(n_nodes, enode)
    colors, ptr, elems = native.element_colors(n_nodes, enode)
    dev.set_pattern(row_ptr, col, emap)
    dev.set_colors(ptr, elems)
    assert np.max(np.diff(row_ptr)) <= 32
    ke = dev.element_stiffness()
    nnz = col.shape[0]
    emap2 = np.asarray(emap).reshape(-1, 64)
    dev.assemble("serial")
    assert np.array_equal(bits(dev.stiffness()), bits(lfo.assemble_serial(ke, emap2, nnz)))
    dev.assemble("coloring")
    assert np.array_equal(bits(dev.stiffness()), bits(colored_sum(ke, emap2, ptr, elems, nnz)))
    serial = lfo.assemble_serial(ke, emap2, nnz)
    dev.assemble("atomic")
    atomic = dev.stiffness()
    assert np.max(np.abs(atomic - serial)) <= 1e-13 * np.max(np.abs(serial))


@pytest.mark.parametrize("name", FULL_CASES)
def test_stages_match_reference(dev, golden_index, name):
    g = load_golden(name)
    meta = golden_index[name]["meta"]
    enode, plane, (colors, ptr, elems) = upload(dev, g, meta)
    nnz = g["col"].shape[0]
    print(f"case {name}: dofs={g['f'].shape[0]} nnz={nnz}", flush=True)
    # element stiffness
    ke = dev.element_stiffness()
    assert np.array_equal(bits(ke), bits(g["ke"]))
    # colouring (bit-identical to the reference's colour-ordered sums)
    dev.assemble("coloring")
    k_col = dev.stiffness()
    assert np.array_equal(bits(k_col), bits(colored_sum(g["ke"], g["emap"], ptr, elems, nnz)))
    if meta["assembly"] == "coloring":
        assert np.array_equal(bits(k_col), bits(g["K"]))
    # atomic
    dev.assemble("atomic")
    k_serial_ref = lfo.assemble_serial(g["ke"], g["emap"], nnz)
    assert np.max(np.abs(dev.stiffness() - k_serial_ref)) <= 1e-13 * np.max(np.abs(k_serial_ref))
    # serial (gather in element order) last: the solve below uses it
    rep = dev.assemble("serial")
    assert rep["jacobian_failures"] == 0
    K = dev.stiffness()
    assert np.array_equal(bits(K), bits(k_serial_ref))
    if meta["assembly"] == "serial":
        assert np.array_equal(bits(K), bits(g["K"]))
    # Dirichlet elimination (on the reference's matrix when it used another strategy)
    dev.set_stiffness(g["K"])
    dev.set_load(g["f"])
    dev.dirichlet(g["cdofs"], g["cvals"])
    kc, fc = dev.system()
    assert np.array_equal(bits(kc), bits(g["Kc"])) and np.array_equal(bits(fc), bits(g["fc"]))
    # preconditioner applications
    assert np.array_equal(bits(dev.precond_apply("jacobi", g["probe"])), bits(g["jacobi_invdiag"] * g["probe"]))
    assert np.array_equal(bits(dev.precond_apply("block", g["probe"])), bits(g["block1_apply_probe"]))
    assert np.array_equal(bits(dev.precond_apply("ic0", g["probe"])), bits(g["ic0_apply_probe"]))
    # PCG SpMV kernels (row-segment streaming 4, 5 and node-slice 6) sum each row in
    # column order from 0.0, as the reference's CSR loop (sparse.cpp:104-120); the
    # row-segment kernels' FP32 copy is the plain rounding of every value
    y_ref = lfo.spmv(g["row_ptr"], g["col"], g["Kc"], g["probe"])
    y32_ref = lfo.spmv(g["row_ptr"], g["col"], g["Kc"].astype(np.float32).astype(np.float64), g["probe"])
    for variant in (4, 5, 6):
        assert np.array_equal(bits(dev.spmv(g["probe"], "fp64", variant)), bits(y_ref)), variant
    for variant in (4, 5):
        assert np.array_equal(bits(dev.spmv(g["probe"], "mixed", variant)), bits(y32_ref)), variant
    check_stencil_spmv(dev, g["row_ptr"], g["col"], g["Kc"], g["probe"])
    # the block storage product (sparse.cpp:128-198): node-slice and exact node-stencil
    # kernels summing each 2x2 block's products first
    info = dev.layout_info()
    if info["node_blocked"]:
        for variant in (6, 7):
            assert np.array_equal(bits(dev.spmv(g["probe"], "fp64", variant, block_order=True)), bits(g["bsr_probe"]))
    if info["stencil_exact_planes"]:
        assert np.array_equal(bits(dev.spmv(g["probe"], "fp64", 9, block_order=True)), bits(g["bsr_probe"]))
    # the case's own subdomain block-Jacobi (any subdomain size)
    if "dof_subdomain" in g:
        dev.set_subdomains(g["dof_subdomain"][0::2], meta["subdomains"])
        assert np.array_equal(bits(dev.precond_apply("subdomain", g["probe"])), bits(g["block_apply_probe"]))
    # PCG with the case's preconditioner and storage and the reference stop rule
    precond = PRECOND_NAME[case_precond(meta)]
    storage = "bsr" if case_storage(meta) else "csr"
    u, rep = dev.pcg(precond, "fp64", "reference", meta["tolerance"], storage=storage)
    assert rep["converged"] == 1
    assert abs(rep["iterations"] - meta["iterations"]) <= 2 + 0.01 * meta["iterations"]
    assert rep["initial_residual"] == pytest.approx(meta["initial_residual"], rel=1e-14)
    assert np.linalg.norm(u - g["u"]) <= 1e-9 * np.linalg.norm(g["u"])
    # reactions and fields from the device solution, against the oracle on that solution
    R = dev.reactions()
    assert np.array_equal(bits(R), bits(lfo.spmv(g["row_ptr"], g["col"], g["K"], u) - g["f"]))
    centre, gp, nodal = dev.postprocess()
    o_centre, o_gp = lfo.element_fields(g["xy"], enode, g["elem_material"], plane, u)
    assert np.array_equal(bits(centre), bits(o_centre)) and np.array_equal(bits(gp), bits(o_gp))
    assert np.array_equal(bits(nodal), bits(lfo.smooth_nodal(g["xy"].shape[0], enode, o_gp.reshape(-1, 36))))
    # and against the reference's own outputs to solver accuracy
    scale = np.max(np.abs(g["R"]))
    assert np.max(np.abs(R - g["R"])) <= 1e-9 * scale
    assert np.max(np.abs(centre - g["elem_centre"])) <= 1e-9 * np.max(np.abs(g["elem_centre"]))


def check_compensated(y, y_ref, scale, row_abs, x):
    """FP32 off-diagonal values with the compensated FP64 diagonal: FP32 rounding of each
    off-diagonal value (2^-24 relative), once in the product and once in the diagonal's
    compensation, plus FP64 summation."""
    bound = 2.0 ** -23 * (scale + row_abs * np.abs(x)) + 1e-14 * scale
    return bool(np.all(np.abs(y - y_ref) <= bound))


def check_translations(dev, row_ptr, col, kc, variant):
    """The compensated FP32 copy multiplies the two nodal translations as K does."""
    n = row_ptr.shape[0] - 1
    for t in (np.tile([1.0, 0.0], n // 2), np.tile([0.0, 1.0], n // 2)):
        kt = lfo.spmv(row_ptr, col, kc, t)
        t_scale = lfo.spmv(row_ptr, col, np.abs(kc), t)
        assert np.all(np.abs(dev.spmv(t, "mixed", variant) - kt) <= 1e-13 * t_scale), variant


def check_stencil_spmv(dev, row_ptr, col, kc, x):
    """Node-slice and node-stencil multiplies where the pattern admits them. Exact
    stencil form: every offset in column order, bit-identical to the CSR row loop.
    Symmetric form: block (i, i-d) is the transpose of block (i-d, i), equal to the
    stored one up to the rounding of Ke's (a,b)/(b,a) entries, so each row is within
    1e-14 of |K||x|. FP32 copies of both layouts: compensated (check_compensated), and
    the nodal translations act exactly."""
    info = dev.layout_info()
    y_ref = lfo.spmv(row_ptr, col, kc, x)
    scale = lfo.spmv(row_ptr, col, np.abs(kc), np.abs(x))
    row_abs = lfo.spmv(row_ptr, col, np.abs(kc), np.ones_like(x))
    if info["node_blocked"]:
        assert check_compensated(dev.spmv(x, "mixed", 6), y_ref, scale, row_abs, x)
        check_translations(dev, row_ptr, col, kc, 6)
    if info["stencil_exact_planes"]:
        assert np.array_equal(bits(dev.spmv(x, "fp64", 9)), bits(y_ref))
    if info["stencil_sym_planes"]:
        assert np.all(np.abs(dev.spmv(x, "fp64", 8) - y_ref) <= 1e-14 * scale)
        assert check_compensated(dev.spmv(x, "mixed", 8), y_ref, scale, row_abs, x)
        check_translations(dev, row_ptr, col, kc, 8)
    if info["sym_slices"]:  # symmetric node slices: lower blocks read as transposed upper ones
        assert np.all(np.abs(dev.spmv(x, "fp64", 15) - y_ref) <= 1e-14 * scale)
        assert check_compensated(dev.spmv(x, "mixed", 15), y_ref, scale, row_abs, x)
        check_translations(dev, row_ptr, col, kc, 15)
    return info


def _c1(dev, assembly="serial"):
    return pipeline.solve(meshgen.cantilever_50(), assembly=assembly, device=dev)


def test_c1_pipeline_matches_reference(dev, golden_index):
    """50x50 plane-stress cantilever: structures and K bit-exact, u and R to 1e-9."""
    ref = golden_index["cantilever50"]
    print("case 50x50 serial jacobi", flush=True)
    out = _c1(dev)
    s = out["system"]
    assert digest(s.row_ptr.astype(np.int64)) == ref["sha256"]["row_ptr"]
    assert digest(s.emap.astype(np.int64)) == ref["sha256"]["emap"]
    assert digest(s.device.stiffness()) == ref["sha256"]["K"]
    kc, fc = s.device.system()
    assert digest(kc) == ref["sha256"]["Kc"] and digest(fc) == ref["sha256"]["fc"]
    gold = load_golden("cantilever50_u")
    assert abs(out["pcg"]["iterations"] - 313) <= 3
    assert np.linalg.norm(out["u"] - gold["u"]) <= 1e-9 * np.linalg.norm(gold["u"])
    assert np.max(np.abs(out["reaction"] - gold["R"])) <= 1e-9 * np.max(np.abs(gold["R"]))
    assert abs(out["reaction"][1::2].sum() - 1.0) <= 1e-10  # reactions balance the load


@pytest.mark.parametrize("precond,iters", [("jacobi", 1240), ("block", 1240), ("ic0", 359)])
def test_preconditioners_200x200(dev, golden_index, precond, iters):
    """200x200 cantilever, RCM, coloring: K digest exact, PCG vs the oracle."""
    print(f"case 200x200 rcm coloring {precond}", flush=True)
    sys_ = pipeline.setup(meshgen.cantilever_200_rcm(), device=dev, assembly="coloring", reordering="rcm")
    ref = golden_index[f"cantilever200_{'block1' if precond == 'block' else precond}"]
    assert digest(dev.stiffness()) == ref["sha256"]["K"]
    assert sys_.assembly_report["coloring_count"] == 4
    kc, fc = dev.system()
    assert digest(kc) == ref["sha256"]["Kc"]
    u, rep = dev.pcg(precond, "fp64", "reference", 1e-10)
    assert rep["converged"] == 1 and abs(rep["iterations"] - iters) <= 2 + 0.01 * iters
    o_u, o_rep = lfo.pcg(sys_.row_ptr, sys_.col, kc, fc, {"jacobi": 0, "block": 1, "ic0": 2}[precond])
    assert o_rep["iterations"] == iters
    assert np.linalg.norm(u - o_u) <= 1e-9 * np.linalg.norm(o_u)
    if precond == "ic0":
        assert rep["breakdowns"] == 0 and rep["levels"] > 0


def test_cantilever200_subdomain_block_and_block_storage(dev, golden_index):
    """200x200, RCM, coloring with the reference's default block preconditioner (rcb,
    64 nodes per subdomain: 632 dense 2x2-node-blocked Cholesky factors) and with block
    storage: the split, the subdomain applications and the block storage products are
    the reference's bit for bit; the solves take its iteration counts (734, 1240) and
    agree with the Jacobi solve of the same system."""
    print("case 200x200 rcm coloring block rcb64 / storage bsr", flush=True)
    sys_ = pipeline.setup(meshgen.cantilever_200_rcm(), device=dev, assembly="coloring", reordering="rcm")
    ref = golden_index["cantilever200_block64"]
    nodes, count = native.subdomain_split(sys_.indexed.xy, "rcb", 64)
    assert count == 632 and digest(np.repeat(nodes.astype(np.int64), 2)) == ref["sha256"]["dof_subdomain"]
    dev.set_subdomains(nodes, count)
    kc, fc = dev.system()
    x = np.random.default_rng(11).standard_normal(fc.shape[0])
    blk = lfo.BlockJacobi(sys_.row_ptr, sys_.col, kc, np.repeat(nodes, 2), count)
    assert np.array_equal(bits(dev.precond_apply("subdomain", x)), bits(blk.apply(x)))
    assert blk.capture_ratio == ref["meta"]["capture_ratio"]
    u_jac, _ = dev.pcg("jacobi", "fp64", "reference", 1e-10)
    u, rep = dev.pcg("subdomain", "fp64", "reference", 1e-10)
    assert rep["converged"] == 1 and abs(rep["iterations"] - ref["meta"]["iterations"]) <= 2 + 0.01 * 734
    assert rep["degenerate_blocks"] == 0
    assert np.linalg.norm(u - u_jac) <= 1e-9 * np.linalg.norm(u_jac)
    # block storage: products bit for bit, the PCG on the node-slice and default layouts
    y_bsr = lfo.bsr_spmv(sys_.row_ptr, sys_.col, kc, x)
    assert np.array_equal(bits(dev.spmv(x, "fp64", 6, block_order=True)), bits(y_bsr))
    ref_bsr = golden_index["cantilever200_bsr_jacobi"]["meta"]
    for variant in (6, 0):
        ub, rb = dev.pcg("jacobi", "fp64", "reference", 1e-10, storage="bsr", spmv_variant=variant)
        assert rb["converged"] == 1 and abs(rb["iterations"] - ref_bsr["iterations"]) <= 2 + 0.01 * 1240
        assert np.linalg.norm(ub - u_jac) <= 1e-9 * np.linalg.norm(u_jac)
    with pytest.raises(native.LightFemError) as err:
        dev.pcg("jacobi", "fp64", "reference", 1e-10, storage="bsr", spmv_variant=4)
    assert err.value.code == native.ERR_UNSUPPORTED


@pytest.mark.parametrize("precond,switch,variant,drop", [
    ("jacobi", 0.0, 0, 0.0), ("block", 0.0, 0, 0.0), ("jacobi", 0.0, 8, 0.0), ("jacobi", 0.0, 8, 0.1),
    ("jacobi", -1.0, 8, 0.1), ("jacobi", 0.0, 6, 0.1), ("jacobi", -1.0, 6, 0.1), ("jacobi", 0.0, 12, 0.0),
    ("block", -1.0, 13, 0.0)])
def test_mixed_precision_matches_fp64(dev, precond, switch, variant, drop):
    """FP32 values (after an FP64 phase, or from the start), FP64 vectors and dots,
    with or without FP64 residual replacements: the same answer as FP64 and a direct
    solve to 1e-9."""
    import scipy.sparse
    import scipy.sparse.linalg
    n = 64
    print(f"case mixed {n}x{n} log-uniform E plane strain {precond} switch={switch} variant={variant} drop={drop}",
          flush=True)
    m = meshgen.loguniform_cantilever(n)
    sys_ = pipeline.setup(m, device=dev, assembly="coloring")
    kc, fc = dev.system()
    u64, r64 = dev.pcg(precond, "fp64", "relative", 1e-10)
    um, rm = dev.pcg(precond, "mixed", "relative", 1e-10, mixed_switch=switch, spmv_variant=variant,
                     replace_drop=drop)
    assert r64["converged"] and rm["converged"]
    assert (rm["residual_replacements"] - rm["refine_passes"] >= 1) == (drop > 0.0)
    # both forms stop on the recurrence residual; the mixed form then refines on the FP64
    # residual b - K u until it meets the target or stops halving (the FP32 products
    # leave it at the FP32 operator's rounding level, a few 1e-9 here, before refinement)
    assert rm["true_residual"] <= 1.0e-9 * rm["initial_residual"]
    assert r64["true_residual"] <= 1.0e-9 * r64["initial_residual"]
    A = scipy.sparse.csr_matrix((kc, sys_.col, sys_.row_ptr))
    exact = scipy.sparse.linalg.spsolve(A.tocsc(), fc)
    for u in (u64, um):
        assert np.linalg.norm(u - exact) <= 1e-9 * np.linalg.norm(exact)


def test_deterministic_and_graph_equals_direct(dev):
    print("case determinism 50x50", flush=True)
    pipeline.setup(meshgen.cantilever_50(), device=dev)
    u1, r1 = dev.pcg("jacobi")
    u2, r2 = dev.pcg("jacobi")
    u3, r3 = dev.pcg("jacobi", direct_launch=True, check_every=7)
    assert np.array_equal(bits(u1), bits(u2)) and np.array_equal(bits(u1), bits(u3))
    assert r1["iterations"] == r2["iterations"] == r3["iterations"]
    for variant in (1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 13, 14, 15):
        u4, r4 = dev.pcg("jacobi", spmv_variant=variant)
        assert r4["converged"] and np.linalg.norm(u4 - u1) <= 1e-12 * np.linalg.norm(u1), variant
    for variant in (10, 12, 13, 14, 15):
        u10, r10 = dev.pcg("jacobi", spmv_variant=variant, direct_launch=True, check_every=7)
        assert np.array_equal(bits(u10), bits(dev.pcg("jacobi", spmv_variant=variant)[0])), variant


@pytest.mark.parametrize("precond", ["jacobi", "block"])
def test_fused_slice_iterations(dev, precond):
    """The fused node-slice iterations on an RCM-numbered log-E mesh (no fixed node offsets):
    one launch (12), two launches (13) and one launch with the sums folded by the next launch
    (14) per iteration give the same bits (the same sums over the same nodes per block, folded
    in the same order, p' of column nodes formed with the owner's arithmetic), also with FP32
    values after the switch and from the first iteration; all match the three-launch
    node-slice iteration (6) to 1e-9 with its iteration count within 1% + 2, in CSR and in
    block storage order; they honour the iteration cap and refuse the sweep preconditioners."""
    print(f"case fused node slices 48x40 rcm log-E {precond}", flush=True)
    m = meshgen.structured_cantilever(48, 40, young=meshgen.loguniform_young(48 * 40, 7), plane_mode="plane_strain")
    pipeline.setup(m, device=dev, assembly="coloring", reordering="rcm")
    assert dev.layout_info()["stencil_sym_planes"] == 0  # RCM: node slices only
    for storage in ("csr", "bsr"):
        u6, r6 = dev.pcg(precond, "fp64", "relative", 1e-10, spmv_variant=6, storage=storage)
        u12, r12 = dev.pcg(precond, "fp64", "relative", 1e-10, spmv_variant=12, storage=storage)
        u13, r13 = dev.pcg(precond, "fp64", "relative", 1e-10, spmv_variant=13, storage=storage)
        u14, r14 = dev.pcg(precond, "fp64", "relative", 1e-10, spmv_variant=14, storage=storage)
        u0, r0 = dev.pcg(precond, "fp64", "relative", 1e-10, storage=storage)  # the default: 14 at this size
        assert np.array_equal(bits(u12), bits(u13)) and r12["iterations"] == r13["iterations"], storage
        assert np.array_equal(bits(u12), bits(u14)) and r12["iterations"] == r14["iterations"], storage
        assert r12["true_residual"] == r14["true_residual"], storage
        assert np.array_equal(bits(u0), bits(u12)), storage
        assert r12["converged"] and abs(r12["iterations"] - r6["iterations"]) <= 2 + 0.01 * r6["iterations"]
        assert np.linalg.norm(u12 - u6) <= 1e-9 * np.linalg.norm(u6), storage
    for switch in (0.0, -1.0):  # FP32 values after the default switch / from the first iteration
        mixed = [dev.pcg(precond, "mixed", "relative", 1e-10, mixed_switch=switch, spmv_variant=v)
                 for v in (12, 13, 14)]
        for um, rm in mixed[1:]:
            assert np.array_equal(bits(um), bits(mixed[0][0])), switch
            assert rm["iterations"] == mixed[0][1]["iterations"] and rm["refine_passes"] == mixed[0][1]["refine_passes"]
        assert mixed[0][1]["converged"] and np.linalg.norm(mixed[0][0] - u6) <= 1e-9 * np.linalg.norm(u6)
    for variant in (12, 13, 14):
        with pytest.raises(native.LightFemError) as err:
            dev.pcg(precond, "fp64", "relative", 1e-10, max_iterations=5, spmv_variant=variant)
        assert err.value.code == native.ERR_NOT_CONVERGED
        _, capped = dev.pcg(precond, "fp64", "relative", 1e-10, max_iterations=5, spmv_variant=variant,
                            raise_on_error=False)
        assert capped["iterations"] == 5 and not capped["converged"], variant
        with pytest.raises(native.LightFemError) as err:
            dev.pcg("ic0", spmv_variant=variant)
        assert err.value.code == native.ERR_UNSUPPORTED


@pytest.mark.parametrize("precond", ["jacobi", "block"])
def test_symmetric_slice_iteration(dev, precond):
    """Variant 15, the two-launch iteration on the symmetric node slices (each node's
    blocks from its diagonal on hold values, the blocks before it are read as the
    transposes of the column nodes' blocks), on an RCM-numbered mesh of more than 1,024
    slices, where it is the default: the multiply within 1e-14 of |K||x| per row of the CSR
    product (its FP32 copy within the compensated bound, nodal translations exact), the PCG
    against the node-slice iteration (6) to 1e-9 with its iteration count within 1% + 2,
    graph and direct launches bit-identical, the iteration cap honoured; mixed precision
    on it within 1e-9 of FP64; 15 refuses block storage and the sweep preconditioners."""
    print(f"case symmetric node slices 262x260 rcm {precond}", flush=True)
    sys_ = pipeline.setup(meshgen.structured_cantilever(262, 260, young=meshgen.loguniform_young(262 * 260, 11)),
                          device=dev, assembly="coloring", reordering="rcm")
    info = dev.layout_info()
    assert info["stencil_sym_planes"] == 0 and info["n_slices"] > 1024 and info["sym_slices"] == 1
    kc, _ = dev.system()
    x = np.random.default_rng(5).standard_normal(dev.n)
    y_ref = lfo.spmv(sys_.row_ptr, sys_.col, kc, x)
    scale = lfo.spmv(sys_.row_ptr, sys_.col, np.abs(kc), np.abs(x))
    assert np.all(np.abs(dev.spmv(x, "fp64", 15) - y_ref) <= 1e-14 * scale)
    row_abs = lfo.spmv(sys_.row_ptr, sys_.col, np.abs(kc), np.ones_like(x))
    assert check_compensated(dev.spmv(x, "mixed", 15), y_ref, scale, row_abs, x)
    check_translations(dev, sys_.row_ptr, sys_.col, kc, 15)
    u6, r6 = dev.pcg(precond, "fp64", "relative", 1e-10, spmv_variant=6)
    u15, r15 = dev.pcg(precond, "fp64", "relative", 1e-10, spmv_variant=15)
    u0, r0 = dev.pcg(precond, "fp64", "relative", 1e-10)  # the FP64 default at this size
    ud, rd = dev.pcg(precond, "fp64", "relative", 1e-10, spmv_variant=15, direct_launch=True, check_every=7)
    assert r15["converged"] and abs(r15["iterations"] - r6["iterations"]) <= 2 + 0.01 * r6["iterations"]
    assert np.linalg.norm(u15 - u6) <= 1e-9 * np.linalg.norm(u6)
    assert np.array_equal(bits(u0), bits(u15)) and np.array_equal(bits(ud), bits(u15))
    assert r15["true_residual"] <= 1e-9 * r15["initial_residual"]
    um, rm = dev.pcg(precond, "mixed", "relative", 1e-10)  # default: 15 with its FP32 copy
    um15, rm15 = dev.pcg(precond, "mixed", "relative", 1e-10, spmv_variant=15)
    assert np.array_equal(bits(um), bits(um15)) and rm["iterations"] == rm15["iterations"]
    assert rm["converged"] and np.linalg.norm(um - u6) <= 1e-9 * np.linalg.norm(u6)
    um32, rm32 = dev.pcg(precond, "mixed", "relative", 1e-10, spmv_variant=15, mixed_switch=-1.0)
    assert rm32["converged"] and np.linalg.norm(um32 - u6) <= 1e-9 * np.linalg.norm(u6)
    _, capped = dev.pcg(precond, "fp64", "relative", 1e-10, max_iterations=5, spmv_variant=15, raise_on_error=False)
    assert capped["iterations"] == 5 and not capped["converged"]
    for kwargs in ({"storage": "bsr"}, {"preconditioner": "ic0"}):
        args = {"preconditioner": precond, "precision": "fp64", **kwargs}
        with pytest.raises(native.LightFemError) as err:
            dev.pcg(args["preconditioner"], args["precision"], "relative", 1e-10, spmv_variant=15,
                    storage=kwargs.get("storage", "csr"))
        assert err.value.code == native.ERR_UNSUPPORTED, kwargs


def test_stencil_layouts(dev):
    """The node-stencil SpMV (variants 8, 9) on structured meshes numbered row by row:
    offsets found from the pattern, multiplies against the CSR row loop, PCG against
    the node-slice PCG (both preconditioners it fuses), run-to-run identical."""
    print("case stencil 50x50 and 64x64 log-E", flush=True)
    sys_ = pipeline.setup(meshgen.cantilever_50(), device=dev)
    info = dev.layout_info()
    assert info["stencil_offsets"] == [-52, -51, -50, -1, 0, 1, 50, 51, 52]
    assert info["stencil_sym_planes"] == 5 and info["stencil_exact_planes"] == 9
    kc, _ = dev.system()
    x = np.random.default_rng(7).standard_normal(dev.n)
    check_stencil_spmv(dev, sys_.row_ptr, sys_.col, kc, x)
    assert np.array_equal(bits(dev.spmv(x, "fp64", 9)), bits(dev.spmv(x, "fp64", 6)))
    pipeline.setup(meshgen.loguniform_cantilever(64), device=dev, assembly="coloring")
    for precond in ("jacobi", "block"):
        u6, r6 = dev.pcg(precond, "fp64", "relative", 1e-10, spmv_variant=6)
        u9, r9 = dev.pcg(precond, "fp64", "relative", 1e-10, spmv_variant=9)
        assert np.array_equal(bits(u6), bits(u9)) and r6["iterations"] == r9["iterations"], precond
        for variant in (8, 10, 11):  # 10, 11: the whole iteration fused into one launch (1D, 2D tiles)
            u8, r8 = dev.pcg(precond, "fp64", "relative", 1e-10, spmv_variant=variant)
            u8b, r8b = dev.pcg(precond, "fp64", "relative", 1e-10, spmv_variant=variant)
            assert np.array_equal(bits(u8), bits(u8b)) and r8["iterations"] == r8b["iterations"]
            assert abs(r8["iterations"] - r6["iterations"]) <= 2 + 0.01 * r6["iterations"], (precond, variant)
            assert np.linalg.norm(u8 - u6) <= 1e-9 * np.linalg.norm(u6), (precond, variant)
    assert dev.layout_info()["grid_row_len"] == 65
    # the fused iterations honour the iteration cap and report the reference's count
    for variant in (10, 11):
        with pytest.raises(native.LightFemError) as err:
            dev.pcg("jacobi", "fp64", "relative", 1e-10, max_iterations=5, spmv_variant=variant)
        assert err.value.code == native.ERR_NOT_CONVERGED
        _, capped = dev.pcg("jacobi", "fp64", "relative", 1e-10, max_iterations=5, spmv_variant=variant,
                            raise_on_error=False)
        assert capped["iterations"] == 5 and not capped["converged"]
    # 2D tiles need rows of at least 64 nodes: the 50x50 mesh (51 per row) has none
    pipeline.setup(meshgen.cantilever_50(), device=dev)
    assert dev.layout_info()["grid_row_len"] == 0
    with pytest.raises(native.LightFemError) as err:
        dev.pcg("jacobi", spmv_variant=11)
    assert err.value.code == native.ERR_UNSUPPORTED
    # an RCM-numbered mesh has no fixed offsets: the default falls back to node slices
    pipeline.setup(meshgen.cantilever_200_rcm(), device=dev, assembly="coloring", reordering="rcm")
    info = dev.layout_info()
    assert info["node_blocked"] == 1 and info["stencil_sym_planes"] == 0 and info["fused_window_pairs"] == 0
    with pytest.raises(native.LightFemError) as err:
        dev.spmv(np.ones(dev.n), "fp64", 8)
    assert err.value.code == native.ERR_UNSUPPORTED


@pytest.mark.parametrize("nx,ny", [(63, 10), (100, 70), (129, 101), (64, 64)])
def test_grid_tiles_on_uneven_meshes(dev, nx, ny):
    """The 2D-tiled iteration (variant 11) on meshes whose rows end inside a 64-node column
    tile and whose last row tile is partial: the node-slice FP64 PCG's displacements (rows
    summed as the reference) within 1e-9 for both fused preconditioners, in FP64 and with
    FP32 values (mixed, with its FP64 refinement); the iteration count of the 1D-tiled
    fused iteration (variant 10: the same arithmetic, partial sums grouped by other tiles)
    within 1% + 2 in FP64; two runs bit-identical. The stand-alone symmetric multiply (variant
    8, on the same 2D tiles here) against the CSR product and its FP32 copy's bound."""
    print(f"case grid tiles {nx}x{ny}", flush=True)
    m = meshgen.structured_cantilever(nx, ny, lx=float(nx), ly=float(ny),
                                      young=meshgen.loguniform_young(nx * ny, nx + ny), plane_mode="plane_strain")
    sys_ = pipeline.setup(m, device=dev, assembly="coloring")
    assert dev.layout_info()["grid_row_len"] == nx + 1
    kc, _ = dev.system()
    check_stencil_spmv(dev, sys_.row_ptr, sys_.col, kc, np.random.default_rng(nx).standard_normal(dev.n))
    for precond in ("jacobi", "block"):
        u6, r6 = dev.pcg(precond, "fp64", "relative", 1e-10, spmv_variant=6)
        _, r10 = dev.pcg(precond, "fp64", "relative", 1e-10, spmv_variant=10)
        for precision in ("fp64", "mixed"):
            u11, r11 = dev.pcg(precond, precision, "relative", 1e-10, spmv_variant=11)
            u11b, _ = dev.pcg(precond, precision, "relative", 1e-10, spmv_variant=11)
            print(f"  {precond} {precision}: iterations 6: {r6['iterations']}, 10: {r10['iterations']}, "
                  f"11: {r11['iterations']}, u vs 6 {np.linalg.norm(u11 - u6) / np.linalg.norm(u6):.2e}", flush=True)
            assert np.array_equal(bits(u11), bits(u11b)), (precond, precision)
            assert np.linalg.norm(u11 - u6) <= 1e-9 * np.linalg.norm(u6), (precond, precision)
            if precision == "fp64":
                assert abs(r11["iterations"] - r10["iterations"]) <= 2 + 0.01 * r10["iterations"], precond


@pytest.mark.parametrize("variant", [10, 11])
def test_fused_iteration_iterate_at_every_cap(dev, variant):
    """The fused stencil iteration adds x += alpha p for two iterations on every second
    launch and completes an odd last launch after the solve: the iterate returned after
    an odd or even iteration cap, or at convergence, is the reference PCG iterate after
    as many iterations (a missing or doubled alpha p term is off by orders of magnitude
    more than the rounding of the symmetric stencil sum order)."""
    print("case fused iteration caps 64x64 log-E", flush=True)
    sys_ = pipeline.setup(meshgen.loguniform_cantilever(64), device=dev, assembly="coloring")
    kc, fc = dev.system()
    for cap in (1, 2, 3, 6, 7, 50, 51, 100000):
        u, rep = dev.pcg("jacobi", "fp64", "relative", 1e-10, max_iterations=cap, spmv_variant=variant,
                         raise_on_error=False)
        ref, rref = lfo.pcg(sys_.row_ptr, sys_.col, kc, fc, lfo.JACOBI, lfo.STOP_RELATIVE, 1e-10, cap)
        assert rep["converged"] == rref["converged"], cap
        if cap < 100000:
            assert rep["iterations"] == rref["iterations"] == cap
        assert np.linalg.norm(u - ref) <= 1e-9 * np.linalg.norm(ref), cap


def test_fused_direction_is_bit_identical(dev):
    """Variant 7 forms p = M^-1 r + beta p inside the node-slice multiply (two launches
    per iteration); variant 6 runs the same multiply and a separate direction kernel.
    Same values in the same order: identical solutions and iteration counts, with and
    without graphs, for both fused preconditioners and both precisions."""
    print("case fused direction 200x200", flush=True)
    pipeline.setup(meshgen.cantilever_200_rcm(), device=dev, assembly="coloring", reordering="rcm")
    for precond in ("jacobi", "block"):
        u6, r6 = dev.pcg(precond, spmv_variant=6)
        u7, r7 = dev.pcg(precond, spmv_variant=7)
        assert np.array_equal(bits(u6), bits(u7)) and r6["iterations"] == r7["iterations"], precond
    u7, r7 = dev.pcg("jacobi", spmv_variant=7, direct_launch=True, check_every=5)
    assert np.array_equal(bits(u7), bits(dev.pcg("jacobi", spmv_variant=6)[0]))
    m6, q6 = dev.pcg("jacobi", "mixed", "relative", 1e-8, spmv_variant=6)
    m7, q7 = dev.pcg("jacobi", "mixed", "relative", 1e-8, spmv_variant=7)
    assert np.array_equal(bits(m6), bits(m7)) and q6["iterations"] == q7["iterations"]


def test_error_paths(dev):
    print("case error paths", flush=True)
    m = meshgen.structured_cantilever(6, 6)
    pipeline.setup(m, device=dev)
    with pytest.raises(native.LightFemError) as err:
        dev.pcg("jacobi", max_iterations=1)
    assert err.value.code == native.ERR_NOT_CONVERGED
    bad = meshgen.structured_cantilever(3, 3)
    bad.enodes[4] = bad.enodes[4][::-1].copy()  # clockwise element
    with pytest.raises(native.LightFemError) as err:
        pipeline.setup(bad, device=dev)
    assert err.value.code == native.ERR_JACOBIAN
    assert "non-positive Jacobian" in str(err.value)
    neg = meshgen.structured_cantilever(3, 3, e_uniform=-1.0)
    pipeline.setup(neg, device=dev)
    with pytest.raises(native.LightFemError) as err:
        dev.pcg("jacobi")
    assert err.value.code in (native.ERR_NOT_SPD, native.ERR_NONFINITE)
    # a new mesh drops the previous mesh's pattern and colours: assembling before they are
    # set again is a state error, not a sum over stale lists
    sys_ = pipeline.setup(m, device=dev)
    ix = sys_.indexed
    dev.set_mesh(ix.xy, ix.enodes, ix.elem_mat, ix.materials, ix.plane_mode)
    for strategy in ("coloring", "serial"):
        with pytest.raises(native.LightFemError) as err:
            dev.assemble(strategy)
        assert err.value.code == native.ERR_STATE, strategy


def test_full_size_assembly_properties(dev):
    """1000x1000 log-uniform-E cantilever: strategies agree, Dirichlet rows are identity."""
    print("case 1000x1000 assembly strategies", flush=True)
    m = meshgen.loguniform_cantilever(1000)
    sys_ = pipeline.setup(m, device=dev, assembly="serial")
    assert sys_.col.shape[0] == 36024004 and sys_.row_ptr.shape[0] - 1 == 2004002
    k_serial = dev.stiffness()
    dev.assemble("coloring")
    k_col = dev.stiffness()
    dev.assemble("atomic")
    k_at = dev.stiffness()
    scale = np.max(np.abs(k_serial))
    assert np.max(np.abs(k_col - k_serial)) <= 1e-13 * scale
    assert np.max(np.abs(k_at - k_serial)) <= 1e-13 * scale
    # symmetry of the assembled matrix on a sample of rows (Ke is symmetric to rounding:
    # its (a,b) and (b,a) entries are summed in different orders, element.cpp:150-157)
    rows = np.arange(0, 2004002, 9973)
    for r in rows:
        for k in range(sys_.row_ptr[r], sys_.row_ptr[r + 1]):
            c = sys_.col[k]
            kk = sys_.row_ptr[c] + np.searchsorted(sys_.col[sys_.row_ptr[c]:sys_.row_ptr[c + 1]], r)
            assert abs(k_serial[k] - k_serial[kk]) <= 1e-14 * scale
    dev.set_stiffness(k_serial)
    dev.set_load(sys_.indexed.load)
    dev.dirichlet(sys_.indexed.cdofs, sys_.indexed.cvals)
    kc, fc = dev.system()
    for d in sys_.indexed.cdofs[::50]:
        seg = slice(sys_.row_ptr[d], sys_.row_ptr[d + 1])
        assert np.array_equal(kc[seg], (sys_.col[seg] == d).astype(float))
    x = np.random.default_rng(1).standard_normal(2004002)
    y = dev.spmv(x)
    import scipy.sparse
    A = scipy.sparse.csr_matrix((kc, sys_.col, sys_.row_ptr), shape=(2004002, 2004002))
    ref = A @ x
    assert np.max(np.abs(y - ref)) <= 1e-12 * np.max(np.abs(ref))
    # the node-slice, node-stencil (exact form) and row-segment kernels all sum rows in column order
    assert np.array_equal(bits(dev.spmv(x, "fp64", 6)), bits(dev.spmv(x, "fp64", 4)))
    assert np.array_equal(bits(dev.spmv(x, "fp64", 9)), bits(dev.spmv(x, "fp64", 4)))
    info = check_stencil_spmv(dev, sys_.row_ptr, sys_.col, kc, x)
    assert info["stencil_sym_planes"] == 5
