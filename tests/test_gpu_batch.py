"""GPU batches: many independent problems with several load cases in one batched
Jacobi-PCG, against FP64 direct solves of each system (oracle assembly and Dirichlet
elimination), single-problem GPU solves, and across launch shapes.

Bars: displacements within 1e-9 relative of the direct solve (FP64, relative tolerance
1e-12); mixed precision (FP64 values, then compensated FP32 values) with the FP64
residual within the bound the FP32 values allow; bit-identical results for every
launch shape and run.
"""
import numpy as np
import pytest

from lightfem_amd import meshgen, native, pipeline
from oracle import lfo

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.int64)


@pytest.fixture(scope="module")
def dev():
    d = native.Device(0)
    yield d
    d.close()


def small_batch(n_problems=7, seed=4242, n_min=3, n_max=14, zero_case=True):
    specs = meshgen.batch_problems(n_problems, seed=seed, n_min=n_min, n_max=n_max)
    models = [meshgen.batch_problem_model(sp) for sp in specs]
    loads = [meshgen.batch_case_loads(m, sp) for m, sp in zip(models, specs)]
    if zero_case:
        loads[2][1] = 0.0  # a case with no load converges at once with u = 0
    return specs, models, loads


def direct_solutions(bs, loads):
    out = []
    for p, ix in enumerate(bs.indexed):
        props = ix.materials[ix.elem_mat]
        ds = lfo.DirectSystem(ix.xy, ix.enodes, props, ix.plane_mode, ix.load, ix.cdofs, ix.cvals)
        sols = []
        for c in range(len(loads[p])):
            _, fc = lfo.dirichlet(ds.row_ptr, ds.col, ds.K, ix.load_vector(np.asarray(loads[p][c])), ix.cdofs,
                                  ix.cvals)
            sols.append(ds.solve(fc))
        out.append(np.array(sols))
    return out


def test_batch_matches_direct_solves(dev):
    print("case batch 7 problems x 4 cases, fp64", flush=True)
    specs, models, loads = small_batch()
    bs = pipeline.batch_setup(models, loads, device=dev)
    assert np.all(bs.node_ptr % 64 == 0)
    u, rep, summary = dev.batch_pcg("fp64", "relative", 1e-12)
    assert rep["converged"].all() and summary["systems_converged"] == 28
    assert summary["block_threads"] in (64, 128, 256, 512) and summary["slices_per_wave"] in (1, 2, 4)
    assert summary["sym_slices"] == 1  # FP64 products on the symmetric node slices
    ref = direct_solutions(bs, loads)
    for p, up in enumerate(bs.split(u)):
        for c in range(4):
            scale = max(np.linalg.norm(ref[p][c]), 1e-300)
            assert np.linalg.norm(up[c] - ref[p][c]) <= 1e-9 * scale, (p, c)
    assert np.all(bs.split(u)[2][1] == 0.0) and rep["iterations"][2, 1] == 0
    # padding nodes stay zero
    for p, ix in enumerate(bs.indexed):
        assert np.all(u[:, 2 * (bs.node_ptr[p] + ix.n_nodes):2 * bs.node_ptr[p + 1]] == 0.0)
    # true residuals reported in FP64
    assert np.all(rep["true_residual"] <= 1e-10 * np.maximum(rep["initial_residual"], 1e-300) + 1e-300)


def test_batch_shapes_and_runs_are_bit_identical(dev):
    print("case batch launch shapes", flush=True)
    _, models, loads = small_batch(n_problems=5, seed=99, n_min=20, n_max=40, zero_case=False)
    pipeline.batch_setup(models, loads, device=dev)
    u0, r0, _ = dev.batch_pcg("fp64", "relative", 1e-10, tune=False, block_threads=256, slices_per_wave=1)
    for threads, per_wave in ((64, 1), (128, 2), (512, 4)):
        u1, r1, s1 = dev.batch_pcg("fp64", "relative", 1e-10, tune=False, block_threads=threads,
                                   slices_per_wave=per_wave)
        assert s1["block_threads"] == threads and s1["slices_per_wave"] == per_wave
        assert np.array_equal(bits(u1), bits(u0)) and np.array_equal(r1["iterations"], r0["iterations"])


def test_batch_equals_single_problem_solves(dev):
    """Each problem of a batch takes the single-problem solver's iteration count (the
    dot products fold in a different order, so within 1% + 2) and solution."""
    print("case batch vs single-problem solves", flush=True)
    specs, models, loads = small_batch(n_problems=4, seed=7, n_min=16, n_max=48, zero_case=False)
    bs = pipeline.batch_setup(models, loads, device=dev)
    u, rep, _ = dev.batch_pcg("fp64", "relative", 1e-10)
    parts = bs.split(u)
    for p, (m, sp) in enumerate(zip(models, specs)):
        single = pipeline.setup(m, device=dev, assembly="coloring")
        for c in range(4):
            dev.set_load(single.indexed.load_vector(np.asarray(loads[p][c])))
            dev.dirichlet(single.indexed.cdofs, single.indexed.cvals)
            us, rs = dev.pcg("jacobi", "fp64", "relative", 1e-10)
            it = rep["iterations"][p, c]
            assert abs(it - rs["iterations"]) <= 2 + 0.01 * rs["iterations"], (p, c)
            assert np.linalg.norm(parts[p][c] - us) <= 1e-8 * np.linalg.norm(us), (p, c)


@pytest.mark.parametrize("switch", [0.0, -1.0])
def test_batch_mixed_precision(dev, switch):
    """Mixed batches: FP64 values until a problem's residuals fall to 1e-3 ||b|| (the default; or FP32
    from the start), then FP32 off-diagonal values with the compensated FP64 diagonal
    (K~ = K + dK). Every system converges on its recurrence residual r; the FP64
    residual b - K u = r + dK (u - u_switch) is bounded by ||r|| + ||dK u|| with
    |dK u| <= 2^-23 (|K| |u| + |K|1 |u|) row by row (rounding of each off-diagonal value
    in the product and in the diagonal's compensation)."""
    print(f"case batch mixed precision switch {switch}", flush=True)
    specs, models, loads = small_batch(n_problems=6, seed=31, n_min=8, n_max=40, zero_case=False)
    bs = pipeline.batch_setup(models, loads, device=dev)
    u, rep, summary = dev.batch_pcg("mixed", "relative", 1e-10, mixed_switch=switch)
    ref = direct_solutions(bs, loads)
    err = max(np.linalg.norm(up[c] - ref[p][c]) / np.linalg.norm(ref[p][c])
              for p, up in enumerate(bs.split(u)) for c in range(4))
    true_rel = rep["true_residual"] / rep["initial_residual"]
    print(f"true residual / ||b|| max {true_rel.max():.3e}, u error max {err:.3e}", flush=True)
    assert rep["converged"].all()
    assert summary["problems_switched"] == len(models)
    if switch < 0:
        assert np.all(rep["switch_iteration"] == 0)
    else:
        assert np.all(rep["switch_iteration"] > 0)
    for p, (ix, up) in enumerate(zip(bs.indexed, bs.split(u))):
        props = ix.materials[ix.elem_mat]
        ds = lfo.DirectSystem(ix.xy, ix.enodes, props, ix.plane_mode, ix.load, ix.cdofs, ix.cvals)
        kc, _ = lfo.dirichlet(ds.row_ptr, ds.col, ds.K, np.zeros(2 * ix.n_nodes), ix.cdofs, ix.cvals)
        row_abs = lfo.spmv(ds.row_ptr, ds.col, np.abs(kc), np.ones(2 * ix.n_nodes))
        for c in range(4):
            x = up[c]
            dk = 2.0 ** -23 * (lfo.spmv(ds.row_ptr, ds.col, np.abs(kc), np.abs(x)) + row_abs * np.abs(x))
            bound = rep["final_residual"][p, c] + np.linalg.norm(dk) + 1e-13 * rep["initial_residual"][p, c]
            assert rep["true_residual"][p, c] <= bound, (p, c)
    if switch >= 0:
        assert true_rel.max() <= 1e-8 and err <= 1e-7
    else:
        assert err <= 1e-6
    u64, rep64, _ = dev.batch_pcg("fp64", "relative", 1e-10)
    assert np.all(rep64["switch_iteration"] == -1)
    print(f"mixed switched {summary['problems_switched']} max iterations {summary['max_iterations']} "
          f"(fp64 {int(rep64['iterations'].max())})", flush=True)


def test_batch_input_errors(dev):
    print("case batch input errors", flush=True)
    _, models, loads = small_batch(n_problems=2, seed=5, n_min=10, n_max=12, zero_case=False)
    bs = pipeline.batch_setup(models, loads, device=dev)
    assert bs.node_ptr[1] >= 128
    with pytest.raises(native.LightFemError) as err:  # not whole slices
        dev.batch_set_problems(np.array([0, 32, bs.node_ptr[-1]]))
    assert err.value.code == native.ERR_ARG
    with pytest.raises(native.LightFemError) as err:  # cuts problem 0's elements
        dev.batch_set_problems(np.array([0, 64, bs.node_ptr[-1]]))
    assert err.value.code == native.ERR_ARG and "spans two problems" in str(err.value)
    with pytest.raises(native.LightFemError) as err:  # too many load cases
        dev.batch_set_loads(np.zeros((5, dev.n)), bs.cdofs, bs.cvals)
    assert err.value.code == native.ERR_ARG
