"""CPU checks of the batch workload generator and the merged-mesh layout rules."""
import numpy as np

from lightfem_amd import meshgen


def test_c5_generator_ranges_and_determinism():
    specs = meshgen.batch_problems(256)
    again = meshgen.batch_problems(256)
    nx = np.array([s["nx"] for s in specs])
    ny = np.array([s["ny"] for s in specs])
    assert nx.min() >= 32 and nx.max() <= 256 and ny.min() >= 32 and ny.max() <= 256
    assert len(set(nx.tolist())) > 100  # uneven sizes
    nu = np.array([s["nu"] for s in specs])
    assert nu.min() >= 0.2 and nu.max() < 0.45
    modes = [s["plane_mode"] for s in specs]
    assert 80 < modes.count("plane_strain") < 176
    for a, b in zip(specs[:5], again[:5]):
        assert a["nx"] == b["nx"] and np.array_equal(a["young"], b["young"])
    e = specs[0]["young"]
    assert e.shape == (specs[0]["nx"] * specs[0]["ny"],) and e.min() >= 1e-2 and e.max() <= 1.0


def test_c5_load_cases_have_unit_total_force():
    spec = meshgen.batch_problems(3, n_min=5, n_max=9)[1]
    model = meshgen.batch_problem_model(spec)
    cases = meshgen.batch_case_loads(model, spec)
    assert cases.shape == (4, spec["ny"] + 1, 2)
    for c, theta in enumerate(spec["theta"]):
        total = cases[c].sum(axis=0)
        assert np.allclose(total, [np.cos(theta), np.sin(theta)], rtol=0, atol=1e-14)
    assert np.allclose(model.xy[:, 0].max(), spec["nx"]) and np.allclose(model.xy[:, 1].max(), spec["ny"])
