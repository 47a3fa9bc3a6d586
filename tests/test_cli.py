"""The `lightfem` CLI on the gfx950 backend against the reference CLI's outputs
(tests/golden/golden.json "cli"): `check` byte-identical, exit codes, and
`solve` tables equal to solver accuracy with identical layout."""
import io
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

CLI = os.path.join(ROOT, "lightfem_amd", "bin", "lightfem")


def run(args, cwd):
    return subprocess.run([CLI, *args], capture_output=True, text=True, cwd=cwd, timeout=300)


CLI_CASES = ["one_element", "scrambled", "extended", "grid6_bsr_block_rcb2", "grid6_block_chunk4", "truss_tension",
             "truss_triangle", "plate_with_bar"]


@pytest.mark.parametrize("name", CLI_CASES)
def test_check_is_identical(golden_index, tmp_path, name):
    ref = golden_index["cli"][name]
    fem = tmp_path / f"{name}.fem"
    fem.write_text(ref["fem"])
    out = run(["check", str(fem)], tmp_path)
    assert out.returncode == ref["check_rc"] == 0
    assert out.stdout.replace(str(fem), "<FEM>") == ref["check_stdout"]


def test_exit_codes(tmp_path):
    assert run(["--version"], tmp_path).stdout == "LightFEM 0.8.0\n"
    assert run([], tmp_path).returncode == 1
    assert run(["bogus", "x.fem"], tmp_path).returncode == 1
    bad = tmp_path / "bad.fem"
    bad.write_text("[nodes]\n1 not_a_number 0\n")
    out = run(["check", str(bad)], tmp_path)
    assert out.returncode == 2 and f"{bad}:2: invalid number for x: not_a_number" in out.stderr
    invalid = tmp_path / "cw.fem"
    invalid.write_text("[materials]\n1 1e6 0.25 1\n[nodes]\n1 0 0\n2 1 0\n3 1 1\n4 0 1\n[elements]\n1 1 4 3 2 1\n"
                       "[boundary]\n1 0 0\n4 0 0\n[loads]\n2 0 -1\n")
    out = run(["check", str(invalid)], tmp_path)
    assert out.returncode == 3 and "E_NONPOSITIVE_JACOBIAN" in out.stdout
    unknown = tmp_path / "unknown.fem"
    unknown.write_text("[settings]\nprecision fp16\n[materials]\n1 1 0.3 1\n")
    out = run(["check", str(unknown)], tmp_path)
    assert out.returncode == 3 and "E_UNSUPPORTED_PRECISION" in out.stdout


def _table(text):
    lines = text.splitlines()
    return lines[0], np.loadtxt(io.StringIO("\n".join(lines[1:])), delimiter=",", ndmin=2)


@pytest.mark.gpu
@pytest.mark.parametrize("name", CLI_CASES)
def test_solve_outputs_match_reference(golden_index, tmp_path, name):
    ref = golden_index["cli"][name]
    fem = tmp_path / f"{name}.fem"
    fem.write_text(ref["fem"])
    print(f"cli solve {name}", flush=True)
    out = run(["solve", str(fem), str(tmp_path / f"{name}.vtk")], tmp_path)
    assert out.returncode == 0, out.stderr
    assert "SOLVED:" in out.stdout and "DEVICE: backend=gfx950" in out.stdout
    for table in ("nodes", "gauss_points"):
        ours = (tmp_path / f"{name}_{table}.csv").read_text()
        head_o, data_o = _table(ours)
        head_r, data_r = _table(ref[f"{table}_csv"])
        assert head_o == head_r and data_o.shape == data_r.shape
        scale = np.maximum(np.abs(data_r).max(axis=0), 1e-30)
        assert np.all(np.abs(data_o - data_r) <= 1e-9 * scale + 1e-14)
    vtk_o = (tmp_path / f"{name}.vtk").read_text().splitlines()
    vtk_r = ref["vtk"].splitlines()
    assert len(vtk_o) == len(vtk_r)
    for a, b in zip(vtk_o, vtk_r):
        if not any(ch.isdigit() for ch in b):
            assert a == b  # section headers and keywords are identical
