"""The `lightfem` CLI on the gfx950 backend against the reference CLI's outputs
(tests/golden/golden.json "cli"): `check` byte-identical, exit codes, and
`solve` tables equal to solver accuracy with identical layout."""
import io
import os
import re
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


def report_lines(text: str) -> dict:
    """The SOLVED / PROFILE / PARALLEL / ITERATIVE lines of a `solve` run (main.cpp:113-157)
    and the model summary's settings line, as {tag: {key: value text}}; PROFILE keys carry
    their before_ / after_ prefix."""
    out = {}
    for line in text.splitlines():
        tag, _, rest = line.strip().partition(": ")
        if tag == "PROFILE":
            rest = re.sub(r"(before|after)\(([^)]*)\)",
                          lambda m: ", ".join(f"{m.group(1)}_{kv}" for kv in m.group(2).split(", ")), rest)
        elif line.startswith("  plane_mode="):
            tag, rest = "SETTINGS", line.strip()
        elif tag not in ("SOLVED", "PARALLEL", "ITERATIVE", "DEVICE"):
            continue
        out[tag] = dict(kv.split("=", 1) for kv in rest.split(", "))
    return out


def check_report(ours_text: str, ref_text: str) -> None:
    """Our solve report against the reference CLI's: sizes, profile, colours and the
    iterative-solver description exact, iteration count within 1% + 2, the initial
    residual to the printed digits, and both final residuals under the stop threshold."""
    ours, ref = report_lines(ours_text), report_lines(ref_text)
    for tag in ("SOLVED", "PROFILE", "PARALLEL", "ITERATIVE"):
        assert tag in ours and tag in ref, tag
    assert ours["PROFILE"] == ref["PROFILE"]
    for key in ("dofs", "nnz"):
        assert ours["SOLVED"][key] == ref["SOLVED"][key], key
    it_o, it_r = int(ours["SOLVED"]["pcg_iterations"]), int(ref["SOLVED"]["pcg_iterations"])
    assert abs(it_o - it_r) <= 0.01 * it_r + 2, (it_o, it_r)
    r0_o, r0_r = float(ours["SOLVED"]["initial_residual"]), float(ref["SOLVED"]["initial_residual"])
    assert abs(r0_o - r0_r) <= 1e-5 * r0_r
    target = float(ref["SETTINGS"]["tolerance"]) * max(1.0, r0_r)
    assert float(ours["SOLVED"]["final_residual"]) <= target and float(ref["SOLVED"]["final_residual"]) <= target
    for key in ("assembly", "requested_threads", "colors"):
        assert ours["PARALLEL"][key] == ref["PARALLEL"][key], key
    for key in ("storage", "preconditioner", "subdomains", "capture_ratio", "matrix_mb_per_matvec"):
        assert ours["ITERATIVE"][key] == ref["ITERATIVE"][key], key


def vtk_sections(text: str):
    """A legacy VTK file as a list of (keyword line, numbers that follow it)."""
    sections = []
    for line in text.splitlines():
        tokens = line.split()
        if not tokens:
            continue
        try:
            values = [float(t) for t in tokens]
        except ValueError:
            sections.append((line, []))
            continue
        if not sections:
            sections.append(("", []))
        sections[-1][1].extend(values)
    return sections


def check_vtk(ours_text: str, ref_text: str) -> None:
    """Keyword lines identical; every numeric block (points, cells, cell types, point and
    cell data) within 1e-9 of the reference relative to the block's largest magnitude,
    plus 1e-14 of the file's largest (a field that is zero in exact arithmetic, such as
    a shear stress under uniform tension, carries rounding of the displacements)."""
    ours, ref = vtk_sections(ours_text), vtk_sections(ref_text)
    assert [k for k, _ in ours] == [k for k, _ in ref]
    file_scale = max((float(np.abs(np.asarray(b)).max()) for _, b in ref if b), default=0.0)
    for (key, a), (_, b) in zip(ours, ref):
        a, b = np.asarray(a), np.asarray(b)
        assert a.shape == b.shape, key
        if b.size:
            scale = float(np.abs(b).max())
            assert np.all(np.abs(a - b) <= 1e-9 * scale + 1e-14 * file_scale), key


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
    check_vtk((tmp_path / f"{name}.vtk").read_text(), ref["vtk"])
    check_report(out.stdout, ref["solve_stdout"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["cantilever50", "cantilever200_rcm_coloring_jacobi", "cantilever200_rcm_atomic_ic0",
                                  "cantilever200_rcm_coloring_block64", "cantilever200_rcm_block1_bsr"])
def test_solve_report_matches_reference(golden_index, tmp_path, name):
    """The report lines of `solve` on the 50x50 and 200x200 cantilevers (the reference's
    stdout kept in golden.json "cli_reports"; the `.fem` regenerated from the generator)."""
    from lightfem_amd import meshgen
    ref = golden_index["cli_reports"][name]
    models = {"cantilever50": meshgen.cantilever_50,
              "cantilever200_rcm_coloring_jacobi": meshgen.cantilever_200_rcm,
              "cantilever200_rcm_atomic_ic0": lambda: meshgen.cantilever_200_rcm("atomic", "ic0"),
              "cantilever200_rcm_coloring_block64": lambda: meshgen.cantilever_200_rcm("coloring", "block"),
              "cantilever200_rcm_block1_bsr": lambda: meshgen.cantilever_200_rcm("coloring", "block")}
    fem = tmp_path / f"{name}.fem"
    models[name]().to_fem(str(fem))
    if ref["extra_settings"]:
        text = fem.read_text().replace("[settings]\n", "[settings]\n" + "\n".join(ref["extra_settings"]) + "\n", 1)
        fem.write_text(text)
    print(f"cli solve {name}", flush=True)
    out = run(["solve", str(fem), str(tmp_path / f"{name}.vtk")], tmp_path)
    assert out.returncode == 0, out.stderr
    check_report(out.stdout, ref["solve_stdout"])


def test_report_checker_rejects_planted_differences(golden_index):
    """The report and VTK comparisons fail on a changed colour count, a changed profile
    figure, an iteration count outside the band and a moved VTK value."""
    ref = golden_index["cli_reports"]["cantilever200_rcm_coloring_jacobi"]["solve_stdout"]
    check_report(ref, ref)
    for old, new in (("colors=4", "colors=5"), ("profile=42949601", "profile=42949602"),
                     ("pcg_iterations=1240", "pcg_iterations=1300"), ("capture_ratio=0,", "capture_ratio=0.1,")):
        assert old in ref
        with pytest.raises(AssertionError):
            check_report(ref.replace(old, new), ref)
    vtk = golden_index["cli"]["grid6_bsr_block_rcb2"]["vtk"]
    check_vtk(vtk, vtk)
    moved = vtk.replace("0.3333328740201125", "0.3333328840201125", 1)  # 1e-8 off
    assert moved != vtk
    with pytest.raises(AssertionError):
        check_vtk(moved, vtk)
