"""`lightfem check` on models full of diagnostics: every ERROR / WARNING line, its
order and the exit code equal the reference CLI's (compiled from its sources in
oracle/_ref; skipped where that is not built)."""
import os
import subprocess

import pytest

from conftest import ROOT

OURS = os.path.join(ROOT, "lightfem_amd", "bin", "lightfem")
REF = os.path.join(ROOT, "oracle", "_ref", "lightfem_ref")

MODELS = {
    "duplicates_and_unknowns": """[settings]
output result.txt
[materials]
1 1 0.3 1
1 2 0.3 1
3 -1 0.6 0
[nodes]
1 0 0
2 1 0
3 1 1
4 0 1
4 0 2
5 1e-15 0
6 5 5
0 9 9
[elements]
1 1 2 3 4 1
1 1 2 3 9 7
2 1 1 3 4 1
3 1 4 3 2 3
[boundary]
1 0 0
1 0 free
12 0 0
[loads]
3 0 -1
3 1 0
99 0 1
1 1 0
""",
    "disconnected_and_rigid": """[materials]
1 1 0.3 1
2 1 0.3 1
[nodes]
1 0 0
2 1 0
3 1 1
4 0 1
5 3 0
6 4 0
7 4 1
8 3 1
9 7 7
[elements]
1 1 2 3 4 1
2 5 6 7 8 1
[boundary]
1 0 free
[loads]
3 0 0
[body_forces]
5 0 1
[edge_loads]
1 5 traction 1 0
7 1 pressure 2
""",
    "settings": """[settings]
tolerance 2
max_iterations 0
deformation_scale -1
solver cg
reordering amd
assembly magic
storage coo
preconditioner ilu
subdomain_split x
subdomain_nodes 0
output out.vtk
[materials]
1 1 0.3 1
[nodes]
1 0 0
2 1 0
3 1 1
4 0 1
[elements]
1 1 2 3 4 1
[boundary]
1 0 0
2 0 0
""",
    "coincident_columns": "[materials]\n1 1 0.3 1\n[nodes]\n" + "".join(
        f"{j * 4 + i + 1} {i} {j}\n" for j in range(4) for i in range(4)) +
    "17 1 1\n18 1.000000000000001 2\n19 1 2\n20 2 3.000000000000001\n21 0 0\n[elements]\n" + "".join(
        f"{j * 3 + i + 1} {j * 4 + i + 1} {j * 4 + i + 2} {j * 4 + i + 6} {j * 4 + i + 5} 1\n"
        for j in range(3) for i in range(3)) + "[boundary]\n1 0 0\n5 0 0\n[loads]\n16 0 -1\n",
}


@pytest.mark.skipif(not os.path.exists(REF), reason="reference CLI not built here")
@pytest.mark.parametrize("name", sorted(MODELS))
def test_check_diagnostics_match_reference(tmp_path, name):
    fem = tmp_path / f"{name}.fem"
    fem.write_text(MODELS[name])
    ours = subprocess.run([OURS, "check", str(fem)], capture_output=True, text=True, timeout=60)
    ref = subprocess.run([REF, "check", str(fem)], capture_output=True, text=True, timeout=60)
    assert ours.returncode == ref.returncode
    assert ours.stdout == ref.stdout
    assert ours.stderr == ref.stderr
