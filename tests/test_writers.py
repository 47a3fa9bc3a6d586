"""The .fem parser and the VTK / CSV writers against the reference's, exactly. The
probe prints a digest of every parsed record (so the parallel fast path of the parser
is compared with the reference's reader), then writes the same
synthetic results (oracle/writer_probe.cpp fills every field of every node and element
with splitmix64 values of wide range, zeros, negative zeros and subnormals) are written
by the reference writer (compiled from its sources in oracle/_ref) and by this
repository's host library, and the files must be identical."""
import os
import subprocess

import pytest

from conftest import ROOT
from lightfem_amd import meshgen

REF_PROBE = os.path.join(ROOT, "oracle", "_ref", "writer_probe")
OUR_PROBE = os.path.join(ROOT, "oracle", "build", "writer_probe")


def fem_cases(golden_index, tmp_path):
    for name in ("one_element", "scrambled", "extended"):
        path = tmp_path / f"{name}.fem"
        path.write_text(golden_index["cli"][name]["fem"])
        yield name, path
    path = tmp_path / "grid.fem"
    meshgen.structured_cantilever(40, 25, young=meshgen.loguniform_young(1000, 3)).to_fem(str(path))
    yield "grid", path


@pytest.mark.skipif(not (os.path.exists(REF_PROBE) and os.path.exists(OUR_PROBE)),
                    reason="writer probes not built (the reference writer builds where its sources are)")
def test_vtk_and_csv_bytes_match_reference(golden_index, tmp_path):
    for name, fem in fem_cases(golden_index, tmp_path):
        for seed in (1, 2):
            outs = {}
            for who, probe in (("ref", REF_PROBE), ("ours", OUR_PROBE)):
                prefix = tmp_path / f"{name}_{seed}_{who}"
                run = subprocess.run([probe, str(fem), str(prefix), str(seed)], capture_output=True, text=True,
                                     timeout=120)
                assert run.returncode == 0, run.stderr
                outs[who] = [run.stdout.encode()] + [open(f"{prefix}{ext}", "rb").read()
                                                     for ext in (".vtk", "_nodes.csv", "_gauss_points.csv")]
            for a, b in zip(outs["ours"], outs["ref"]):
                assert a == b, name


@pytest.mark.skipif(not (os.path.exists(REF_PROBE) and os.path.exists(OUR_PROBE)),
                    reason="writer probes not built (the reference writer builds where its sources are)")
def test_parser_edge_tokens_match_reference(tmp_path):
    """Rows the fast path hands to the line parser (signs, exponents, hex, tiny and
    huge values, comments, blank lines, CR line ends, repeated sections) parse to the
    same records as the reference's reader."""
    text = ("# edge tokens\r\n[settings]\nplane_mode plane_stress\n[materials]\n1 +1.5e+00 0.3 1\n"
            "2 0x1p-2 .25 1.\n[nodes]\n1 0 0 # c\n2 1.0 -0.0\n3 1 1\n4 0 1e0\n\n"
            "[elements]\n1 1 2 3 4 1\n[nodes]\n5 2 0\n6 2 1\n[elements]\n2 +2 5 6 3 2\n"
            "[boundary]\n1 0 0\n4 fixed\n[loads]\n5 0 -1e-3\n6 0 -1E-3\n")
    fem = tmp_path / "edge.fem"
    fem.write_text(text)
    digests = []
    for probe in (REF_PROBE, OUR_PROBE):
        run = subprocess.run([probe, str(fem), str(tmp_path / "edge"), "1"], capture_output=True, text=True,
                             timeout=60)
        assert run.returncode == 0, run.stderr
        digests.append(run.stdout)
    assert digests[0] == digests[1]
