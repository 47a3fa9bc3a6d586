"""Host stages of the product (CSR pattern + element map, RCM, colouring,
profile, numbering) against the golden vectors and the oracle, bit for bit, and
the C ABI export check. CPU only."""
import os
import re

import numpy as np
import pytest

from conftest import FULL_CASES, ROOT, load_golden
from lightfem_amd import meshgen, native, pipeline
from oracle import lfo


def _declared(header: str) -> list:
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lf[gh]_[a-z0-9_]+)\s*\(", text)))


def test_c_abi_exports_every_declared_symbol():
    gpu = native.gpu_lib()
    host = native.host_lib()
    declared_gpu = _declared("lightfem_gpu.h")
    declared_host = _declared("lightfem_host.h")
    assert declared_gpu and declared_host
    for name in declared_gpu:
        assert hasattr(gpu, name), name
    for name in declared_host:
        assert hasattr(host, name), name
    assert sorted(native.GPU_SYMBOLS) == declared_gpu
    assert sorted(native.HOST_SYMBOLS) == declared_host


def test_gpu_context_fails_cleanly_without_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(native.LightFemError):
        native.Device(0)


@pytest.mark.parametrize("name", FULL_CASES)
def test_pattern_and_colors_match_reference(name):
    g = load_golden(name)
    enode = g["edofs"][:, 0::2] // 2
    n_nodes = g["xy"].shape[0]
    row_ptr, col, emap = native.csr_pattern(n_nodes, enode)
    assert np.array_equal(row_ptr, g["row_ptr"]) and np.array_equal(col, g["col"])
    assert np.array_equal(emap, g["emap"])
    colors, ptr, elems = native.element_colors(n_nodes, enode)
    assert np.array_equal(colors, g["colors"])
    assert ptr[-1] == enode.shape[0]
    for c in range(ptr.shape[0] - 1):
        group = elems[ptr[c]:ptr[c + 1]]
        assert np.all(np.diff(group) > 0) and np.all(colors[group] == c)


@pytest.mark.parametrize("name", ["scrambled_rcm", "grid16_rcm_coloring"])
def test_rcm_matches_reference(name):
    g = load_golden(name)
    enode_ids = g["node_ids"][g["edofs"][:, 0::2] // 2]
    order = native.rcm(np.sort(g["node_ids"]), enode_ids)
    assert np.array_equal(order, g["rcm_order"])


def test_index_model_matches_reference_numbering():
    """Dofs by ascending node id (test_input_validation.cpp:55-67), then RCM order."""
    m = meshgen.structured_cantilever(16, 16)
    g = load_golden("grid16_rcm_coloring")
    ix = pipeline.index_model(m, "rcm")
    assert np.array_equal(ix.node_ids, g["node_ids"])
    assert np.array_equal(ix.xy, g["xy"])
    assert np.array_equal(2 * ix.enodes, g["edofs"][:, 0::2])
    assert np.array_equal(ix.load, g["f"])
    assert np.array_equal(ix.cdofs, g["cdofs"]) and np.array_equal(ix.cvals, g["cvals"])
    # scrambled ids: numbering by ascending id
    perm = np.random.default_rng(3).permutation(m.n_nodes)
    m2 = meshgen.structured_cantilever(4, 3)
    m2.node_ids = m2.node_ids[::-1].copy()
    ix2 = pipeline.index_model(m2)
    assert np.array_equal(ix2.node_ids, np.sort(m2.node_ids))
    assert perm.shape[0] == m.n_nodes


@pytest.mark.parametrize("nx,ny", [(1, 1), (7, 3), (40, 25)])
def test_pattern_matches_oracle_and_closed_form(nx, ny):
    m = meshgen.structured_cantilever(nx, ny)
    enode = m.enodes - 1
    row_ptr, col, emap = native.csr_pattern(m.n_nodes, enode)
    edofs = np.stack([2 * enode, 2 * enode + 1], axis=2).reshape(-1, 8)
    o_row, o_col, o_map = lfo.pattern(2 * m.n_nodes, edofs)
    assert np.array_equal(row_ptr, o_row) and np.array_equal(col, o_col) and np.array_equal(emap, o_map)
    if nx == ny:
        assert col.shape[0] == 4 * (3 * nx + 1) ** 2


def test_pattern_unused_node_keeps_diagonal():
    # node 4 (index 4) touches no element: its rows hold only the diagonal (sparse.cpp:18)
    enode = np.array([[0, 1, 2, 3]])
    row_ptr, col, emap = native.csr_pattern(5, enode)
    edofs = np.stack([2 * enode, 2 * enode + 1], axis=2).reshape(-1, 8)
    o_row, o_col, o_map = lfo.pattern(10, edofs)
    assert np.array_equal(row_ptr, o_row) and np.array_equal(col, o_col) and np.array_equal(emap, o_map)


def test_rcm_and_colors_match_oracle_on_irregular_numbering():
    m = meshgen.structured_cantilever(13, 9)
    rng = np.random.default_rng(11)
    new_ids = rng.permutation(m.n_nodes) * 3 + 5  # scrambled, non-contiguous ids
    remap = dict(zip(m.node_ids.tolist(), new_ids.tolist()))
    enode_ids = np.vectorize(remap.get)(m.enodes)
    sorted_ids = np.sort(new_ids)
    assert np.array_equal(native.rcm(sorted_ids, enode_ids), lfo.rcm(sorted_ids, enode_ids))
    colors, _, _ = native.element_colors(m.n_nodes, m.enodes - 1)
    o_colors, _ = lfo.colors(enode_ids)
    assert np.array_equal(colors, o_colors)


def test_profile_readme_200():
    """README.md:125 structure through the product host stages."""
    m = meshgen.structured_cantilever(200, 200)
    ix = pipeline.index_model(m)
    row_ptr, col, _ = native.csr_pattern(ix.n_nodes, ix.enodes)
    assert native.csr_profile(row_ptr, col) == {"nonzeros": 1444804, "bandwidth": 405, "profile": 32522001}
    ix = pipeline.index_model(m, "rcm")
    row_ptr, col, _ = native.csr_pattern(ix.n_nodes, ix.enodes)
    assert native.csr_profile(row_ptr, col) == {"nonzeros": 1444804, "bandwidth": 803, "profile": 42949601}


def test_c2_integer_structures_match_reference_digests(golden_index):
    import hashlib
    dig = lambda a: hashlib.sha256(np.ascontiguousarray(a, dtype="<i8").tobytes()).hexdigest()  # noqa: E731
    ref = golden_index["cantilever200_jacobi"]["sha256"]
    ix = pipeline.index_model(meshgen.cantilever_200_rcm(), "rcm")
    assert dig(ix.node_ids) == ref["node_ids"]
    row_ptr, col, emap = native.csr_pattern(ix.n_nodes, ix.enodes)
    assert dig(row_ptr) == ref["row_ptr"] and dig(col) == ref["col"] and dig(emap) == ref["emap"]
    colors, ptr, _ = native.element_colors(ix.n_nodes, ix.enodes)
    assert dig(colors) == ref["colors"] and ptr.shape[0] - 1 == 4
