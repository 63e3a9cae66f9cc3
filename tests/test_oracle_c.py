"""The C restatement (CPU baseline) agrees with the numpy oracle."""

import numpy as np
import pytest

from oracle import assembly, clib, element, materials, mesh

J2 = [210000.0, 0.3, 250.0, 1000.0, 0.0, 0.0]


@pytest.mark.parametrize("mat,props", [(0, [210000.0, 0.3]), (1, J2), (1, [210000.0, 0.15, 355.0, 1000.0, 200.0, 1400.0])])
def test_c_element_loop_matches_numpy_oracle(mat, props):
    m = mesh.box_mesh(4, 3, 5, jitter=0.1, seed=3)
    conn = m["conn"]
    E, N = conn.shape[0], m["coords"].shape[0]
    ns = 13 if mat == 1 else 12
    dU = 3e-3 * np.random.RandomState(3).uniform(-1, 1, 3 * N)
    el = element.HexElements("C3D8", m["coords"], conn)
    omat = materials.VonMises(props) if mat == 1 else materials.LinearElastic(props)
    Ke, Pe, st = el.compute(dU[el.dof_map], np.zeros((E, 8, ns)), omat)
    V, Pe2, st2, fails = clib.elements_hexa8(m["coords"], conn, np.array(props), mat, dU, np.zeros((E, 8, ns)))
    assert fails == 0
    assert np.abs(V - Ke.reshape(-1)).max() <= 1e-13 * np.abs(Ke).max()
    assert np.abs(Pe2 - Pe).max() <= 1e-13 * np.abs(Pe).max()
    assert np.abs(st2 - st).max() <= 1e-13 * np.abs(st).max()


def test_c_pattern_xmap_and_scatter_match():
    m = mesh.box_mesh(5, 4, 3, jitter=0.1)
    el = element.HexElements("C3D8", m["coords"], m["conn"])
    D = 3 * m["coords"].shape[0]
    I, J = assembly.vij_pattern(el.dof_map)
    ip, ix = assembly.csr_pattern(I, J, D)
    ip2, ix2 = clib.csr_pattern(D, el.dof_map)
    assert np.array_equal(ip, ip2) and np.array_equal(ix, ix2)
    x = assembly.x_map(I, J, ip, ix)
    assert np.array_equal(x, clib.x_map(el.dof_map, ip2, ix2))
    V = np.random.RandomState(0).randn(I.size)
    assert np.array_equal(assembly.update_csr(V, x, ix.size), clib.update_csr(V, x, ix.size))


def test_c_pcg_solves_spd_system():
    import scipy.sparse.linalg as spla

    m = mesh.box_mesh(4, 4, 4, jitter=0.1)
    el = element.HexElements("C3D8", m["coords"], m["conn"])
    D = 3 * m["coords"].shape[0]
    Ke, _, _ = el.compute(np.zeros((el.conn.shape[0], 24)), np.zeros((el.conn.shape[0], 8, 12)),
                          materials.LinearElastic([1.0, 0.3]))
    I, J = assembly.vij_pattern(el.dof_map)
    ip, ix = assembly.csr_pattern(I, J, D)
    data = assembly.update_csr(Ke.reshape(-1), assembly.x_map(I, J, ip, ix), ix.size)
    K = assembly.csr_matrix(data, ip, ix)
    K = K + 1e-2 * __import__("scipy.sparse", fromlist=["eye"]).eye(D, format="csr")  # remove rigid modes
    K.sort_indices()
    b = np.random.RandomState(1).randn(D)
    x, its, res, _ = clib.pcg(K.indptr, K.indices, K.data, b, 1e-12, 0.0, 10000, 60.0)
    assert its > 0 and res <= 1e-12 * np.linalg.norm(b) * 1.0001
    assert np.allclose(x, spla.spsolve(K, b), rtol=0, atol=1e-8 * np.abs(x).max())
