"""Pins the oracle materials by closed forms, invariants and finite differences.

VonMises follows edelweissfe/materials/vonmises/vonmises.py:186-254.  Drucker-Prager has
no reference implementation; it is pinned here by the textbook properties of
de Souza Neto, Peric & Owen (2008), section 8.3.3.
"""

import numpy as np
import pytest

from oracle import materials
from oracle.exceptions import CutbackRequest

J2 = [210000.0, 0.3, 250.0, 1000.0, 0.0, 0.0]
J2NL = [210000.0, 0.15, 355.0, 1000.0, 200.0, 1400.0]


def _call(mat, stress, kappa, de):
    s, C, m = mat.compute(np.atleast_2d(stress).astype(float), np.atleast_2d(kappa).astype(float).reshape(-1, 1),
                          np.atleast_2d(de).astype(float))
    return s[0], C[0], m[0, 0]


def _dev_norm(s):
    d = s.copy()
    d[:3] -= s[:3].mean()
    return np.sqrt((d[:3] ** 2).sum() + 2 * (d[3:] ** 2).sum())


def test_linear_elastic_matches_hooke():
    E, nu = 210000.0, 0.3
    mat = materials.LinearElastic([E, nu])
    de = np.array([1e-4, -2e-5, 3e-5, 4e-5, -1e-5, 2e-5])
    s, C, _ = mat.compute(np.zeros((1, 6)), np.zeros((1, 0)), de[None])
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    G = E / (2 * (1 + nu))
    eps = de.copy()
    eps[3:] /= 2
    expect = np.concatenate([lam * eps[:3].sum() + 2 * G * eps[:3], 2 * G * eps[3:]])
    assert np.allclose(s[0], expect, rtol=1e-13)


def test_vonmises_closed_form_linear_hardening():
    mat = materials.VonMises(J2)
    E, nu, fy0, H = 210000.0, 0.3, 250.0, 1000.0
    G = E / (2 * (1 + nu))
    de = np.array([-3e-3, 1e-3, 1e-3, 0.0, 0.0, 0.0])
    s, C, k = _call(mat, np.zeros(6), 0.0, de)
    Ce = materials.elasticity_matrix(E, nu)[0]
    trial = Ce @ de
    ftr = _dev_norm(trial) - np.sqrt(2 / 3) * fy0
    assert ftr > 0
    dk = ftr / (np.sqrt(6) * G + np.sqrt(2 / 3) * H)
    assert abs(k - dk) < 1e-12 * dk
    # consistency: ||s|| = sqrt(2/3) fy(kappa)
    assert abs(_dev_norm(s) - np.sqrt(2 / 3) * (fy0 + H * k)) < 1e-9
    # pressure is elastic
    assert abs(s[:3].mean() - trial[:3].mean()) < 1e-9


@pytest.mark.parametrize("props", [J2, J2NL])
def test_vonmises_consistent_tangent_finite_differences(props):
    mat = materials.VonMises(props)
    s0 = np.array([120.0, -40.0, 30.0, 25.0, -10.0, 15.0])
    k0 = 2e-3
    de = np.array([2.0e-3, -1.2e-3, 0.4e-3, 0.9e-3, -0.3e-3, 0.5e-3])
    s, C, k = _call(mat, s0, k0, de)
    assert k > k0
    h = 1e-8
    Cfd = np.zeros((6, 6))
    for j in range(6):
        e = np.zeros(6)
        e[j] = h
        Cfd[:, j] = (_call(mat, s0, k0, de + e)[0] - _call(mat, s0, k0, de - e)[0]) / (2 * h)
    assert np.abs(C - Cfd).max() < 1e-5 * np.abs(C).max()
    assert np.allclose(C, C.T, atol=1e-8 * np.abs(C).max())


def test_vonmises_zero_increment_shortcut():
    mat = materials.VonMises(J2)
    s0 = np.array([500.0, 0, 0, 0, 0, 0])  # would be far outside the yield surface
    s, C, k = _call(mat, s0, 0.0, np.full(6, 1e-16))
    assert np.array_equal(s, s0) and k == 0.0
    assert np.array_equal(C, materials.elasticity_matrix(210000.0, 0.3)[0])


def test_vonmises_local_newton_failure_requests_cutback():
    # extremely steep saturation hardening needs more than 15 Newton steps
    mat = materials.VonMises([210000.0, 0.3, 250.0, 0.0, 1e5, 1e6])
    with pytest.raises(CutbackRequest) as e:
        _call(mat, np.zeros(6), 0.0, np.array([-0.05, 0.02, 0.02, 0, 0, 0]))
    assert e.value.cutbackSize == 0.5


DP = [30000.0, 0.2, 30.0, 10.0, 10.0, 0.0]


def _dp_yield(props, s, e):
    eta, xi, _ = materials.dp_parameters(props[2], props[3], props[4])
    p = s[:3].mean()
    return _dev_norm(s) / np.sqrt(2) + eta * p - xi * (props[4] + props[5] * e)


@pytest.mark.parametrize("props", [DP, [30000.0, 0.2, 30.0, 10.0, 10.0, 500.0]])
def test_drucker_prager_cone_return_consistency_and_fd_tangent(props):
    mat = materials.DruckerPrager(props)
    s0 = np.array([-20.0, -10.0, -15.0, 4.0, -2.0, 3.0])
    de = np.array([-2e-3, 1.5e-3, 0.5e-3, 1.0e-3, -0.5e-3, 0.8e-3])
    s, C, e = _call(mat, s0, 0.0, de)
    assert e > 0
    assert abs(_dp_yield(props, s, e)) < 1e-9 * np.abs(s).max()
    h = 1e-8
    Cfd = np.zeros((6, 6))
    for j in range(6):
        d = np.zeros(6)
        d[j] = h
        Cfd[:, j] = (_call(mat, s0, 0.0, de + d)[0] - _call(mat, s0, 0.0, de - d)[0]) / (2 * h)
    assert np.abs(C - Cfd).max() < 1e-5 * np.abs(C).max()
    assert not np.allclose(C, C.T, atol=1e-6 * np.abs(C).max())  # non-associative => non-symmetric


def test_drucker_prager_associative_tangent_is_symmetric():
    props = [30000.0, 0.2, 30.0, 30.0, 10.0, 0.0]
    mat = materials.DruckerPrager(props)
    s, C, e = _call(mat, np.array([-20.0, -10.0, -15.0, 4.0, -2.0, 3.0]), 0.0,
                    np.array([-2e-3, 1.5e-3, 0.5e-3, 1.0e-3, -0.5e-3, 0.8e-3]))
    assert e > 0
    assert np.allclose(C, C.T, atol=1e-9 * np.abs(C).max())


def test_drucker_prager_zero_friction_reduces_to_perfect_j2():
    """phi = psi = 0: eta = etabar = 0, xi = 2/sqrt(3); sqrt(J2) = xi c is von Mises with sigma_y = 2c."""
    c = 10.0
    dp = materials.DruckerPrager([30000.0, 0.2, 0.0, 0.0, c, 0.0])
    vm = materials.VonMises([30000.0, 0.2, 2 * c, 0.0, 0.0, 0.0])
    de = np.array([-2e-3, 1.5e-3, 0.5e-3, 1.0e-3, -0.5e-3, 0.8e-3])
    s1, C1, _ = _call(dp, np.zeros(6), 0.0, de)
    s2, C2, _ = _call(vm, np.zeros(6), 0.0, de)
    assert np.allclose(s1, s2, rtol=1e-12, atol=1e-12)
    assert np.allclose(C1, C2, rtol=1e-9, atol=1e-6)


def test_drucker_prager_hydrostatic_tension_returns_to_apex():
    props = DP
    eta, xi, _ = materials.dp_parameters(props[2], props[3], props[4])
    mat = materials.DruckerPrager(props)
    s, C, e = _call(mat, np.zeros(6), 0.0, np.array([1e-2, 1e-2, 1e-2, 0, 0, 0]))
    assert np.allclose(s[3:], 0) and np.allclose(s[:3], s[0])
    assert abs(s[0] - props[4] * xi / eta) < 1e-10 * abs(s[0])
    assert np.allclose(C, 0.0)  # perfect plasticity: zero apex tangent


def test_drucker_prager_apex_without_dilatancy_asks_for_a_cutback():
    """psi = 0: a trial state in hydrostatic tension beyond the apex has no return (the flow is
    purely deviatoric and cannot reduce p), so the material asks for a smaller increment instead
    of dividing by etabar = 0; a smaller tension inside the cone still returns to the cone."""
    from oracle.exceptions import CutbackRequest

    props = [30000.0, 0.2, 30.0, 0.0, 10.0, 0.0]
    mat = materials.DruckerPrager(props)
    with pytest.raises(CutbackRequest) as ei:
        _call(mat, np.zeros(6), 0.0, np.array([1e-2, 1e-2, 1e-2, 0, 0, 0]))
    assert ei.value.cutbackSize == 0.5
    s, C, e = _call(mat, np.zeros(6), 0.0, np.array([-2e-3, 1.5e-3, 0.5e-3, 1.0e-3, -0.5e-3, 0.8e-3]))
    assert np.all(np.isfinite(s)) and np.all(np.isfinite(C)) and e > 0


def test_sections_update_each_element_with_its_own_material():
    """Sections (sectionbase.py:86-128: one material per element set): each QP takes exactly
    the update of its element's material run alone; a state column a material does not use
    stays as it was."""
    rng = np.random.default_rng(5)
    nq, E = 8, 6
    of = np.array([0, 1, 2, 1, 0, 2])
    mats = [materials.VonMises(J2), materials.LinearElastic([30000.0, 0.25]), materials.DruckerPrager(DP)]
    sec = materials.Sections(mats, of, nq)
    n = E * nq
    stress = rng.uniform(-50, 50, (n, 6))
    mstate = np.column_stack([rng.uniform(0, 1e-3, n)])
    dstrain = rng.uniform(-3e-3, 3e-3, (n, 6))
    s, C, m = sec.compute(stress, mstate, dstrain)
    of_qp = np.repeat(of, nq)
    for k, mat in enumerate(mats):
        sel = of_qp == k
        s_k, C_k, m_k = mat.compute(stress[sel], mstate[sel][:, :mat.n_state], dstrain[sel])
        assert np.array_equal(s[sel], s_k) and np.array_equal(C[sel], C_k)
        assert np.array_equal(m[sel][:, :mat.n_state], m_k)
        assert np.array_equal(m[sel][:, mat.n_state:], mstate[sel][:, mat.n_state:])
    assert (m[of_qp == 0][:, 0] > mstate[of_qp == 0][:, 0]).any()  # J2 QPs went plastic
