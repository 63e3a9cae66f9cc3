"""Hypoelastic small-strain materials (oracle side), vectorised over QPs.

Voigt order 11,22,33,12,13,23 with engineering shear strains, as in
edelweissfe/materials/vonmises/vonmises.py:35-54.

* LinearElastic   edelweissfe/materials/linearelastic/linearelastic.py:95-117,185-210.
* VonMises        edelweissfe/materials/vonmises/vonmises.py:105-254 (J2, nonlinear
                  isotropic hardening fy(k) = fy0 + H k + dfy (1 - exp(-delta k)),
                  radial return with a scalar Newton on dKappa, absolute tol 1e-12,
                  15 iterations, CutbackRequest(.., 0.5) on failure, elastic
                  tangent and untouched stress when ||dEps|| < 1e-14).
* DruckerPrager   no reference implementation.  Textbook non-associative
                  Drucker-Prager with linear cohesion hardening: de Souza Neto,
                  Peric & Owen, "Computational Methods for Plasticity" (2008),
                  section 8.3.3 (return to the smooth cone, return to the apex,
                  consistent tangents, Boxes 8.8-8.9).  Parity unpinned by the
                  reference; pinned by closed forms and finite differences in
                  tests/test_oracle_materials.py.
"""

import numpy as np

from .exceptions import CutbackRequest

IDEV = np.array([
    [2 / 3, -1 / 3, -1 / 3, 0, 0, 0],
    [-1 / 3, 2 / 3, -1 / 3, 0, 0, 0],
    [-1 / 3, -1 / 3, 2 / 3, 0, 0, 0],
    [0, 0, 0, 1, 0, 0],
    [0, 0, 0, 0, 1, 0],
    [0, 0, 0, 0, 0, 1],
])
IDEV_HALF_SHEAR = np.array([
    [2 / 3, -1 / 3, -1 / 3, 0, 0, 0],
    [-1 / 3, 2 / 3, -1 / 3, 0, 0, 0],
    [-1 / 3, -1 / 3, 2 / 3, 0, 0, 0],
    [0, 0, 0, 0.5, 0, 0],
    [0, 0, 0, 0, 0.5, 0],
    [0, 0, 0, 0, 0, 0.5],
])

VONMISES_MAXITER = 15
VONMISES_TOL = 1e-12


def elasticity_matrix(E, v):
    """(M,6,6) for arrays E, v of shape (M,)."""
    E = np.atleast_1d(np.asarray(E, dtype=float))
    v = np.atleast_1d(np.asarray(v, dtype=float))
    f = E / ((1 + v) * (1 - 2 * v))
    C = np.zeros((E.size, 6, 6))
    for i in range(3):
        for j in range(3):
            C[:, i, j] = f * (v if i != j else (1 - v))
    for i in range(3, 6):
        C[:, i, i] = f * ((1 - 2 * v) / 2)
    return C


def _per_qp(props, m, n):
    p = np.asarray(props, dtype=float)
    if p.ndim == 1:
        p = np.broadcast_to(p, (m, n))
    return p


class LinearElastic:
    n_state = 0

    def __init__(self, props):
        self.props = np.asarray(props, dtype=float)

    def compute(self, stress, mstate, dstrain):
        m = stress.shape[0]
        p = _per_qp(self.props, m, 2)
        Ei = elasticity_matrix(p[:, 0], p[:, 1])
        if Ei.shape[0] == 1:
            Ei = np.broadcast_to(Ei, (m, 6, 6))
        stress = stress + np.einsum("mij,mj->mi", Ei, dstrain)
        return stress, np.array(Ei), mstate


class VonMises:
    """props: [E, nu, fy0, HLin, dfy, delta], one row or one row per QP."""

    n_state = 1

    def __init__(self, props):
        self.props = np.asarray(props, dtype=float)

    def compute(self, stress, mstate, dstrain):
        m = stress.shape[0]
        p = _per_qp(self.props, m, 6)
        E, v, fy0, H, dfy, delta = (p[:, i] for i in range(6))
        G = E / (2 * (1.0 + v))
        Ei = elasticity_matrix(E, v)
        if Ei.shape[0] == 1:
            Ei = np.broadcast_to(Ei, (m, 6, 6))
        kappa = mstate[:, 0].copy()

        def fy(k):
            return fy0 + H * k + dfy * (1.0 - np.exp(-delta * k))

        def dfy_dk(k):
            return H + dfy * delta * np.exp(-delta * k)

        zero = np.linalg.norm(dstrain, axis=1) < 10**-14
        trial = stress + np.einsum("mj,mji->mi", dstrain, Ei)
        dev = trial @ IDEV.T
        devNorm = np.sqrt(np.sum(np.square(dev[:, 0:3]), axis=1) + 2 * np.sum(np.square(dev[:, 3:6]), axis=1))
        f = devNorm - np.sqrt(2 / 3) * fy(kappa)
        plastic = (f > 0.0) & ~zero

        new_stress = np.where(zero[:, None], stress, trial)
        C = np.array(Ei)
        if plastic.any():
            dk = np.zeros(m)
            counter = np.zeros(m, dtype=int)

            def R(dk_):
                return devNorm - np.sqrt(6) * G * dk_ - np.sqrt(2 / 3) * fy(kappa + dk_)

            r = R(dk)
            active = plastic & (np.abs(r) > VONMISES_TOL)
            while active.any():
                if (counter[active] == VONMISES_MAXITER).any():
                    raise CutbackRequest("Von Mises Newton failed.", 0.5)
                dR = -np.sqrt(6) * G - np.sqrt(2 / 3) * dfy_dk(kappa + dk)
                dk = np.where(active, dk - r / dR, dk)
                counter = counter + active
                r = R(dk)
                active = plastic & (np.abs(r) > VONMISES_TOL)
            with np.errstate(invalid="ignore", divide="ignore"):
                n = dev / devNorm[:, None]
            dLambda = np.sqrt(3 / 2) * dk
            kap_new = kappa + dk
            s_pl = trial - 2.0 * G[:, None] * dLambda[:, None] * n
            a = 2.0 * G * (1.0 / (1.0 + dfy_dk(kap_new) / (3.0 * G)) - 2.0 * G * dLambda / devNorm)
            b = 4.0 * G**2 * dLambda / devNorm
            C_pl = Ei - a[:, None, None] * np.einsum("mi,mj->mij", n, n) - b[:, None, None] * IDEV_HALF_SHEAR
            new_stress = np.where(plastic[:, None], s_pl, new_stress)
            C = np.where(plastic[:, None, None], C_pl, C)
            kappa = np.where(plastic, kap_new, kappa)
        mstate = mstate.copy()
        mstate[:, 0] = kappa
        return new_stress, C, mstate


def dp_parameters(phi_deg, psi_deg, c):
    """Outer-edge Mohr-Coulomb match (de Souza Neto et al. 2008, eq. 8.89):
    eta = 6 sin(phi) / (sqrt(3) (3 - sin(phi))), xi = 6 cos(phi) / (sqrt(3) (3 - sin(phi))),
    etabar as eta with the dilation angle."""
    ph, ps = np.radians(phi_deg), np.radians(psi_deg)
    eta = 6 * np.sin(ph) / (np.sqrt(3) * (3 - np.sin(ph)))
    xi = 6 * np.cos(ph) / (np.sqrt(3) * (3 - np.sin(ph)))
    etab = 6 * np.sin(ps) / (np.sqrt(3) * (3 - np.sin(ps)))
    return eta, xi, etab


class DruckerPrager:
    """Non-associative Drucker-Prager, linear cohesion hardening c(e) = c0 + H e.

    props: [E, nu, phi_deg, psi_deg, c0, H].  Yield Phi = sqrt(J2(s)) + eta p - xi c,
    potential sqrt(J2(s)) + etabar p, p = tr(sigma)/3 (tension positive).
    State: accumulated equivalent plastic strain e.
    """

    n_state = 1

    def __init__(self, props):
        self.props = np.asarray(props, dtype=float)

    def compute(self, stress, mstate, dstrain):
        m = stress.shape[0]
        p6 = _per_qp(self.props, m, 6)
        E, v, phi, psi, c0, H = (p6[:, i] for i in range(6))
        G = E / (2 * (1.0 + v))
        K = E / (3 * (1.0 - 2 * v))
        eta, xi, etab = dp_parameters(phi, psi, c0)
        Ei = elasticity_matrix(E, v)
        if Ei.shape[0] == 1:
            Ei = np.broadcast_to(Ei, (m, 6, 6))
        e = mstate[:, 0].copy()
        trial = stress + np.einsum("mij,mj->mi", Ei, dstrain)
        ptr = trial[:, 0:3].sum(axis=1) / 3.0
        s = trial.copy()
        s[:, 0:3] -= ptr[:, None]
        snorm = np.sqrt(np.sum(s[:, 0:3] ** 2, axis=1) + 2 * np.sum(s[:, 3:6] ** 2, axis=1))
        sqJ2 = snorm / np.sqrt(2.0)
        coh = c0 + H * e
        Phi = sqJ2 + eta * ptr - xi * coh
        plastic = Phi > 0.0

        new_stress = trial.copy()
        C = np.array(Ei)
        one = np.array([1, 1, 1, 0, 0, 0], dtype=float)
        Id = IDEV_HALF_SHEAR
        if plastic.any():
            A = 1.0 / (G + K * eta * etab + xi**2 * H)
            dgam = Phi * A
            cone = plastic & (sqJ2 - G * dgam >= 0.0)
            apex = plastic & ~cone
            if (apex & ~(etab > 0.0)).any():
                # a trial state beyond the apex of a non-dilatant cone (psi = 0): the plastic
                # flow is purely deviatoric and cannot bring p back to the apex, so no return
                # exists; ask for a smaller increment, as the von Mises local Newton does
                raise CutbackRequest("Drucker-Prager: trial stress beyond the apex with zero dilatancy, no return.",
                                     0.5)
            with np.errstate(invalid="ignore", divide="ignore"):
                n = s / snorm[:, None]  # unit deviatoric direction (tensor norm)
            # smooth cone: s = (1 - G dgam / sqrt(J2)) s_tr, p = p_tr - K etabar dgam, e += xi dgam
            fac = np.where(cone, 1.0 - G * dgam / np.where(sqJ2 > 0, sqJ2, 1.0), 1.0)
            p_c = ptr - K * etab * dgam
            s_c = s * fac[:, None]
            sig_c = s_c + p_c[:, None] * one[None, :]
            # consistent tangent (Box 8.9, smooth portion); the book's ||eps_d trial|| = ||s_tr|| / 2G
            r2 = np.sqrt(2.0)
            gsn = r2 * G * dgam / np.where(snorm > 0, snorm, 1.0)
            t1 = 2 * G * (1.0 - gsn)
            t2 = 2 * G * (gsn - G * A)
            t3 = r2 * G * A * K
            t4 = K * (1.0 - K * eta * etab * A)
            nn = np.einsum("mi,mj->mij", n, n)
            nI = np.einsum("mi,j->mij", n, one)
            In = np.einsum("i,mj->mij", one, n)
            II = np.einsum("i,j->ij", one, one)
            D_c = (t1[:, None, None] * Id + t2[:, None, None] * nn
                   - t3[:, None, None] * (eta[:, None, None] * nI + etab[:, None, None] * In)
                   + t4[:, None, None] * II)
            # apex: solve xi/etab-scaled volumetric return (linear hardening => closed form);
            # evaluated for every point and selected below, so a point on the cone with
            # etab = 0 (no apex return exists) must not warn here
            with np.errstate(invalid="ignore", divide="ignore"):
                alpha = xi / etab
                beta = xi / eta
                # r(dev) = beta c(e + alpha dev) - p_tr + K dev = 0
                dev_v = (ptr - beta * coh) / (alpha * beta * H + K)
                p_a = ptr - K * dev_v
                sig_a = p_a[:, None] * one[None, :]
                D_a = (K * (1.0 - K / (K + alpha * beta * H)))[:, None, None] * II
            new_stress = np.where(cone[:, None], sig_c, new_stress)
            new_stress = np.where(apex[:, None], sig_a, new_stress)
            C = np.where(cone[:, None, None], D_c, C)
            C = np.where(apex[:, None, None], D_a, C)
            e = np.where(cone, e + xi * dgam, e)
            e = np.where(apex, e + alpha * dev_v, e)
        mstate = mstate.copy()
        mstate[:, 0] = e
        return new_stress, C, mstate


class Sections:
    """Several materials in one model, as sections assign a material to the elements of their
    element sets (sections/base/sectionbase.py:86-128, assignSectionPropertiesToElement):
    materials[k] is the material of the elements with material_of_element[e] == k.  compute()
    runs every material on the QPs of its elements, in the flattened (element, QP) order of
    HexElements.compute; the state width is the largest of the materials' (the extra columns of
    a material with fewer state variables stay as they are)."""

    def __init__(self, materials, material_of_element, n_qp):
        self.materials = list(materials)
        self.of_qp = np.repeat(np.asarray(material_of_element, dtype=np.int64), n_qp)
        self.n_state = max(m.n_state for m in self.materials)

    def compute(self, stress, mstate, dstrain):
        if stress.shape[0] != self.of_qp.size:
            raise ValueError("Sections.compute covers the whole element set")
        out = stress.copy()
        C = np.empty((stress.shape[0], 6, 6))
        ms = np.array(mstate, dtype=float, copy=True)
        for k, mat in enumerate(self.materials):
            sel = self.of_qp == k
            if not sel.any():
                continue
            s_k, C_k, m_k = mat.compute(stress[sel], mstate[sel][:, :mat.n_state], dstrain[sel])
            out[sel] = s_k
            C[sel] = C_k
            ms[np.ix_(sel, np.arange(mat.n_state))] = m_k
        return out, C, ms


MATERIALS = {"linearelastic": LinearElastic, "vonmises": VonMises, "druckerprager": DruckerPrager}
