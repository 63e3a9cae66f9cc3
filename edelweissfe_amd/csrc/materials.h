// Quadrature-point constitutive updates, FP64, Voigt order 11,22,33,12,13,23 with
// engineering shear strains.
//
// Every tangent of these materials has the form
//   C = A 1(x)1 + B Idev + cnn n(x)n + cn1 n(x)1 + c1n 1(x)n
// with 1 = (1,1,1,0,0,0), Idev the Voigt deviatoric projector (diagonal 2/3, 2/3, 2/3,
// 1/2, 1/2, 1/2, normal off-diagonal -1/3) and n a unit deviatoric direction in stress
// components.  When cc is given, the update also returns [A, B, cnn, cn1, c1n, n0..n5]:
// the 11 numbers from which the matrix-free operator rebuilds C.
#pragma once

#include <hip/hip_runtime.h>

namespace efe {

// Isotropic elasticity matrix (linearelastic.py:95-117, vonmises.py:126-149).
__device__ __forceinline__ void elastic_matrix(double E, double v, double C[36]) {
    const double f = E / ((1.0 + v) * (1.0 - 2.0 * v));
    const double d = f * (1.0 - v), o = f * v, s = f * ((1.0 - 2.0 * v) / 2.0);
#pragma unroll
    for (int i = 0; i < 36; ++i) C[i] = 0.0;
    C[0] = d; C[1] = o; C[2] = o;
    C[6] = o; C[7] = d; C[8] = o;
    C[12] = o; C[13] = o; C[14] = d;
    C[21] = s; C[28] = s; C[35] = s;
}

__device__ __forceinline__ void compact_elastic(double E, double v, double* cc) {
    if (!cc) return;
    cc[0] = E / (3.0 * (1.0 - 2.0 * v));  // K
    cc[1] = E / (1.0 + v);                // 2G
#pragma unroll
    for (int i = 2; i < 11; ++i) cc[i] = 0.0;
}

// LinearElasticMaterial.computeStress (linearelastic.py:185-210): C = Ei, stress += Ei dEps.
__device__ __forceinline__ int le_update(const double* p, double s[6], const double de[6], double C[36],
                                         double* cc = nullptr) {
    elastic_matrix(p[0], p[1], C);
    compact_elastic(p[0], p[1], cc);
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < 6; ++j) acc += C[i * 6 + j] * de[j];
        s[i] += acc;
    }
    return 0;
}

// VonMisesMaterial.computeStress (vonmises.py:186-254): J2 radial return with
// fy(k) = fy0 + H k + dfy (1 - exp(-delta k)), scalar Newton on dKappa (absolute
// tolerance 1e-12, at most 15 steps, vonmises.py:55-56,229-236).  Returns 1 when
// the local Newton fails (the reference raises CutbackRequest(.., 0.5)).
__device__ __forceinline__ int vm_update(const double* p, double s[6], double& kappa, const double de[6],
                                         double C[36], double* cc = nullptr) {
    const double E = p[0], v = p[1], fy0 = p[2], H = p[3], dfy = p[4], delta = p[5];
    const double G = E / (2.0 * (1.0 + v));
    elastic_matrix(E, v, C);
    compact_elastic(E, v, cc);
    const double nde = sqrt(de[0] * de[0] + de[1] * de[1] + de[2] * de[2] + de[3] * de[3] + de[4] * de[4] +
                            de[5] * de[5]);
    if (nde < 1e-14) return 0;  // stress untouched, elastic tangent (vonmises.py:211-213)
    double t[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < 6; ++j) acc += de[j] * C[j * 6 + i];
        t[i] = s[i] + acc;
    }
    const double c23 = 2.0 / 3.0, c13 = 1.0 / 3.0;
    double dv[6];
    dv[0] = c23 * t[0] - c13 * t[1] - c13 * t[2];
    dv[1] = -c13 * t[0] + c23 * t[1] - c13 * t[2];
    dv[2] = -c13 * t[0] - c13 * t[1] + c23 * t[2];
    dv[3] = t[3];
    dv[4] = t[4];
    dv[5] = t[5];
    const double devNorm = sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2] +
                                2.0 * (dv[3] * dv[3] + dv[4] * dv[4] + dv[5] * dv[5]));
    const double s23 = sqrt(2.0 / 3.0), s6 = sqrt(6.0);
    const double k0 = kappa;
    const double f = devNorm - s23 * (fy0 + H * k0 + dfy * (1.0 - exp(-delta * k0)));
    if (!(f > 0.0)) {
#pragma unroll
        for (int i = 0; i < 6; ++i) s[i] = t[i];
        return 0;
    }
    double dk = 0.0;
    int counter = 0;
    double R = devNorm - s6 * G * dk - s23 * (fy0 + H * (k0 + dk) + dfy * (1.0 - exp(-delta * (k0 + dk))));
    while (fabs(R) > 1e-12) {
        if (counter == 15) return 1;
        const double dR = -s6 * G - s23 * (H + dfy * delta * exp(-delta * (k0 + dk)));
        dk -= R / dR;
        ++counter;
        R = devNorm - s6 * G * dk - s23 * (fy0 + H * (k0 + dk) + dfy * (1.0 - exp(-delta * (k0 + dk))));
    }
    const double dLambda = sqrt(3.0 / 2.0) * dk;
    kappa = k0 + dk;
    double n[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) n[i] = dv[i] / devNorm;
#pragma unroll
    for (int i = 0; i < 6; ++i) s[i] = t[i] - 2.0 * G * dLambda * n[i];
    const double hk = H + dfy * delta * exp(-delta * kappa);
    const double a = 2.0 * G * (1.0 / (1.0 + hk / (3.0 * G)) - 2.0 * G * dLambda / devNorm);
    const double b = 4.0 * G * G * dLambda / devNorm;
    if (cc) {  // C = K 1(x)1 + (2G - b) Idev - a n(x)n
        cc[1] = 2.0 * G - b;
        cc[2] = -a;
#pragma unroll
        for (int i = 0; i < 6; ++i) cc[5 + i] = n[i];
    }
    const double idh[6] = {c23, c23, c23, 0.5, 0.5, 0.5};
#pragma unroll
    for (int i = 0; i < 6; ++i) {
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            double id = (i == j) ? idh[i] : ((i < 3 && j < 3) ? -c13 : 0.0);
            C[i * 6 + j] = C[i * 6 + j] - a * n[i] * n[j] - b * id;
        }
    }
    return 0;
}

// Non-associative Drucker-Prager with linear cohesion hardening c = c0 + H e.
// No reference implementation exists; this follows de Souza Neto, Peric & Owen,
// "Computational Methods for Plasticity" (2008), section 8.3.3: outer-edge Mohr-
// Coulomb match for eta, xi, etabar; return to the smooth cone, else to the apex;
// consistent tangents (Box 8.9 and the apex tangent).  p = tr(sigma)/3, tension positive.
// props [E, nu, phi_deg, psi_deg, c0, H]; state: equivalent plastic strain e.
__device__ __forceinline__ int dp_update(const double* p, double s[6], double& e, const double de[6],
                                         double C[36], double* cc = nullptr) {
    const double E = p[0], v = p[1], c0 = p[4], H = p[5];
    const double pi = 3.14159265358979323846;
    const double ph = p[2] * pi / 180.0, ps = p[3] * pi / 180.0;
    const double r3 = sqrt(3.0), r2 = sqrt(2.0);
    const double eta = 6.0 * sin(ph) / (r3 * (3.0 - sin(ph)));
    const double xi = 6.0 * cos(ph) / (r3 * (3.0 - sin(ph)));
    const double etab = 6.0 * sin(ps) / (r3 * (3.0 - sin(ps)));
    const double G = E / (2.0 * (1.0 + v)), K = E / (3.0 * (1.0 - 2.0 * v));
    elastic_matrix(E, v, C);
    compact_elastic(E, v, cc);
    double t[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < 6; ++j) acc += C[i * 6 + j] * de[j];
        t[i] = s[i] + acc;
    }
    const double ptr = (t[0] + t[1] + t[2]) / 3.0;
    double sd[6] = {t[0] - ptr, t[1] - ptr, t[2] - ptr, t[3], t[4], t[5]};
    const double snorm = sqrt(sd[0] * sd[0] + sd[1] * sd[1] + sd[2] * sd[2] +
                              2.0 * (sd[3] * sd[3] + sd[4] * sd[4] + sd[5] * sd[5]));
    const double sqJ2 = snorm / r2;
    const double coh = c0 + H * e;
    const double Phi = sqJ2 + eta * ptr - xi * coh;
    if (!(Phi > 0.0)) {
#pragma unroll
        for (int i = 0; i < 6; ++i) s[i] = t[i];
        return 0;
    }
    const double A = 1.0 / (G + K * eta * etab + xi * xi * H);
    const double dgam = Phi * A;
    const double one[6] = {1, 1, 1, 0, 0, 0};
    if (sqJ2 - G * dgam >= 0.0) {
        const double fac = 1.0 - G * dgam / (sqJ2 > 0.0 ? sqJ2 : 1.0);
        const double pn = ptr - K * etab * dgam;
#pragma unroll
        for (int i = 0; i < 6; ++i) s[i] = sd[i] * fac + pn * one[i];
        e = e + xi * dgam;
        const double sn = snorm > 0.0 ? snorm : 1.0;
        double n[6];
#pragma unroll
        for (int i = 0; i < 6; ++i) n[i] = sd[i] / sn;
        const double gsn = r2 * G * dgam / sn;
        const double t1 = 2.0 * G * (1.0 - gsn), t2 = 2.0 * G * (gsn - G * A);
        const double t3 = r2 * G * A * K, t4 = K * (1.0 - K * eta * etab * A);
        const double c23 = 2.0 / 3.0, c13 = 1.0 / 3.0;
        const double idh[6] = {c23, c23, c23, 0.5, 0.5, 0.5};
#pragma unroll
        for (int i = 0; i < 6; ++i) {
#pragma unroll
            for (int j = 0; j < 6; ++j) {
                double id = (i == j) ? idh[i] : ((i < 3 && j < 3) ? -c13 : 0.0);
                C[i * 6 + j] = t1 * id + t2 * n[i] * n[j] - t3 * (eta * n[i] * one[j] + etab * one[i] * n[j]) +
                               t4 * one[i] * one[j];
            }
        }
        if (cc) {
            cc[0] = t4;
            cc[1] = t1;
            cc[2] = t2;
            cc[3] = -t3 * eta;
            cc[4] = -t3 * etab;
#pragma unroll
            for (int i = 0; i < 6; ++i) cc[5 + i] = n[i];
        }
    } else {
        // beyond the apex: with zero dilatancy (etab = 0) the flow is purely deviatoric and no
        // return to the apex exists; report it like a failed local Newton (a cutback request)
        if (!(etab > 0.0)) {
#pragma unroll
            for (int i = 0; i < 6; ++i) s[i] = t[i];
            return 2;
        }
        const double alpha = xi / etab, beta = xi / eta;
        const double dv = (ptr - beta * coh) / (alpha * beta * H + K);
        const double pn = ptr - K * dv;
#pragma unroll
        for (int i = 0; i < 6; ++i) s[i] = pn * one[i];
        e = e + alpha * dv;
        const double ka = K * (1.0 - K / (K + alpha * beta * H));
#pragma unroll
        for (int i = 0; i < 6; ++i)
#pragma unroll
            for (int j = 0; j < 6; ++j) C[i * 6 + j] = ka * one[i] * one[j];
        if (cc) {
            cc[0] = ka;
            cc[1] = 0.0;
        }
    }
    return 0;
}

}  // namespace efe
