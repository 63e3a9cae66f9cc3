/*
 * C restatement of the reference's CPU path for hexa8 models (oracle side, TEST
 * INFRASTRUCTURE and CPU baseline only).
 *
 *  - element loop: OpenMP over elements with schedule(dynamic), as
 *    edelweissfe/solvers/nonlinearimplicitstaticparallelmk2.pyx:157-185; per element
 *    the DisplacementElement.computeYourself algorithm
 *    (edelweissfe/elements/displacementelement/element.py:290-346): B per QP at the
 *    exchanged evaluation point (eta_i, xi_i, zeta_i), dEps = B dU, material update,
 *    K += B^T C B detJ w (dense), P -= B^T sigma detJ w, strain += dEps; the block
 *    is written row-major into V as the reference does (element.py:318);
 *  - materials: LinearElasticMaterial (linearelastic.py:185-210) and
 *    VonMisesMaterial.computeStress (vonmises.py:186-254);
 *  - CSR pattern + COO->CSR position map by binary search (csrgenerator.pyx:47-98)
 *    and the serial scatter updateCSR (csrgenerator.pyx:100-115);
 *  - P / F accumulation in element order (nonlinearimplicitstatic.py:843-844);
 *  - CSR SpMV and 3x3 block-Jacobi PCG (OpenMP), the CPU stand-in for the direct
 *    solver on meshes where a direct solve is infeasible.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static const double NAT8[8][3] = {{-1, -1, -1}, {-1, -1, 1}, {1, -1, 1}, {1, -1, -1},
                                  {-1, 1, -1},  {-1, 1, 1},  {1, 1, 1}, {1, 1, -1}};

static void elasticity(double E, double v, double C[36]) {
    double f = E / ((1 + v) * (1 - 2 * v));
    memset(C, 0, 36 * sizeof(double));
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) C[i * 6 + j] = f * (i == j ? (1 - v) : v);
    for (int i = 3; i < 6; ++i) C[i * 6 + i] = f * ((1 - 2 * v) / 2);
}

/* returns 1 on local Newton failure (CutbackRequest) */
static int vonmises(const double* p, double s[6], double* kappa, const double de[6], double C[36]) {
    double E = p[0], v = p[1], fy0 = p[2], H = p[3], dfy = p[4], delta = p[5];
    double G = E / (2 * (1.0 + v));
    elasticity(E, v, C);
    double nde = 0;
    for (int i = 0; i < 6; ++i) nde += de[i] * de[i];
    if (sqrt(nde) < 1e-14) return 0;
    double t[6];
    for (int i = 0; i < 6; ++i) {
        double a = 0;
        for (int j = 0; j < 6; ++j) a += de[j] * C[j * 6 + i];
        t[i] = s[i] + a;
    }
    double m = (t[0] + t[1] + t[2]) / 3.0;
    double dv[6] = {t[0] - m, t[1] - m, t[2] - m, t[3], t[4], t[5]};
    double dn = sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2] + 2 * (dv[3] * dv[3] + dv[4] * dv[4] + dv[5] * dv[5]));
    double k0 = *kappa, s23 = sqrt(2.0 / 3.0), s6 = sqrt(6.0);
    if (!(dn - s23 * (fy0 + H * k0 + dfy * (1 - exp(-delta * k0))) > 0)) {
        memcpy(s, t, sizeof(t));
        return 0;
    }
    double dk = 0;
    int it = 0;
    double R = dn - s6 * G * dk - s23 * (fy0 + H * (k0 + dk) + dfy * (1 - exp(-delta * (k0 + dk))));
    while (fabs(R) > 1e-12) {
        if (it == 15) return 1;
        double dR = -s6 * G - s23 * (H + dfy * delta * exp(-delta * (k0 + dk)));
        dk -= R / dR;
        ++it;
        R = dn - s6 * G * dk - s23 * (fy0 + H * (k0 + dk) + dfy * (1 - exp(-delta * (k0 + dk))));
    }
    double dl = sqrt(1.5) * dk;
    *kappa = k0 + dk;
    double n[6];
    for (int i = 0; i < 6; ++i) n[i] = dv[i] / dn;
    for (int i = 0; i < 6; ++i) s[i] = t[i] - 2 * G * dl * n[i];
    double hk = H + dfy * delta * exp(-delta * *kappa);
    double a = 2 * G * (1.0 / (1.0 + hk / (3 * G)) - 2 * G * dl / dn), b = 4 * G * G * dl / dn;
    double idh[6] = {2.0 / 3, 2.0 / 3, 2.0 / 3, 0.5, 0.5, 0.5};
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) {
            double id = i == j ? idh[i] : (i < 3 && j < 3 ? -1.0 / 3 : 0.0);
            C[i * 6 + j] -= a * n[i] * n[j] + b * id;
        }
    return 0;
}

/*
 * state AoS [e][q][ns], ns = 12 (+1 for von Mises); V [e][24*24] row-major Ke; Pe [e][24].
 * Returns the number of QPs whose local Newton failed.
 */
int oracle_elements_hexa8(int E, const double* coords, const int32_t* conn, const double* props, int per_element,
                          int material, const double* dU, const double* sRef, double* sTmp, double* V, double* Pe) {
    const double s3 = 1.0 / sqrt(3.0);
    double xi[8], eta[8], zeta[8];
    const double t8[4] = {-1, -1, 1, 1}, z8[4] = {-1, 1, 1, -1};
    for (int q = 0; q < 8; ++q) {
        xi[q] = s3 * (q < 4 ? -1.0 : 1.0);
        eta[q] = s3 * t8[q % 4];
        zeta[q] = s3 * z8[q % 4];
    }
    const int ns = material == 1 ? 13 : 12, np = material == 1 ? 6 : 2;
    int failures = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : failures)
    for (int e = 0; e < E; ++e) {
        double X[8][3], u[24];
        for (int a = 0; a < 8; ++a) {
            int nd = conn[(int64_t)e * 8 + a];
            for (int c = 0; c < 3; ++c) {
                X[a][c] = coords[(int64_t)nd * 3 + c];
                u[a * 3 + c] = dU[(int64_t)nd * 3 + c];
            }
        }
        const double* p = per_element ? props + (int64_t)e * np : props;
        double* K = V + (int64_t)e * 576;
        double* P = Pe + (int64_t)e * 24;
        memset(K, 0, 576 * sizeof(double));
        memset(P, 0, 24 * sizeof(double));
        for (int q = 0; q < 8; ++q) {
            /* evaluation point exchanged as the reference's polynomial-fit Jacobian does */
            double ps[3] = {eta[q], xi[q], zeta[q]};
            double dN[8][3], J[9] = {0};
            for (int a = 0; a < 8; ++a) {
                double ax = NAT8[a][0], ay = NAT8[a][1], az = NAT8[a][2];
                dN[a][0] = ax * (1 + ay * ps[1]) * (1 + az * ps[2]) / 8;
                dN[a][1] = ay * (1 + ax * ps[0]) * (1 + az * ps[2]) / 8;
                dN[a][2] = az * (1 + ax * ps[0]) * (1 + ay * ps[1]) / 8;
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c) J[r * 3 + c] += dN[a][r] * X[a][c];
            }
            double det = J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) +
                         J[2] * (J[3] * J[7] - J[4] * J[6]);
            double iJ[9] = {(J[4] * J[8] - J[5] * J[7]) / det, (J[2] * J[7] - J[1] * J[8]) / det,
                            (J[1] * J[5] - J[2] * J[4]) / det, (J[5] * J[6] - J[3] * J[8]) / det,
                            (J[0] * J[8] - J[2] * J[6]) / det, (J[2] * J[3] - J[0] * J[5]) / det,
                            (J[3] * J[7] - J[4] * J[6]) / det, (J[1] * J[6] - J[0] * J[7]) / det,
                            (J[0] * J[4] - J[1] * J[3]) / det};
            double B[6][24];
            memset(B, 0, sizeof(B));
            for (int a = 0; a < 8; ++a) {
                double g[3];
                for (int c = 0; c < 3; ++c) g[c] = iJ[c * 3] * dN[a][0] + iJ[c * 3 + 1] * dN[a][1] + iJ[c * 3 + 2] * dN[a][2];
                B[0][3 * a] = g[0];
                B[1][3 * a + 1] = g[1];
                B[2][3 * a + 2] = g[2];
                B[3][3 * a] = g[1];
                B[3][3 * a + 1] = g[0];
                B[4][3 * a] = g[2];
                B[4][3 * a + 2] = g[0];
                B[5][3 * a + 1] = g[2];
                B[5][3 * a + 2] = g[1];
            }
            double de[6];
            for (int i = 0; i < 6; ++i) {
                double a = 0;
                for (int j = 0; j < 24; ++j) a += B[i][j] * u[j];
                de[i] = a;
            }
            const double* sr = sRef + ((int64_t)e * 8 + q) * ns;
            double* st = sTmp + ((int64_t)e * 8 + q) * ns;
            double s[6], kap = ns > 12 ? sr[12] : 0.0, C[36];
            memcpy(s, sr, sizeof(s));
            if (material == 1) {
                failures += vonmises(p, s, &kap, de, C);
            } else {
                elasticity(p[0], p[1], C);
                for (int i = 0; i < 6; ++i)
                    for (int j = 0; j < 6; ++j) s[i] += C[i * 6 + j] * de[j];
            }
            memcpy(st, s, sizeof(s));
            for (int i = 0; i < 6; ++i) st[6 + i] = sr[6 + i] + de[i];
            if (ns > 12) st[12] = kap;
            /* K += B^T C B detJ w (w = 1), dense as B.T @ C @ B */
            double CB[6][24];
            for (int i = 0; i < 6; ++i)
                for (int j = 0; j < 24; ++j) {
                    double a = 0;
                    for (int k = 0; k < 6; ++k) a += C[i * 6 + k] * B[k][j];
                    CB[i][j] = a;
                }
            for (int r = 0; r < 24; ++r)
                for (int c = 0; c < 24; ++c) {
                    double a = 0;
                    for (int k = 0; k < 6; ++k) a += B[k][r] * CB[k][c];
                    K[r * 24 + c] += a * det;
                }
            for (int r = 0; r < 24; ++r) {
                double a = 0;
                for (int k = 0; k < 6; ++k) a += B[k][r] * s[k];
                P[r] -= a * det;
            }
        }
    }
    return failures;
}

static int cmp_i32(const void* a, const void* b) {
    int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
    return (x > y) - (x < y);
}

/* CSR pattern of the VIJ pairs (rows idcs[f mod n], cols idcs[f div n] per element block):
 * sorted unique columns per row.  indptr[nDof+1] out; returns nnz, indices allocated with malloc. */
int64_t oracle_csr_pattern(int nDof, int E, int n, const int32_t* dofmap, int32_t* indptr, int32_t** indices_out) {
    int64_t* cnt = calloc((size_t)nDof + 1, sizeof(int64_t));
    for (int64_t e = 0; e < E; ++e)
        for (int r = 0; r < n; ++r) cnt[dofmap[e * n + r] + 1] += n;
    for (int i = 0; i < nDof; ++i) cnt[i + 1] += cnt[i];
    int32_t* cols = malloc(sizeof(int32_t) * (size_t)cnt[nDof]);
    int64_t* fill = malloc(sizeof(int64_t) * (size_t)nDof);
    memcpy(fill, cnt, sizeof(int64_t) * nDof);
    for (int64_t e = 0; e < E; ++e)
        for (int r = 0; r < n; ++r) {
            int row = dofmap[e * n + r];
            for (int c = 0; c < n; ++c) cols[fill[row]++] = dofmap[e * n + c];
        }
    int64_t nnz = 0;
    indptr[0] = 0;
    for (int i = 0; i < nDof; ++i) {
        int32_t* seg = cols + cnt[i];
        int64_t len = cnt[i + 1] - cnt[i];
        qsort(seg, (size_t)len, sizeof(int32_t), cmp_i32);
        int64_t u = 0;
        for (int64_t k = 0; k < len; ++k)
            if (k == 0 || seg[k] != seg[k - 1]) cols[nnz + u++] = seg[k];
        nnz += u;
        indptr[i + 1] = (int32_t)nnz;
    }
    *indices_out = realloc(cols, sizeof(int32_t) * (size_t)nnz);
    free(cnt);
    free(fill);
    return nnz;
}

void oracle_free(void* p) { free(p); }

/* x[k] for COO pair k of element block e: binary search as csrgenerator.pyx:86-98 */
void oracle_x_map(int E, int n, const int32_t* dofmap, const int32_t* indptr, const int32_t* indices, int32_t* x) {
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < E; ++e)
        for (int f = 0; f < n * n; ++f) {
            int32_t row = dofmap[e * n + f % n], col = dofmap[e * n + f / n];
            int32_t lo = indptr[row], hi = indptr[row + 1] - 1;
            while (lo < hi) {
                int32_t mid = (lo + hi) / 2;
                if (indices[mid] < col) lo = mid + 1;
                else hi = mid;
            }
            x[e * n * n + f] = lo;
        }
}

/* updateCSR: data[:] = 0; data[x[k]] += V[k] sequentially (csrgenerator.pyx:110-113) */
void oracle_update_csr(int64_t nV, const int32_t* x, const double* V, double* data, int64_t nnz) {
    memset(data, 0, sizeof(double) * (size_t)nnz);
    for (int64_t k = 0; k < nV; ++k) data[x[k]] += V[k];
}

/* P[el] += Pe; F[el] += |Pe| in element order */
void oracle_gather_pf(int E, int n, const int32_t* dofmap, const double* Pe, double* P, double* F, int nDof) {
    memset(P, 0, sizeof(double) * nDof);
    memset(F, 0, sizeof(double) * nDof);
    for (int64_t e = 0; e < E; ++e)
        for (int r = 0; r < n; ++r) {
            P[dofmap[e * n + r]] += Pe[e * n + r];
            F[dofmap[e * n + r]] += fabs(Pe[e * n + r]);
        }
}

static void spmv(int n, const int32_t* ip, const int32_t* ix, const double* a, const double* x, double* y) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        double s = 0;
        for (int32_t k = ip[i]; k < ip[i + 1]; ++k) s += a[k] * x[ix[k]];
        y[i] = s;
    }
}

static double dot(int n, const double* a, const double* b) {
    double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

/* 3x3 block-Jacobi PCG on a symmetric CSR matrix (Dirichlet rows/columns eliminated by the caller).
 * Stops at ||r|| <= max(rtol ||b||, atol), at max_iter, or after budget_s seconds.
 * out[0] = iterations, out[1] = final ||r||, out[2] = seconds. */
void oracle_pcg(int n, const int32_t* ip, const int32_t* ix, const double* a, const double* b, double* x,
                double rtol, double atol, int max_iter, double budget_s, double* out) {
    double t0 = omp_get_wtime();
    int nb = n / 3;
    double* Minv = malloc(sizeof(double) * 9 * nb);
    double *r = malloc(sizeof(double) * n), *z = malloc(sizeof(double) * n), *p = malloc(sizeof(double) * n),
           *q = malloc(sizeof(double) * n);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nb; ++k) {
        double m[9] = {0};
        for (int i = 0; i < 3; ++i)
            for (int32_t j = ip[3 * k + i]; j < ip[3 * k + i + 1]; ++j)
                if (ix[j] >= 3 * k && ix[j] < 3 * k + 3) m[i * 3 + ix[j] - 3 * k] = a[j];
        double d = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) + m[2] * (m[3] * m[7] - m[4] * m[6]);
        double* o = Minv + 9 * k;
        o[0] = (m[4] * m[8] - m[5] * m[7]) / d; o[1] = (m[2] * m[7] - m[1] * m[8]) / d; o[2] = (m[1] * m[5] - m[2] * m[4]) / d;
        o[3] = (m[5] * m[6] - m[3] * m[8]) / d; o[4] = (m[0] * m[8] - m[2] * m[6]) / d; o[5] = (m[2] * m[3] - m[0] * m[5]) / d;
        o[6] = (m[3] * m[7] - m[4] * m[6]) / d; o[7] = (m[1] * m[6] - m[0] * m[7]) / d; o[8] = (m[0] * m[4] - m[1] * m[3]) / d;
    }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        x[i] = 0;
        r[i] = b[i];
    }
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nb; ++k)
        for (int i = 0; i < 3; ++i) {
            z[3 * k + i] = Minv[9 * k + 3 * i] * r[3 * k] + Minv[9 * k + 3 * i + 1] * r[3 * k + 1] + Minv[9 * k + 3 * i + 2] * r[3 * k + 2];
            p[3 * k + i] = z[3 * k + i];
        }
    double bn = sqrt(dot(n, b, b)), tol = fmax(rtol * bn, atol);
    double rz = dot(n, r, z), rr = bn * bn;
    int it = 0;
    while (it < max_iter && sqrt(rr) > tol && omp_get_wtime() - t0 < budget_s) {
        spmv(n, ip, ix, a, p, q);
        double alpha = rz / dot(n, p, q);
        double rzn = 0, rrn = 0;
#pragma omp parallel for schedule(static) reduction(+ : rzn, rrn)
        for (int k = 0; k < nb; ++k) {
            double rv[3];
            for (int i = 0; i < 3; ++i) {
                x[3 * k + i] += alpha * p[3 * k + i];
                rv[i] = r[3 * k + i] - alpha * q[3 * k + i];
                r[3 * k + i] = rv[i];
                rrn += rv[i] * rv[i];
            }
            for (int i = 0; i < 3; ++i) {
                double zi = Minv[9 * k + 3 * i] * rv[0] + Minv[9 * k + 3 * i + 1] * rv[1] + Minv[9 * k + 3 * i + 2] * rv[2];
                z[3 * k + i] = zi;
                rzn += rv[i] * zi;
            }
        }
        double beta = rzn / rz;
        rz = rzn;
        rr = rrn;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
        ++it;
    }
    out[0] = it;
    out[1] = sqrt(rr);
    out[2] = omp_get_wtime() - t0;
    free(Minv); free(r); free(z); free(p); free(q);
}

int oracle_threads(void) { return omp_get_max_threads(); }

/* ---- GMRES(m) + ILU(0) phases on the CPU (timed baseline of the non-symmetric solve) ----
 * Scalar ILU(0) in natural row order, IKJ variant (Saad, Iterative Methods for Sparse
 * Linear Systems, 2nd ed., Algorithm 10.4), on a CSR matrix with sorted column indices
 * and a nonzero diagonal; lu holds a copy of the values and is factorised in place.
 * Rows are factorised in order until budget_s has elapsed; returns the rows done. */
int oracle_ilu0(int n, const int32_t* ip, const int32_t* ix, double* lu, int32_t* diag, double budget_s) {
    double t0 = omp_get_wtime();
    for (int i = 0; i < n; ++i) {
        diag[i] = -1;
        for (int32_t p = ip[i]; p < ip[i + 1]; ++p)
            if (ix[p] == i) diag[i] = p;
    }
    int i = 0;
    for (; i < n; ++i) {
        if ((i & 1023) == 0 && omp_get_wtime() - t0 > budget_s) break;
        for (int32_t kb = ip[i]; kb < ip[i + 1] && ix[kb] < i; ++kb) {
            const int k = ix[kb];
            lu[kb] /= lu[diag[k]];
            int32_t jb = kb + 1, kj = diag[k] + 1;
            while (jb < ip[i + 1] && kj < ip[k + 1]) {
                if (ix[jb] == ix[kj]) {
                    lu[jb] -= lu[kb] * lu[kj];
                    ++jb;
                    ++kj;
                } else if (ix[jb] < ix[kj]) {
                    ++jb;
                } else {
                    ++kj;
                }
            }
        }
    }
    return i;
}

/* z = (LU)^{-1} r: forward substitution with the unit lower factor, backward with the upper. */
static void ilu0_apply(int n, const int32_t* ip, const int32_t* ix, const double* lu, const int32_t* diag,
                       const double* r, double* z) {
    for (int i = 0; i < n; ++i) {
        double s = r[i];
        for (int32_t p = ip[i]; p < diag[i]; ++p) s -= lu[p] * z[ix[p]];
        z[i] = s;
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = z[i];
        for (int32_t p = diag[i] + 1; p < ip[i + 1]; ++p) s -= lu[p] * z[ix[p]];
        z[i] = s / lu[diag[i]];
    }
}

/* Seconds per call of the kernels of one GMRES iteration, averaged over reps calls:
 * out[0] y = A x (OpenMP CSR), out[1] ILU(0) application (serial substitutions),
 * out[2] classical Gram-Schmidt of one vector against j basis vectors (OpenMP dots
 * and one combined update), out[3] the norm and scaling of the new basis vector. */
void oracle_gmres_phase_times(int n, const int32_t* ip, const int32_t* ix, const double* a, const double* lu,
                              const int32_t* diag, int j, int reps, double* out) {
    double* x = (double*)malloc(sizeof(double) * n);
    double* y = (double*)malloc(sizeof(double) * n);
    double* V = (double*)malloc(sizeof(double) * (size_t)n * (j + 1));
    double* h = (double*)malloc(sizeof(double) * (j + 1));
    for (int i = 0; i < n; ++i) x[i] = 1.0 / (1.0 + i % 7);
    for (size_t i = 0; i < (size_t)n * (j + 1); ++i) V[i] = 1.0 / (1.0 + i % 11);
    double t = omp_get_wtime();
    for (int r = 0; r < reps; ++r) spmv(n, ip, ix, a, x, y);
    out[0] = (omp_get_wtime() - t) / reps;
    t = omp_get_wtime();
    for (int r = 0; r < reps; ++r) ilu0_apply(n, ip, ix, lu, diag, y, x);
    out[1] = (omp_get_wtime() - t) / reps;
    t = omp_get_wtime();
    for (int r = 0; r < reps; ++r) {
        for (int q = 0; q < j; ++q) h[q] = dot(n, V + (size_t)q * n, y);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) {
            double s = y[i];
            for (int q = 0; q < j; ++q) s -= h[q] * V[(size_t)q * n + i];
            y[i] = s;
        }
    }
    out[2] = (omp_get_wtime() - t) / reps;
    t = omp_get_wtime();
    for (int r = 0; r < reps; ++r) {
        const double nr = sqrt(dot(n, y, y));
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) V[(size_t)j * n + i] = y[i] / (nr > 0 ? nr : 1.0);
    }
    out[3] = (omp_get_wtime() - t) / reps;
    free(x);
    free(y);
    free(V);
    free(h);
}
