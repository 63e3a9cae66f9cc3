// Krylov solve of a scalar CSR system, for the linear-solver plugin contract
// f(A: csr_matrix, b) -> x of config/linsolve.py:46-86 (called by NIST.linearSolve,
// nonlinearimplicitstatic.py:727-751).  The matrix is the one NIST hands to its linear solver:
// K after applyDirichletK (nonlinearimplicitstatic.py:559-593), whose constrained rows are unit
// rows (zeroed, diagonal 1, eliminate_zeros) while their columns stay.  Rows whose only entry is
// a unit diagonal are taken as prescribed: x = b there, and their columns move to the right-hand
// side, so the Krylov method works on the free block K_ff (symmetric when K is):
//   K_ff x_f = b_f - K_fd b_d.
// Methods: Jacobi-preconditioned CG (symmetric K), BiCGStab (van der Vorst 1992, right
// preconditioned) and restarted GMRES (right preconditioned) for non-symmetric K.
//
// Data: CSR as given (int32 indptr/indices, FP64 values).  SpMV: 16 lanes per row, each lane
// striding the row by 16, then a fixed-order shuffle reduction; rows of a wavefront are
// consecutive, so its value and index loads are contiguous runs.  Dot products: per-workgroup
// partials, re-reduced in a fixed order by the kernel that needs the scalar, so results are
// bitwise reproducible.  The scalar recursion (alpha, beta, omega, the residual norm) stays on
// the device; the host reads a convergence flag every check_every iterations.
#include <cmath>
#include <vector>

#include "efe_internal.h"

namespace {

constexpr int kBS = 256;            // threads per workgroup
constexpr int kLanesPerRow = 16;    // SpMV lanes per row
constexpr int kMaxWG = kCsrPartials;  // partials per dot product (vector kernels stride)

__device__ __forceinline__ double wg_sum(double v, double* sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = kBS / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// sum of nwg partials in a fixed order, by every workgroup that needs it
__device__ __forceinline__ double sum_partials(const double* __restrict__ part, int nwg, double* sh) {
    double v = 0.0;
    for (int i = threadIdx.x; i < nwg; i += kBS) v += part[i];
    return wg_sum(v, sh);
}

// Prescribed rows: a single stored entry, on the diagonal, equal to 1.
__global__ void csr_classify_kernel(int n, const int32_t* __restrict__ ip, const int32_t* __restrict__ ix,
                                    const double* __restrict__ v, uint8_t* __restrict__ fixed,
                                    double* __restrict__ dinv) {
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i >= n) return;
    const int a = ip[i], b = ip[i + 1];
    const bool f = (b - a == 1) && ix[a] == i && v[a] == 1.0;
    fixed[i] = f ? 1 : 0;
    double d = 0.0;
    for (int k = a; k < b; ++k)
        if (ix[k] == i) d += v[k];
    dinv[i] = (f || d == 0.0) ? 1.0 : 1.0 / d;
}

// y = A_ff x (free rows, free columns); 0 on prescribed rows.  With bx (the prescribed
// values): y = b - A_fd bx on free rows instead, the right-hand side of the free block.
__global__ __launch_bounds__(kBS) void csr_spmv_kernel(int n, const int32_t* __restrict__ ip,
                                                       const int32_t* __restrict__ ix,
                                                       const double* __restrict__ v,
                                                       const uint8_t* __restrict__ fixed,
                                                       const double* __restrict__ x, double* __restrict__ y,
                                                       const double* __restrict__ b) {
    const int g = (blockIdx.x * kBS + threadIdx.x) / kLanesPerRow;
    const int l = threadIdx.x & (kLanesPerRow - 1);
    const int row = g;
    double acc = 0.0;
    const bool live = row < n && !fixed[row];
    if (live) {
        const int a = ip[row], e = ip[row + 1];
        for (int k = a + l; k < e; k += kLanesPerRow) {
            const int c = ix[k];
            if (b) {  // right-hand side: the prescribed columns only
                if (fixed[c]) acc += v[k] * x[c];
            } else if (!fixed[c]) {
                acc += v[k] * x[c];
            }
        }
    }
#pragma unroll
    for (int o = kLanesPerRow / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, kLanesPerRow);
    if (row < n && l == 0) {
        if (b) y[row] = live ? b[row] - acc : 0.0;
        else y[row] = live ? acc : 0.0;
    }
}

// part[wg] = sum over the workgroup's rows of u.w (grid-stride, fixed assignment)
__global__ __launch_bounds__(kBS) void csr_dot_kernel(int n, const double* __restrict__ u,
                                                      const double* __restrict__ w, double* __restrict__ part) {
    __shared__ double sh[kBS];
    double s = 0.0;
    for (int i = blockIdx.x * kBS + threadIdx.x; i < n; i += gridDim.x * kBS) s += u[i] * w[i];
    s = wg_sum(s, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// CG scalars, sc: [0] r.z, [1] alpha, [2] r.r, [3] tol^2, [4] done flag (as double), [5] iterations
// z = Dinv r; partials of r.z (slot 0) and r.r (slot 1)
__global__ __launch_bounds__(kBS) void cg_precond_kernel(int n, const double* __restrict__ dinv,
                                                         const double* __restrict__ r, double* __restrict__ z,
                                                         double* __restrict__ part) {
    __shared__ double sh[kBS];
    double rz = 0.0, rr = 0.0;
    for (int i = blockIdx.x * kBS + threadIdx.x; i < n; i += gridDim.x * kBS) {
        const double zi = dinv[i] * r[i];
        z[i] = zi;
        rz += r[i] * zi;
        rr += r[i] * r[i];
    }
    rz = wg_sum(rz, sh);
    rr = wg_sum(rr, sh);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = rz;
        part[kMaxWG + blockIdx.x] = rr;
    }
}

// p = z + beta p with beta = rz_new / rz_old (rz_old in sc[0]); first: p = z
__global__ __launch_bounds__(kBS) void cg_direction_kernel(int n, const double* __restrict__ z, double* __restrict__ p,
                                                           const double* __restrict__ part, int nwg,
                                                           const double* __restrict__ sc, int first) {
    __shared__ double sh[kBS];
    if (sc[4] != 0.0) return;
    const double rz = sum_partials(part, nwg, sh);
    const double beta = first ? 0.0 : rz / sc[0];
    // first: p = z, whatever the workspace held before
    for (int i = blockIdx.x * kBS + threadIdx.x; i < n; i += gridDim.x * kBS) p[i] = first ? z[i] : z[i] + beta * p[i];
}

// one workgroup: rz_old <- rz (after the direction), rr and the stop test; count: the call
// ends an iteration (sc[5] counts the iterations run)
__global__ void cg_scalars_kernel(const double* __restrict__ part, int nwg, double* __restrict__ sc, int count) {
    __shared__ double sh[kBS];
    if (sc[4] != 0.0) return;
    const double rz = sum_partials(part, nwg, sh);
    const double rr = sum_partials(part + kMaxWG, nwg, sh);
    if (threadIdx.x == 0) {
        sc[0] = rz;
        sc[2] = rr;
        if (count) sc[5] += 1.0;
        if (!(rr > sc[3])) sc[4] = 1.0;  // converged (or NaN)
    }
}

// x += alpha p, r -= alpha q with alpha = rz / (p.q)
__global__ __launch_bounds__(kBS) void cg_update_kernel(int n, double* __restrict__ x, double* __restrict__ r,
                                                        const double* __restrict__ p, const double* __restrict__ q,
                                                        const double* __restrict__ part, int nwg,
                                                        const double* __restrict__ sc) {
    __shared__ double sh[kBS];
    if (sc[4] != 0.0) return;
    const double pq = sum_partials(part, nwg, sh);
    const double alpha = sc[0] / pq;
    for (int i = blockIdx.x * kBS + threadIdx.x; i < n; i += gridDim.x * kBS) {
        x[i] += alpha * p[i];
        r[i] -= alpha * q[i];
    }
}

// x on prescribed rows = b
__global__ void csr_finish_kernel(int n, const uint8_t* __restrict__ fixed, const double* __restrict__ b,
                                  double* __restrict__ x) {
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i < n && fixed[i]) x[i] = b[i];
}

__global__ void zero_kernel(int n, double* __restrict__ x) {
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i < n) x[i] = 0.0;
}

__global__ void copy_kernel(int n, const double* __restrict__ a, double* __restrict__ b) {
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i < n) b[i] = a[i];
}

__global__ void set_scalars_kernel(double* __restrict__ sc, double tol2) {
    if (threadIdx.x == 0) {
        for (int k = 0; k < 16; ++k) sc[k] = 0.0;
        sc[3] = tol2;
    }
}

// BiCGStab scalars sc: [0] rho, [1] alpha, [2] omega, [3] tol^2, [4] done, [5] iterations,
// [6] r.r, [7] rho_new
// t = (r0hat . v) partials in part, alpha = rho / (r0hat.v); s = r - alpha v (into r), x += alpha phat
__global__ __launch_bounds__(kBS) void bicg_alpha_kernel(int n, double* __restrict__ x, double* __restrict__ r,
                                                         const double* __restrict__ phat,
                                                         const double* __restrict__ v,
                                                         const double* __restrict__ part, int nwg,
                                                         double* __restrict__ sc) {
    __shared__ double sh[kBS];
    if (sc[4] != 0.0) return;
    const double rv = sum_partials(part, nwg, sh);
    const double alpha = sc[0] / rv;
    if (blockIdx.x == 0 && threadIdx.x == 0) sc[1] = alpha;
    for (int i = blockIdx.x * kBS + threadIdx.x; i < n; i += gridDim.x * kBS) {
        x[i] += alpha * phat[i];
        r[i] -= alpha * v[i];
    }
}

// partials of t.s (slot 0) and t.t (slot 1), s in r
__global__ __launch_bounds__(kBS) void bicg_ts_kernel(int n, const double* __restrict__ t, const double* __restrict__ s,
                                                      double* __restrict__ part, const double* __restrict__ sc) {
    __shared__ double sh[kBS];
    if (sc[4] != 0.0) return;
    double ts = 0.0, tt = 0.0;
    for (int i = blockIdx.x * kBS + threadIdx.x; i < n; i += gridDim.x * kBS) {
        ts += t[i] * s[i];
        tt += t[i] * t[i];
    }
    ts = wg_sum(ts, sh);
    tt = wg_sum(tt, sh);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = ts;
        part[kMaxWG + blockIdx.x] = tt;
    }
}

// omega = t.s / t.t; x += omega shat; r = s - omega t; partials of r0hat.r (slot 0) and r.r (slot 1)
__global__ __launch_bounds__(kBS) void bicg_omega_kernel(int n, double* __restrict__ x, double* __restrict__ r,
                                                         const double* __restrict__ shat,
                                                         const double* __restrict__ t,
                                                         const double* __restrict__ r0,
                                                         const double* __restrict__ part_in, int nwg,
                                                         double* __restrict__ part_out, double* __restrict__ sc) {
    __shared__ double sh[kBS];
    if (sc[4] != 0.0) return;
    const double ts = sum_partials(part_in, nwg, sh);
    const double tt = sum_partials(part_in + kMaxWG, nwg, sh);
    const double omega = tt > 0.0 ? ts / tt : 0.0;
    if (blockIdx.x == 0 && threadIdx.x == 0) sc[2] = omega;
    double rr0 = 0.0, rr = 0.0;
    for (int i = blockIdx.x * kBS + threadIdx.x; i < n; i += gridDim.x * kBS) {
        x[i] += omega * shat[i];
        const double ri = r[i] - omega * t[i];
        r[i] = ri;
        rr0 += r0[i] * ri;
        rr += ri * ri;
    }
    rr0 = wg_sum(rr0, sh);
    rr = wg_sum(rr, sh);
    if (threadIdx.x == 0) {
        part_out[blockIdx.x] = rr0;
        part_out[kMaxWG + blockIdx.x] = rr;
    }
}

// one workgroup: rho_new, r.r, stop test, beta into sc[7] = rho_new, sc[1] beta-ready
__global__ void bicg_scalars_kernel(const double* __restrict__ part, int nwg, double* __restrict__ sc) {
    __shared__ double sh[kBS];
    if (sc[4] != 0.0) return;
    const double rho = sum_partials(part, nwg, sh);
    const double rr = sum_partials(part + kMaxWG, nwg, sh);
    if (threadIdx.x == 0) {
        sc[6] = rr;
        sc[5] += 1.0;
        if (!(rr > sc[3])) {
            sc[4] = 1.0;
        } else {
            sc[8] = (rho / sc[0]) * (sc[1] / sc[2]);  // beta
            sc[0] = rho;
        }
    }
}

// p = r + beta (p - omega v); first: p = r
__global__ __launch_bounds__(kBS) void bicg_p_kernel(int n, const double* __restrict__ r, double* __restrict__ p,
                                                     const double* __restrict__ v, const double* __restrict__ sc,
                                                     int first) {
    if (sc[4] != 0.0) return;
    const double beta = first ? 0.0 : sc[8], omega = sc[2];
    for (int i = blockIdx.x * kBS + threadIdx.x; i < n; i += gridDim.x * kBS)
        p[i] = first ? r[i] : r[i] + beta * (p[i] - omega * v[i]);
}

__global__ void jacobi_plain_kernel(int n, const double* __restrict__ dinv, const double* __restrict__ x,
                                    double* __restrict__ y) {
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i < n) y[i] = dinv[i] * x[i];
}

inline void jacobi_free_kernel(int n, const double* dinv, const double* x, double* y, hipStream_t st) {
    jacobi_plain_kernel<<<(n + kBS - 1) / kBS, kBS, 0, st>>>(n, dinv, x, y);
}

// elementwise y = dinv x
__global__ void jacobi_kernel(int n, const double* __restrict__ dinv, const double* __restrict__ x,
                              double* __restrict__ y, const double* __restrict__ sc) {
    if (sc[4] != 0.0) return;
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i < n) y[i] = dinv[i] * x[i];
}

// first rho = r0.r and rr, stop test
__global__ void bicg_init_scalars_kernel(const double* __restrict__ part, int nwg, double* __restrict__ sc) {
    __shared__ double sh[kBS];
    const double rho = sum_partials(part, nwg, sh);
    const double rr = sum_partials(part + kMaxWG, nwg, sh);
    if (threadIdx.x == 0) {
        sc[0] = rho;
        sc[6] = rr;
        if (!(rr > sc[3])) sc[4] = 1.0;
    }
}

inline int grid_of(int n) { return (n + kBS - 1) / kBS; }

// flag <- 1 when indptr is not non-decreasing from 0 to nnz = indptr[n] or an index is outside
// [0, n); a row's indices are read only when its range lies inside [0, nnz)
__global__ void csr_check_kernel(int n, int nnz, const int32_t* __restrict__ ip, const int32_t* __restrict__ ix,
                                 int* __restrict__ flag) {
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i >= n) return;
    const int a = ip[i], b = ip[i + 1];
    if ((i == 0 && a != 0) || a < 0 || b < a || b > nnz) {
        *flag = 1;
        return;
    }
    for (int k = a; k < b; ++k)
        if (ix[k] < 0 || ix[k] >= n) {
            *flag = 1;
            return;
        }
}

}  // namespace

// Validates a device CSR pattern before any kernel indexes through it; *bad = 1 when invalid.
// flag: one device int of scratch.
int efe_csr_check_pattern(int32_t n, const int32_t* ip, const int32_t* ix, int* flag, hipStream_t st, int* bad) {
    int h = 0;
    EFE_HIP(hipMemcpyAsync(flag, &h, sizeof(int), hipMemcpyHostToDevice, st));
    // nnz = indptr[n] first: every row's range is checked against [0, nnz) before it is walked
    int32_t last = 0;
    EFE_HIP(hipMemcpyAsync(&last, ip + n, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    EFE_HIP(hipStreamSynchronize(st));
    if (last < 0) {
        *bad = 1;
        return EFE_OK;
    }
    csr_check_kernel<<<grid_of(n), kBS, 0, st>>>(n, last, ip, ix, flag);
    EFE_HIP(hipGetLastError());
    EFE_HIP(hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, st));
    EFE_HIP(hipStreamSynchronize(st));
    *bad = h;
    return EFE_OK;
}

size_t efe_csr_workspace_doubles(int32_t n, int32_t solver, int32_t restart) {
    // CG: x r z p q (5 n); BiCGStab: x r r0 p phat v shat t (8 n); GMRES(m): V (m+1) n, w z x rhs
    // (4 n); plus Dinv (n), fixed flags (n bytes, rounded up), partials (4 kMaxWG, GMRES
    // (m+2) kMaxWG), 16 scalars and, for GMRES, m+2 coefficients
    if (solver == EFE_SOLVER_GMRES) {
        const size_t m = restart > 0 ? restart : 50;
        return (m + 1 + 4 + 1) * (size_t)n + ((size_t)n + 7) / 8 + (m + 2) * kMaxWG + 16 + (m + 2);
    }
    const size_t vec = (solver == EFE_SOLVER_CG ? 5 : 8) + 1;
    return vec * (size_t)n + ((size_t)n + 7) / 8 + 4 * kMaxWG + 16;
}

namespace {

// part[i][wg] = partial of V_i . w over the workgroup's rows, i < cnt (grid-stride)
__global__ __launch_bounds__(kBS) void csr_multidot_kernel(int n, const double* __restrict__ V, int cnt,
                                                           const double* __restrict__ w, double* __restrict__ part) {
    __shared__ double sh[kBS];
    for (int i = 0; i < cnt; ++i) {
        const double* vi = V + (size_t)i * n;
        double s = 0.0;
        for (int r = blockIdx.x * kBS + threadIdx.x; r < n; r += gridDim.x * kBS) s += vi[r] * w[r];
        s = wg_sum(s, sh);
        if (threadIdx.x == 0) part[(size_t)i * kMaxWG + blockIdx.x] = s;
    }
}

// w -= sum_i c_i V_i, i < cnt
__global__ void csr_sub_combination_kernel(int n, const double* __restrict__ V, int cnt, const double* __restrict__ c,
                                           double* __restrict__ w) {
    const int r = blockIdx.x * kBS + threadIdx.x;
    if (r >= n) return;
    double acc = w[r];
    for (int i = 0; i < cnt; ++i) acc -= c[i] * V[(size_t)i * n + r];
    w[r] = acc;
}

// x += Dinv (sum_i c_i V_i), i < cnt
__global__ void csr_add_precond_combination_kernel(int n, const double* __restrict__ V, int cnt,
                                                   const double* __restrict__ c, const double* __restrict__ dinv,
                                                   double* __restrict__ x) {
    const int r = blockIdx.x * kBS + threadIdx.x;
    if (r >= n) return;
    double acc = 0.0;
    for (int i = 0; i < cnt; ++i) acc += c[i] * V[(size_t)i * n + r];
    x[r] += dinv[r] * acc;
}

// y = a x
__global__ void csr_scale_kernel(int n, double a, const double* __restrict__ x, double* __restrict__ y) {
    const int r = blockIdx.x * kBS + threadIdx.x;
    if (r < n) y[r] = a * x[r];
}

// y = a - b
__global__ void csr_sub_kernel(int n, const double* __restrict__ a, const double* __restrict__ b,
                               double* __restrict__ y) {
    const int r = blockIdx.x * kBS + threadIdx.x;
    if (r < n) y[r] = a[r] - b[r];
}

// GMRES(m) (Saad & Schultz 1986) with right Jacobi preconditioning and classical Gram-Schmidt
// applied twice; the Hessenberg least-squares problem (Givens rotations) lives on the host.
// Partials are summed on the host in a fixed order, so the iteration is reproducible.
int csr_gmres(int32_t n, const int32_t* ip, const int32_t* ix, const double* v, const uint8_t* fixed,
              const double* dinv, double* x, double* rhs, double* V, double* w, double* z, double* part,
              double* coef, int m, double tol, int maxit, int* iters, double* resid, hipStream_t st) {
    const int gd = grid_of(n);
    const int nwg = gd < kMaxWG ? gd : kMaxWG;
    const int gs = (int)(((int64_t)n * kLanesPerRow + kBS - 1) / kBS);
    std::vector<double> hp((size_t)(m + 1) * kMaxWG), H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), y(m), hc(m + 1);
    auto norm = [&](const double* a, double* out) -> int {
        csr_dot_kernel<<<nwg, kBS, 0, st>>>(n, a, a, part);
        EFE_HIP(hipMemcpyAsync(hp.data(), part, nwg * sizeof(double), hipMemcpyDeviceToHost, st));
        EFE_HIP(hipStreamSynchronize(st));
        double s = 0.0;
        for (int i = 0; i < nwg; ++i) s += hp[i];
        *out = std::sqrt(s);
        return EFE_OK;
    };
    int it = 0;
    double beta = 0.0;
    while (true) {
        // r = rhs - A x into w
        csr_spmv_kernel<<<gs, kBS, 0, st>>>(n, ip, ix, v, fixed, x, w, nullptr);
        csr_sub_kernel<<<gd, kBS, 0, st>>>(n, rhs, w, w);
        if (int rc = norm(w, &beta)) return rc;
        *resid = beta;
        if (!(beta > tol) || it >= maxit) break;
        csr_scale_kernel<<<gd, kBS, 0, st>>>(n, 1.0 / beta, w, V);
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        int j = 0;
        for (; j < m; ++j) {
            jacobi_free_kernel(n, dinv, V + (size_t)j * n, z, st);
            csr_spmv_kernel<<<gs, kBS, 0, st>>>(n, ip, ix, v, fixed, z, w, nullptr);
            std::fill(hc.begin(), hc.end(), 0.0);
            for (int pass = 0; pass < 2; ++pass) {
                csr_multidot_kernel<<<nwg, kBS, 0, st>>>(n, V, j + 1, w, part);
                for (int i = 0; i <= j; ++i)
                    EFE_HIP(hipMemcpyAsync(hp.data() + (size_t)i * kMaxWG, part + (size_t)i * kMaxWG,
                                           nwg * sizeof(double), hipMemcpyDeviceToHost, st));
                EFE_HIP(hipStreamSynchronize(st));
                std::vector<double> c(j + 1);
                for (int i = 0; i <= j; ++i) {
                    double s = 0.0;
                    for (int k = 0; k < nwg; ++k) s += hp[(size_t)i * kMaxWG + k];
                    c[i] = s;
                    hc[i] += s;
                }
                EFE_HIP(hipMemcpyAsync(coef, c.data(), (j + 1) * sizeof(double), hipMemcpyHostToDevice, st));
                csr_sub_combination_kernel<<<gd, kBS, 0, st>>>(n, V, j + 1, coef, w);
                EFE_HIP(hipStreamSynchronize(st));  // c must outlive the copy
            }
            double hn = 0.0;
            if (int rc = norm(w, &hn)) return rc;
            hc[j + 1] = hn;
            ++it;
            // Givens rotations of the new column
            for (int i = 0; i < j; ++i) {
                const double t = cs[i] * hc[i] + sn[i] * hc[i + 1];
                hc[i + 1] = -sn[i] * hc[i] + cs[i] * hc[i + 1];
                hc[i] = t;
            }
            const double den = std::hypot(hc[j], hc[j + 1]);
            cs[j] = den > 0.0 ? hc[j] / den : 1.0;
            sn[j] = den > 0.0 ? hc[j + 1] / den : 0.0;
            hc[j] = den;
            hc[j + 1] = 0.0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            for (int i = 0; i <= j; ++i) H[(size_t)i * m + j] = hc[i];
            *resid = std::fabs(g[j + 1]);
            const bool stop = !(*resid > tol) || it >= maxit || !(hn > 0.0) || !(hn == hn);
            if (!stop) csr_scale_kernel<<<gd, kBS, 0, st>>>(n, 1.0 / hn, w, V + (size_t)(j + 1) * n);
            if (stop) {
                ++j;
                break;
            }
        }
        // y = H^-1 g (upper triangular j x j), x += Dinv V y
        for (int i = j - 1; i >= 0; --i) {
            double s = g[i];
            for (int k = i + 1; k < j; ++k) s -= H[(size_t)i * m + k] * y[k];
            y[i] = H[(size_t)i * m + i] != 0.0 ? s / H[(size_t)i * m + i] : 0.0;
        }
        EFE_HIP(hipMemcpyAsync(coef, y.data(), j * sizeof(double), hipMemcpyHostToDevice, st));
        csr_add_precond_combination_kernel<<<gd, kBS, 0, st>>>(n, V, j, coef, dinv, x);
        EFE_HIP(hipGetLastError());
        EFE_HIP(hipStreamSynchronize(st));  // y must outlive the copy
        if (!(beta == beta)) break;
    }
    *iters = it;
    return EFE_OK;
}

}  // namespace

namespace {
int csr_gmres_solve(int32_t n, const int32_t* ip, const int32_t* ix, const double* v, const double* b, double* xout,
                    const efe_solve_opts* o, efe_solve_info* info, double* ws, hipStream_t st) {
    const int m = o->restart > 0 ? o->restart : 50;
    double* V = ws;
    double* w = V + (size_t)(m + 1) * n;
    double* z = w + n;
    double* x = z + n;
    double* rhs = x + n;
    double* dinv = rhs + n;
    uint8_t* fixed = (uint8_t*)(dinv + n);
    double* part = dinv + n + ((size_t)n + 7) / 8;
    double* coef = part + (size_t)(m + 2) * kMaxWG + 16;
    const int gd = grid_of(n);
    const int gs = (int)(((int64_t)n * kLanesPerRow + kBS - 1) / kBS);
    csr_classify_kernel<<<gd, kBS, 0, st>>>(n, ip, ix, v, fixed, dinv);
    csr_spmv_kernel<<<gs, kBS, 0, st>>>(n, ip, ix, v, fixed, b, rhs, b);  // right-hand side of the free block
    zero_kernel<<<gd, kBS, 0, st>>>(n, x);
    EFE_HIP(hipGetLastError());
    const int nwg = gd < kMaxWG ? gd : kMaxWG;
    csr_dot_kernel<<<nwg, kBS, 0, st>>>(n, rhs, rhs, part);
    std::vector<double> hp(nwg);
    EFE_HIP(hipMemcpyAsync(hp.data(), part, nwg * sizeof(double), hipMemcpyDeviceToHost, st));
    EFE_HIP(hipStreamSynchronize(st));
    double bb = 0.0;
    for (int i = 0; i < nwg; ++i) bb += hp[i];
    const double bnorm = std::sqrt(bb);
    info->rhs_norm = bnorm;
    info->iterations = 0;
    info->status = EFE_OK;
    info->residual = bnorm;
    if (!(bnorm == bnorm)) {
        info->status = EFE_NAN;
    } else if (bnorm > 0.0) {
        const double tol = std::fmax(o->rtol * bnorm, o->atol);
        const int maxit = o->max_iter > 0 ? o->max_iter : 10000;
        int it = 0;
        double res = bnorm;
        if (int rc = csr_gmres(n, ip, ix, v, fixed, dinv, x, rhs, V, w, z, part, coef, m, tol, maxit, &it, &res, st))
            return rc;
        info->iterations = it;
        info->residual = res;
        if (!(res == res)) info->status = EFE_NAN;
        else if (res > tol) info->status = EFE_NOT_CONVERGED;
    }
    csr_finish_kernel<<<gd, kBS, 0, st>>>(n, fixed, b, x);
    EFE_HIP(hipMemcpyAsync(xout, x, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, st));
    EFE_HIP(hipGetLastError());
    EFE_HIP(hipStreamSynchronize(st));
    return EFE_OK;
}
}  // namespace

// The solve on device arrays with the caller's workspace (efe_csr_workspace_doubles doubles).
int efe_csr_solve_ws(int32_t n, const int32_t* ip, const int32_t* ix, const double* v, const double* b, double* xout,
                     const efe_solve_opts* o, efe_solve_info* info, double* ws, hipStream_t st) {
    if (o->solver != EFE_SOLVER_CG && o->solver != EFE_SOLVER_BICGSTAB && o->solver != EFE_SOLVER_GMRES) {
        efe_set_error("CSR solve: unknown solver");
        return EFE_ERR_ARG;
    }
    if (o->solver == EFE_SOLVER_GMRES) return csr_gmres_solve(n, ip, ix, v, b, xout, o, info, ws, st);
    const bool cg = o->solver == EFE_SOLVER_CG;
    const int nvec = cg ? 5 : 8;
    double* vecs = ws;
    double* dinv = ws + (size_t)nvec * n;
    uint8_t* fixed = (uint8_t*)(dinv + n);
    double* part = dinv + n + ((size_t)n + 7) / 8;
    double* sc = part + 4 * kMaxWG;
    const int gd = grid_of(n);
    const int nwg = gd < kMaxWG ? gd : kMaxWG;  // workgroups of the grid-stride vector kernels
    const int gs = (int)(((int64_t)n * kLanesPerRow + kBS - 1) / kBS);
    double hs[16];
    csr_classify_kernel<<<gd, kBS, 0, st>>>(n, ip, ix, v, fixed, dinv);
    double* x = vecs;
    double* r = vecs + n;
    // right-hand side of the free block
    csr_spmv_kernel<<<gs, kBS, 0, st>>>(n, ip, ix, v, fixed, b, r, b);
    zero_kernel<<<gd, kBS, 0, st>>>(n, x);
    // ||b_f'||
    csr_dot_kernel<<<nwg, kBS, 0, st>>>(n, r, r, part);
    set_scalars_kernel<<<1, 64, 0, st>>>(sc, 0.0);
    EFE_HIP(hipGetLastError());
    std::vector<double> hp(nwg);
    EFE_HIP(hipMemcpyAsync(hp.data(), part, nwg * sizeof(double), hipMemcpyDeviceToHost, st));
    EFE_HIP(hipStreamSynchronize(st));
    double bb = 0.0;
    for (int i = 0; i < nwg; ++i) bb += hp[i];
    const double bnorm = std::sqrt(bb);
    info->rhs_norm = bnorm;
    info->iterations = 0;
    info->status = EFE_OK;
    info->residual = bnorm;
    const int maxit = o->max_iter > 0 ? o->max_iter : 10000;
    const int check = o->check_every > 0 ? o->check_every : 32;
    if (!(bnorm == bnorm)) {
        info->status = EFE_NAN;
    } else if (bnorm > 0.0) {
        const double tol = std::fmax(o->rtol * bnorm, o->atol);
        set_scalars_kernel<<<1, 64, 0, st>>>(sc, tol * tol);
        int it = 0;
        if (cg) {
            double *z = vecs + 2 * (size_t)n, *p = vecs + 3 * (size_t)n, *q = vecs + 4 * (size_t)n;
            cg_precond_kernel<<<nwg, kBS, 0, st>>>(n, dinv, r, z, part);
            cg_direction_kernel<<<nwg, kBS, 0, st>>>(n, z, p, part, nwg, sc, 1);
            cg_scalars_kernel<<<1, kBS, 0, st>>>(part, nwg, sc, 0);
            while (true) {
                for (int j = 0; j < check; ++j) {
                    csr_spmv_kernel<<<gs, kBS, 0, st>>>(n, ip, ix, v, fixed, p, q, nullptr);
                    csr_dot_kernel<<<nwg, kBS, 0, st>>>(n, p, q, part);
                    cg_update_kernel<<<nwg, kBS, 0, st>>>(n, x, r, p, q, part, nwg, sc);
                    cg_precond_kernel<<<nwg, kBS, 0, st>>>(n, dinv, r, z, part);
                    cg_direction_kernel<<<nwg, kBS, 0, st>>>(n, z, p, part, nwg, sc, 0);
                    cg_scalars_kernel<<<1, kBS, 0, st>>>(part, nwg, sc, 1);
                }
                EFE_HIP(hipGetLastError());
                EFE_HIP(hipMemcpyAsync(hs, sc, 16 * sizeof(double), hipMemcpyDeviceToHost, st));
                EFE_HIP(hipStreamSynchronize(st));
                it = (int)hs[5];
                if (hs[4] != 0.0) break;
                if (it >= maxit) {
                    info->status = EFE_NOT_CONVERGED;
                    break;
                }
            }
            info->residual = std::sqrt(hs[2]);
        } else {
            double *r0 = vecs + 2 * (size_t)n, *p = vecs + 3 * (size_t)n, *ph = vecs + 4 * (size_t)n;
            double *vv = vecs + 5 * (size_t)n, *sh = vecs + 6 * (size_t)n, *t = vecs + 7 * (size_t)n;
            double* part2 = part + 2 * kMaxWG;  // the omega kernel reads part while it writes these
            copy_kernel<<<gd, kBS, 0, st>>>(n, r, r0);
            cg_precond_kernel<<<nwg, kBS, 0, st>>>(n, dinv, r, ph, part);  // slot 0 unused, slot 1 = r.r
            csr_dot_kernel<<<nwg, kBS, 0, st>>>(n, r0, r, part);
            bicg_init_scalars_kernel<<<1, kBS, 0, st>>>(part, nwg, sc);
            bicg_p_kernel<<<nwg, kBS, 0, st>>>(n, r, p, vv, sc, 1);
            while (true) {
                for (int j = 0; j < check; ++j) {
                    jacobi_kernel<<<gd, kBS, 0, st>>>(n, dinv, p, ph, sc);
                    csr_spmv_kernel<<<gs, kBS, 0, st>>>(n, ip, ix, v, fixed, ph, vv, nullptr);
                    csr_dot_kernel<<<nwg, kBS, 0, st>>>(n, r0, vv, part);
                    bicg_alpha_kernel<<<nwg, kBS, 0, st>>>(n, x, r, ph, vv, part, nwg, sc);
                    jacobi_kernel<<<gd, kBS, 0, st>>>(n, dinv, r, sh, sc);
                    csr_spmv_kernel<<<gs, kBS, 0, st>>>(n, ip, ix, v, fixed, sh, t, nullptr);
                    bicg_ts_kernel<<<nwg, kBS, 0, st>>>(n, t, r, part, sc);
                    bicg_omega_kernel<<<nwg, kBS, 0, st>>>(n, x, r, sh, t, r0, part, nwg, part2, sc);
                    bicg_scalars_kernel<<<1, kBS, 0, st>>>(part2, nwg, sc);
                    bicg_p_kernel<<<nwg, kBS, 0, st>>>(n, r, p, vv, sc, 0);
                }
                EFE_HIP(hipGetLastError());
                EFE_HIP(hipMemcpyAsync(hs, sc, 16 * sizeof(double), hipMemcpyDeviceToHost, st));
                EFE_HIP(hipStreamSynchronize(st));
                it = (int)hs[5];
                if (hs[4] != 0.0) break;
                if (it >= maxit) {
                    info->status = EFE_NOT_CONVERGED;
                    break;
                }
            }
            info->residual = std::sqrt(hs[6]);
        }
        info->iterations = it;
        if (!(info->residual == info->residual)) info->status = EFE_NAN;
    }
    csr_finish_kernel<<<gd, kBS, 0, st>>>(n, fixed, b, x);
    EFE_HIP(hipMemcpyAsync(xout, x, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, st));
    EFE_HIP(hipGetLastError());
    EFE_HIP(hipStreamSynchronize(st));
    return EFE_OK;
}

