// GMRES(m) with multicolor block ILU(0) or block-Jacobi right preconditioning, for the
// non-symmetric tangents of non-associative plasticity, in place of the reference's
// unsymmetric direct solve (PARDISO mtype 11, linsolve/pardiso/pardiso.pyx:104-151).
//
// Block ILU(0): the zero-fill incomplete factorisation on the 3x3-block pattern of K, in
// the elimination order (color, node) of a greedy coloring of the node graph.  The
// factor is kept in block CSR with rows in that order, so every color is one contiguous
// range of mutually independent rows: the factorisation and both triangular sweeps take
// one launch per color (8 on a hexa8 grid) instead of one per dependency level.
//
// GMRES: right preconditioning (the Arnoldi residual is the true residual of K x = b),
// classical Gram-Schmidt with one re-orthogonalisation (all projections of a step in one
// pass over the basis), Givens rotations on the host with one readback per step.
#include <algorithm>
#include <array>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <vector>

#include "efe_internal.h"

struct GmresWork {
    int D = 0, N = 0, m = 0, nwg = 0;
    // ILU structure
    bool ilu_ready = false;
    int ncolors = 0;
    std::vector<int> cstart;
    std::vector<int> clow, cup;  // per color: most blocks before / after the diagonal in a row
    int32_t *fptr = nullptr, *fcol = nullptr, *fidx = nullptr, *dslot = nullptr, *perm = nullptr;
    int64_t nbF = 0;
    int maxUpper = 0;  // most blocks in a row's upper part
    int maxRow = 0;    // most blocks in a row
    double *F64 = nullptr, *Dinv64 = nullptr;
    float *F32 = nullptr, *Dinv32 = nullptr;
    int factored_version[2] = {-1, -1};
    // Krylov basis and work vectors (one precision at a time)
    void* V = nullptr;   // [(m+1)][D]
    void* w = nullptr;
    void* z = nullptr;
    void* x = nullptr;
    void* t = nullptr;   // permuted vector for the ILU sweeps
    int vprec = -1;
    double* part = nullptr;  // [(m+1)][nwg]
    double* hdev = nullptr;  // [3][m+2]: h1, h2, norm
    double* hhost = nullptr; // pinned
    hipStream_t ks = nullptr;
    hipEvent_t evA = nullptr, evB = nullptr;
    hipEvent_t t0 = nullptr, t1 = nullptr, t2 = nullptr;  // timing: solve start, factor done, end
    // one captured hipGraph per Arnoldi step j (the launches between two readbacks) and the
    // pointers and sizes it was captured with
    std::vector<hipGraphExec_t> stepGraph;
    std::vector<std::array<int64_t, 8>> stepKey;
};

namespace {
void drop_step_graphs(GmresWork* g) {
    for (hipGraphExec_t& e : g->stepGraph)
        if (e) hipGraphExecDestroy(e);
    g->stepGraph.clear();
    g->stepKey.clear();
}
}  // namespace

namespace {

constexpr int BS = 256;
constexpr int kFactorCap = 128;  // upper-part blocks of a row the LDS factor kernel stages

__device__ __forceinline__ int xcd_block_g() {
    const int per = gridDim.x >> 3;
    return (blockIdx.x & 7) * per + (blockIdx.x >> 3);
}

inline int round8g(int n) { return (n + 7) & ~7; }

__device__ __forceinline__ double bsum(double v, double* sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = BS / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    double r = sh[0];
    __syncthreads();
    return r;
}

template <typename T>
__device__ __forceinline__ void mat3mul(const T* a, const T* b, T* o) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) o[i * 3 + j] = a[i * 3] * b[j] + a[i * 3 + 1] * b[3 + j] + a[i * 3 + 2] * b[6 + j];
}

__device__ __forceinline__ void inv3d(const double m[9], double o[9]) {
    const double det = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
                       m[2] * (m[3] * m[7] - m[4] * m[6]);
    const double id = 1.0 / det;
    o[0] = (m[4] * m[8] - m[5] * m[7]) * id;
    o[1] = (m[2] * m[7] - m[1] * m[8]) * id;
    o[2] = (m[1] * m[5] - m[2] * m[4]) * id;
    o[3] = (m[5] * m[6] - m[3] * m[8]) * id;
    o[4] = (m[0] * m[8] - m[2] * m[6]) * id;
    o[5] = (m[2] * m[3] - m[0] * m[5]) * id;
    o[6] = (m[3] * m[7] - m[4] * m[6]) * id;
    o[7] = (m[1] * m[6] - m[0] * m[7]) * id;
    o[8] = (m[0] * m[4] - m[1] * m[3]) * id;
}

// F <- K in the factor's block-CSR order (thread per operator node row).
template <typename T>
__global__ __launch_bounds__(BS) void ilu_copy_kernel(int N, const int32_t* __restrict__ bptr,
                                                      const int32_t* __restrict__ slice_off,
                                                      const int32_t* __restrict__ nodeSlot,
                                                      const T* __restrict__ K, const int32_t* __restrict__ fidx,
                                                      T* __restrict__ F) {
    const int n = blockIdx.x * BS + threadIdx.x;
    if (n >= N) return;
    const int ps = nodeSlot[n];
    const int s = ps / kSlice, l = ps - (ps / kSlice) * kSlice;
    const int64_t vb = (int64_t)slice_off[s] * 9 + l;
    const int b0 = bptr[n], nb = bptr[n + 1] - b0;
    // several blocks in flight per lane; measured on the 48^3 hexa20 operator: FP64 2.25 / 1.98
    // / 2.66 ms at 1 / 4 / 8 blocks, FP32 1.44 / 1.20 / 1.04 ms
    constexpr int kUnroll = sizeof(T) == 8 ? 4 : 8;
#pragma unroll kUnroll
    for (int k = 0; k < nb; ++k) {
        T* dst = F + (int64_t)fidx[b0 + k] * 9;
#pragma unroll
        for (int v = 0; v < 9; ++v) dst[v] = K[vb + (int64_t)(k * 9 + v) * kSlice];
    }
}

// Rows [r0, r1) of one color: IKJ block ILU(0).  Rows k < i in elimination order are final.
// A group of G lanes works on one row; the row's block at position p belongs to lane
// (p - b0) % G, which alone reads and writes it.  For each k in the row's lower part
// (in order) the owner of (i, k) forms L_ik = A_ik Dinv_k and broadcasts it; every lane
// then updates its own blocks (i, j), j after k, for which (k, j) is in row k's upper part
// (binary search: a row's columns are sorted in elimination order).
template <typename T, int G>
__global__ __launch_bounds__(BS) void ilu_factor_kernel(int r0, int r1, const int32_t* __restrict__ fptr,
                                                        const int32_t* __restrict__ fcol,
                                                        const int32_t* __restrict__ dslot, T* __restrict__ F,
                                                        T* __restrict__ Dinv) {
    const int tid = blockIdx.x * BS + threadIdx.x;
    const int i = r0 + tid / G, lane = tid % G;
    if (i >= r1) return;  // a whole group leaves together
    const int b0 = fptr[i], b1 = fptr[i + 1], d = b0 + dslot[i];
    for (int kb = b0; kb < d; ++kb) {
        const int k = fcol[kb];
        const int owner = (kb - b0) % G;
        T lik[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (lane == owner) {
            T a[9];
#pragma unroll
            for (int v = 0; v < 9; ++v) a[v] = F[(int64_t)kb * 9 + v];
            mat3mul(a, Dinv + (int64_t)k * 9, lik);
#pragma unroll
            for (int v = 0; v < 9; ++v) F[(int64_t)kb * 9 + v] = lik[v];
        }
#pragma unroll
        for (int v = 0; v < 9; ++v) lik[v] = __shfl(lik[v], owner, G);
        const int u0 = fptr[k] + dslot[k] + 1, u1 = fptr[k + 1];
        const int first = kb + 1 + ((lane - (kb + 1 - b0) % G) % G + G) % G;  // first owned position > kb
        for (int jb = first; jb < b1; jb += G) {
            const int cj = fcol[jb];
            int lo = u0, hi = u1;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (fcol[mid] < cj) lo = mid + 1;
                else hi = mid;
            }
            if (lo < u1 && fcol[lo] == cj) {
                T u[9], p[9];
#pragma unroll
                for (int v = 0; v < 9; ++v) u[v] = F[(int64_t)lo * 9 + v];
                mat3mul(lik, u, p);
#pragma unroll
                for (int v = 0; v < 9; ++v) F[(int64_t)jb * 9 + v] -= p[v];
            }
        }
    }
    if (lane == (d - b0) % G) {
        double dd[9], di[9];
#pragma unroll
        for (int v = 0; v < 9; ++v) dd[v] = (double)F[(int64_t)d * 9 + v];
        inv3d(dd, di);
#pragma unroll
        for (int v = 0; v < 9; ++v) Dinv[(int64_t)i * 9 + v] = (T)di[v];
    }
}

// ilu_factor_kernel with row k's upper column list staged in LDS: the G lanes of a row load it
// together (coalesced), then each searches its columns there instead of through dependent
// global loads.  The same updates in the same order: the factor is bitwise identical.
// Needs every row's upper part to hold at most CAP blocks.
template <typename T, int G, int CAP>
__global__ __launch_bounds__(BS) void ilu_factor_lds_kernel(int r0, int r1, const int32_t* __restrict__ fptr,
                                                            const int32_t* __restrict__ fcol,
                                                            const int32_t* __restrict__ dslot, T* __restrict__ F,
                                                            T* __restrict__ Dinv) {
    __shared__ int32_t cols[(BS / G) * CAP];
    const int tid = blockIdx.x * BS + threadIdx.x;
    const int i = r0 + tid / G, lane = tid % G;
    int32_t* kc = cols + (threadIdx.x / G) * CAP;  // this row group's list (groups lie inside one wave)
    if (i >= r1) return;  // a whole group leaves together
    const int b0 = fptr[i], b1 = fptr[i + 1], d = b0 + dslot[i];
    for (int kb = b0; kb < d; ++kb) {
        const int k = fcol[kb];
        const int owner = (kb - b0) % G;
        const int u0 = fptr[k] + dslot[k] + 1, nu = fptr[k + 1] - u0;
        for (int t = lane; t < nu; t += G) kc[t] = fcol[u0 + t];
        T lik[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (lane == owner) {
            T a[9];
#pragma unroll
            for (int v = 0; v < 9; ++v) a[v] = F[(int64_t)kb * 9 + v];
            mat3mul(a, Dinv + (int64_t)k * 9, lik);
#pragma unroll
            for (int v = 0; v < 9; ++v) F[(int64_t)kb * 9 + v] = lik[v];
        }
#pragma unroll
        for (int v = 0; v < 9; ++v) lik[v] = __shfl(lik[v], owner, G);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int first = kb + 1 + ((lane - (kb + 1 - b0) % G) % G + G) % G;  // first owned position > kb
        for (int jb = first; jb < b1; jb += G) {
            const int cj = fcol[jb];
            int lo = 0, hi = nu;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (kc[mid] < cj) lo = mid + 1;
                else hi = mid;
            }
            if (lo < nu && kc[lo] == cj) {
                T u[9], p[9];
                const int64_t ub = (int64_t)(u0 + lo) * 9;
#pragma unroll
                for (int v = 0; v < 9; ++v) u[v] = F[ub + v];
                mat3mul(lik, u, p);
#pragma unroll
                for (int v = 0; v < 9; ++v) F[(int64_t)jb * 9 + v] -= p[v];
            }
        }
        // every lane's searches are done before the list is overwritten
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    if (lane == (d - b0) % G) {
        double dd[9], di[9];
#pragma unroll
        for (int v = 0; v < 9; ++v) dd[v] = (double)F[(int64_t)d * 9 + v];
        inv3d(dd, di);
#pragma unroll
        for (int v = 0; v < 9; ++v) Dinv[(int64_t)i * 9 + v] = (T)di[v];
    }
}

// ilu_factor_lds_kernel with the row's own blocks held in registers: the G lanes of a row own
// its blocks b0 + lane + t G (t < NT) from the start, update them there and write them once at
// the end.  What remains per elimination step k is k's upper column list (LDS), Dinv_k and the
// U_kj blocks the row meets.  The same updates in the same order as ilu_factor_kernel: the
// factor is bitwise identical.  Needs rows of at most G NT blocks and upper parts of at most
// CAP blocks.
template <typename T, int G, int NT, int CAP>
__global__ __launch_bounds__(BS) void ilu_factor_reg_kernel(int r0, int r1, const int32_t* __restrict__ fptr,
                                                            const int32_t* __restrict__ fcol,
                                                            const int32_t* __restrict__ dslot, T* __restrict__ F,
                                                            T* __restrict__ Dinv) {
    __shared__ int32_t cols[(BS / G) * CAP];
    const int tid = blockIdx.x * BS + threadIdx.x;
    const int i = r0 + tid / G, lane = tid % G;
    int32_t* kc = cols + (threadIdx.x / G) * CAP;
    if (i >= r1) return;  // a whole group leaves together
    const int b0 = fptr[i], b1 = fptr[i + 1], d = b0 + dslot[i];
    T blk[NT][9];
    int col[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int jb = b0 + lane + t * G;
        col[t] = jb < b1 ? fcol[jb] : -1;
#pragma unroll
        for (int v = 0; v < 9; ++v) blk[t][v] = jb < b1 ? F[(int64_t)jb * 9 + v] : (T)0;
    }
    for (int kb = b0; kb < d; ++kb) {
        const int rel = kb - b0, owner = rel % G, tk = rel / G;
        int ck = 0;
#pragma unroll
        for (int t = 0; t < NT; ++t)
            if (t == tk) ck = col[t];
        const int k = __shfl(ck, owner, G);
        const int u0 = fptr[k] + dslot[k] + 1, nu = fptr[k + 1] - u0;
        for (int t = lane; t < nu; t += G) kc[t] = fcol[u0 + t];
        T lik[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (lane == owner) {
            T a[9];
#pragma unroll
            for (int t = 0; t < NT; ++t)
                if (t == tk)
#pragma unroll
                    for (int v = 0; v < 9; ++v) a[v] = blk[t][v];
            mat3mul(a, Dinv + (int64_t)k * 9, lik);
#pragma unroll
            for (int t = 0; t < NT; ++t)
                if (t == tk)
#pragma unroll
                    for (int v = 0; v < 9; ++v) blk[t][v] = lik[v];
        }
#pragma unroll
        for (int v = 0; v < 9; ++v) lik[v] = __shfl(lik[v], owner, G);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int jb = b0 + lane + t * G;
            if (jb <= kb || jb >= b1) continue;
            const int cj = col[t];
            int lo = 0, hi = nu;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (kc[mid] < cj) lo = mid + 1;
                else hi = mid;
            }
            if (lo < nu && kc[lo] == cj) {
                T u[9], p[9];
                const int64_t ub = (int64_t)(u0 + lo) * 9;
#pragma unroll
                for (int v = 0; v < 9; ++v) u[v] = F[ub + v];
                mat3mul(lik, u, p);
#pragma unroll
                for (int v = 0; v < 9; ++v) blk[t][v] -= p[v];
            }
        }
        // every lane's searches are done before the list is overwritten
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int jb = b0 + lane + t * G;
        if (jb >= b1) continue;
#pragma unroll
        for (int v = 0; v < 9; ++v) F[(int64_t)jb * 9 + v] = blk[t][v];
        if (jb == d) {
            double dd[9], di[9];
#pragma unroll
            for (int v = 0; v < 9; ++v) dd[v] = (double)blk[t][v];
            inv3d(dd, di);
#pragma unroll
            for (int v = 0; v < 9; ++v) Dinv[(int64_t)i * 9 + v] = (T)di[v];
        }
    }
}

// t = perm(v): t[factor row i] = v[node perm[i]]
template <typename T>
__global__ void permute_in_kernel(int N, const int32_t* __restrict__ perm, const T* __restrict__ v,
                                  T* __restrict__ t) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int n = perm[i];
    t[3 * i] = v[3 * n];
    t[3 * i + 1] = v[3 * n + 1];
    t[3 * i + 2] = v[3 * n + 2];
}

template <typename T>
__global__ void permute_out_kernel(int N, const int32_t* __restrict__ perm, const T* __restrict__ t,
                                   T* __restrict__ v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int n = perm[i];
    v[3 * n] = t[3 * i];
    v[3 * n + 1] = t[3 * i + 1];
    v[3 * n + 2] = t[3 * i + 2];
}

// Sum of the 3-vectors of a group of G lanes (fixed butterfly order).
template <typename T, int G>
__device__ __forceinline__ void group_sum3(T& a, T& b, T& c) {
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) {
        a += __shfl_xor(a, off, G);
        b += __shfl_xor(b, off, G);
        c += __shfl_xor(c, off, G);
    }
}

// forward sweep of one color: t_i -= sum_{k before i} L_ik t_k (unit block-lower);
// G lanes per row, each over every G-th block of the row's lower part.  A lane takes U of its
// blocks at a time: their column indices, then the t entries they name, then the blocks are
// loaded together, so one pass costs two dependent memory round trips, not two per block.
// Each lane adds its blocks in the same order as one at a time.
// With Dinv given (the last color, whose rows have no blocks after them), the backward sweep's
// t_i = Dinv_i (t_i - 0) follows in the same launch.
template <typename T, int G, int U>
__global__ __launch_bounds__(BS) void ilu_lower_kernel(int r0, int r1, const int32_t* __restrict__ fptr,
                                                       const int32_t* __restrict__ fcol,
                                                       const int32_t* __restrict__ dslot, const T* __restrict__ F,
                                                       T* __restrict__ t, const T* __restrict__ Dinv) {
    const int tid = blockIdx.x * BS + threadIdx.x;
    const int i = r0 + tid / G, lane = tid % G;
    if (i >= r1) return;
    const int b0 = fptr[i], d = b0 + dslot[i];
    T y0 = 0, y1 = 0, y2 = 0;
    for (int kb0 = b0 + lane; kb0 < d; kb0 += G * U) {
        int kk[U];
#pragma unroll
        for (int u = 0; u < U; ++u) kk[u] = kb0 + u * G < d ? fcol[kb0 + u * G] : -1;
        T x[U][3];
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int c = 0; c < 3; ++c) x[u][c] = kk[u] >= 0 ? t[3 * kk[u] + c] : (T)0;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (kk[u] < 0) break;
            const T* a = F + (int64_t)(kb0 + u * G) * 9;
            y0 += a[0] * x[u][0] + a[1] * x[u][1] + a[2] * x[u][2];
            y1 += a[3] * x[u][0] + a[4] * x[u][1] + a[5] * x[u][2];
            y2 += a[6] * x[u][0] + a[7] * x[u][1] + a[8] * x[u][2];
        }
    }
    group_sum3<T, G>(y0, y1, y2);
    if (lane == 0) {
        y0 = t[3 * i] - y0;
        y1 = t[3 * i + 1] - y1;
        y2 = t[3 * i + 2] - y2;
        if (Dinv) {
            const T* m = Dinv + (int64_t)i * 9;
            const T z0 = y0 - (T)0, z1 = y1 - (T)0, z2 = y2 - (T)0;  // as the backward sweep forms them
            y0 = m[0] * z0 + m[1] * z1 + m[2] * z2;
            y1 = m[3] * z0 + m[4] * z1 + m[5] * z2;
            y2 = m[6] * z0 + m[7] * z1 + m[8] * z2;
        }
        t[3 * i] = y0;
        t[3 * i + 1] = y1;
        t[3 * i + 2] = y2;
    }
}

// backward sweep of one color: t_i = Dinv_i (t_i - sum_{j after i} U_ij t_j); blocks taken U
// at a time per lane as in the forward sweep
template <typename T, int G, int U>
__global__ __launch_bounds__(BS) void ilu_upper_kernel(int r0, int r1, const int32_t* __restrict__ fptr,
                                                       const int32_t* __restrict__ fcol,
                                                       const int32_t* __restrict__ dslot, const T* __restrict__ F,
                                                       const T* __restrict__ Dinv, T* __restrict__ t) {
    const int tid = blockIdx.x * BS + threadIdx.x;
    const int i = r0 + tid / G, lane = tid % G;
    if (i >= r1) return;
    const int e = fptr[i + 1];
    T y0 = 0, y1 = 0, y2 = 0;
    for (int jb0 = fptr[i] + dslot[i] + 1 + lane; jb0 < e; jb0 += G * U) {
        int jj[U];
#pragma unroll
        for (int u = 0; u < U; ++u) jj[u] = jb0 + u * G < e ? fcol[jb0 + u * G] : -1;
        T x[U][3];
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int c = 0; c < 3; ++c) x[u][c] = jj[u] >= 0 ? t[3 * jj[u] + c] : (T)0;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (jj[u] < 0) break;
            const T* a = F + (int64_t)(jb0 + u * G) * 9;
            y0 += a[0] * x[u][0] + a[1] * x[u][1] + a[2] * x[u][2];
            y1 += a[3] * x[u][0] + a[4] * x[u][1] + a[5] * x[u][2];
            y2 += a[6] * x[u][0] + a[7] * x[u][1] + a[8] * x[u][2];
        }
    }
    group_sum3<T, G>(y0, y1, y2);
    if (lane == 0) {
        y0 = t[3 * i] - y0;
        y1 = t[3 * i + 1] - y1;
        y2 = t[3 * i + 2] - y2;
        const T* m = Dinv + (int64_t)i * 9;
        t[3 * i] = m[0] * y0 + m[1] * y1 + m[2] * y2;
        t[3 * i + 1] = m[3] * y0 + m[4] * y1 + m[5] * y2;
        t[3 * i + 2] = m[6] * y0 + m[7] * y1 + m[8] * y2;
    }
}

template <typename T>
__global__ void bjacobi_kernel(int N, const T* __restrict__ Minv, const T* __restrict__ v, T* __restrict__ z) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const T* M = Minv + (int64_t)n * 9;
    const T a = v[3 * n], b = v[3 * n + 1], c = v[3 * n + 2];
    z[3 * n] = M[0] * a + M[1] * b + M[2] * c;
    z[3 * n + 1] = M[3] * a + M[4] * b + M[5] * c;
    z[3 * n + 2] = M[6] * a + M[7] * b + M[8] * c;
}

// y = K x (sliced ELL), optionally y = b - K x
template <typename T>
__global__ __launch_bounds__(BS) void gm_spmv_kernel(int N, const int32_t* __restrict__ slice_off,
                                                     const int32_t* __restrict__ scol,
                                                     const int32_t* __restrict__ slotNode, const T* __restrict__ K,
                                                     const T* __restrict__ x, T* __restrict__ y,
                                                     const double* __restrict__ b) {
    const int ps = xcd_block_g() * BS + threadIdx.x;  // row slot
    if (ps >= N) return;
    const int n = slotNode[ps];
    const int s = ps / kSlice, l = ps - (ps / kSlice) * kSlice;
    const int o0 = slice_off[s], deg = (slice_off[s + 1] - o0) / kSlice;
    const int32_t* cp = scol + o0 + l;
    const T* kp = K + (int64_t)o0 * 9 + l;
    T y0 = 0, y1 = 0, y2 = 0;
    // unrolled so that the loads of several blocks are in flight together; measured on the 48^3
    // hexa20 operator: FP64 455 / 438 / 491 / 452 us at 1 / 2 / 3 / 4 blocks, FP32 243 / 214 /
    // 209 / 210 us
    constexpr int kUnroll = sizeof(T) == 8 ? 2 : 3;
#pragma unroll kUnroll
    for (int k = 0; k < deg; ++k) {
        const int m = __builtin_nontemporal_load(cp + k * kSlice);
        const T x0 = x[3 * m], x1 = x[3 * m + 1], x2 = x[3 * m + 2];
        const T* v = kp + (int64_t)k * 9 * kSlice;
        T a[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) a[i] = __builtin_nontemporal_load(v + i * kSlice);
        y0 += a[0] * x0 + a[1] * x1 + a[2] * x2;
        y1 += a[3] * x0 + a[4] * x1 + a[5] * x2;
        y2 += a[6] * x0 + a[7] * x1 + a[8] * x2;
    }
    if (b) {
        y[3 * n] = (T)b[3 * n] - y0;
        y[3 * n + 1] = (T)b[3 * n + 1] - y1;
        y[3 * n + 2] = (T)b[3 * n + 2] - y2;
    } else {
        y[3 * n] = y0;
        y[3 * n + 1] = y1;
        y[3 * n + 2] = y2;
    }
}

// partial dots of w with basis vectors V_0..V_j (blockIdx.y = vector)
template <typename T>
__global__ __launch_bounds__(BS) void mdot_kernel(int D, const T* __restrict__ V, const T* __restrict__ w,
                                                  double* __restrict__ part, int nwg) {
    __shared__ double sh[BS];
    const int blk = blockIdx.x;
    const T* vi = V + (int64_t)blockIdx.y * D;
    double s = 0.0;
#pragma unroll 4
    for (int i = blk * BS + threadIdx.x; i < D; i += nwg * BS) s += (double)vi[i] * (double)w[i];
    s = bsum(s, sh);
    if (threadIdx.x == 0) part[(int64_t)blockIdx.y * nwg + blk] = s;
}

__global__ __launch_bounds__(BS) void mdot_reduce_kernel(const double* __restrict__ part, int nwg,
                                                         double* __restrict__ out, int accumulate) {
    __shared__ double sh[BS];
    double s = 0.0;
    for (int i = threadIdx.x; i < nwg; i += BS) s += part[(int64_t)blockIdx.x * nwg + i];
    s = bsum(s, sh);
    if (threadIdx.x == 0) out[blockIdx.x] = accumulate ? out[blockIdx.x] + s : s;
}

// w -= sum_i h_i V_i  (sign = -1) or w = sum_i h_i V_i (sign = 0)
template <typename T>
__global__ void combine_kernel(int D, int nv, const T* __restrict__ V, const double* __restrict__ h,
                               T* __restrict__ w, int overwrite) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D) return;
    double acc = overwrite ? 0.0 : (double)w[i];
#pragma unroll 4
    for (int k = 0; k < nv; ++k) acc += (overwrite ? 1.0 : -1.0) * h[k] * (double)V[(int64_t)k * D + i];
    w[i] = (T)acc;
}

template <typename T>
__global__ void scale_kernel(int D, const T* __restrict__ w, const double* __restrict__ nrm2, T* __restrict__ v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D) return;
    v[i] = (T)((double)w[i] / sqrt(nrm2[0]));
}

template <typename T>
__global__ void axpy_kernel(int D, const T* __restrict__ z, T* __restrict__ x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < D) x[i] += z[i];
}

template <typename T>
__global__ void zero_kernel(int D, T* __restrict__ x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < D) x[i] = 0;
}

template <typename T>
__global__ void finish_kernel(int D, const T* __restrict__ x, const double* __restrict__ b,
                              const uint8_t* __restrict__ dirflag, int ndir, double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D) return;
    out[i] = (ndir && dirflag[i]) ? b[i] : (double)x[i];
}

int grid_of(int n) { return (n + BS - 1) / BS; }

int build_ilu(efe_ctx* c, GmresWork* g) {
    IluPattern ip;
    std::string err;
    if (!efe_build_ilu_pattern(c->hp, ip, err)) {
        efe_set_error(err);
        return EFE_ERR_ARG;
    }
    const int N = c->N;
    // block CSR of the factor in elimination order; fidx: operator block -> factor block
    std::vector<int32_t> fptr(N + 1, 0), fcol, fidx(c->hp.bptr[N]), dslot(N);
    for (int i = 0; i < N; ++i) fptr[i + 1] = fptr[i] + ip.rowlen[i];
    fcol.resize(fptr[N]);
    for (int i = 0; i < N; ++i) {
        const int n = ip.perm[i];
        const int s = i / kSlice, l = i % kSlice;
        dslot[i] = ip.dslot[i];
        std::vector<std::pair<int32_t, int32_t>> row;
        for (int32_t b = c->hp.bptr[n]; b < c->hp.bptr[n + 1]; ++b) row.emplace_back(ip.iperm[c->hp.bcol[b]], b);
        std::sort(row.begin(), row.end());
        for (int k = 0; k < ip.rowlen[i]; ++k) {
            fcol[fptr[i] + k] = row[k].first;
            fidx[row[k].second] = fptr[i] + k;
        }
        (void)s;
        (void)l;
    }
    g->maxUpper = 0;
    g->maxRow = 0;
    for (int i = 0; i < N; ++i) {
        g->maxUpper = std::max(g->maxUpper, fptr[i + 1] - fptr[i] - dslot[i] - 1);
        g->maxRow = std::max(g->maxRow, fptr[i + 1] - fptr[i]);
    }
    g->ncolors = ip.ncolors;
    g->cstart.assign(ip.color_start.begin(), ip.color_start.end());
    g->clow.assign(g->ncolors, 0);
    g->cup.assign(g->ncolors, 0);
    for (int col = 0; col < g->ncolors; ++col)
        for (int i = g->cstart[col]; i < g->cstart[col + 1]; ++i) {
            g->clow[col] = std::max(g->clow[col], dslot[i]);
            g->cup[col] = std::max(g->cup[col], fptr[i + 1] - fptr[i] - dslot[i] - 1);
        }
    g->nbF = fptr[N];
    EFE_HIP(hipMalloc(&g->fptr, fptr.size() * sizeof(int32_t)));
    EFE_HIP(hipMalloc(&g->fcol, fcol.size() * sizeof(int32_t)));
    EFE_HIP(hipMalloc(&g->fidx, fidx.size() * sizeof(int32_t)));
    EFE_HIP(hipMalloc(&g->dslot, dslot.size() * sizeof(int32_t)));
    EFE_HIP(hipMalloc(&g->perm, (size_t)N * sizeof(int32_t)));
    EFE_HIP(hipMemcpy(g->fptr, fptr.data(), fptr.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    EFE_HIP(hipMemcpy(g->fcol, fcol.data(), fcol.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    EFE_HIP(hipMemcpy(g->fidx, fidx.data(), fidx.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    EFE_HIP(hipMemcpy(g->dslot, dslot.data(), dslot.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    EFE_HIP(hipMemcpy(g->perm, ip.perm.data(), (size_t)N * sizeof(int32_t), hipMemcpyHostToDevice));
    g->ilu_ready = true;
    return EFE_OK;
}

template <typename T>
int ilu_factor(efe_ctx* c, GmresWork* g, const T* K, T*& F, T*& Dinv, int pi, hipStream_t st) {
    if (!F) {
        EFE_HIP(hipMalloc(&F, (size_t)g->nbF * 9 * sizeof(T)));
        EFE_HIP(hipMalloc(&Dinv, (size_t)c->N * 9 * sizeof(T)));
    }
    if (g->factored_version[pi] == c->Kversion) return EFE_OK;
    ilu_copy_kernel<T><<<grid_of(c->N), BS, 0, st>>>(c->N, c->bptr, c->slice_off, c->nodeSlot, K, g->fidx, F);
    for (int col = 0; col < g->ncolors; ++col) {
        const int r0 = g->cstart[col], r1 = g->cstart[col + 1];
        if (r1 <= r0) continue;
        // the row in registers: hexa8 rows have up to 27 blocks, hexa20 rows up to 81.  (Issuing
        // the next steps' row-k loads ahead, with a second LDS list buffer, was measured: no
        // gain at 48^3 hexa20, the extra registers cost as much as the hidden latency.)
        if (g->maxUpper <= kFactorCap && g->maxRow <= 32) {
            ilu_factor_reg_kernel<T, 16, 2, kFactorCap><<<grid_of((r1 - r0) * 16), BS, 0, st>>>(
                r0, r1, g->fptr, g->fcol, g->dslot, F, Dinv);
        } else if (g->maxUpper <= kFactorCap && g->maxRow <= 96) {
            ilu_factor_reg_kernel<T, 32, 3, kFactorCap><<<grid_of((r1 - r0) * 32), BS, 0, st>>>(
                r0, r1, g->fptr, g->fcol, g->dslot, F, Dinv);
        } else if (g->maxUpper <= kFactorCap) {  // row k's upper columns staged in LDS
            if (c->NN == 8)
                ilu_factor_lds_kernel<T, 8, kFactorCap><<<grid_of((r1 - r0) * 8), BS, 0, st>>>(
                    r0, r1, g->fptr, g->fcol, g->dslot, F, Dinv);
            else
                ilu_factor_lds_kernel<T, 16, kFactorCap><<<grid_of((r1 - r0) * 16), BS, 0, st>>>(
                    r0, r1, g->fptr, g->fcol, g->dslot, F, Dinv);
        } else if (c->NN == 8) {
            ilu_factor_kernel<T, 8><<<grid_of((r1 - r0) * 8), BS, 0, st>>>(r0, r1, g->fptr, g->fcol, g->dslot, F, Dinv);
        } else {
            ilu_factor_kernel<T, 16><<<grid_of((r1 - r0) * 16), BS, 0, st>>>(r0, r1, g->fptr, g->fcol, g->dslot, F,
                                                                             Dinv);
        }
    }
    EFE_HIP(hipGetLastError());
    g->factored_version[pi] = c->Kversion;
    return EFE_OK;
}

// z = M^{-1} v
template <typename T>
void apply_precond(efe_ctx* c, GmresWork* g, int precond, const T* F, const T* Dinv, const T* Minv, const T* v, T* z,
                   hipStream_t st) {
    const int N = c->N;
    if (precond == EFE_PRECOND_ILU0) {
        T* t = (T*)g->t;
        permute_in_kernel<T><<<grid_of(N), BS, 0, st>>>(N, g->perm, v, t);
        const bool h8 = c->NN == 8;
        const int last = g->ncolors - 1;
        // the last color's backward sweep rides on its forward sweep when its rows have no
        // blocks after them; a color whose rows have no blocks before them (the first) has no
        // forward sweep
        const bool fuseLast = last >= 0 && g->cup[last] == 0;
        for (int col = 0; col < g->ncolors; ++col) {
            const int r0 = g->cstart[col], r1 = g->cstart[col + 1];
            const bool fuse = fuseLast && col == last;
            if (r1 <= r0 || (g->clow[col] == 0 && !fuse)) continue;
            const T* D = fuse ? Dinv : nullptr;
            if (h8)
                ilu_lower_kernel<T, 8, 4><<<grid_of((r1 - r0) * 8), BS, 0, st>>>(r0, r1, g->fptr, g->fcol, g->dslot, F, t,
                                                                              D);
            else
                ilu_lower_kernel<T, 16, 4><<<grid_of((r1 - r0) * 16), BS, 0, st>>>(r0, r1, g->fptr, g->fcol, g->dslot, F,
                                                                                t, D);
        }
        for (int col = g->ncolors - 1; col >= 0; --col) {
            const int r0 = g->cstart[col], r1 = g->cstart[col + 1];
            if (r1 <= r0 || (fuseLast && col == last)) continue;
            if (h8)
                ilu_upper_kernel<T, 8, 4><<<grid_of((r1 - r0) * 8), BS, 0, st>>>(r0, r1, g->fptr, g->fcol, g->dslot, F,
                                                                              Dinv, t);
            else
                ilu_upper_kernel<T, 16, 4><<<grid_of((r1 - r0) * 16), BS, 0, st>>>(r0, r1, g->fptr, g->fcol, g->dslot,
                                                                                F, Dinv, t);
        }
        permute_out_kernel<T><<<grid_of(N), BS, 0, st>>>(N, g->perm, t, z);
    } else {
        bjacobi_kernel<T><<<grid_of(N), BS, 0, st>>>(N, Minv, v, z);
    }
}

// Arnoldi step j: w = K M^{-1} V_j, two classical Gram-Schmidt passes against V_0..V_j
// (projections into hdev[0..], hdev[m+2..]), ||w||^2 into hdev[2(m+2)], V_{j+1} = w / ||w||.
template <typename T>
void gmres_step_launch(efe_ctx* c, GmresWork* g, int precond, const T* F, const T* Dinv, const T* Minv, const T* K,
                       int j, hipStream_t st) {
    const int D = c->D, N = c->N, m = g->m;
    const int gd = grid_of(D), gs = round8g(grid_of(N));
    T* V = (T*)g->V;
    T* w = (T*)g->w;
    T* z = (T*)g->z;
    apply_precond<T>(c, g, precond, F, Dinv, Minv, V + (int64_t)j * D, z, st);
    gm_spmv_kernel<T><<<gs, BS, 0, st>>>(N, c->slice_off, c->scol, c->slotNode, K, z, w, nullptr);
    for (int pass = 0; pass < 2; ++pass) {
        mdot_kernel<T><<<dim3(g->nwg, j + 1), BS, 0, st>>>(D, V, w, g->part, g->nwg);
        mdot_reduce_kernel<<<j + 1, BS, 0, st>>>(g->part, g->nwg, g->hdev + pass * (m + 2), 0);
        combine_kernel<T><<<gd, BS, 0, st>>>(D, j + 1, V, g->hdev + pass * (m + 2), w, 0);
    }
    mdot_kernel<T><<<dim3(g->nwg, 1), BS, 0, st>>>(D, w, w, g->part, g->nwg);
    mdot_reduce_kernel<<<1, BS, 0, st>>>(g->part, g->nwg, g->hdev + 2 * (m + 2), 0);
    scale_kernel<T><<<gd, BS, 0, st>>>(D, w, g->hdev + 2 * (m + 2), V + (int64_t)(j + 1) * D);
}

// Step j as one replayed hipGraph (captured on first use for these buffers); with
// EFE_GMRES_NO_GRAPH set, the launches go to the stream one by one.
template <typename T>
int run_step(efe_ctx* c, GmresWork* g, int precond, const T* F, const T* Dinv, const T* Minv, const T* K, int j,
             hipStream_t st) {
    static const bool graphs = std::getenv("EFE_GMRES_NO_GRAPH") == nullptr;
    if (!graphs) {
        gmres_step_launch<T>(c, g, precond, F, Dinv, Minv, K, j, st);
        EFE_HIP(hipGetLastError());
        return EFE_OK;
    }
    if ((int)g->stepGraph.size() < g->m) {
        g->stepGraph.resize(g->m, nullptr);
        g->stepKey.resize(g->m);
    }
    const std::array<int64_t, 8> key = {precond,
                                        (int64_t)sizeof(T),
                                        (int64_t)(intptr_t)F,
                                        (int64_t)(intptr_t)Dinv,
                                        (int64_t)(intptr_t)Minv,
                                        (int64_t)(intptr_t)K,
                                        (int64_t)(intptr_t)g->V,
                                        g->ncolors};
    if (!g->stepGraph[j] || g->stepKey[j] != key) {
        if (g->stepGraph[j]) hipGraphExecDestroy(g->stepGraph[j]);
        g->stepGraph[j] = nullptr;
        hipGraph_t gr;
        EFE_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
        gmres_step_launch<T>(c, g, precond, F, Dinv, Minv, K, j, st);
        EFE_HIP(hipStreamEndCapture(st, &gr));
        EFE_HIP(hipGraphInstantiate(&g->stepGraph[j], gr, nullptr, nullptr, 0));
        EFE_HIP(hipGraphDestroy(gr));
        g->stepKey[j] = key;
    }
    EFE_HIP(hipGraphLaunch(g->stepGraph[j], st));
    return EFE_OK;
}

template <typename T>
int gmres_solve(efe_ctx* c, GmresWork* g, const double* b, double* xout, const efe_solve_opts* o,
                efe_solve_info* info, const T* K, const T* Minv, int pi) {
    hipStream_t st = g->ks;
    const int D = c->D, N = c->N, m = g->m;
    EFE_HIP(hipEventRecord(g->evA, c->stream));
    EFE_HIP(hipStreamWaitEvent(st, g->evA, 0));
    EFE_HIP(hipEventRecord(g->t0, st));
    T *F = nullptr, *Dinv = nullptr;
    if (o->precond == EFE_PRECOND_ILU0) {
        if (!g->ilu_ready) {
            int rc = build_ilu(c, g);
            if (rc) return rc;
        }
        T*& Fr = pi ? (T*&)g->F32 : (T*&)g->F64;
        T*& Dr = pi ? (T*&)g->Dinv32 : (T*&)g->Dinv64;
        int rc = ilu_factor<T>(c, g, K, Fr, Dr, pi, st);
        if (rc) return rc;
        F = Fr;
        Dinv = Dr;
    }
    EFE_HIP(hipEventRecord(g->t1, st));
    T* V = (T*)g->V;
    T* w = (T*)g->w;
    T* z = (T*)g->z;
    T* x = (T*)g->x;
    const int gd = grid_of(D);
    const int gs = round8g(grid_of(N));
    double* hh = g->hhost;
    // ||b||
    zero_kernel<T><<<gd, BS, 0, st>>>(D, x);
    gm_spmv_kernel<T><<<gs, BS, 0, st>>>(N, c->slice_off, c->scol, c->slotNode, K, x, w, b);
    mdot_kernel<T><<<dim3(g->nwg, 1), BS, 0, st>>>(D, w, w, g->part, g->nwg);
    mdot_reduce_kernel<<<1, BS, 0, st>>>(g->part, g->nwg, g->hdev + 2 * (m + 2), 0);
    EFE_HIP(hipMemcpyAsync(hh, g->hdev + 2 * (m + 2), sizeof(double), hipMemcpyDeviceToHost, st));
    EFE_HIP(hipStreamSynchronize(st));
    const double bnorm = sqrt(hh[0]);
    info->rhs_norm = bnorm;
    info->iterations = 0;
    info->status = EFE_OK;
    info->residual = bnorm;
    const double tol = fmax(o->rtol * bnorm, o->atol);
    const int maxit = o->max_iter > 0 ? o->max_iter : 10000;
    std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), gv(m + 1), y(m);
    double beta = bnorm;
    int it = 0;
    bool done = !(bnorm > tol), bad = !(bnorm == bnorm);
    while (!done && !bad) {
        // r = b - K x is in w; V_0 = r / beta
        scale_kernel<T><<<gd, BS, 0, st>>>(D, w, g->hdev + 2 * (m + 2), V);
        std::fill(gv.begin(), gv.end(), 0.0);
        gv[0] = beta;
        int jend = -1;
        for (int j = 0; j < m; ++j) {
            if (int rc = run_step<T>(c, g, o->precond, F, Dinv, Minv, K, j, st)) return rc;
            EFE_HIP(hipMemcpyAsync(hh, g->hdev, 3 * (m + 2) * sizeof(double), hipMemcpyDeviceToHost, st));
            EFE_HIP(hipStreamSynchronize(st));
            ++it;
            double* col = &H[(size_t)j * (m + 1)];
            for (int i = 0; i <= j; ++i) col[i] = hh[i] + hh[(m + 2) + i];
            col[j + 1] = sqrt(hh[2 * (m + 2)]);
            for (int i = 0; i < j; ++i) {
                const double a = cs[i] * col[i] + sn[i] * col[i + 1];
                col[i + 1] = -sn[i] * col[i] + cs[i] * col[i + 1];
                col[i] = a;
            }
            const double r = std::hypot(col[j], col[j + 1]);
            cs[j] = r > 0 ? col[j] / r : 1.0;
            sn[j] = r > 0 ? col[j + 1] / r : 0.0;
            col[j] = r;
            col[j + 1] = 0.0;
            gv[j + 1] = -sn[j] * gv[j];
            gv[j] = cs[j] * gv[j];
            const double res = fabs(gv[j + 1]);
            info->residual = res;
            jend = j;
            if (!(res == res)) {
                bad = true;
                break;
            }
            if (res <= tol) {
                done = true;
                break;
            }
            if (it >= maxit) break;
        }
        if (bad || jend < 0) break;
        // y = H^{-1} g, u = V y, x += M^{-1} u
        for (int i = jend; i >= 0; --i) {
            double s = gv[i];
            for (int k = i + 1; k <= jend; ++k) s -= H[(size_t)k * (m + 1) + i] * y[k];
            y[i] = s / H[(size_t)i * (m + 1) + i];
        }
        EFE_HIP(hipMemcpyAsync(g->hdev, y.data(), (jend + 1) * sizeof(double), hipMemcpyHostToDevice, st));
        combine_kernel<T><<<gd, BS, 0, st>>>(D, jend + 1, V, g->hdev, w, 1);
        apply_precond<T>(c, g, o->precond, F, Dinv, Minv, w, z, st);
        axpy_kernel<T><<<gd, BS, 0, st>>>(D, z, x);
        if (done || it >= maxit) break;
        // restart: w = b - K x
        gm_spmv_kernel<T><<<gs, BS, 0, st>>>(N, c->slice_off, c->scol, c->slotNode, K, x, w, b);
        mdot_kernel<T><<<dim3(g->nwg, 1), BS, 0, st>>>(D, w, w, g->part, g->nwg);
        mdot_reduce_kernel<<<1, BS, 0, st>>>(g->part, g->nwg, g->hdev + 2 * (m + 2), 0);
        EFE_HIP(hipMemcpyAsync(hh, g->hdev + 2 * (m + 2), sizeof(double), hipMemcpyDeviceToHost, st));
        EFE_HIP(hipStreamSynchronize(st));
        beta = sqrt(hh[0]);
        info->residual = beta;
        if (!(beta == beta)) bad = true;
        if (beta <= tol) done = true;
    }
    EFE_HIP(hipEventRecord(g->t2, st));
    finish_kernel<T><<<gd, BS, 0, st>>>(D, x, b, c->dirflag, c->ndir, xout);
    EFE_HIP(hipGetLastError());
    EFE_HIP(hipEventRecord(g->evB, st));
    EFE_HIP(hipStreamWaitEvent(c->stream, g->evB, 0));
    EFE_HIP(hipStreamSynchronize(st));
    info->iterations = it;
    float fms = 0, ims = 0;
    hipEventElapsedTime(&fms, g->t0, g->t1);
    hipEventElapsedTime(&ims, g->t1, g->t2);
    c->t_kry = it > 0 ? ims / it : 0.0;
    c->tot[EFE_T_KRYLOV] += ims;
    c->tot[EFE_T_KRYLOV_ITERS] += it;
    c->tot[EFE_T_SOLVES] += 1;
    c->tot[EFE_T_PRECOND] += fms;
    if (bad) info->status = EFE_NAN;
    else if (!done) info->status = EFE_NOT_CONVERGED;
    return EFE_OK;
}

int ensure_gmres(efe_ctx* c, int m, int prec) {
    GmresWork* g = c->gw;
    if (!g) {
        g = new GmresWork();
        c->gw = g;
        g->D = c->D;
        g->N = c->N;
        EFE_HIP(hipStreamCreateWithFlags(&g->ks, hipStreamNonBlocking));
        EFE_HIP(hipEventCreateWithFlags(&g->evA, hipEventDisableTiming));
        EFE_HIP(hipEventCreateWithFlags(&g->evB, hipEventDisableTiming));
        EFE_HIP(hipEventCreate(&g->t0));
        EFE_HIP(hipEventCreate(&g->t1));
        EFE_HIP(hipEventCreate(&g->t2));
        g->nwg = 512;
        EFE_HIP(hipHostMalloc(&g->hhost, 3 * 258 * sizeof(double), hipHostMallocDefault));
    }
    if (g->m != m || g->vprec != prec) {
        if (m > 256) {
            efe_set_error("GMRES restart length above 256");
            return EFE_ERR_ARG;
        }
        drop_step_graphs(g);
        hipFree(g->V); hipFree(g->w); hipFree(g->z); hipFree(g->x); hipFree(g->t);
        hipFree(g->part); hipFree(g->hdev);
        const size_t es = prec == EFE_FP32 ? sizeof(float) : sizeof(double);
        const size_t D = (size_t)c->D;
        EFE_HIP(hipMalloc(&g->V, (size_t)(m + 1) * D * es));
        EFE_HIP(hipMalloc(&g->w, D * es));
        EFE_HIP(hipMalloc(&g->z, D * es));
        EFE_HIP(hipMalloc(&g->x, D * es));
        EFE_HIP(hipMalloc(&g->t, D * es));
        EFE_HIP(hipMalloc(&g->part, (size_t)(m + 1) * g->nwg * sizeof(double)));
        EFE_HIP(hipMalloc(&g->hdev, 3 * (size_t)(m + 2) * sizeof(double)));
        g->m = m;
        g->vprec = prec;
    }
    return EFE_OK;
}

}  // namespace

int efe_gmres_solve(efe_ctx* c, const double* b, double* x, const efe_solve_opts* o, efe_solve_info* info) {
    const int m = o->restart > 0 ? o->restart : 50;
    int rc = ensure_gmres(c, m, o->precision);
    if (rc) return rc;
    if (o->precision == EFE_FP32) return gmres_solve<float>(c, c->gw, b, x, o, info, c->K32, c->Minv32, 1);
    return gmres_solve<double>(c, c->gw, b, x, o, info, c->K64, c->Minv64, 0);
}

void efe_gmres_free(efe_ctx* c) {
    GmresWork* g = c->gw;
    if (!g) return;
    void* bufs[] = {g->fptr, g->fcol, g->fidx, g->dslot, g->perm, g->F64, g->Dinv64, g->F32, g->Dinv32,
                    g->V, g->w, g->z, g->x, g->t, g->part, g->hdev};
    for (void* p : bufs)
        if (p) hipFree(p);
    drop_step_graphs(g);
    if (g->hhost) hipHostFree(g->hhost);
    if (g->ks) hipStreamDestroy(g->ks);
    if (g->evA) hipEventDestroy(g->evA);
    if (g->evB) hipEventDestroy(g->evB);
    for (hipEvent_t e : {g->t0, g->t1, g->t2})
        if (e) hipEventDestroy(e);
    delete g;
    c->gw = nullptr;
}
