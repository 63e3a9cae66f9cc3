// Krylov solvers for the assembled tangent on gfx950.
//
// Operator layout: sliced ELL of 3x3 blocks, slice height 64 (one wavefront).
// Thread (slice s, lane l) owns row slot p = 64 s + l, the row of node slotNode[p] (rows
// sorted by length inside windows of nodes); block slot k of that row
// stores its 9 values at Kval[slice_off[s]*9 + (9k+v)*64 + l], so every value
// load of a wavefront is one contiguous 512 B (FP64) / 256 B (FP32) segment and
// the block column index (1 int per 9 values) is shared by the three rows of a
// node.  Padding slots hold zeros and point at the row itself.
//
// PCG with 3x3 block-Jacobi, in place of the reference's direct solve
// (nonlinearimplicitstatic.py:727-751) on symmetric tangents.  One iteration is
// three kernels:
//   A: q = K p, per-workgroup partials of p.q
//   B: alpha = rz/pq; x += alpha p; r -= alpha q; z = Minv r; partials of r.z, r.r
//   C: beta = rz_new/rz; p = z + beta p; convergence test
// Global scalars are never stored by one workgroup for the others: every
// workgroup re-reduces the (few thousand) partials of the previous kernel in the
// same fixed order, so all workgroups hold bitwise identical scalars and the
// whole solve is deterministic.  A runs on the product's grid, B and C on a smaller
// grid of their own (their partials re-reduced by every workgroup cost less).  The two
// r.z partial buffers alternate with the iteration parity, so a batch of an even number
// of iterations is a fixed sequence of launches that is captured once into a hipGraph
// and replayed.
// With the hexa8 brick operator (matrix-free FP32), A is the brick product alone, its
// partials the bricks' shares of p.Kp (Dirichlet DOFs masked), and B gathers each node's
// brick partials itself (q is never stored): p.q = the bricks' sum + the sum of p^2 over the
// Dirichlet DOFs, whose partials C (and the init) form with p.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "efe_internal.h"

struct KrylovWork {
    int D = 0, N = 0, nwg = 0, block = 0;
    // rows per thread: a workgroup owns kr * block consecutive rows; [0] FP64, [1] FP32
    int krp[2] = {1, 1}, nwgp[2] = {0, 0};
    int kr = 1;  // of the precision being solved
    // the vector kernels (init, update, direction) on their own, smaller grid: krv rows per
    // thread, nwgv workgroups; [0] FP64, [1] FP32
    int krvp[2] = {1, 1}, nwgvp[2] = {0, 0};
    int krv = 1, nwgv = 0;
    int ps = 0;  // row stride of the partials array (the largest workgroup count)
    // FP64 and FP32 vector sets
    double *x64 = nullptr, *r64 = nullptr, *z64 = nullptr, *p64 = nullptr, *q64 = nullptr;
    float *x32 = nullptr, *r32 = nullptr, *z32 = nullptr, *p32 = nullptr, *q32 = nullptr;
    double* part = nullptr;  // [6][ps]: pq, rz0, rz1, rr, bb, p.p over the Dirichlet DOFs
    int* flags = nullptr;    // device: [done, nan, iters_local, pad]
    double* dscal = nullptr; // device: [tol2, scratch sum]
    int* hflags = nullptr;   // pinned host copy
    double* hpart = nullptr; // pinned host scalars
    hipStream_t ks = nullptr;
    hipEvent_t evA = nullptr, evB = nullptr, t0 = nullptr, t1 = nullptr;
    hipGraphExec_t graph[3] = {nullptr, nullptr, nullptr};  // FP64, FP32 assembled, FP32 matrix-free
    int graph_batch[3] = {0, 0, 0};
    int graph_block[3] = {0, 0, 0};
    // the operator buffers a batch graph was captured with (a graph keeps their addresses)
    const void* graph_ops[3][4] = {};
};

namespace {

// Workgroups of a launch are dealt round-robin to the 8 XCDs of the MI355X, each with
// its own L2.  Renumbering them so that every XCD owns one contiguous range of node
// rows keeps the x gathers of an XCD inside a slab of the mesh (plus halo), which
// fits its L2.  The grid size of these launches is a multiple of 8.
__device__ __forceinline__ int xcd_block() {
    const int per = gridDim.x >> 3;
    return (blockIdx.x & 7) * per + (blockIdx.x >> 3);
}

__host__ __forceinline__ int round8(int n) { return (n + 7) & ~7; }

// Deterministic workgroup sums.  Within a wavefront an xor butterfly: the two lanes of a
// pair add the same two values (a + b == b + a exactly), so after every step all lanes
// hold bitwise identical sums.  Across the BS/64 wavefronts the per-wave sums are added in
// wave order.  One barrier pair per call, however many values are summed together.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <int BS, int K>
__device__ __forceinline__ void block_sums(double (&v)[K], double* sh) {
    constexpr int NW = BS / 64;
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
    if (NW == 1) return;
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) sh[k * NW + w] = v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double r = sh[k * NW];
        for (int i = 1; i < NW; ++i) r += sh[k * NW + i];
        v[k] = r;
    }
    __syncthreads();
}

template <int BS>
__device__ __forceinline__ double block_sum(double v, double* sh) {
    double a[1] = {v};
    block_sums<BS, 1>(a, sh);
    return a[0];
}

// Sums of K partial arrays (each n long, at part + offs[k]) in a fixed order, computed
// identically by every workgroup.
template <int BS, int K>
__device__ __forceinline__ void reduce_partials_n(const double* __restrict__ part, const int (&offs)[K], int n,
                                                  double (&out)[K], double* sh) {
#pragma unroll
    for (int k = 0; k < K; ++k) out[k] = 0.0;
    for (int i = threadIdx.x; i < n; i += BS) {
#pragma unroll
        for (int k = 0; k < K; ++k) out[k] += part[offs[k] + i];
    }
    block_sums<BS, K>(out, sh);
}

// reduce_partials_n with a count per array; each array's sum is formed in the same order as
// reduce_partials_n would form it.
template <int BS, int K>
__device__ __forceinline__ void reduce_partials_nv(const double* __restrict__ part, const int (&offs)[K],
                                                   const int (&ns)[K], double (&out)[K], double* sh) {
    int nmax = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        out[k] = 0.0;
        nmax = ns[k] > nmax ? ns[k] : nmax;
    }
    for (int i = threadIdx.x; i < nmax; i += BS) {
#pragma unroll
        for (int k = 0; k < K; ++k)
            if (i < ns[k]) out[k] += part[offs[k] + i];
    }
    block_sums<BS, K>(out, sh);
}

template <int BS>
__device__ __forceinline__ double reduce_partials(const double* __restrict__ part, int n, double* sh) {
    const int offs[1] = {0};
    double out[1];
    reduce_partials_n<BS, 1>(part, offs, n, out, sh);
    return out[0];
}

template <typename T, int BS>
__global__ __launch_bounds__(BS) void spmv_kernel(int N, const int32_t* __restrict__ slice_off,
                                                  const int32_t* __restrict__ scol,
                                                  const int32_t* __restrict__ slotNode, const T* __restrict__ K,
                                                  const T* __restrict__ x, T* __restrict__ y,
                                                  double* __restrict__ part, const int* __restrict__ flags,
                                                  int kr) {
    __shared__ double sh[BS];
    if (flags && flags[0]) return;
    const int blk = xcd_block();
    double dot = 0.0;
    for (int it = 0; it < kr; ++it) {
        const int ps = (blk * kr + it) * BS + threadIdx.x;  // row slot
        if (ps >= N) break;
        const int n = slotNode[ps];
        const int s = ps / kSlice, l = ps - (ps / kSlice) * kSlice;
        const int o0 = slice_off[s], deg = (slice_off[s + 1] - o0) / kSlice;
        const int32_t* cp = scol + o0 + l;
        const T* kp = K + (int64_t)o0 * 9 + l;
        T y0 = 0, y1 = 0, y2 = 0;
        // matrix values and indices are streamed once per product: non-temporal loads
        // keep them from displacing the gathered x in L2
#pragma unroll 3
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
        y[3 * n] = y0;
        y[3 * n + 1] = y1;
        y[3 * n + 2] = y2;
        if (part) dot += (double)x[3 * n] * y0 + (double)x[3 * n + 1] * y1 + (double)x[3 * n + 2] * y2;
    }
    if (part) {
        double s = block_sum<BS>(dot, sh);
        if (threadIdx.x == 0) part[blk] = s;
    }
}

// x = 0, r = b (b_dir = b on Dirichlet DOFs is part of the system), z = Minv r, p = z,
// partials r.z -> part[1], b.b -> part[4].
template <typename T, int BS, bool DP2 = false>
__global__ __launch_bounds__(BS) void pcg_init_kernel(int N, const double* __restrict__ b, const T* __restrict__ Minv,
                                                      T* __restrict__ x, T* __restrict__ r, T* __restrict__ z,
                                                      T* __restrict__ p, double* __restrict__ part, int ps,
                                                      int* __restrict__ flags, int kr,
                                                      const uint8_t* __restrict__ dirflag = nullptr) {
    __shared__ double sh[BS];
    const int blk = xcd_block();
    double rz = 0, bb = 0, pd = 0;
    for (int it = 0; it < kr; ++it) {
        const int n = (blk * kr + it) * BS + threadIdx.x;
        if (n >= N) break;
        T rv[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            rv[i] = (T)b[3 * n + i];
            x[3 * n + i] = 0;
            r[3 * n + i] = rv[i];
            bb += (double)b[3 * n + i] * b[3 * n + i];
        }
        const T* M = Minv + (int64_t)n * 9;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            T zi = M[3 * i] * rv[0] + M[3 * i + 1] * rv[1] + M[3 * i + 2] * rv[2];
            z[3 * n + i] = zi;
            p[3 * n + i] = zi;
            rz += (double)rv[i] * zi;
            if (DP2 && dirflag[3 * n + i]) pd += (double)zi * zi;
        }
    }
    constexpr int NS = DP2 ? 3 : 2;
    double sums[NS];
    sums[0] = rz;
    sums[1] = bb;
    if constexpr (DP2) sums[2] = pd;
    block_sums<BS, NS>(sums, sh);
    const double s1 = sums[0], s2 = sums[1];
    if (threadIdx.x == 0) {
        part[1 * ps + blk] = s1;
        part[4 * ps + blk] = s2;
        if constexpr (DP2) part[5 * ps + blk] = sums[NS - 1];
        if (blk == 0) {
            flags[0] = 0;
            flags[1] = 0;
            flags[2] = 0;
        }
    }
}

template <typename T, int BS>
__global__ __launch_bounds__(BS) void pcg_update_kernel(int N, const T* __restrict__ Minv, T* __restrict__ x,
                                                        T* __restrict__ r, T* __restrict__ z,
                                                        const T* __restrict__ p, const T* __restrict__ q,
                                                        double* __restrict__ part, int ps, int nop, int nvec,
                                                        int par, const int* __restrict__ flags, int kr) {
    __shared__ double sh[BS];
    if (flags[0]) return;
    // p.q: nop partials of the operator's grid; r.z: nvec partials of this grid
    const int offs[2] = {0, (1 + par) * ps};
    const int ns[2] = {nop, nvec};
    double sc[2];
    reduce_partials_nv<BS, 2>(part, offs, ns, sc, sh);
    const double pq = sc[0], rz = sc[1];
    const T alpha = (T)(rz / pq);
    const int blk = xcd_block();
    double rzn = 0, rr = 0;
    for (int it = 0; it < kr; ++it) {
        const int n = (blk * kr + it) * BS + threadIdx.x;
        if (n >= N) break;
        T rv[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            x[3 * n + i] += alpha * p[3 * n + i];
            rv[i] = r[3 * n + i] - alpha * q[3 * n + i];
            r[3 * n + i] = rv[i];
            rr += (double)rv[i] * rv[i];
        }
        const T* M = Minv + (int64_t)n * 9;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            T zi = M[3 * i] * rv[0] + M[3 * i + 1] * rv[1] + M[3 * i + 2] * rv[2];
            z[3 * n + i] = zi;
            rzn += (double)rv[i] * zi;
        }
    }
    double sums[2] = {rzn, rr};
    block_sums<BS, 2>(sums, sh);
    const double s1 = sums[0], s2 = sums[1];
    if (threadIdx.x == 0) {
        part[(2 - par) * ps + blk] = s1;
        part[3 * ps + blk] = s2;
    }
}

template <typename T, int BS, bool DP2 = false>
__global__ __launch_bounds__(BS) void pcg_direction_kernel(int N, const T* __restrict__ z, T* __restrict__ p,
                                                           double* __restrict__ part, int ps, int nvec,
                                                           int par, int jlocal, const double* __restrict__ dtol2,
                                                           int* __restrict__ flags, int kr,
                                                           const uint8_t* __restrict__ dirflag = nullptr) {
    __shared__ double sh[BS];
    if (flags[0]) return;
    const int offs[3] = {(1 + par) * ps, (2 - par) * ps, 3 * ps};
    double sc[3];
    reduce_partials_n<BS, 3>(part, offs, nvec, sc, sh);
    const double rzo = sc[0], rzn = sc[1], rr = sc[2];
    const bool bad = !(rr == rr) || !(rzn == rzn);
    const bool done = bad || rr <= dtol2[0];
    if (done) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            flags[2] = jlocal + 1;
            flags[1] = bad ? 1 : 0;
            flags[0] = 1;
        }
        return;
    }
    const T beta = (T)(rzn / rzo);
    const int blk = xcd_block();
    double pd = 0;
    for (int it = 0; it < kr; ++it) {
        const int n = (blk * kr + it) * BS + threadIdx.x;
        if (n >= N) break;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const T pv = z[3 * n + i] + beta * p[3 * n + i];
            p[3 * n + i] = pv;
            if (DP2 && dirflag[3 * n + i]) pd += (double)pv * pv;
        }
    }
    if (DP2) {
        const double s = block_sum<BS>(pd, sh);
        if (threadIdx.x == 0) part[5 * ps + blk] = s;
    }
}

// B for the hexa8 brick operator (FP32): p.q = the bricks' partials (nop of them) + the
// Dirichlet p.p partials; q of node n = the sum of its brick partials (in brick order), p on
// its Dirichlet DOFs, formed here and not stored.
template <int BS>
__global__ __launch_bounds__(BS) void pcg_update_mfb_kernel(int N, const float* __restrict__ Minv,
                                                            float* __restrict__ x, float* __restrict__ r,
                                                            float* __restrict__ z, const float* __restrict__ p,
                                                            const int32_t* __restrict__ gptr,
                                                            const float* __restrict__ gF,
                                                            const uint8_t* __restrict__ dirflag,
                                                            double* __restrict__ part, int ps, int nop, int nvec,
                                                            int par, const int* __restrict__ flags, int kr) {
    __shared__ double sh[BS];
    if (flags[0]) return;
    const int offs[3] = {0, 5 * ps, (1 + par) * ps};
    const int ns[3] = {nop, nvec, nvec};
    double sc[3];
    reduce_partials_nv<BS, 3>(part, offs, ns, sc, sh);
    const double pq = sc[0] + sc[1], rz = sc[2];
    const float alpha = (float)(rz / pq);
    const int blk = xcd_block();
    double rzn = 0, rr = 0;
    for (int it = 0; it < kr; ++it) {
        const int n = (blk * kr + it) * BS + threadIdx.x;
        if (n >= N) break;
        float qv[3] = {0.0f, 0.0f, 0.0f};
        const int g0 = gptr[n], g1 = gptr[n + 1];
        for (int c = g0; c < g1; ++c) {
            qv[0] += gF[3 * (int64_t)c];
            qv[1] += gF[3 * (int64_t)c + 1];
            qv[2] += gF[3 * (int64_t)c + 2];
        }
        float rv[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const float pv = p[3 * n + i];
            const float q = dirflag[3 * n + i] ? pv : qv[i];
            x[3 * n + i] += alpha * pv;
            rv[i] = r[3 * n + i] - alpha * q;
            r[3 * n + i] = rv[i];
            rr += (double)rv[i] * rv[i];
        }
        const float* M = Minv + (int64_t)n * 9;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const float zi = M[3 * i] * rv[0] + M[3 * i + 1] * rv[1] + M[3 * i + 2] * rv[2];
            z[3 * n + i] = zi;
            rzn += (double)rv[i] * zi;
        }
    }
    double sums[2] = {rzn, rr};
    block_sums<BS, 2>(sums, sh);
    if (threadIdx.x == 0) {
        part[(2 - par) * ps + blk] = sums[0];
        part[3 * ps + blk] = sums[1];
    }
}

// x_out = x (converted), with Dirichlet DOFs taken from b exactly.
template <typename T>
__global__ void pcg_finish_kernel(int D, const T* __restrict__ x, const double* __restrict__ b,
                                  const uint8_t* __restrict__ dirflag, int ndir, double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D) return;
    out[i] = (ndir && dirflag[i]) ? b[i] : (double)x[i];
}

// Convergence quantities: max|R| (first index on ties), max|ddU|, sum|F|, NaN flag.
template <int BS>
__global__ __launch_bounds__(BS) void norms_stage1(int D, const double* __restrict__ R, const double* __restrict__ ddU,
                                                   const double* __restrict__ F, double* __restrict__ part,
                                                   int nwg) {
    __shared__ double sv[BS];
    __shared__ long long si[BS];
    __shared__ double sh[BS];
    double mr = -1.0, md = 0.0, sf = 0.0, nan = 0.0;
    long long ir = 0x7fffffffffffffffLL;
    for (int i = blockIdx.x * BS + threadIdx.x; i < D; i += BS * gridDim.x) {
        const double a = fabs(R[i]);
        if (!(a == a)) nan = 1.0;
        if (a > mr) {
            mr = a;
            ir = i;
        }
        if (ddU) {
            const double d = fabs(ddU[i]);
            if (!(d == d)) nan = 1.0;
            md = d > md ? d : md;
        }
        sf += fabs(F[i]);
    }
    sv[threadIdx.x] = mr;
    si[threadIdx.x] = ir;
    __syncthreads();
    for (int s = BS / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            const double o = sv[threadIdx.x + s];
            const long long oi = si[threadIdx.x + s];
            if (o > sv[threadIdx.x] || (o == sv[threadIdx.x] && oi < si[threadIdx.x])) {
                sv[threadIdx.x] = o;
                si[threadIdx.x] = oi;
            }
        }
        __syncthreads();
    }
    const double smax = sv[0];
    const long long simax = si[0];
    __syncthreads();
    sh[threadIdx.x] = md;
    __syncthreads();
    for (int s = BS / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    const double dmax = sh[0];
    __syncthreads();
    const double fsum = block_sum<BS>(sf, sh);
    const double nn = block_sum<BS>(nan, sh);
    if (threadIdx.x == 0) {
        part[0 * nwg + blockIdx.x] = smax;
        part[1 * nwg + blockIdx.x] = (double)simax;
        part[2 * nwg + blockIdx.x] = dmax;
        part[3 * nwg + blockIdx.x] = fsum;
        part[4 * nwg + blockIdx.x] = nn;
    }
}

// Fixed-order sum of n partials by one 256-thread workgroup.
__global__ __launch_bounds__(256) void sum_partials_kernel(const double* __restrict__ part, int n,
                                                           double* __restrict__ out) {
    __shared__ double sh[256];
    const double s = reduce_partials<256>(part, n, sh);
    if (threadIdx.x == 0) *out = s;
}

// Final reduction of the per-workgroup norms by one 256-thread workgroup (fixed order).
__global__ __launch_bounds__(256) void norms_stage2(const double* __restrict__ part, int nwg,
                                                    double* __restrict__ out) {
    __shared__ double sv[256];
    __shared__ double si[256];
    __shared__ double sh[256];
    double mr = -1.0, ir = 1e300, md = 0;
    for (int w = threadIdx.x; w < nwg; w += 256) {
        const double v = part[w];
        if (v > mr || (v == mr && part[nwg + w] < ir)) {
            mr = v;
            ir = part[nwg + w];
        }
        md = fmax(md, part[2 * nwg + w]);
    }
    sv[threadIdx.x] = mr;
    si[threadIdx.x] = ir;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            const double o = sv[threadIdx.x + s], oi = si[threadIdx.x + s];
            if (o > sv[threadIdx.x] || (o == sv[threadIdx.x] && oi < si[threadIdx.x])) {
                sv[threadIdx.x] = o;
                si[threadIdx.x] = oi;
            }
        }
        __syncthreads();
    }
    const double rmax = sv[0], rarg = si[0];
    __syncthreads();
    sh[threadIdx.x] = md;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    const double dmax = sh[0];
    __syncthreads();
    const double sf = reduce_partials<256>(part + 3 * nwg, nwg, sh);
    const double nn = reduce_partials<256>(part + 4 * nwg, nwg, sh);
    if (threadIdx.x == 0) {
        out[0] = rmax < 0 ? 0.0 : rmax;
        out[1] = rmax < 0 ? 0.0 : rarg;
        out[2] = dmax;
        out[3] = sf;
        out[4] = nn > 0 ? 1.0 : 0.0;
    }
}

// Per-specimen norms (batched contexts): one workgroup per specimen over its DOF range.
template <int BS>
__global__ __launch_bounds__(BS) void norms_batched_kernel(const int64_t* __restrict__ dofoff,
                                                           const double* __restrict__ R,
                                                           const double* __restrict__ ddU,
                                                           const double* __restrict__ F, double* __restrict__ out) {
    __shared__ double sh[BS];
    const int sp = blockIdx.x;
    double mr = 0, md = 0, sf = 0, nan = 0;
    for (int64_t i = dofoff[sp] + threadIdx.x; i < dofoff[sp + 1]; i += BS) {
        const double a = fabs(R[i]);
        if (!(a == a)) nan = 1.0;
        mr = fmax(mr, a);
        if (ddU) {
            const double d = fabs(ddU[i]);
            if (!(d == d)) nan = 1.0;
            md = fmax(md, d);
        }
        sf += fabs(F[i]);
    }
    sh[threadIdx.x] = mr;
    __syncthreads();
    for (int s = BS / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    const double rmax = sh[0];
    __syncthreads();
    sh[threadIdx.x] = md;
    __syncthreads();
    for (int s = BS / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    const double dmax = sh[0];
    __syncthreads();
    const double fs = block_sum<BS>(sf, sh);
    const double nn = block_sum<BS>(nan, sh);
    if (threadIdx.x == 0) {
        out[sp * 4 + 0] = rmax;
        out[sp * 4 + 1] = dmax;
        out[sp * 4 + 2] = fs;
        out[sp * 4 + 3] = nn > 0 ? 1.0 : 0.0;
    }
}

template <typename T, int BS>
void launch_iter(efe_ctx* c, KrylovWork* w, hipStream_t st, T* x, T* r, T* z, T* p, T* q, const T* K,
                 const T* Minv, int par, int j) {
    const int grid = w->nwg;
    const int kr = w->kr;
    const int gv = w->nwgv, krv = w->krv;
    if constexpr (sizeof(T) == 4) {
        if (!K && c->mfb) {  // hexa8 bricks: the product's partials gathered by the update
            efe_mf_launch_apply_partials(c, st, p, w->part, w->flags);
            pcg_update_mfb_kernel<BS><<<gv, BS, 0, st>>>(c->N, Minv, x, r, z, p, c->mfbPptr, c->mfbF, c->dirflag,
                                                         w->part, w->ps, round8(c->mfbG), gv, par, w->flags, krv);
            pcg_direction_kernel<T, BS, true><<<gv, BS, 0, st>>>(c->N, z, p, w->part, w->ps, gv, par, j, w->dscal,
                                                                 w->flags, krv, c->dirflag);
            return;
        }
    }
    if (K)
        spmv_kernel<T, BS><<<grid, BS, 0, st>>>(c->N, c->slice_off, c->scol, c->slotNode, K, p, q, w->part, w->flags,
                                                kr);
    else  // matrix-free FP32 operator
        efe_mf_launch_operator(c, st, (const float*)p, (float*)q, w->part, w->flags, grid, BS, kr);
    pcg_update_kernel<T, BS><<<gv, BS, 0, st>>>(c->N, Minv, x, r, z, p, q, w->part, w->ps, grid, gv, par, w->flags,
                                                krv);
    pcg_direction_kernel<T, BS><<<gv, BS, 0, st>>>(c->N, z, p, w->part, w->ps, gv, par, j, w->dscal, w->flags, krv);
}

template <typename T>
void launch_iter_bs(efe_ctx* c, KrylovWork* w, hipStream_t st, T* x, T* r, T* z, T* p, T* q, const T* K,
                    const T* Minv, int par, int j) {
    switch (w->block) {
        case 64: launch_iter<T, 64>(c, w, st, x, r, z, p, q, K, Minv, par, j); break;
        case 128: launch_iter<T, 128>(c, w, st, x, r, z, p, q, K, Minv, par, j); break;
        case 512: launch_iter<T, 512>(c, w, st, x, r, z, p, q, K, Minv, par, j); break;
        default: launch_iter<T, 256>(c, w, st, x, r, z, p, q, K, Minv, par, j); break;
    }
}

template <typename T>
void launch_init_bs(efe_ctx* c, KrylovWork* w, hipStream_t st, const double* b, const T* Minv, T* x, T* r, T* z,
                    T* p, bool dp2) {
    if (dp2) {  // the brick operator's Dirichlet p.p partials
        switch (w->block) {
            case 64: pcg_init_kernel<T, 64, true><<<w->nwgv, 64, 0, st>>>(c->N, b, Minv, x, r, z, p, w->part, w->ps, w->flags, w->krv, c->dirflag); break;
            case 128: pcg_init_kernel<T, 128, true><<<w->nwgv, 128, 0, st>>>(c->N, b, Minv, x, r, z, p, w->part, w->ps, w->flags, w->krv, c->dirflag); break;
            case 512: pcg_init_kernel<T, 512, true><<<w->nwgv, 512, 0, st>>>(c->N, b, Minv, x, r, z, p, w->part, w->ps, w->flags, w->krv, c->dirflag); break;
            default: pcg_init_kernel<T, 256, true><<<w->nwgv, 256, 0, st>>>(c->N, b, Minv, x, r, z, p, w->part, w->ps, w->flags, w->krv, c->dirflag); break;
        }
        return;
    }
    switch (w->block) {
        case 64: pcg_init_kernel<T, 64><<<w->nwgv, 64, 0, st>>>(c->N, b, Minv, x, r, z, p, w->part, w->ps, w->flags, w->krv); break;
        case 128: pcg_init_kernel<T, 128><<<w->nwgv, 128, 0, st>>>(c->N, b, Minv, x, r, z, p, w->part, w->ps, w->flags, w->krv); break;
        case 512: pcg_init_kernel<T, 512><<<w->nwgv, 512, 0, st>>>(c->N, b, Minv, x, r, z, p, w->part, w->ps, w->flags, w->krv); break;
        default: pcg_init_kernel<T, 256><<<w->nwgv, 256, 0, st>>>(c->N, b, Minv, x, r, z, p, w->part, w->ps, w->flags, w->krv); break;
    }
}

int ensure_work(efe_ctx* c) {
    int block = c->kblock;
    if (block != 64 && block != 128 && block != 256 && block != 512) block = 256;
    KrylovWork* w = c->kw;
    // the brick operator writes one p.Kp partial per brick into the partials array
    const int need = c->mfb ? round8(c->mfbG) : 0;
    if (w && w->block == block && w->ps >= need) return EFE_OK;
    if (!w) {
        w = new KrylovWork();
        c->kw = w;
        w->D = c->D;
        w->N = c->N;
        size_t D = (size_t)c->D;
        EFE_HIP(hipMalloc(&w->x64, D * sizeof(double)));
        EFE_HIP(hipMalloc(&w->r64, D * sizeof(double)));
        EFE_HIP(hipMalloc(&w->z64, D * sizeof(double)));
        EFE_HIP(hipMalloc(&w->p64, D * sizeof(double)));
        EFE_HIP(hipMalloc(&w->q64, D * sizeof(double)));
        EFE_HIP(hipMalloc(&w->x32, D * sizeof(float)));
        EFE_HIP(hipMalloc(&w->r32, D * sizeof(float)));
        EFE_HIP(hipMalloc(&w->z32, D * sizeof(float)));
        EFE_HIP(hipMalloc(&w->p32, D * sizeof(float)));
        EFE_HIP(hipMalloc(&w->q32, D * sizeof(float)));
        EFE_HIP(hipMalloc(&w->flags, 4 * sizeof(int)));
        EFE_HIP(hipMalloc(&w->dscal, 4 * sizeof(double)));
        EFE_HIP(hipHostMalloc(&w->hflags, 4 * sizeof(int), hipHostMallocDefault));
        EFE_HIP(hipHostMalloc(&w->hpart, 8 * sizeof(double), hipHostMallocDefault));
        EFE_HIP(hipStreamCreateWithFlags(&w->ks, hipStreamNonBlocking));
        EFE_HIP(hipEventCreateWithFlags(&w->evA, hipEventDisableTiming));
        EFE_HIP(hipEventCreateWithFlags(&w->evB, hipEventDisableTiming));
        EFE_HIP(hipEventCreate(&w->t0));
        EFE_HIP(hipEventCreate(&w->t1));
    }
    for (int i = 0; i < 3; ++i) {
        if (w->graph[i]) hipGraphExecDestroy(w->graph[i]);
        w->graph[i] = nullptr;
    }
    if (w->part) hipFree(w->part);
    w->block = block;
    // bounded workgroup counts: every workgroup re-reduces all partials in a fixed order,
    // so their number must stay bounded at any mesh size.  Measured on MI355X: FP64 is
    // fastest with at most ~2048 workgroups (fewer starve its SpMV of waves), the FP32
    // operators with at most ~1024; uncapped, the reductions dominate at 96^3.
    const int groups = (c->N + block - 1) / block;
    const int cap[2] = {2048, 1024};
    // The vector kernels (init, update, direction) stream less per row than the product and
    // re-reduce every partial of the previous kernel in every workgroup, so they take at most
    // 512 workgroups of several rows per thread.  Measured at 64^3 (history Krylov iteration,
    // FP64 / mixed): the product's grid 125.6 / 48.6 us, at most 512 123.7 / 47.4, 256 123.1 /
    // 47.4, 128 125.9 / 49.7.
    const int vcap = 512;
    w->ps = 0;
    for (int i = 0; i < 2; ++i) {
        w->krp[i] = (groups + cap[i] - 1) / cap[i];
        w->nwgp[i] = round8((c->N + block * w->krp[i] - 1) / (block * w->krp[i]));
        w->krvp[i] = (groups + vcap - 1) / vcap;
        w->nwgvp[i] = round8((c->N + block * w->krvp[i] - 1) / (block * w->krvp[i]));
        w->ps = std::max(w->ps, std::max(w->nwgp[i], w->nwgvp[i]));
    }
    w->ps = std::max(w->ps, need);
    w->kr = w->krp[0];
    w->nwg = w->nwgp[0];
    w->krv = w->krvp[0];
    w->nwgv = w->nwgvp[0];
    EFE_HIP(hipMalloc(&w->part, (size_t)6 * w->ps * sizeof(double)));
    return EFE_OK;
}

template <typename T>
int pcg_solve(efe_ctx* c, const double* b, double* xout, const efe_solve_opts* o, efe_solve_info* info, T* x, T* r,
              T* z, T* p, T* q, const T* K, const T* Minv, int gi) {
    KrylovWork* w = c->kw;
    hipStream_t st = w->ks;
    w->kr = w->krp[sizeof(T) == 4 ? 1 : 0];
    w->nwg = w->nwgp[sizeof(T) == 4 ? 1 : 0];
    w->krv = w->krvp[sizeof(T) == 4 ? 1 : 0];
    w->nwgv = w->nwgvp[sizeof(T) == 4 ? 1 : 0];
    // order after the caller's stream
    EFE_HIP(hipEventRecord(w->evA, c->stream));
    EFE_HIP(hipStreamWaitEvent(st, w->evA, 0));
    launch_init_bs<T>(c, w, st, b, Minv, x, r, z, p, !K && c->mfb);
    EFE_HIP(hipGetLastError());
    // b.b (host) to fix the absolute tolerance
    sum_partials_kernel<<<1, 256, 0, st>>>(w->part + 4 * w->ps, w->nwgv, w->dscal + 1);
    EFE_HIP(hipMemcpyAsync(w->hpart, w->dscal + 1, sizeof(double), hipMemcpyDeviceToHost, st));
    EFE_HIP(hipStreamSynchronize(st));
    const double bnorm = sqrt(w->hpart[0]);
    info->rhs_norm = bnorm;
    info->iterations = 0;
    info->status = EFE_OK;
    info->residual = bnorm;
    const int D = c->D;
    if (bnorm == 0.0 || !(bnorm == bnorm)) {
        pcg_finish_kernel<T><<<(D + 255) / 256, 256, 0, st>>>(D, x, b, c->dirflag, c->ndir, xout);
        EFE_HIP(hipEventRecord(w->evB, st));
        EFE_HIP(hipStreamWaitEvent(c->stream, w->evB, 0));
        if (!(bnorm == bnorm)) info->status = EFE_NAN;
        return EFE_OK;
    }
    const double tol = fmax(o->rtol * bnorm, o->atol);
    const double tol2 = tol * tol;
    int batch = o->check_every > 0 ? o->check_every : 16;
    if (batch & 1) batch += 1;
    const int maxit = o->max_iter > 0 ? o->max_iter : 10000;
    w->hpart[1] = tol2;
    EFE_HIP(hipMemcpyAsync(w->dscal, w->hpart + 1, sizeof(double), hipMemcpyHostToDevice, st));
    // (re)build the captured batch when its shape or an operator buffer changed
    const void* ops[4] = {K, Minv, c->mfC, c->mfbF};
    bool same = true;
    for (int i = 0; i < 4; ++i) same = same && w->graph_ops[gi][i] == ops[i];
    if (!w->graph[gi] || w->graph_batch[gi] != batch || w->graph_block[gi] != w->block || !same) {
        if (w->graph[gi]) hipGraphExecDestroy(w->graph[gi]);
        w->graph[gi] = nullptr;
        hipGraph_t g;
        EFE_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
        for (int j = 0; j < batch; ++j) launch_iter_bs<T>(c, w, st, x, r, z, p, q, K, Minv, j & 1, j);
        EFE_HIP(hipStreamEndCapture(st, &g));
        EFE_HIP(hipGraphInstantiate(&w->graph[gi], g, nullptr, nullptr, 0));
        EFE_HIP(hipGraphDestroy(g));
        w->graph_batch[gi] = batch;
        w->graph_block[gi] = w->block;
        for (int i = 0; i < 4; ++i) w->graph_ops[gi][i] = ops[i];
    }
    int it = 0;
    EFE_HIP(hipEventRecord(w->t0, st));
    while (true) {
        EFE_HIP(hipGraphLaunch(w->graph[gi], st));
        EFE_HIP(hipMemcpyAsync(w->hflags, w->flags, 3 * sizeof(int), hipMemcpyDeviceToHost, st));
        EFE_HIP(hipStreamSynchronize(st));
        if (w->hflags[0]) {
            it += w->hflags[2];
            if (w->hflags[1]) info->status = EFE_NAN;
            break;
        }
        it += batch;
        if (it >= maxit) {
            info->status = EFE_NOT_CONVERGED;
            break;
        }
    }
    EFE_HIP(hipEventRecord(w->t1, st));
    pcg_finish_kernel<T><<<(D + 255) / 256, 256, 0, st>>>(D, x, b, c->dirflag, c->ndir, xout);
    EFE_HIP(hipGetLastError());
    EFE_HIP(hipEventRecord(w->evB, st));
    EFE_HIP(hipStreamWaitEvent(c->stream, w->evB, 0));
    // final residual of the recursion
    sum_partials_kernel<<<1, 256, 0, st>>>(w->part + 3 * w->ps, w->nwgv, w->dscal + 1);
    EFE_HIP(hipMemcpyAsync(w->hpart, w->dscal + 1, sizeof(double), hipMemcpyDeviceToHost, st));
    EFE_HIP(hipStreamSynchronize(st));
    info->residual = sqrt(w->hpart[0]);
    info->iterations = it;
    float ms = 0;
    hipEventElapsedTime(&ms, w->t0, w->t1);
    c->t_kry = it > 0 ? ms / it : 0.0;
    c->tot[EFE_T_KRYLOV] += ms;
    c->tot[EFE_T_KRYLOV_ITERS] += it;
    c->tot[EFE_T_SOLVES] += 1;
    return EFE_OK;
}

}  // namespace

int efe_krylov_solve(efe_ctx* c, const double* b, double* x, const efe_solve_opts* o, efe_solve_info* info) {
    if (c->Kprec < 0) {
        efe_set_error("no assembled operator: call efe_compute_elements first");
        return EFE_ERR_STATE;
    }
    if (o->precision != c->Kprec) {
        efe_set_error("solve precision differs from the precision K was assembled in");
        return EFE_ERR_ARG;
    }
    if (o->precision == EFE_FP32_MATFREE && (o->solver != EFE_SOLVER_CG || o->precond != EFE_PRECOND_BLOCK_JACOBI)) {
        efe_set_error("the matrix-free operator runs with block-Jacobi CG");
        return EFE_ERR_ARG;
    }
    if (o->solver == EFE_SOLVER_GMRES) return efe_gmres_solve(c, b, x, o, info);
    if (o->solver != EFE_SOLVER_CG || o->precond != EFE_PRECOND_BLOCK_JACOBI) {
        efe_set_error("CG runs with the block-Jacobi preconditioner (ILU(0) is available with GMRES)");
        return EFE_ERR_ARG;
    }
    int rc = ensure_work(c);
    if (rc != EFE_OK) return rc;
    KrylovWork* w = c->kw;
    if (o->precision == EFE_FP32_MATFREE)
        return pcg_solve<float>(c, b, x, o, info, w->x32, w->r32, w->z32, w->p32, w->q32, nullptr, c->Minv32, 2);
    if (o->precision == EFE_FP32)
        return pcg_solve<float>(c, b, x, o, info, w->x32, w->r32, w->z32, w->p32, w->q32, c->K32, c->Minv32, 1);
    return pcg_solve<double>(c, b, x, o, info, w->x64, w->r64, w->z64, w->p64, w->q64, c->K64, c->Minv64, 0);
}

int efe_launch_spmv64(efe_ctx* c, const double* x, double* y) {
    if (c->Kprec != EFE_FP64) {
        efe_set_error("efe_spmv needs an FP64 operator");
        return EFE_ERR_STATE;
    }
    spmv_kernel<double, 256><<<round8((c->N + 255) / 256), 256, 0, c->stream>>>(
        c->N, c->slice_off, c->scol, c->slotNode, c->K64, x, y, nullptr, nullptr, 1);
    EFE_HIP(hipGetLastError());
    return EFE_OK;
}

int efe_launch_norms(efe_ctx* c, const double* R, const double* ddU, const double* F, double* out) {
    const int BS = 256;
    int nwg = (c->D + BS - 1) / BS;
    if (nwg > 1024) nwg = 1024;
    norms_stage1<BS><<<nwg, BS, 0, c->stream>>>(c->D, R, ddU, F, c->red + 8, nwg);
    norms_stage2<<<1, 256, 0, c->stream>>>(c->red + 8, nwg, c->red);
    EFE_HIP(hipGetLastError());
    EFE_HIP(hipMemcpyAsync(out, c->red, 5 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    EFE_HIP(hipStreamSynchronize(c->stream));
    return EFE_OK;
}


void efe_krylov_free(efe_ctx* c) {
    KrylovWork* w = c->kw;
    if (!w) return;
    for (int i = 0; i < 3; ++i)
        if (w->graph[i]) hipGraphExecDestroy(w->graph[i]);
    hipFree(w->x64); hipFree(w->r64); hipFree(w->z64); hipFree(w->p64); hipFree(w->q64);
    hipFree(w->x32); hipFree(w->r32); hipFree(w->z32); hipFree(w->p32); hipFree(w->q32);
    hipFree(w->part); hipFree(w->flags); hipFree(w->dscal);
    hipHostFree(w->hflags); hipHostFree(w->hpart);
    if (w->ks) hipStreamDestroy(w->ks);
    if (w->evA) hipEventDestroy(w->evA);
    if (w->evB) hipEventDestroy(w->evB);
    if (w->t0) hipEventDestroy(w->t0);
    if (w->t1) hipEventDestroy(w->t1);
    delete w;
    c->kw = nullptr;
}
