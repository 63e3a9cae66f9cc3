// Batched independent specimens: one block-diagonal operator, one element launch, and a
// PCG whose scalars, tolerance and convergence are per specimen.
//
// Specimens are contiguous node ranges.  Per-specimen reductions are deterministic and
// size independent.  Inside the PCG iteration each specimen's node range is cut into
// 256-node chunks that start at the specimen's first node; a workgroup sums one chunk in a
// fixed order and one wavefront sums a specimen's chunk partials in a fixed order, so the
// order depends only on positions relative to the specimen.  The set-up reductions sum
// per-node values over the whole range in one workgroup per specimen.
// A converged specimen is frozen (its rows are skipped), so each specimen follows exactly
// the iteration of a PCG run on its own system.
#include <cmath>
#include <vector>

#include "efe_internal.h"

struct BatchWork {
    int N = 0, S = 0, prec = -1;
    void *x = nullptr, *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr;
    double *nv0 = nullptr, *nv1 = nullptr;     // per-node values
    double* sc = nullptr;                      // per specimen: [0] rz, [1] rtol2, [2] rr, [3] tol2, [4] beta
    int *active = nullptr, *iters = nullptr;   // per specimen
    int* nactive = nullptr;                    // device count
    int* acc = nullptr;                        // per specimen accept flags
    double* nout = nullptr;                    // per specimen norms [S][4]
    int C = 0;                                 // 256-node chunks over all specimens
    int32_t *chunkNode0 = nullptr, *chunkSpec = nullptr;  // [C] first node, specimen
    int32_t* specChunk = nullptr;              // [S+1] first chunk of each specimen
    double *pqp = nullptr, *rzp = nullptr, *rrp = nullptr;  // [C] chunk partials of p.q, r.z, r.r
    int* hbuf = nullptr;                       // pinned
    hipStream_t ks = nullptr;
    hipEvent_t evA = nullptr, evB = nullptr;
    hipGraphExec_t graph = nullptr;
    int graph_batch = 0;
    int graph_op = -1;  // precision code the captured graph was built for (operator choice)
    // device-resident Newton rounds (efe_newton_round_batched): per specimen its Newton state
    // (ITERATING 0, CONVERGED 1, FAILED 2, IDLE 3), iteration counter, residual growth count and
    // previous flux residual, failure reason, first-iteration flag, limits, and work counters;
    // per DOF the prescribed values of first iterations, the right-hand-side correction, R - corr
    // and the Newton correction
    int *nst = nullptr, *nit = nullptr, *ngrow = nullptr, *nreason = nullptr, *nfirst = nullptr;
    int *nextrap = nullptr, *nhasdir = nullptr, *nmaxit = nullptr, *nmaxgrow = nullptr, *nsolve = nullptr;
    int *npass = nullptr, *nsolves = nullptr, *nkry = nullptr, *ncount = nullptr;
    double *ngprev = nullptr, *nacc = nullptr;
    double *dirv = nullptr, *corr = nullptr, *rhs = nullptr, *ddu = nullptr;
    hipGraphExec_t ngraph = nullptr;            // one captured Newton iteration
    const void* ngraph_key[6] = {};             // dU, P, F, PExt, deltaT, precision it was built for
};

namespace {

constexpr int BS = 256;
enum { RZ = 0, RTOL2 = 1, RR = 2, TOL2 = 3, BETA = 4, NSC = 5 };

// Fixed-order workgroup sums: an xor butterfly inside each wavefront (all lanes end with
// bitwise identical sums), then the per-wave sums in wave order.
template <int K>
__device__ __forceinline__ void wsums(double (&v)[K], double* sh) {
    constexpr int NW = BS / 64;
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_xor(v[k], o, 64);
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) sh[k * NW + (threadIdx.x >> 6)] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double r = sh[k * NW];
        for (int i = 1; i < NW; ++i) r += sh[k * NW + i];
        v[k] = r;
    }
    __syncthreads();
}

// Fixed-order sums of K per-node arrays over [n0, n1).  The order depends only on the
// positions relative to n0, so a specimen sums the same way alone as inside a batch.
// Four independent accumulators per lane keep four loads of each array in flight.
template <int K>
__device__ __forceinline__ void range_sums(const double* const (&v)[K], int n0, int n1, double (&out)[K],
                                           double* sh) {
    double a[4][K];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int k = 0; k < K; ++k) a[j][k] = 0.0;
    for (int n = n0 + threadIdx.x; n < n1; n += 4 * BS) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int m = n + j * BS;
            if (m < n1)
#pragma unroll
                for (int k = 0; k < K; ++k) a[j][k] += v[k][m];
        }
    }
#pragma unroll
    for (int k = 0; k < K; ++k) out[k] = (a[0][k] + a[1][k]) + (a[2][k] + a[3][k]);
    wsums<K>(out, sh);
}

// Sum of chunk partials v[k0, k1) by one wavefront: lane j adds k0 + j, k0 + j + 64, ...,
// then an xor butterfly; every lane ends with the same bits.
__device__ __forceinline__ double chunk_sum(const double* __restrict__ v, int k0, int k1) {
    const int lane = threadIdx.x & 63;
    double a = 0.0;
    for (int k = k0 + lane; k < k1; k += 64) a += v[k];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
    return a;
}

template <typename T>
__global__ __launch_bounds__(BS) void b_init_kernel(int N, const int32_t* __restrict__ spec,
                                                    const int* __restrict__ active, const double* __restrict__ b,
                                                    const T* __restrict__ Minv, T* __restrict__ x, T* __restrict__ r,
                                                    T* __restrict__ z, T* __restrict__ p, double* __restrict__ nrz,
                                                    double* __restrict__ nrr) {
    const int n = blockIdx.x * BS + threadIdx.x;
    if (n >= N) return;
    const bool on = active[spec[n]] != 0;
    T rv[3];
    double rr = 0, rz = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        rv[i] = on ? (T)b[3 * n + i] : (T)0;
        x[3 * n + i] = 0;
        r[3 * n + i] = rv[i];
        rr += (double)rv[i] * rv[i];
    }
    const T* M = Minv + (int64_t)n * 9;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const T zi = M[3 * i] * rv[0] + M[3 * i + 1] * rv[1] + M[3 * i + 2] * rv[2];
        z[3 * n + i] = zi;
        p[3 * n + i] = zi;
        rz += (double)rv[i] * zi;
    }
    nrz[n] = rz;
    nrr[n] = rr;
}

// per specimen: rz, rr; converged specimens (or r = 0) are frozen at iteration 0
__global__ __launch_bounds__(BS) void b_init_reduce_kernel(const int64_t* __restrict__ noff,
                                                           const double* __restrict__ nrz,
                                                           const double* __restrict__ nrr,
                                                           double* __restrict__ sc, int* __restrict__ active,
                                                           int* __restrict__ iters) {
    __shared__ double sh[BS];
    const int s = blockIdx.x;
    const double* const arr[2] = {nrz, nrr};
    double sums[2];
    range_sums<2>(arr, (int)noff[s], (int)noff[s + 1], sums, sh);
    const double rz = sums[0], rr = sums[1];
    if (threadIdx.x == 0) {
        // stop rule ||r_s|| <= max(atol_s, rtol_s ||b_s||)
        const double t2 = fmax(sc[s * NSC + TOL2], sc[s * NSC + RTOL2] * rr);
        sc[s * NSC + TOL2] = t2;
        sc[s * NSC + RZ] = rz;
        sc[s * NSC + RR] = rr;
        iters[s] = 0;
        if (active[s] && (rr <= t2 || !(rr == rr))) active[s] = 0;
    }
}

template <typename T>
__global__ __launch_bounds__(BS) void b_spmv_kernel(int N, const int32_t* __restrict__ spec,
                                                    const int* __restrict__ active,
                                                    const int32_t* __restrict__ slice_off,
                                                    const int32_t* __restrict__ scol,
                                                    const int32_t* __restrict__ slotNode, const T* __restrict__ K,
                                                    const T* __restrict__ p, T* __restrict__ q,
                                                    double* __restrict__ npq) {
    const int ps = blockIdx.x * BS + threadIdx.x;  // row slot
    if (ps >= N) return;
    const int n = slotNode[ps];
    if (!active[spec[n]]) {
        npq[n] = 0.0;
        return;
    }
    const int s = ps / kSlice, l = ps - (ps / kSlice) * kSlice;
    const int o0 = slice_off[s], deg = (slice_off[s + 1] - o0) / kSlice;
    const int32_t* cp = scol + o0 + l;
    const T* kp = K + (int64_t)o0 * 9 + l;
    T y0 = 0, y1 = 0, y2 = 0;
    for (int k = 0; k < deg; ++k) {
        const int m = __builtin_nontemporal_load(cp + k * kSlice);
        const T x0 = p[3 * m], x1 = p[3 * m + 1], x2 = p[3 * m + 2];
        const T* v = kp + (int64_t)k * 9 * kSlice;
        T a[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) a[i] = __builtin_nontemporal_load(v + i * kSlice);
        y0 += a[0] * x0 + a[1] * x1 + a[2] * x2;
        y1 += a[3] * x0 + a[4] * x1 + a[5] * x2;
        y2 += a[6] * x0 + a[7] * x1 + a[8] * x2;
    }
    q[3 * n] = y0;
    q[3 * n + 1] = y1;
    q[3 * n + 2] = y2;
    npq[n] = (double)p[3 * n] * y0 + (double)p[3 * n + 1] * y1 + (double)p[3 * n + 2] * y2;
}

// p.q of one chunk (workgroup per chunk; frozen specimens skip)
__global__ __launch_bounds__(BS) void b_pq_chunk_kernel(const int32_t* __restrict__ chunkNode0,
                                                        const int32_t* __restrict__ chunkSpec,
                                                        const int64_t* __restrict__ noff,
                                                        const int* __restrict__ active,
                                                        const double* __restrict__ npq, double* __restrict__ pqp) {
    __shared__ double sh[BS / 64];
    const int c = blockIdx.x;
    const int s = chunkSpec[c];
    if (!active[s]) return;
    const int n = chunkNode0[c] + threadIdx.x;
    double v[1] = {n < noff[s + 1] ? npq[n] : 0.0};
    wsums<1>(v, sh);
    if (threadIdx.x == 0) pqp[c] = v[0];
}

// x += alpha p, r -= alpha q, z = Minv r with alpha = rz / (sum of the specimen's p.q chunk
// partials), and the chunk partials of r.z and r.r (workgroup per chunk)
template <typename T>
__global__ __launch_bounds__(BS) void b_update_kernel(const int32_t* __restrict__ chunkNode0,
                                                      const int32_t* __restrict__ chunkSpec,
                                                      const int32_t* __restrict__ specChunk,
                                                      const int64_t* __restrict__ noff,
                                                      const int* __restrict__ active, const double* __restrict__ sc,
                                                      const double* __restrict__ pqp, const T* __restrict__ Minv,
                                                      T* __restrict__ x, T* __restrict__ r, T* __restrict__ z,
                                                      const T* __restrict__ p, const T* __restrict__ q,
                                                      double* __restrict__ rzp, double* __restrict__ rrp) {
    __shared__ double sh[2 * (BS / 64)];
    const int c = blockIdx.x;
    const int s = chunkSpec[c];
    if (!active[s]) return;
    const double pq = chunk_sum(pqp, specChunk[s], specChunk[s + 1]);
    const T alpha = (T)(sc[s * NSC + RZ] / pq);
    const int n = chunkNode0[c] + threadIdx.x;
    double part[2] = {0.0, 0.0};  // r.z, r.r
    if (n < noff[s + 1]) {
        T rv[3];
        double rr = 0, rz = 0;
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
            const T zi = M[3 * i] * rv[0] + M[3 * i + 1] * rv[1] + M[3 * i + 2] * rv[2];
            z[3 * n + i] = zi;
            rz += (double)rv[i] * zi;
        }
        part[0] = rz;
        part[1] = rr;
    }
    wsums<2>(part, sh);
    if (threadIdx.x == 0) {
        rzp[c] = part[0];
        rrp[c] = part[1];
    }
}

// beta = rz_new / rz, convergence per specimen (one wavefront per specimen)
__global__ __launch_bounds__(64) void b_beta_kernel(const int32_t* __restrict__ specChunk,
                                                    const double* __restrict__ rzp, const double* __restrict__ rrp,
                                                    double* __restrict__ sc, int* __restrict__ active,
                                                    int* __restrict__ iters) {
    const int s = blockIdx.x;
    if (!active[s]) return;
    const double rz = chunk_sum(rzp, specChunk[s], specChunk[s + 1]);
    const double rr = chunk_sum(rrp, specChunk[s], specChunk[s + 1]);
    if (threadIdx.x == 0) {
        sc[s * NSC + BETA] = rz / sc[s * NSC + RZ];
        sc[s * NSC + RZ] = rz;
        sc[s * NSC + RR] = rr;
        iters[s] += 1;
        if (rr <= sc[s * NSC + TOL2] || !(rr == rr) || !(rz == rz)) active[s] = 0;
    }
}

template <typename T>
__global__ __launch_bounds__(BS) void b_direction_kernel(int N, const int32_t* __restrict__ spec,
                                                         const int* __restrict__ active,
                                                         const double* __restrict__ sc, const T* __restrict__ z,
                                                         T* __restrict__ p) {
    const int n = blockIdx.x * BS + threadIdx.x;
    if (n >= N) return;
    const int s = spec[n];
    if (!active[s]) return;
    const T beta = (T)sc[s * NSC + BETA];
#pragma unroll
    for (int i = 0; i < 3; ++i) p[3 * n + i] = z[3 * n + i] + beta * p[3 * n + i];
}

__global__ void b_count_kernel(int S, const int* __restrict__ active, int* __restrict__ nactive) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int c = 0;
    for (int s = 0; s < S; ++s) c += active[s] ? 1 : 0;
    *nactive = c;
}

template <typename T>
__global__ void b_finish_kernel(int D, const T* __restrict__ x, const double* __restrict__ b,
                                const uint8_t* __restrict__ dirflag, int ndir, double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D) return;
    out[i] = (ndir && dirflag[i]) ? b[i] : (double)x[i];
}

int gridn(int n) { return (n + BS - 1) / BS; }

template <typename T>
void launch_b_batch(efe_ctx* c, BatchWork* w, hipStream_t st, const T* K, const T* Minv, int batch) {
    const int N = c->N, S = c->nspec;
    T *x = (T*)w->x, *r = (T*)w->r, *z = (T*)w->z, *p = (T*)w->p, *q = (T*)w->q;
    for (int j = 0; j < batch; ++j) {
        if (K)
            b_spmv_kernel<T><<<gridn(N), BS, 0, st>>>(N, c->specOfNode, w->active, c->slice_off, c->scol,
                                                      c->slotNode, K, p, q, w->nv0);
        else  // matrix-free FP32 operator
            efe_mf_launch_operator_batched(c, st, (const float*)p, (float*)q, w->nv0, w->active);
        b_pq_chunk_kernel<<<w->C, BS, 0, st>>>(w->chunkNode0, w->chunkSpec, c->specNodeOff, w->active, w->nv0,
                                               w->pqp);
        b_update_kernel<T><<<w->C, BS, 0, st>>>(w->chunkNode0, w->chunkSpec, w->specChunk, c->specNodeOff, w->active,
                                                 w->sc, w->pqp, Minv, x, r, z, p, q, w->rzp, w->rrp);
        b_beta_kernel<<<S, 64, 0, st>>>(w->specChunk, w->rzp, w->rrp, w->sc, w->active, w->iters);
        b_direction_kernel<T><<<gridn(N), BS, 0, st>>>(N, c->specOfNode, w->active, w->sc, z, p);
    }
    b_count_kernel<<<1, 1, 0, st>>>(S, w->active, w->nactive);
}

int ensure_batch(efe_ctx* c, int prec) {
    BatchWork* w = c->bw;
    if (!w) {
        w = new BatchWork();
        c->bw = w;
        w->N = c->N;
        w->S = c->nspec;
        EFE_HIP(hipMalloc(&w->nv0, (size_t)c->N * sizeof(double)));
        EFE_HIP(hipMalloc(&w->nv1, (size_t)c->N * sizeof(double)));
        EFE_HIP(hipMalloc(&w->sc, (size_t)c->nspec * NSC * sizeof(double)));
        EFE_HIP(hipMalloc(&w->active, (size_t)c->nspec * sizeof(int)));
        EFE_HIP(hipMalloc(&w->iters, (size_t)c->nspec * sizeof(int)));
        EFE_HIP(hipMalloc(&w->nactive, sizeof(int)));
        EFE_HIP(hipMalloc(&w->acc, (size_t)c->nspec * sizeof(int)));
        EFE_HIP(hipMalloc(&w->nout, (size_t)c->nspec * 4 * sizeof(double)));
        EFE_HIP(hipHostMalloc(&w->hbuf, sizeof(int) * 4, hipHostMallocDefault));
        EFE_HIP(hipStreamCreateWithFlags(&w->ks, hipStreamNonBlocking));
        // 256-node chunks of every specimen, starting at its first node
        std::vector<int64_t> off((size_t)c->nspec + 1);
        EFE_HIP(hipMemcpyAsync(off.data(), c->specNodeOff, off.size() * sizeof(int64_t), hipMemcpyDeviceToHost,
                               c->stream));
        EFE_HIP(hipStreamSynchronize(c->stream));
        std::vector<int32_t> n0, cs, sco(1, 0);
        for (int sp = 0; sp < c->nspec; ++sp) {
            for (int64_t n = off[sp]; n < off[sp + 1]; n += BS) {
                n0.push_back((int32_t)n);
                cs.push_back(sp);
            }
            sco.push_back((int32_t)n0.size());
        }
        w->C = (int)n0.size();
        const size_t nc = n0.size() ? n0.size() : 1;
        EFE_HIP(hipMalloc(&w->chunkNode0, nc * sizeof(int32_t)));
        EFE_HIP(hipMalloc(&w->chunkSpec, nc * sizeof(int32_t)));
        EFE_HIP(hipMalloc(&w->specChunk, sco.size() * sizeof(int32_t)));
        EFE_HIP(hipMalloc(&w->pqp, nc * sizeof(double)));
        EFE_HIP(hipMalloc(&w->rzp, nc * sizeof(double)));
        EFE_HIP(hipMalloc(&w->rrp, nc * sizeof(double)));
        if (w->C) {
            EFE_HIP(hipMemcpyAsync(w->chunkNode0, n0.data(), n0.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                                   c->stream));
            EFE_HIP(hipMemcpyAsync(w->chunkSpec, cs.data(), cs.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                                   c->stream));
        }
        EFE_HIP(hipMemcpyAsync(w->specChunk, sco.data(), sco.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                               c->stream));
        EFE_HIP(hipStreamSynchronize(c->stream));
        EFE_HIP(hipEventCreateWithFlags(&w->evA, hipEventDisableTiming));
        EFE_HIP(hipEventCreateWithFlags(&w->evB, hipEventDisableTiming));
    }
    if (prec >= 0 && w->prec != prec) {
        hipFree(w->x); hipFree(w->r); hipFree(w->z); hipFree(w->p); hipFree(w->q);
        const size_t es = prec == EFE_FP32 ? sizeof(float) : sizeof(double), D = (size_t)c->D;
        EFE_HIP(hipMalloc(&w->x, D * es));
        EFE_HIP(hipMalloc(&w->r, D * es));
        EFE_HIP(hipMalloc(&w->z, D * es));
        EFE_HIP(hipMalloc(&w->p, D * es));
        EFE_HIP(hipMalloc(&w->q, D * es));
        if (w->graph) hipGraphExecDestroy(w->graph);
        w->graph = nullptr;
        w->prec = prec;
    }
    return EFE_OK;
}

// The batched PCG from the tolerances sc[TOL2], sc[RTOL2] and the active flags already on the
// device (after hipStreamWaitEvent of the context stream); info, spec_iters and xout as
// efe_linsolve_batched.
template <typename T>
int batched_pcg_run(efe_ctx* c, const double* b, double* xout, const efe_solve_opts* o, int32_t* spec_iters,
                    efe_solve_info* info, const T* K, const T* Minv) {
    BatchWork* w = c->bw;
    hipStream_t st = w->ks;
    const int N = c->N, S = c->nspec, D = c->D;
    std::vector<double> sc((size_t)S * NSC, 0.0);
    b_init_kernel<T><<<gridn(N), BS, 0, st>>>(N, c->specOfNode, w->active, b, Minv, (T*)w->x, (T*)w->r, (T*)w->z,
                                              (T*)w->p, w->nv0, w->nv1);
    b_init_reduce_kernel<<<S, BS, 0, st>>>(c->specNodeOff, w->nv0, w->nv1, w->sc, w->active, w->iters);
    EFE_HIP(hipGetLastError());
    int batch = o->check_every > 0 ? o->check_every : 16;
    const int maxit = o->max_iter > 0 ? o->max_iter : 10000;
    if (!w->graph || w->graph_batch != batch || w->graph_op != o->precision) {
        if (w->graph) hipGraphExecDestroy(w->graph);
        w->graph = nullptr;
        hipGraph_t g;
        EFE_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
        launch_b_batch<T>(c, w, st, K, Minv, batch);
        EFE_HIP(hipStreamEndCapture(st, &g));
        EFE_HIP(hipGraphInstantiate(&w->graph, g, nullptr, nullptr, 0));
        EFE_HIP(hipGraphDestroy(g));
        w->graph_batch = batch;
        w->graph_op = o->precision;
    }
    int it = 0;
    info->status = EFE_OK;
    while (true) {
        EFE_HIP(hipGraphLaunch(w->graph, st));
        EFE_HIP(hipMemcpyAsync(w->hbuf, w->nactive, sizeof(int), hipMemcpyDeviceToHost, st));
        EFE_HIP(hipStreamSynchronize(st));
        it += batch;
        if (w->hbuf[0] == 0) break;
        if (it >= maxit) {
            info->status = EFE_NOT_CONVERGED;
            break;
        }
    }
    b_finish_kernel<T><<<(D + 255) / 256, 256, 0, st>>>(D, (const T*)w->x, b, c->dirflag, c->ndir, xout);
    EFE_HIP(hipGetLastError());
    std::vector<int> its(S);
    EFE_HIP(hipMemcpyAsync(its.data(), w->iters, S * sizeof(int), hipMemcpyDeviceToHost, st));
    EFE_HIP(hipMemcpyAsync(sc.data(), w->sc, sc.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    EFE_HIP(hipEventRecord(w->evB, st));
    EFE_HIP(hipStreamWaitEvent(c->stream, w->evB, 0));
    EFE_HIP(hipStreamSynchronize(st));
    int mx = 0;
    double worst = 0.0;
    for (int s = 0; s < S; ++s) {
        if (spec_iters) spec_iters[s] = its[s];
        mx = its[s] > mx ? its[s] : mx;
        const double rr = sc[(size_t)s * NSC + RR];
        if (!(rr == rr)) info->status = EFE_NAN;
        worst = rr > worst ? rr : worst;
    }
    info->iterations = mx;
    info->residual = sqrt(worst);
    info->rhs_norm = 0.0;
    return EFE_OK;
}

template <typename T>
int batched_pcg(efe_ctx* c, const double* b, double* xout, const efe_solve_opts* o, const double* spec_tol,
                const double* spec_rtol, const uint8_t* spec_active, int32_t* spec_iters, efe_solve_info* info,
                const T* K, const T* Minv) {
    BatchWork* w = c->bw;
    hipStream_t st = w->ks;
    const int S = c->nspec;
    std::vector<double> sc((size_t)S * NSC, 0.0);
    std::vector<int> act(S);
    for (int s = 0; s < S; ++s) {
        sc[(size_t)s * NSC + TOL2] = spec_tol[s] * spec_tol[s];
        const double rt = spec_rtol ? spec_rtol[s] : o->rtol;
        sc[(size_t)s * NSC + RTOL2] = rt * rt;
        act[s] = spec_active ? (spec_active[s] ? 1 : 0) : 1;
    }
    EFE_HIP(hipEventRecord(w->evA, c->stream));
    EFE_HIP(hipStreamWaitEvent(st, w->evA, 0));
    EFE_HIP(hipMemcpyAsync(w->sc, sc.data(), sc.size() * sizeof(double), hipMemcpyHostToDevice, st));
    EFE_HIP(hipMemcpyAsync(w->active, act.data(), act.size() * sizeof(int), hipMemcpyHostToDevice, st));
    int rc = batched_pcg_run<T>(c, b, xout, o, spec_iters, info, K, Minv);
    EFE_HIP(hipStreamSynchronize(st));  // sc and act must outlive the copies
    return rc;
}

}  // namespace

int efe_batched_linsolve(efe_ctx* c, const double* b, double* x, const efe_solve_opts* o, const double* spec_tol,
                         const double* spec_rtol, const uint8_t* spec_active, int32_t* spec_iters,
                         efe_solve_info* info) {
    if (!c->specOfNode || c->nspec < 1) {
        efe_set_error("batched solve needs a context created with specimens");
        return EFE_ERR_STATE;
    }
    if (o->precision != c->Kprec) {
        efe_set_error("solve precision differs from the precision K was assembled in");
        return EFE_ERR_ARG;
    }
    if (o->solver != EFE_SOLVER_CG || o->precond != EFE_PRECOND_BLOCK_JACOBI) {
        efe_set_error("batched solves use block-Jacobi PCG");
        return EFE_ERR_ARG;
    }
    const int vp = o->precision == EFE_FP64 ? EFE_FP64 : EFE_FP32;  // vector precision
    int rc = ensure_batch(c, vp);
    if (rc) return rc;
    if (o->precision == EFE_FP32_MATFREE)
        return batched_pcg<float>(c, b, x, o, spec_tol, spec_rtol, spec_active, spec_iters, info, nullptr,
                                  c->Minv32);
    if (o->precision == EFE_FP32)
        return batched_pcg<float>(c, b, x, o, spec_tol, spec_rtol, spec_active, spec_iters, info, c->K32,
                                  c->Minv32);
    return batched_pcg<double>(c, b, x, o, spec_tol, spec_rtol, spec_active, spec_iters, info, c->K64, c->Minv64);
}

// per-specimen convergence quantities: [max|R|, ||ddU||_inf, ||F||_1, NaN]
__global__ __launch_bounds__(BS) void b_norms_kernel(const int64_t* __restrict__ noff, const double* __restrict__ R,
                                                     const double* __restrict__ ddU, const double* __restrict__ F,
                                                     double* __restrict__ out) {
    __shared__ double sh[BS];
    const int s = blockIdx.x;
    const int64_t d0 = 3 * noff[s], d1 = 3 * noff[s + 1];
    double mr = 0, md = 0, sf = 0, nan = 0;
    for (int64_t i = d0 + threadIdx.x; i < d1; i += BS) {
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
    for (int k = BS / 2; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + k]);
        __syncthreads();
    }
    const double rmax = sh[0];
    __syncthreads();
    sh[threadIdx.x] = md;
    __syncthreads();
    for (int k = BS / 2; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + k]);
        __syncthreads();
    }
    const double dmax = sh[0];
    __syncthreads();
    double sums[2] = {sf, nan};
    wsums<2>(sums, sh);
    const double fs = sums[0], nn = sums[1];
    if (threadIdx.x == 0) {
        out[s * 4 + 0] = rmax;
        out[s * 4 + 1] = dmax;
        out[s * 4 + 2] = fs;
        out[s * 4 + 3] = nn > 0 ? 1.0 : 0.0;
    }
}

// Ref <- Temp for the elements of accepted specimens
__global__ void b_accept_kernel(int64_t EQ, int Q, int nstate, const int32_t* __restrict__ specOfElem,
                                const int* __restrict__ acc, const double* __restrict__ tmp, double* __restrict__ ref) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= EQ) return;
    if (!acc[specOfElem[i / Q]]) return;
    for (int c = 0; c < nstate; ++c) ref[c * EQ + i] = tmp[c * EQ + i];
}

int efe_launch_norms_batched(efe_ctx* c, const double* R, const double* ddU, const double* F, double* out) {
    if (!c->specOfNode) {
        efe_set_error("batched norms need a context created with specimens");
        return EFE_ERR_STATE;
    }
    int rc = ensure_batch(c, -1);
    if (rc) return rc;
    double* dout = c->bw->nout;
    b_norms_kernel<<<c->nspec, BS, 0, c->stream>>>(c->specNodeOff, R, ddU, F, dout);
    EFE_HIP(hipGetLastError());
    EFE_HIP(hipMemcpyAsync(out, dout, (size_t)c->nspec * 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    EFE_HIP(hipStreamSynchronize(c->stream));
    return EFE_OK;
}

int efe_launch_accept_batched(efe_ctx* c, const uint8_t* accept) {
    int rc = ensure_batch(c, -1);
    if (rc) return rc;
    std::vector<int> a(c->nspec);
    for (int s = 0; s < c->nspec; ++s) a[s] = accept[s] ? 1 : 0;
    int* d = c->bw->acc;
    EFE_HIP(hipMemcpyAsync(d, a.data(), a.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    const int64_t EQ = (int64_t)c->E * c->Q;
    b_accept_kernel<<<(unsigned)((EQ + 255) / 256), 256, 0, c->stream>>>(EQ, c->Q, c->nstate, c->specOfElem, d,
                                                                         c->stateTmp, c->stateRef);
    EFE_HIP(hipGetLastError());
    EFE_HIP(hipStreamSynchronize(c->stream));
    return EFE_OK;
}

void efe_batched_free(efe_ctx* c) {
    BatchWork* w = c->bw;
    if (!w) return;
    void* bufs[] = {w->x,      w->r,      w->z,      w->p,          w->q,         w->nv0,
                    w->nv1,    w->sc,     w->active, w->iters,      w->nactive,   w->acc,
                    w->nout,   w->pqp,    w->rzp,    w->rrp,        w->chunkNode0, w->chunkSpec,
                    w->specChunk, w->nst, w->nit, w->ngrow, w->nreason, w->nfirst, w->nextrap, w->nhasdir,
                    w->nmaxit, w->nmaxgrow, w->nsolve, w->npass, w->nsolves, w->nkry, w->ncount, w->ngprev,
                    w->nacc, w->dirv, w->corr, w->rhs, w->ddu};
    for (void* p : bufs)
        if (p) hipFree(p);
    if (w->ngraph) hipGraphExecDestroy(w->ngraph);
    if (w->hbuf) hipHostFree(w->hbuf);
    if (w->graph) hipGraphExecDestroy(w->graph);
    if (w->ks) hipStreamDestroy(w->ks);
    if (w->evA) hipEventDestroy(w->evA);
    if (w->evB) hipEventDestroy(w->evB);
    delete w;
    c->bw = nullptr;
}
