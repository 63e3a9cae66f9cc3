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
#include <cstring>
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
    // device-resident Newton rounds (efe_newton_round_batched): the per-specimen Newton table
    // ni [S][NI] (inputs of the round, state, counters) and nd [S][ND], the options nopt, the
    // pass count nreadout [2] (specimens to solve; PCG still active), per-specimen norms of
    // the last correction nout2 [S][4], and the round's DOF vectors
    int* ni = nullptr;
    double *nd = nullptr, *nopt = nullptr, *nout2 = nullptr;
    int* ncount = nullptr;
    double *du = nullptr, *nP = nullptr, *nF = nullptr, *pext = nullptr, *dtv = nullptr;
    double *dirv = nullptr, *corr = nullptr, *rhs = nullptr, *ddu = nullptr;
    hipGraphExec_t ngraph[3] = {};  // pass variants: first pass, first pass with prescribed increments, later
    int64_t nkey[3][10] = {};       // launch parameters each was captured with
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
        for (hipGraphExec_t& g : w->ngraph) {
            if (g) hipGraphExecDestroy(g);
            g = nullptr;
        }
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

// ---- device-resident Newton rounds ----------------------------------------------------------
namespace {

// per-specimen Newton table ni [S][NI]
enum { NI_LIVE = 0, NI_EXTRAP, NI_HASDIR, NI_MAXIT, NI_MAXGROW, NI_STATE, NI_IT, NI_GROW, NI_REASON, NI_FIRST,
       NI_GO, NI_PASSES, NI_SOLVES, NI_KRYLOV, NI };
enum { ND_NACC = 0, ND_PREV, ND };  // nd [S][ND]: average-flux divisor, previous flux residual
// options nopt
enum { OP_RTOL = 0, OP_RTOL_FIRST, OP_ATOLF, OP_FLUX, OP_FLUX_ALT, OP_CORR, OP_FLUX_FLOOR, OP_AVG_FLOOR,
       OP_MIXED_FIRST, OP_ALT_AFTER, OP_KRY_MAXIT, NOPT };

// max as numpy.maximum: NaN when either is NaN
__device__ __forceinline__ double np_max(double a, double b) {
    if (!(a == a)) return a;
    if (!(b == b)) return b;
    return a > b ? a : b;
}

// Head of a later pass: the krylov failures of the last solve (its iteration limit, or a
// correction with an infinite entry when some specimen's PCG produced NaN), it += 1 and the work
// counters of the specimens that solved.  One workgroup.
__global__ __launch_bounds__(BS) void nr_post_kernel(int S, int* __restrict__ ni, const double* __restrict__ sc,
                                                     const int* __restrict__ iters,
                                                     const double* __restrict__ dnorm,
                                                     const double* __restrict__ nopt) {
    __shared__ int anyNaN;
    if (threadIdx.x == 0) anyNaN = 0;
    __syncthreads();
    int mine = 0;
    for (int s = threadIdx.x; s < S; s += BS) {
        const double rr = sc[s * NSC + RR];
        if (!(rr == rr)) mine = 1;
    }
    if (mine) atomicOr(&anyNaN, 1);
    __syncthreads();
    const int kmax = (int)nopt[OP_KRY_MAXIT];
    for (int s = threadIdx.x; s < S; s += BS) {
        int* t = ni + (size_t)s * NI;
        if (!t[NI_GO]) continue;
        const int its = iters[s];
        const double dmax = dnorm[s * 4 + 0];
        if (its >= kmax || (anyNaN && !isfinite(dmax))) {
            t[NI_STATE] = EFE_NEWTON_FAILED;
            t[NI_REASON] = EFE_FAIL_KRYLOV;
        }
        t[NI_IT] += 1;
        t[NI_SOLVES] += 1;
        t[NI_KRYLOV] += its;
    }
}

// dU += ddU
__global__ void nr_add_kernel(int64_t D, double* __restrict__ du, const double* __restrict__ ddu) {
    const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
    if (i < D) du[i] = du[i] + ddu[i];
}

// Start of a pass: the specimens of the element pass (those iterating), which of them take
// their first iteration (prescribed increment applied), cleared status words.  One workgroup.
__global__ __launch_bounds__(BS) void nr_pass_kernel(int S, int* __restrict__ ni, int* __restrict__ specActive,
                                                     int* __restrict__ status, int* __restrict__ specStatus) {
    if (threadIdx.x == 0) status[0] = 0;
    for (int s = threadIdx.x; s < S; s += BS) {
        int* t = ni + (size_t)s * NI;
        const bool it = t[NI_STATE] == EFE_NEWTON_ITERATING;
        t[NI_FIRST] = it && t[NI_IT] == 0 && !t[NI_EXTRAP] && t[NI_HASDIR];
        specActive[s] = it ? 1 : 0;
        specStatus[s] = 0;
        if (it) t[NI_PASSES] += 1;
    }
}

// prescribed values of first iterations: deltaT on the specimen's DOFs, 0 elsewhere; and a
// cleared correction
__global__ void nr_dirv_kernel(int64_t D, const int32_t* __restrict__ specOfNode, const int* __restrict__ ni,
                               const double* __restrict__ dtv, double* __restrict__ dirv, double* __restrict__ corr) {
    const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
    if (i >= D) return;
    const int s = specOfNode[i / 3];
    dirv[i] = ni[(size_t)s * NI + NI_FIRST] ? dtv[i] : 0.0;
    corr[i] = 0.0;
}

// R = P + PExt, prescribed DOFs: the first iterations' prescribed value, else 0
__global__ void nr_residual_kernel(int64_t D, const double* __restrict__ P, const double* __restrict__ pext,
                                   const uint8_t* __restrict__ dirflag, int ndir, const double* __restrict__ dirv,
                                   double* __restrict__ R) {
    const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
    if (i >= D) return;
    double r = pext ? P[i] + pext[i] : P[i];
    if (ndir && dirflag[i]) r = dirv ? dirv[i] : 0.0;
    R[i] = r;
}

// b = R - correction
__global__ void nr_sub_kernel(int64_t D, double* __restrict__ R, const double* __restrict__ corr) {
    const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
    if (i < D) R[i] = R[i] - corr[i];
}

// The convergence test of every specimen (nonlinearimplicitstatic.py:625-725 with the loop's
// divergence and iteration limits :405-458): cutback requests of the element pass, NaN, converged
// (correction < tol and flux <= max(tol_flux avg, floor)), growing residual or maxIter; then the
// specimens still iterating solve, with the PCG tolerances of the batched solve.  `hasDdu`: a
// correction exists (every pass after the round's first).  One workgroup; cnt[0] = solving count.
__global__ __launch_bounds__(BS) void nr_decide_kernel(int S, int hasDdu, int* __restrict__ ni,
                                                       double* __restrict__ nd, const double* __restrict__ nout,
                                                       const int* __restrict__ specStatus,
                                                       const double* __restrict__ nopt, int* __restrict__ active,
                                                       double* __restrict__ sc, int* __restrict__ cnt) {
    __shared__ int count;
    if (threadIdx.x == 0) count = 0;
    __syncthreads();
    int mine = 0;
    for (int s = threadIdx.x; s < S; s += BS) {
        int* t = ni + (size_t)s * NI;
        double* u = nd + (size_t)s * ND;
        int st = t[NI_STATE];
        if (st == EFE_NEWTON_ITERATING && specStatus[s]) {
            st = EFE_NEWTON_FAILED;
            t[NI_REASON] = EFE_FAIL_CUTBACK;
        }
        const int it = t[NI_IT];
        const double avg = np_max(nopt[OP_AVG_FLOOR], nout[s * 4 + 2] / u[ND_NACC]);
        const double tolF = it < (int)nopt[OP_ALT_AFTER] ? nopt[OP_FLUX] : nopt[OP_FLUX_ALT];
        const double thr = np_max(tolF * avg, nopt[OP_FLUX_FLOOR]);
        const int first = t[NI_FIRST];
        if (st == EFE_NEWTON_ITERATING && !first) {
            const double flux = nout[s * 4 + 0];
            const double correction = hasDdu ? nout[s * 4 + 1] : 0.0;
            const double prev = u[ND_PREV];
            const int grew = t[NI_GROW] + (flux > prev ? 1 : 0);
            u[ND_PREV] = flux;
            t[NI_GROW] = grew;
            if (nout[s * 4 + 3] != 0.0) {
                st = EFE_NEWTON_FAILED;
                t[NI_REASON] = EFE_FAIL_NAN;
            } else if (correction < nopt[OP_CORR] && flux <= thr) {
                st = EFE_NEWTON_CONVERGED;
            } else if (grew > t[NI_MAXGROW] || it == t[NI_MAXIT]) {
                st = EFE_NEWTON_FAILED;
                t[NI_REASON] = EFE_FAIL_ITERATIONS;
            }
        }
        const int go = st == EFE_NEWTON_ITERATING;
        t[NI_STATE] = st;
        t[NI_GO] = go;
        active[s] = go;
        const double a = nopt[OP_ATOLF] * thr;
        sc[s * NSC + TOL2] = a * a;
        const double rt = (nopt[OP_MIXED_FIRST] != 0.0 && first) ? nopt[OP_RTOL_FIRST] : nopt[OP_RTOL];
        sc[s * NSC + RTOL2] = rt * rt;
        mine += go;
    }
    if (mine) atomicAdd(&count, mine);
    __syncthreads();
    if (threadIdx.x == 0) cnt[0] = count;
}

int ensure_newton(efe_ctx* c) {
    BatchWork* w = c->bw;
    if (w->ni) return EFE_OK;
    const size_t S = c->nspec, D = c->D;
    EFE_HIP(hipMalloc(&w->ni, S * NI * sizeof(int)));
    EFE_HIP(hipMalloc(&w->nd, S * ND * sizeof(double)));
    EFE_HIP(hipMalloc(&w->nopt, NOPT * sizeof(double)));
    EFE_HIP(hipMalloc(&w->nout2, S * 4 * sizeof(double)));
    EFE_HIP(hipMalloc(&w->ncount, 2 * sizeof(int)));
    double** vecs[] = {&w->du, &w->nP, &w->nF, &w->pext, &w->dtv, &w->dirv, &w->corr, &w->rhs, &w->ddu};
    for (double** v : vecs) {
        EFE_HIP(hipMalloc(v, D * sizeof(double)));
        EFE_HIP(hipMemsetAsync(*v, 0, D * sizeof(double), c->stream));
    }
    EFE_HIP(hipStreamSynchronize(c->stream));
    return EFE_OK;
}

// The launches of one Newton pass on `st`: variant 0 the round's first pass, 1 the first pass
// with prescribed increments (right-hand-side correction), 2 a later pass (which first applies
// the previous solve).  Every launch reads its data from device memory, so the captured graph
// replays for any round.
template <typename T>
int newton_pass_launch(efe_ctx* c, hipStream_t st, int variant, int prec, bool hasPExt, int batch, const T* K,
                       const T* Minv) {
    BatchWork* w = c->bw;
    const int S = c->nspec, N = c->N;
    const int64_t D = c->D;
    const unsigned gd = (unsigned)((D + BS - 1) / BS);
    if (variant == 2) {
        b_finish_kernel<T><<<gd, BS, 0, st>>>((int)D, (const T*)w->x, w->rhs, c->dirflag, c->ndir, w->ddu);
        b_norms_kernel<<<S, BS, 0, st>>>(c->specNodeOff, w->ddu, nullptr, w->ddu, w->nout2);
        nr_post_kernel<<<1, BS, 0, st>>>(S, w->ni, w->sc, w->iters, w->nout2, w->nopt);
        nr_add_kernel<<<gd, BS, 0, st>>>(D, w->du, w->ddu);
    }
    nr_pass_kernel<<<1, BS, 0, st>>>(S, w->ni, c->specActive, c->status, c->specStatus);
    if (variant == 1) nr_dirv_kernel<<<gd, BS, 0, st>>>(D, c->specOfNode, w->ni, w->dtv, w->dirv, w->corr);
    EFE_HIP(hipGetLastError());
    if (int rc = efe_launch_elements(c, w->du, prec)) return rc;
    const double* dv = variant == 1 ? w->dirv : nullptr;
    double* cr = variant == 1 ? w->corr : nullptr;
    if (prec == EFE_FP32_MATFREE) {
        if (int rc = efe_launch_mf_assembly(c, w->nP, w->nF, dv, cr)) return rc;
    } else if (int rc = efe_launch_assembly(c, w->nP, w->nF, dv, cr, prec)) {
        return rc;
    }
    nr_residual_kernel<<<gd, BS, 0, st>>>(D, w->nP, hasPExt ? w->pext : nullptr, c->dirflag, c->ndir, dv, w->rhs);
    b_norms_kernel<<<S, BS, 0, st>>>(c->specNodeOff, w->rhs, variant == 2 ? w->ddu : nullptr, w->nF, w->nout);
    nr_decide_kernel<<<1, BS, 0, st>>>(S, variant == 2, w->ni, w->nd, w->nout, c->specStatus, w->nopt, w->active,
                                       w->sc, w->ncount);
    if (variant == 1) nr_sub_kernel<<<gd, BS, 0, st>>>(D, w->rhs, w->corr);
    b_init_kernel<T><<<gridn(N), BS, 0, st>>>(N, c->specOfNode, w->active, w->rhs, Minv, (T*)w->x, (T*)w->r,
                                              (T*)w->z, (T*)w->p, w->nv0, w->nv1);
    b_init_reduce_kernel<<<S, BS, 0, st>>>(c->specNodeOff, w->nv0, w->nv1, w->sc, w->active, w->iters);
    launch_b_batch<T>(c, w, st, K, Minv, batch);
    EFE_HIP(hipGetLastError());
    return EFE_OK;
}

template <typename T>
int capture(efe_ctx* c, hipGraphExec_t* out, hipStream_t st, int (*fn)(efe_ctx*, hipStream_t, int, int, bool, int,
                                                                        const T*, const T*),
            int variant, int prec, bool hasPExt, int batch, const T* K, const T* Minv) {
    hipGraph_t g;
    EFE_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
    const int rc = fn(c, st, variant, prec, hasPExt, batch, K, Minv);
    const hipError_t e = hipStreamEndCapture(st, &g);
    if (rc) {
        if (e == hipSuccess) hipGraphDestroy(g);
        return rc;
    }
    EFE_HIP(e);
    EFE_HIP(hipGraphInstantiate(out, g, nullptr, nullptr, 0));
    EFE_HIP(hipGraphDestroy(g));
    return EFE_OK;
}

template <typename T>
int newton_round(efe_ctx* c, const efe_newton_opts* o, bool hasPExt, bool anyFirst, int maxPasses, int* passes,
                 const T* K, const T* Minv) {
    BatchWork* w = c->bw;
    hipStream_t st = c->stream;
    const int batch = o->check_every > 0 ? o->check_every : 16;
    const int maxit = o->krylov_max_iter > 0 ? o->krylov_max_iter : 10000;
    // the PCG batch graph of later batches (also the one efe_linsolve_batched replays)
    if (!w->graph || w->graph_batch != batch || w->graph_op != o->precision) {
        if (w->graph) hipGraphExecDestroy(w->graph);
        w->graph = nullptr;
        hipGraph_t g;
        EFE_HIP(hipStreamBeginCapture(w->ks, hipStreamCaptureModeRelaxed));
        launch_b_batch<T>(c, w, w->ks, K, Minv, batch);
        EFE_HIP(hipStreamEndCapture(w->ks, &g));
        EFE_HIP(hipGraphInstantiate(&w->graph, g, nullptr, nullptr, 0));
        EFE_HIP(hipGraphDestroy(g));
        w->graph_batch = batch;
        w->graph_op = o->precision;
    }
    for (int pass = 0;; ++pass) {
        if (pass > maxPasses) {
            efe_set_error("efe_newton_round_batched: more passes than any specimen's iteration limit allows");
            return EFE_ERR_STATE;
        }
        const int variant = pass == 0 ? (anyFirst ? 1 : 0) : 2;
        // everything a graph bakes in by value
        const int64_t key[10] = {1,
                                 o->precision,
                                 hasPExt,
                                 batch,
                                 c->ndir,
                                 c->epw,
                                 (int64_t)(intptr_t)c->stateRef,
                                 (int64_t)(intptr_t)c->stateTmp,
                                 (int64_t)(intptr_t)K,
                                 (int64_t)(intptr_t)Minv};
        if (!w->ngraph[variant] || memcmp(key, w->nkey[variant], sizeof(key)) != 0) {
            if (w->ngraph[variant]) hipGraphExecDestroy(w->ngraph[variant]);
            w->ngraph[variant] = nullptr;
            if (int rc = capture<T>(c, &w->ngraph[variant], st, newton_pass_launch<T>, variant, o->precision,
                                    hasPExt, batch, K, Minv))
                return rc;
            memcpy(w->nkey[variant], key, sizeof(key));
        }
        EFE_HIP(hipGraphLaunch(w->ngraph[variant], st));
        ++*passes;
        EFE_HIP(hipMemcpyAsync(w->hbuf, w->ncount, sizeof(int), hipMemcpyDeviceToHost, st));
        EFE_HIP(hipMemcpyAsync(w->hbuf + 1, w->nactive, sizeof(int), hipMemcpyDeviceToHost, st));
        EFE_HIP(hipStreamSynchronize(st));
        if (w->hbuf[0] == 0) break;  // no specimen iterates: the round is over
        int it = batch;
        while (w->hbuf[1] != 0 && it < maxit) {
            EFE_HIP(hipGraphLaunch(w->graph, st));
            EFE_HIP(hipMemcpyAsync(w->hbuf + 1, w->nactive, sizeof(int), hipMemcpyDeviceToHost, st));
            EFE_HIP(hipStreamSynchronize(st));
            it += batch;
        }
    }
    return EFE_OK;
}

}  // namespace

int efe_newton_round_batched_impl(efe_ctx* c, double* dU, double* P, double* F, const double* PExt,
                                  const double* deltaT, const uint8_t* live, const uint8_t* extrap,
                                  const uint8_t* hasDir, const int32_t* maxIter, const int32_t* maxGrow,
                                  const double* nAcc, const efe_newton_opts* o, int32_t* state, int32_t* iterations,
                                  int32_t* reason, int32_t* work) {
    const int S = c->nspec;
    const int vp = o->precision == EFE_FP64 ? EFE_FP64 : EFE_FP32;
    if (int rc = efe_prepare_elements(c, o->precision)) return rc;
    if (int rc = ensure_batch(c, vp)) return rc;
    if (int rc = ensure_newton(c)) return rc;
    BatchWork* w = c->bw;
    // the round runs on the batch stream (a graph cannot be captured on the legacy default
    // stream), ordered after the context stream's work and before its later work
    hipStream_t st = w->ks, home = c->stream;
    EFE_HIP(hipEventRecord(w->evA, home));
    EFE_HIP(hipStreamWaitEvent(st, w->evA, 0));
    std::vector<int> ni((size_t)S * NI, 0);
    std::vector<double> nd((size_t)S * ND, 0.0);
    bool anyFirst = false;
    int maxPasses = 1;
    for (int s = 0; s < S; ++s) {
        int* t = ni.data() + (size_t)s * NI;
        t[NI_LIVE] = live[s] ? 1 : 0;
        t[NI_EXTRAP] = extrap[s] ? 1 : 0;
        t[NI_HASDIR] = hasDir[s] ? 1 : 0;
        t[NI_MAXIT] = maxIter[s];
        t[NI_MAXGROW] = maxGrow[s];
        t[NI_STATE] = live[s] ? EFE_NEWTON_ITERATING : EFE_NEWTON_IDLE;
        nd[(size_t)s * ND + ND_NACC] = nAcc[s];
        if (live[s] && !extrap[s] && hasDir[s]) anyFirst = true;
        if (live[s] && maxIter[s] + 2 > maxPasses) maxPasses = maxIter[s] + 2;
    }
    double opt[NOPT] = {o->rtol,           o->rtol_first, o->atol_factor,  o->flux_tol,
                        o->flux_tol_alt,   o->correction_tol, o->flux_floor, o->avg_floor,
                        (double)(o->mixed_first != 0), (double)o->alt_after,
                        (double)(o->krylov_max_iter > 0 ? o->krylov_max_iter : 10000)};
    const size_t Db = (size_t)c->D * sizeof(double);
    EFE_HIP(hipMemcpyAsync(w->ni, ni.data(), ni.size() * sizeof(int), hipMemcpyHostToDevice, st));
    EFE_HIP(hipMemcpyAsync(w->nd, nd.data(), nd.size() * sizeof(double), hipMemcpyHostToDevice, st));
    EFE_HIP(hipMemcpyAsync(w->nopt, opt, sizeof(opt), hipMemcpyHostToDevice, st));
    EFE_HIP(hipMemcpyAsync(w->du, dU, Db, hipMemcpyDeviceToDevice, st));
    EFE_HIP(hipMemcpyAsync(w->dtv, deltaT, Db, hipMemcpyDeviceToDevice, st));
    if (PExt) EFE_HIP(hipMemcpyAsync(w->pext, PExt, Db, hipMemcpyDeviceToDevice, st));
    const bool wasActive = c->useActive;
    c->useActive = true;  // the passes cover the iterating specimens (specActive, set on the device)
    c->stream = st;       // the element and assembly launches follow
    int passes = 0, rc;
    if (o->precision == EFE_FP32_MATFREE)
        rc = newton_round<float>(c, o, PExt != nullptr, anyFirst, maxPasses, &passes, nullptr, c->Minv32);
    else if (o->precision == EFE_FP32)
        rc = newton_round<float>(c, o, PExt != nullptr, anyFirst, maxPasses, &passes, c->K32, c->Minv32);
    else
        rc = newton_round<double>(c, o, PExt != nullptr, anyFirst, maxPasses, &passes, c->K64, c->Minv64);
    c->stream = home;
    c->useActive = wasActive;
    if (rc) {
        hipStreamSynchronize(st);
        return rc;
    }
    c->kscrValid = o->precision != EFE_FP32_MATFREE;
    c->pscrValid = true;
    c->Kprec = o->precision;
    c->Kversion++;
    c->tot[EFE_T_ELEMENT_PASSES] += passes;
    EFE_HIP(hipMemcpyAsync(dU, w->du, Db, hipMemcpyDeviceToDevice, st));
    if (P) EFE_HIP(hipMemcpyAsync(P, w->nP, Db, hipMemcpyDeviceToDevice, st));
    if (F) EFE_HIP(hipMemcpyAsync(F, w->nF, Db, hipMemcpyDeviceToDevice, st));
    EFE_HIP(hipMemcpyAsync(ni.data(), w->ni, ni.size() * sizeof(int), hipMemcpyDeviceToHost, st));
    EFE_HIP(hipEventRecord(w->evB, st));
    EFE_HIP(hipStreamWaitEvent(home, w->evB, 0));
    EFE_HIP(hipStreamSynchronize(st));
    for (int s = 0; s < S; ++s) {
        const int* t = ni.data() + (size_t)s * NI;
        state[s] = t[NI_STATE];
        iterations[s] = t[NI_IT];
        reason[s] = t[NI_REASON];
        work[3 * s + 0] = t[NI_PASSES];
        work[3 * s + 1] = t[NI_SOLVES];
        work[3 * s + 2] = t[NI_KRYLOV];
    }
    return EFE_OK;
}

void efe_batched_free(efe_ctx* c) {
    BatchWork* w = c->bw;
    if (!w) return;
    void* bufs[] = {w->x,      w->r,      w->z,      w->p,          w->q,          w->nv0,
                    w->nv1,    w->sc,     w->active, w->iters,      w->nactive,    w->acc,
                    w->nout,   w->pqp,    w->rzp,    w->rrp,        w->chunkNode0, w->chunkSpec,
                    w->specChunk, w->ni,  w->nd,     w->nopt,       w->nout2,      w->ncount,
                    w->du,     w->nP,     w->nF,     w->pext,       w->dtv,        w->dirv,
                    w->corr,   w->rhs,    w->ddu};
    for (void* p : bufs)
        if (p) hipFree(p);
    for (hipGraphExec_t g : w->ngraph)
        if (g) hipGraphExecDestroy(g);
    if (w->hbuf) hipHostFree(w->hbuf);
    if (w->graph) hipGraphExecDestroy(w->graph);
    if (w->ks) hipStreamDestroy(w->ks);
    if (w->evA) hipEventDestroy(w->evA);
    if (w->evB) hipEventDestroy(w->evB);
    delete w;
    c->bw = nullptr;
}
