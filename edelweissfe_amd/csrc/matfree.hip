// Matrix-free FP32 operator for the mixed-precision form (EFE_FP32_MATFREE).
//
// The element pass stores, per element and QP, the compact tangent
// w detJ [A, B, cnn, (cn1, c1n,)] and the deviatoric direction n as n_11, n_22, n_12, n_13,
// n_23 (n_33 = -n_11 - n_22): 8 floats, 10 for Drucker-Prager, of
// C = A 1(x)1 + B Idev + cnn n(x)n + cn1 n(x)1 + c1n 1(x)n, in [QP][field][element] order
// so that consecutive lanes read consecutive elements.  y = K p is then, per element (one
// thread):
//   J   = sum_a dN_a/dxi (x) (x_a - x_0)              (rebuilt: cheaper than storing J^-1)
//   H   = sum_a p_a (x) dN_a/dxi                      (3x3)
//   eps = sym(H J^-1)                                 (engineering Voigt)
//   sig = C eps                                       (compact form, ~30 flops)
//   M_ik = sum_j sig_ij J^-1_jk
//   f_a = M dN_a/dxi                                  -> node-sorted force slots
// followed by a per-node sum of its slots in element order (the order of the assembled
// path's row sums), so the product is deterministic.  Dirichlet DOFs are
// eliminated symmetrically, as in the assembled operator: p is read as 0 on them and their
// rows return p (unit diagonal).  The same two kernels with the prescribed values as p and
// no elimination give the right-hand-side correction K[:, dir] dir_values.
//
// Algorithmic bytes per application: 32..40 per element-QP pair (tangent), 4 per
// element-node (connectivity) and 2 x 4 per element-node (slot index, force slot written and
// read), plus the node vectors; against 9 x 4 + 4 bytes per 3x3 block of the assembled FP32
// operator.  It also removes the FP32 assembly from every Newton iteration.
#include <algorithm>
#include <type_traits>
#include <vector>

#include "efe_internal.h"

namespace {

// slots staged per chunk by the gather: 2048 x 12 B (+1/8 skew) = 27 KiB of LDS, one chunk
// for the 256 nodes of a workgroup of an interior hexa8 region
constexpr int kGatherChunk = 2048;

__device__ __forceinline__ int xcd_block_mf() {
    const int per = gridDim.x >> 3;
    return (blockIdx.x & 7) * per + (blockIdx.x >> 3);
}

// Deterministic workgroup sum: xor butterfly inside each wavefront (all lanes end with
// bitwise identical sums), then the per-wave sums in wave order.
template <int BS>
__device__ __forceinline__ double block_sum_mf(double v, double* sh) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    constexpr int NW = BS / 64;
    if (NW == 1) return v;
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = sh[0];
    for (int i = 1; i < NW; ++i) r += sh[i];
    __syncthreads();
    return r;
}

// Sum of the force slots of node n0 + threadIdx.x, for the BS consecutive nodes
// n0 .. n0 + BS - 1 of a workgroup (block-uniform call).  Their slots are one contiguous
// range of the node-sorted force list; it is staged through LDS in chunks with coalesced
// loads (consecutive lanes, consecutive words), then every lane sums its own node's slots
// in slot order, the same order as a direct per-node loop.  LDS layout: component-major
// with a skew of one word per 8 slots (lane l starts near slot 8 l, so the skewed word
// index 9 l + .. falls into distinct banks) and a row pitch of CHP words (CHP = 1 mod 64,
// so the three components of one slot, stored by neighbouring lanes, differ in bank too).
// A workgroup in which no lane wants its sum (frozen specimens) reads nothing.
template <int BS>
__device__ __forceinline__ void gather_slots(int n0, int N, const int32_t* __restrict__ nptr,
                                             const float* __restrict__ Fe, bool want, float* buf, float (&y)[3]) {
    constexpr int CH = kGatherChunk, CHP = kGatherChunk + kGatherChunk / 8 + 1;
    constexpr int R = (3 * CH + BS - 1) / BS;  // staged words per lane
    y[0] = y[1] = y[2] = 0.0f;
    if (!__syncthreads_or(want)) return;
    const int n = n0 + threadIdx.x;
    const int s0 = nptr[n0], s1 = nptr[min(n0 + BS, N)];
    const int mb = want ? nptr[n] : 0, me = want ? nptr[n + 1] : 0;
    for (int c0 = s0; c0 < s1; c0 += CH) {
        const int cnt = min(CH, s1 - c0);
        const float* src = Fe + (int64_t)c0 * 3;
        // all loads first (independent, so they are in flight together), then the stores
        float tmp[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = threadIdx.x + r * BS;
            tmp[r] = i < 3 * cnt ? src[i] : 0.0f;
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = threadIdx.x + r * BS;
            if (i < 3 * cnt) {
                const int c = i / 3, k = i - 3 * (i / 3);
                buf[k * CHP + c + (c >> 3)] = tmp[r];
            }
        }
        __syncthreads();
        const int lo = max(mb, c0), hi = min(me, c0 + cnt);
        for (int c = lo; c < hi; ++c) {
            const int w = (c - c0) + ((c - c0) >> 3);
            y[0] += buf[w];
            y[1] += buf[CHP + w];
            y[2] += buf[2 * CHP + w];
        }
        __syncthreads();
    }
}

constexpr int kGatherLds = 3 * (kGatherChunk + kGatherChunk / 8 + 1);

// gather_slots for short lists (brick partials, ~2.4 per node): each lane reads its node's
// entries directly, in the same order, so the sums are bitwise those of gather_slots.
__device__ __forceinline__ void gather_direct(int n, const int32_t* __restrict__ nptr, const float* __restrict__ Fe,
                                              bool want, float (&y)[3]) {
    y[0] = y[1] = y[2] = 0.0f;
    if (!want) return;
    const int b = nptr[n], e = nptr[n + 1];
    for (int c = b; c < e; ++c) {
        y[0] += Fe[3 * (int64_t)c];
        y[1] += Fe[3 * (int64_t)c + 1];
        y[2] += Fe[3 * (int64_t)c + 2];
    }
}

// One thread per element: element force f_e = K_e p_e.  J is rebuilt from the node
// coordinates relative to the element's first node (formed in FP64, so that FP32 keeps the
// element's own scale), the compact tangent is read per QP, and f_a is written to the
// node-sorted slot pos[a][e] of the force list, so the gather reads contiguous memory.
template <int NN, int Q, int NC, typename TP>
__global__ __launch_bounds__(256) void mf_apply_kernel(int E, int N, const int32_t* __restrict__ conn,
                                                       const double* __restrict__ coords,
                                                       const float* __restrict__ dnxi,
                                                       const float* __restrict__ mfC, const TP* __restrict__ p,
                                                       const uint8_t* __restrict__ dirflag, int mask,
                                                       const int* __restrict__ flags,
                                                       const int32_t* __restrict__ pos, float* __restrict__ Fs,
                                                       const int* __restrict__ active,
                                                       const int32_t* __restrict__ specOfElem) {
    if (flags && flags[0]) return;
    const int e = xcd_block_mf() * 256 + threadIdx.x;
    if (e >= E) return;
    if (active && !active[specOfElem[e]]) return;  // frozen specimen of a batched solve
    float pe[NN][3], fe[NN][3], xe[NN][3];
    double x0[3];
    // the tangents are streamed once per product: non-temporal loads keep them from
    // evicting the gathered p and coordinates from L2 (measured: 131 us instead of 151 us
    // at 96^3; for the connectivity and slot indices plain loads are faster)
#pragma unroll
    for (int a = 0; a < NN; ++a) {
        const int nd = conn[(int64_t)a * E + e];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            float v = (float)p[3 * nd + i];
            if (mask && dirflag[3 * nd + i]) v = 0.0f;
            pe[a][i] = v;
            fe[a][i] = 0.0f;
            const double xc = coords[(int64_t)i * N + nd];
            if (a == 0) x0[i] = xc;
            xe[a][i] = (float)(xc - x0[i]);
        }
    }
    for (int q = 0; q < Q; ++q) {
        const float* dn = dnxi + q * NN * 3;
        const float* dj = dnxi + (Q + q) * NN * 3;  // the table J is formed with
        float H[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // H[i*3+k] = sum_a p_a,i dN_a/dxi_k
        float J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // J[k*3+c] = sum_a dJ_a/dxi_k x_a,c
#pragma unroll
        for (int a = 0; a < NN; ++a) {
            const float d[3] = {dn[a * 3], dn[a * 3 + 1], dn[a * 3 + 2]};
            const float dk[3] = {dj[a * 3], dj[a * 3 + 1], dj[a * 3 + 2]};
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    H[i * 3 + k] += pe[a][i] * d[k];
                    J[k * 3 + i] += dk[k] * xe[a][i];
                }
        }
        const float det = J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) +
                          J[2] * (J[3] * J[7] - J[4] * J[6]);
        const float id = 1.0f / det;
        const float iJ[9] = {(J[4] * J[8] - J[5] * J[7]) * id, (J[2] * J[7] - J[1] * J[8]) * id,
                             (J[1] * J[5] - J[2] * J[4]) * id, (J[5] * J[6] - J[3] * J[8]) * id,
                             (J[0] * J[8] - J[2] * J[6]) * id, (J[2] * J[3] - J[0] * J[5]) * id,
                             (J[3] * J[7] - J[4] * J[6]) * id, (J[1] * J[6] - J[0] * J[7]) * id,
                             (J[0] * J[4] - J[1] * J[3]) * id};
        // du_i/dx_j = sum_k H_ik iJ_jk
        float g[9];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                g[i * 3 + j] = H[i * 3] * iJ[j * 3] + H[i * 3 + 1] * iJ[j * 3 + 1] + H[i * 3 + 2] * iJ[j * 3 + 2];
        const float ep[6] = {g[0], g[4], g[8], g[1] + g[3], g[2] + g[6], g[5] + g[7]};
        float cf[NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) cf[k] = __builtin_nontemporal_load(mfC + (int64_t)(q * NC + k) * E + e);
        const float A = cf[0], B = cf[1], cnn = cf[2];
        const float cn1 = NC == 10 ? cf[3] : 0.0f, c1n = NC == 10 ? cf[4] : 0.0f;
        const float* nc = cf + (NC == 10 ? 5 : 3);
        const float n[6] = {nc[0], nc[1], -(nc[0] + nc[1]), nc[2], nc[3], nc[4]};  // n deviatoric
        const float tr = ep[0] + ep[1] + ep[2];
        const float ne = n[0] * ep[0] + n[1] * ep[1] + n[2] * ep[2] + n[3] * ep[3] + n[4] * ep[4] + n[5] * ep[5];
        const float t3 = tr * (1.0f / 3.0f);
        const float sn = cnn * ne + cn1 * tr;     // coefficient of n
        const float s1 = A * tr + c1n * ne;       // coefficient of 1
        float sg[6];
#pragma unroll
        for (int i = 0; i < 3; ++i) sg[i] = s1 + B * (ep[i] - t3) + sn * n[i];
#pragma unroll
        for (int i = 3; i < 6; ++i) sg[i] = 0.5f * B * ep[i] + sn * n[i];
        // tensor sigma (Voigt 11,22,33,12,13,23); M_ik = sum_j sig_ij iJ_jk
        const float S[9] = {sg[0], sg[3], sg[4], sg[3], sg[1], sg[5], sg[4], sg[5], sg[2]};
        float M[9];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int k = 0; k < 3; ++k)
                M[i * 3 + k] = S[i * 3] * iJ[k] + S[i * 3 + 1] * iJ[3 + k] + S[i * 3 + 2] * iJ[6 + k];
#pragma unroll
        for (int a = 0; a < NN; ++a) {
            const float d0 = dn[a * 3], d1 = dn[a * 3 + 1], d2 = dn[a * 3 + 2];
#pragma unroll
            for (int i = 0; i < 3; ++i) fe[a][i] += M[i * 3] * d0 + M[i * 3 + 1] * d1 + M[i * 3 + 2] * d2;
        }
    }
#pragma unroll
    for (int a = 0; a < NN; ++a) {
        float* dst = Fs + (int64_t)pos[(int64_t)a * E + e] * 3;
#pragma unroll
        for (int i = 0; i < 3; ++i) dst[i] = fe[a][i];
    }
}

// Hexa8 with 2x2x2 Gauss points: the same product with the trilinear interpolation written
// in monomials.  A nodal field v gives v(x, y, z) = c0 + c_x x + c_y y + c_z z + c_xy xy +
// c_yz yz + c_xz xz + c_xyz xyz with c = (1/8) H v (H: the +-1 Walsh matrix of the node
// signs, applied as a 3-stage butterfly), so its natural gradient at a point costs 3 FMA per
// direction instead of a sum over the 8 nodes; the element forces f_a = sum_q M_q grad N_a(q)
// are accumulated in the same monomial basis and mapped back to the nodes by the transposed
// butterfly.  J (from the coordinates) and H = grad p use the first form, the forces the
// second: 90 instead of 216 multiply-adds per QP.  The evaluation point of QP q is the
// reference's exchanged point (eta_q, xi_q, zeta_q).
//
// Node order (xi, eta, zeta) signs: n0 (-,-,-) n1 (-,-,+) n2 (+,-,+) n3 (+,-,-) n4 (-,+,-)
// n5 (-,+,+) n6 (+,+,+) n7 (+,+,-); canonical butterfly index k = [xi>0] + 2 [eta>0] +
// 4 [zeta>0]:
__constant__ int8_t kHex8Canon[8] = {0, 3, 4, 7, 1, 2, 5, 6};  // node of canonical index k

// in place: w[k] (canonical order) -> r[k] = sum_j prod_{i in bits(k)} s_i(j) w[j]
__device__ __forceinline__ void walsh3(float (&w)[8]) {
#pragma unroll
    for (int h = 1; h < 8; h <<= 1)
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (!(k & h)) {
                const float lo = w[k], hi = w[k | h];
                w[k] = lo + hi;
                w[k | h] = hi - lo;
            }
}

// in place: A[k] (monomial coefficients, A[0] = 0) -> f[k] = sum_S A_S prod_{i in S} s_i(k)
__device__ __forceinline__ void walsh3_t(float (&w)[8]) {
#pragma unroll
    for (int h = 1; h < 8; h <<= 1)
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (!(k & h)) {
                const float lo = w[k], hi = w[k | h];
                w[k] = lo - hi;
                w[k | h] = lo + hi;
            }
}

template <int NC, typename TP>
__global__ __launch_bounds__(256) void mf_apply_hex8_kernel(int E, int N, const int32_t* __restrict__ conn,
                                                            const double* __restrict__ coords,
                                                            const float* __restrict__ dnxi,
                                                            const float* __restrict__ mfC, const TP* __restrict__ p,
                                                            const uint8_t* __restrict__ dirflag, int mask,
                                                            const int* __restrict__ flags,
                                                            const int32_t* __restrict__ pos, float* __restrict__ Fs,
                                                            const int* __restrict__ active,
                                                            const int32_t* __restrict__ specOfElem) {
    (void)dnxi;
    if (flags && flags[0]) return;
    const int e = xcd_block_mf() * 256 + threadIdx.x;
    if (e >= E) return;
    if (active && !active[specOfElem[e]]) return;  // frozen specimen of a batched solve
    // monomial coefficients (1/8 included) of p and of the coordinates relative to node 0
    // (formed in FP64, so that FP32 keeps the element's own scale)
    float cp[3][8], cx[3][8];
    double x0[3];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int a = kHex8Canon[k];
        const int nd = conn[(int64_t)a * E + e];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            float v = (float)p[3 * nd + i];
            if (mask && dirflag[3 * nd + i]) v = 0.0f;
            cp[i][k] = v;
            const double xc = coords[(int64_t)i * N + nd];
            if (k == 0) x0[i] = xc;
            cx[i][k] = (float)(xc - x0[i]);
        }
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        walsh3(cp[i]);
        walsh3(cx[i]);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            cp[i][k] *= 0.125f;
            cx[i][k] *= 0.125f;
        }
    }
    // force coefficients in the same basis, x 8
    float fa[3][8];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 8; ++k) fa[i][k] = 0.0f;
    const float g = 0.57735026918962576451f;
    // not unrolled: the unrolled loop hoists the tangent loads of all QPs and needs 154
    // VGPRs (3 waves per SIMD instead of 4), measured slower; so is loading the next QP's
    // tangent ahead
#pragma unroll 1
    for (int q = 0; q < 8; ++q) {
        // evaluation point (eta_q, xi_q, zeta_q) of the QP table xi = (-,-,-,-,+,+,+,+),
        // eta = (-,-,+,+,-,-,+,+), zeta = (-,+,+,-,-,+,+,-) (times 1/sqrt 3)
        const int r = q & 3;
        const float x = (r < 2 ? -g : g), y = (q < 4 ? -g : g), z = ((r == 1 || r == 2) ? g : -g);
        const float xy = x * y, yz = y * z, xz = x * z;
        // natural gradients: d/dx = c1 + c3 y + c5 z + c7 yz, d/dy = c2 + c3 x + c6 z + c7 xz,
        // d/dz = c4 + c5 x + c6 y + c7 xy   (index = canonical monomial bits)
        float H[9], J[9];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            H[i * 3 + 0] = cp[i][1] + cp[i][3] * y + cp[i][5] * z + cp[i][7] * yz;
            H[i * 3 + 1] = cp[i][2] + cp[i][3] * x + cp[i][6] * z + cp[i][7] * xz;
            H[i * 3 + 2] = cp[i][4] + cp[i][5] * x + cp[i][6] * y + cp[i][7] * xy;
            J[0 * 3 + i] = cx[i][1] + cx[i][3] * y + cx[i][5] * z + cx[i][7] * yz;
            J[1 * 3 + i] = cx[i][2] + cx[i][3] * x + cx[i][6] * z + cx[i][7] * xz;
            J[2 * 3 + i] = cx[i][4] + cx[i][5] * x + cx[i][6] * y + cx[i][7] * xy;
        }
        const float det = J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) +
                          J[2] * (J[3] * J[7] - J[4] * J[6]);
        const float id = 1.0f / det;
        const float iJ[9] = {(J[4] * J[8] - J[5] * J[7]) * id, (J[2] * J[7] - J[1] * J[8]) * id,
                             (J[1] * J[5] - J[2] * J[4]) * id, (J[5] * J[6] - J[3] * J[8]) * id,
                             (J[0] * J[8] - J[2] * J[6]) * id, (J[2] * J[3] - J[0] * J[5]) * id,
                             (J[3] * J[7] - J[4] * J[6]) * id, (J[1] * J[6] - J[0] * J[7]) * id,
                             (J[0] * J[4] - J[1] * J[3]) * id};
        float gr[9];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                gr[i * 3 + j] = H[i * 3] * iJ[j * 3] + H[i * 3 + 1] * iJ[j * 3 + 1] + H[i * 3 + 2] * iJ[j * 3 + 2];
        const float ep[6] = {gr[0], gr[4], gr[8], gr[1] + gr[3], gr[2] + gr[6], gr[5] + gr[7]};
        float cf[NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) cf[k] = __builtin_nontemporal_load(mfC + (int64_t)(q * NC + k) * E + e);
        const float A = cf[0], B = cf[1], cnn = cf[2];
        const float cn1 = NC == 10 ? cf[3] : 0.0f, c1n = NC == 10 ? cf[4] : 0.0f;
        const float* nc = cf + (NC == 10 ? 5 : 3);
        const float n[6] = {nc[0], nc[1], -(nc[0] + nc[1]), nc[2], nc[3], nc[4]};  // n deviatoric
        const float tr = ep[0] + ep[1] + ep[2];
        const float ne = n[0] * ep[0] + n[1] * ep[1] + n[2] * ep[2] + n[3] * ep[3] + n[4] * ep[4] + n[5] * ep[5];
        const float t3 = tr * (1.0f / 3.0f);
        const float sn = cnn * ne + cn1 * tr;
        const float s1 = A * tr + c1n * ne;
        float sg[6];
#pragma unroll
        for (int i = 0; i < 3; ++i) sg[i] = s1 + B * (ep[i] - t3) + sn * n[i];
#pragma unroll
        for (int i = 3; i < 6; ++i) sg[i] = 0.5f * B * ep[i] + sn * n[i];
        const float S[9] = {sg[0], sg[3], sg[4], sg[3], sg[1], sg[5], sg[4], sg[5], sg[2]};
        // f_a,i += sum_k M_ik dN_a/dxi_k with M_ik = sum_j sig_ij iJ_jk, in monomials:
        // fa[1] += M0, fa[2] += M1, fa[4] += M2, fa[3] += M0 y + M1 x, fa[5] += M0 z + M2 x,
        // fa[6] += M1 z + M2 y, fa[7] += M0 yz + M1 xz + M2 xy   (M_k = column k, x 1/8)
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const float m0 = S[i * 3] * iJ[0] + S[i * 3 + 1] * iJ[3] + S[i * 3 + 2] * iJ[6];
            const float m1 = S[i * 3] * iJ[1] + S[i * 3 + 1] * iJ[4] + S[i * 3 + 2] * iJ[7];
            const float m2 = S[i * 3] * iJ[2] + S[i * 3 + 1] * iJ[5] + S[i * 3 + 2] * iJ[8];
            fa[i][1] += m0;
            fa[i][2] += m1;
            fa[i][4] += m2;
            fa[i][3] += m0 * y + m1 * x;
            fa[i][5] += m0 * z + m2 * x;
            fa[i][6] += m1 * z + m2 * y;
            fa[i][7] += m0 * yz + m1 * xz + m2 * xy;
        }
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) walsh3_t(fa[i]);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int a = kHex8Canon[k];
        float* dst = Fs + (int64_t)pos[(int64_t)a * E + e] * 3;
#pragma unroll
        for (int i = 0; i < 3; ++i) dst[i] = 0.125f * fa[i][k];
    }
}

// Hexa8 product by workgroup bricks: workgroup g takes the run of up to 256 consecutive elements
// e0s[g] .. e0s[g + 1] - 1 (every 256 elements; in batched contexts a run never crosses a
// specimen boundary).  Its local nodes (the distinct nodes of those elements, ascending) are
// staged once into LDS with coalesced loads: p (Dirichlet DOFs masked) and the FP64
// coordinates.  Each thread takes its element's 8 nodes from LDS (local indices lidx), forms the
// element forces exactly as mf_apply_hex8_kernel, and writes them to LDS; then each local node's
// forces are summed in element order (incidence list linc) into one partial per (workgroup,
// node), stored at its slot lpos of the node-ordered partial list.  The gather then sums a
// node's partials (1 to 8, ~2.4 on a box) in workgroup order: the element-order sum of the
// per-slot path, grouped by workgroup.  Versus one slot per (element, node): 8 p and 8 coordinate
// gathers per element become ~2.5 staged loads, and 8 scattered 12-byte stores become ~2.5.
constexpr int kBrick = 256;       // elements per workgroup
constexpr int kBrickNodes = 768;  // local nodes a workgroup can stage (an interior hexa8 pencil: 650)

template <int NC, typename TP>
__global__ __launch_bounds__(kBrick) void mfb_apply_hex8_kernel(
    int E, int N, const int32_t* __restrict__ lptr, const int32_t* __restrict__ lnode,
    const int32_t* __restrict__ lpos, const int32_t* __restrict__ lincp, const uint16_t* __restrict__ linc,
    const uint16_t* __restrict__ lidx, const double* __restrict__ coords, const float* __restrict__ mfC,
    const TP* __restrict__ p, const uint8_t* __restrict__ dirflag, int mask, const int* __restrict__ flags,
    float* __restrict__ Fpart, int G, const int32_t* __restrict__ e0s, const int* __restrict__ active,
    const int32_t* __restrict__ specOfElem, double* __restrict__ part) {
    // LDS, reused: first the staged node data (p 3 x 768 floats, x 3 x 768 doubles), then the
    // element forces (256 x 8 x 3 floats)
    __shared__ double lds[3 * kBrickNodes + (3 * kBrickNodes + 1) / 2];
    __shared__ double red[kBrick / 64];
    double* sx = lds;
    float* sp = (float*)(lds + 3 * kBrickNodes);
    float* sf = (float*)lds;
    if (flags && flags[0]) return;
    const int g = xcd_block_mf();
    if (g >= G) {  // the grid is rounded up to a multiple of 8
        if (part && threadIdx.x == 0) part[g] = 0.0;
        return;
    }
    const int eb = e0s[g], ee = e0s[g + 1];
    // batched solves: a brick of a frozen specimen leaves its partials alone (the gather skips
    // that specimen's nodes)
    if (active && !active[specOfElem[eb]]) return;
    const int t = threadIdx.x;
    const int l0 = lptr[g], nl = lptr[g + 1] - l0;
    for (int i = t; i < nl; i += kBrick) {
        const int nd = lnode[l0 + i];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            float v = (float)p[3 * nd + c];
            if (mask && dirflag[3 * nd + c]) v = 0.0f;
            sp[c * kBrickNodes + i] = v;
            sx[c * kBrickNodes + i] = coords[(int64_t)c * N + nd];
        }
    }
    __syncthreads();
    const int e = eb + t;
    const bool live = e < ee;
    float cp[3][8], cx[3][8];
    if (live) {
        // 8 local indices in one 16-byte load
        const uint4 li4 = *reinterpret_cast<const uint4*>(lidx + (int64_t)e * 8);
        const uint32_t lw[4] = {li4.x, li4.y, li4.z, li4.w};
        int loc[8];
#pragma unroll
        for (int a = 0; a < 8; ++a) loc[a] = (int)((lw[a >> 1] >> (16 * (a & 1))) & 0xffffu);
        double x0[3];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int l = loc[kHex8Canon[k]];
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                cp[i][k] = sp[i * kBrickNodes + l];
                const double xc = sx[i * kBrickNodes + l];
                if (k == 0) x0[i] = xc;
                cx[i][k] = (float)(xc - x0[i]);
            }
        }
    }
    __syncthreads();  // the staging area becomes the force area
    // with part given: the brick's sum of p_e . f_e (its share of p . K p), formed in the
    // monomial basis: f = H^T fa / 8 and cp = H p / 8 with H H = 8 I give p . f = cp . fa
    double en = 0.0;
    if (live) {
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            walsh3(cp[i]);
            walsh3(cx[i]);
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                cp[i][k] *= 0.125f;
                cx[i][k] *= 0.125f;
            }
        }
        float fa[3][8];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int k = 0; k < 8; ++k) fa[i][k] = 0.0f;
        const float gq = 0.57735026918962576451f;
#pragma unroll 1
        for (int q = 0; q < 8; ++q) {
            const int r = q & 3;
            const float x = (r < 2 ? -gq : gq), y = (q < 4 ? -gq : gq), z = ((r == 1 || r == 2) ? gq : -gq);
            const float xy = x * y, yz = y * z, xz = x * z;
            float H[9], J[9];
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                H[i * 3 + 0] = cp[i][1] + cp[i][3] * y + cp[i][5] * z + cp[i][7] * yz;
                H[i * 3 + 1] = cp[i][2] + cp[i][3] * x + cp[i][6] * z + cp[i][7] * xz;
                H[i * 3 + 2] = cp[i][4] + cp[i][5] * x + cp[i][6] * y + cp[i][7] * xy;
                J[0 * 3 + i] = cx[i][1] + cx[i][3] * y + cx[i][5] * z + cx[i][7] * yz;
                J[1 * 3 + i] = cx[i][2] + cx[i][3] * x + cx[i][6] * z + cx[i][7] * xz;
                J[2 * 3 + i] = cx[i][4] + cx[i][5] * x + cx[i][6] * y + cx[i][7] * xy;
            }
            const float det = J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) +
                              J[2] * (J[3] * J[7] - J[4] * J[6]);
            const float id = 1.0f / det;
            const float iJ[9] = {(J[4] * J[8] - J[5] * J[7]) * id, (J[2] * J[7] - J[1] * J[8]) * id,
                                 (J[1] * J[5] - J[2] * J[4]) * id, (J[5] * J[6] - J[3] * J[8]) * id,
                                 (J[0] * J[8] - J[2] * J[6]) * id, (J[2] * J[3] - J[0] * J[5]) * id,
                                 (J[3] * J[7] - J[4] * J[6]) * id, (J[1] * J[6] - J[0] * J[7]) * id,
                                 (J[0] * J[4] - J[1] * J[3]) * id};
            float gr[9];
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j)
                    gr[i * 3 + j] = H[i * 3] * iJ[j * 3] + H[i * 3 + 1] * iJ[j * 3 + 1] + H[i * 3 + 2] * iJ[j * 3 + 2];
            const float ep[6] = {gr[0], gr[4], gr[8], gr[1] + gr[3], gr[2] + gr[6], gr[5] + gr[7]};
            float cf[NC];
#pragma unroll
            for (int k = 0; k < NC; ++k) cf[k] = __builtin_nontemporal_load(mfC + (int64_t)(q * NC + k) * E + e);
            const float A = cf[0], B = cf[1], cnn = cf[2];
            const float cn1 = NC == 10 ? cf[3] : 0.0f, c1n = NC == 10 ? cf[4] : 0.0f;
            const float* nc = cf + (NC == 10 ? 5 : 3);
            const float n[6] = {nc[0], nc[1], -(nc[0] + nc[1]), nc[2], nc[3], nc[4]};  // n deviatoric
            const float tr = ep[0] + ep[1] + ep[2];
            const float ne = n[0] * ep[0] + n[1] * ep[1] + n[2] * ep[2] + n[3] * ep[3] + n[4] * ep[4] + n[5] * ep[5];
            const float t3 = tr * (1.0f / 3.0f);
            const float sn = cnn * ne + cn1 * tr;
            const float s1 = A * tr + c1n * ne;
            float sg[6];
#pragma unroll
            for (int i = 0; i < 3; ++i) sg[i] = s1 + B * (ep[i] - t3) + sn * n[i];
#pragma unroll
            for (int i = 3; i < 6; ++i) sg[i] = 0.5f * B * ep[i] + sn * n[i];
            const float S[9] = {sg[0], sg[3], sg[4], sg[3], sg[1], sg[5], sg[4], sg[5], sg[2]};
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const float m0 = S[i * 3] * iJ[0] + S[i * 3 + 1] * iJ[3] + S[i * 3 + 2] * iJ[6];
                const float m1 = S[i * 3] * iJ[1] + S[i * 3 + 1] * iJ[4] + S[i * 3 + 2] * iJ[7];
                const float m2 = S[i * 3] * iJ[2] + S[i * 3 + 1] * iJ[5] + S[i * 3 + 2] * iJ[8];
                fa[i][1] += m0;
                fa[i][2] += m1;
                fa[i][4] += m2;
                fa[i][3] += m0 * y + m1 * x;
                fa[i][5] += m0 * z + m2 * x;
                fa[i][6] += m1 * z + m2 * y;
                fa[i][7] += m0 * yz + m1 * xz + m2 * xy;
            }
        }
        if (part)
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int k = 0; k < 8; ++k) en += (double)cp[i][k] * (double)fa[i][k];
#pragma unroll
        for (int i = 0; i < 3; ++i) walsh3_t(fa[i]);
        // forces [i][t * 8 + a] (component-major: lanes t write 8 apart, then the node sums read
        // runs of one component)
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int a = kHex8Canon[k];
#pragma unroll
            for (int i = 0; i < 3; ++i) sf[i * (kBrick * 8) + t * 8 + a] = 0.125f * fa[i][k];
        }
    }
    __syncthreads();
    // one partial per local node: its incidences in element order
    for (int i = t; i < nl; i += kBrick) {
        float f0 = 0.0f, f1 = 0.0f, f2 = 0.0f;
        for (int c = lincp[l0 + i]; c < lincp[l0 + i + 1]; ++c) {
            const int s = linc[c];
            f0 += sf[s];
            f1 += sf[kBrick * 8 + s];
            f2 += sf[2 * kBrick * 8 + s];
        }
        float* dst = Fpart + (int64_t)lpos[l0 + i] * 3;
        dst[0] = f0;
        dst[1] = f1;
        dst[2] = f2;
    }
    if (part) {
        const double sum = block_sum_mf<kBrick>(en, red);
        if (threadIdx.x == 0) part[g] = sum;
    }
}

// Batched solves: q = K p on the nodes of active specimens and the per-node p.q values
// that the per-specimen reductions sum; nodes of frozen specimens are left alone (p.q 0).
template <bool DIRECT>
__global__ __launch_bounds__(256) void mf_gather_batched_kernel(int N, const int32_t* __restrict__ nptr,
                                                                const float* __restrict__ Fe,
                                                                const float* __restrict__ p,
                                                                const uint8_t* __restrict__ dirflag,
                                                                const int32_t* __restrict__ spec,
                                                                const int* __restrict__ active, float* __restrict__ q,
                                                                double* __restrict__ npq) {
    __shared__ float buf[DIRECT ? 1 : kGatherLds];
    const int n0 = blockIdx.x * 256, n = n0 + threadIdx.x;
    const bool valid = n < N;
    const bool act = valid && active[spec[n]];
    float yv[3];
    if (DIRECT) gather_direct(n, nptr, Fe, act, yv);
    else gather_slots<256>(n0, N, nptr, Fe, act, buf, yv);
    if (!valid) return;
    if (!act) {
        npq[n] = 0.0;
        return;
    }
    double dot = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float pv = p[3 * n + i];
        if (dirflag[3 * n + i]) yv[i] = pv;
        q[3 * n + i] = yv[i];
        dot += (double)pv * yv[i];
    }
    npq[n] = dot;
}

// MODE 0 (operator): y = sum of the node's element forces in element order on free DOFs,
// y = p on Dirichlet DOFs, per-workgroup partials of p.y.  MODE 1 (correction): y = the
// sum on free DOFs, 0 on Dirichlet DOFs.
template <int BS, int MODE, typename TP, typename TY, bool DIRECT = false>
__global__ __launch_bounds__(BS) void mf_gather_kernel(int N, const int32_t* __restrict__ nptr,
                                                       const int32_t* __restrict__ nlist,
                                                       const float* __restrict__ Fe, const TP* __restrict__ p,
                                                       const uint8_t* __restrict__ dirflag,
                                                       TY* __restrict__ y, double* __restrict__ part,
                                                       const int* __restrict__ flags, int kr) {
    __shared__ float buf[DIRECT ? 1 : kGatherLds];
    __shared__ double sh[BS / 64];
    if (flags && flags[0]) return;
    const int blk = xcd_block_mf();
    double dot = 0.0;
    for (int it = 0; it < kr; ++it) {
        const int n0 = (blk * kr + it) * BS;
        if (n0 >= N) break;
        const int n = n0 + threadIdx.x;
        const bool valid = n < N;
        float yv[3];
        if (DIRECT) gather_direct(n, nptr, Fe, valid, yv);
        else gather_slots<BS>(n0, N, nptr, Fe, valid, buf, yv);
        if (!valid) continue;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const bool d = dirflag[3 * n + i];  // always allocated; not ndir, which a captured graph would freeze
            if (MODE == 0) {
                const float pv = (float)p[3 * n + i];
                if (d) yv[i] = pv;
                dot += (double)pv * yv[i];
            } else if (d) {
                yv[i] = 0.0f;
            }
            y[3 * n + i] = (TY)yv[i];
        }
    }
    if (MODE == 0 && part) {
        const double s = block_sum_mf<BS>(dot, sh);
        if (threadIdx.x == 0) part[blk] = s;
    }
}

// Diagonal 3x3 blocks (element order), symmetric Dirichlet elimination inside the block,
// block-Jacobi inverse, and the P / F gather of the element residuals.
__global__ __launch_bounds__(256) void mf_diag_kernel(
    int N, const int32_t* __restrict__ nptr, const int32_t* __restrict__ nlist, const float* __restrict__ Kscr,
    const double* __restrict__ Pscr, const uint8_t* __restrict__ dirflag, int ndir, float* __restrict__ Minv,
    double* __restrict__ P, double* __restrict__ F, const int* __restrict__ specActive,
    const int32_t* __restrict__ specOfNode) {
    // XCD-aware numbering: each XCD sums a contiguous node range, so the element blocks that
    // neighbouring node rows share are fetched into one L2, not into several
    const int n = xcd_block_mf() * blockDim.x + threadIdx.x;
    if (n >= N) return;
    if (specActive && !specActive[specOfNode[n]]) return;  // a skipped specimen keeps its rows
    // K_nn = sum of the K_aa blocks [e][a][9] of the (element, local node) pairs of node n,
    // in element order (the order of the diagonal block's contributions in the assembly)
    float acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll 4
    for (int c = nptr[n]; c < nptr[n + 1]; ++c) add_block(Kscr + (int64_t)nlist[c] * kBlockPitch, acc);
    double dg[9];
#pragma unroll
    for (int v = 0; v < 9; ++v) dg[v] = acc[v];
    if (ndir) {
        const bool fr[3] = {(bool)dirflag[3 * n], (bool)dirflag[3 * n + 1], (bool)dirflag[3 * n + 2]};
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                if (fr[i] || fr[j]) dg[i * 3 + j] = (i == j) ? 1.0 : 0.0;
    }
    const double det = dg[0] * (dg[4] * dg[8] - dg[5] * dg[7]) - dg[1] * (dg[3] * dg[8] - dg[5] * dg[6]) +
                       dg[2] * (dg[3] * dg[7] - dg[4] * dg[6]);
    const double id = 1.0 / det;
    const double iv[9] = {(dg[4] * dg[8] - dg[5] * dg[7]) * id, (dg[2] * dg[7] - dg[1] * dg[8]) * id,
                          (dg[1] * dg[5] - dg[2] * dg[4]) * id, (dg[5] * dg[6] - dg[3] * dg[8]) * id,
                          (dg[0] * dg[8] - dg[2] * dg[6]) * id, (dg[2] * dg[3] - dg[0] * dg[5]) * id,
                          (dg[3] * dg[7] - dg[4] * dg[6]) * id, (dg[1] * dg[6] - dg[0] * dg[7]) * id,
                          (dg[0] * dg[4] - dg[1] * dg[3]) * id};
#pragma unroll
    for (int v = 0; v < 9; ++v) Minv[(int64_t)n * 9 + v] = (float)iv[v];
    double p0 = 0, p1 = 0, p2 = 0, f0 = 0, f1 = 0, f2 = 0;
#pragma unroll 4
    for (int c = nptr[n]; c < nptr[n + 1]; ++c) {
        const double* src = Pscr + (int64_t)nlist[c] * 3;
        p0 += src[0];
        p1 += src[1];
        p2 += src[2];
        f0 += fabs(src[0]);
        f1 += fabs(src[1]);
        f2 += fabs(src[2]);
    }
    P[3 * n] = p0;
    P[3 * n + 1] = p1;
    P[3 * n + 2] = p2;
    F[3 * n] = f0;
    F[3 * n + 1] = f1;
    F[3 * n + 2] = f2;
}

inline int round8_mf(int n) { return (n + 7) & ~7; }

#define MF_ARGS \
    c->E, c->N, c->conn, c->coords, c->dnxi32, c->mfC, p, c->dirflag, mask, flags, c->elemPos, c->mfF, active, \
        c->specOfElem

template <typename TP>
void launch_apply(efe_ctx* c, hipStream_t st, const TP* p, int mask, const int* flags,
                  const int* active = nullptr, double* part = nullptr) {
    const int grid = round8_mf((c->E + 255) / 256);
    const bool dp = c->mfNC == 10;
    if (c->mfb) {  // bricks (hexa8); grid: the bricks, rounded up to a multiple of 8
        const int gb = round8_mf(c->mfbG);
#define MFB_ARGS                                                                                                 \
    c->E, c->N, c->mfbLptr, c->mfbLnode, c->mfbLpos, c->mfbLincp, c->mfbLinc, c->mfbLidx, c->coords, c->mfC, p, \
        c->dirflag, mask, flags, c->mfbF, c->mfbG, c->mfbE0, active, c->specOfElem, part
        if (dp) mfb_apply_hex8_kernel<10, TP><<<gb, kBrick, 0, st>>>(MFB_ARGS);
        else mfb_apply_hex8_kernel<8, TP><<<gb, kBrick, 0, st>>>(MFB_ARGS);
#undef MFB_ARGS
        return;
    }
    switch (c->eltype) {
        case EFE_C3D8:
            if (dp) mf_apply_hex8_kernel<10, TP><<<grid, 256, 0, st>>>(MF_ARGS);
            else mf_apply_hex8_kernel<8, TP><<<grid, 256, 0, st>>>(MF_ARGS);
            break;
        case EFE_C3D20:
            if (dp) mf_apply_kernel<20, 27, 10, TP><<<grid, 256, 0, st>>>(MF_ARGS);
            else mf_apply_kernel<20, 27, 8, TP><<<grid, 256, 0, st>>>(MF_ARGS);
            break;
        default:
            if (dp) mf_apply_kernel<20, 8, 10, TP><<<grid, 256, 0, st>>>(MF_ARGS);
            else mf_apply_kernel<20, 8, 8, TP><<<grid, 256, 0, st>>>(MF_ARGS);
            break;
    }
}

// the per-node force list the gather sums: brick partials, or one slot per (element, node)
inline const int32_t* gather_ptr(const efe_ctx* c) { return c->mfb ? c->mfbPptr : c->nptr; }
inline const float* gather_forces(const efe_ctx* c) { return c->mfb ? c->mfbF : c->mfF; }

template <int BS>
void launch_gather_op(efe_ctx* c, hipStream_t st, const float* p, float* q, double* part, const int* flags,
                      int nwg, int kr) {
    if (c->mfb)
        mf_gather_kernel<BS, 0, float, float, true><<<nwg, BS, 0, st>>>(c->N, gather_ptr(c), c->nlist,
                                                                        gather_forces(c), p, c->dirflag, q, part,
                                                                        flags, kr);
    else
        mf_gather_kernel<BS, 0, float, float><<<nwg, BS, 0, st>>>(c->N, gather_ptr(c), c->nlist,
                                                                  gather_forces(c), p, c->dirflag, q, part,
                                                                  flags, kr);
}

}  // namespace

namespace {
// Brick tables of the hexa8 operator; left off when a workgroup's elements touch more than
// kBrickNodes nodes (a mesh in an order without locality).  Bricks are runs of up to kBrick
// consecutive elements; in batched contexts every specimen's elements start a new brick, so
// that a specimen's partial sums are grouped the same alone and in any batch (its elements
// must then be one contiguous run per specimen, else the per-slot operator stays).
int mfb_build(efe_ctx* c) {
    if (c->eltype != EFE_C3D8 || c->mfb) return EFE_OK;
    const int E = c->E, N = c->N;
    std::vector<int32_t> e0s(1, 0);
    if (c->specOfNode) {
        std::vector<int32_t> se(E);
        EFE_HIP(hipMemcpy(se.data(), c->specOfElem, (size_t)E * sizeof(int32_t), hipMemcpyDeviceToHost));
        for (int e = 1; e < E; ++e) {
            if (se[e] < se[e - 1]) return EFE_OK;  // specimens' elements not in runs
            if (se[e] != se[e - 1] || e - e0s.back() == kBrick) e0s.push_back(e);
        }
    } else {
        for (int e = kBrick; e < E; e += kBrick) e0s.push_back(e);
    }
    e0s.push_back(E);
    const int G = (int)e0s.size() - 1;
    std::vector<int32_t> conn((size_t)8 * E);  // [a][e]
    EFE_HIP(hipMemcpy(conn.data(), c->conn, conn.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    std::vector<int32_t> lptr(G + 1, 0), lnode, lincp, pcount(N, 0);
    std::vector<uint16_t> linc((size_t)8 * E), lidx((size_t)8 * E);
    lnode.reserve((size_t)3 * E);
    lincp.reserve((size_t)3 * E + 1);
    std::vector<int32_t> mark(N, -1);
    std::vector<int32_t> nodes, cnt, fill;
    for (int g = 0; g < G; ++g) {
        const int e0 = e0s[g], e1 = e0s[g + 1];
        nodes.clear();
        for (int e = e0; e < e1; ++e)
            for (int a = 0; a < 8; ++a) nodes.push_back(conn[(size_t)a * E + e]);
        std::sort(nodes.begin(), nodes.end());
        nodes.erase(std::unique(nodes.begin(), nodes.end()), nodes.end());
        const int nl = (int)nodes.size();
        if (nl > kBrickNodes) return EFE_OK;  // no locality: keep the per-slot operator
        for (int i = 0; i < nl; ++i) mark[nodes[i]] = i;
        cnt.assign(nl + 1, 0);
        for (int e = e0; e < e1; ++e)
            for (int a = 0; a < 8; ++a) {
                const int l = mark[conn[(size_t)a * E + e]];
                lidx[(size_t)e * 8 + a] = (uint16_t)l;
                ++cnt[l + 1];
            }
        for (int i = 0; i < nl; ++i) cnt[i + 1] += cnt[i];
        const int base = e0 * 8;  // linc entries of workgroup g start at its first element's
        fill.assign(cnt.begin(), cnt.end() - 1);
        for (int e = e0; e < e1; ++e)  // element order inside every local node's list
            for (int a = 0; a < 8; ++a) {
                const int l = mark[conn[(size_t)a * E + e]];
                linc[(size_t)base + fill[l]++] = (uint16_t)((e - e0) * 8 + a);
            }
        for (int i = 0; i < nl; ++i) {
            lnode.push_back(nodes[i]);
            lincp.push_back(base + cnt[i]);
            ++pcount[nodes[i]];
            mark[nodes[i]] = -1;
        }
        lptr[g + 1] = (int32_t)lnode.size();
    }
    lincp.push_back(8 * E);
    // partial slots in node order, the workgroups of a node in ascending order
    const int64_t L = (int64_t)lnode.size();
    std::vector<int32_t> pptr(N + 1, 0), lpos(L);
    for (int n = 0; n < N; ++n) pptr[n + 1] = pptr[n] + pcount[n];
    std::vector<int32_t> next(pptr.begin(), pptr.end() - 1);
    for (int64_t i = 0; i < L; ++i) lpos[i] = next[lnode[i]]++;
    // the end of each local node's incidence list is the start of the next one's; the last
    // local node of a workgroup ends where the next workgroup's first begins (lincp[L] = 8 E)
    auto up = [&](auto** dst, const auto& v) -> int {
        using T = typename std::remove_reference<decltype(v[0])>::type;
        EFE_HIP(hipMalloc((void**)dst, (v.size() ? v.size() : 1) * sizeof(T)));
        if (!v.empty()) EFE_HIP(hipMemcpy((void*)*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
        return EFE_OK;
    };
    int rc;
    if ((rc = up(&c->mfbLptr, lptr)) || (rc = up(&c->mfbLnode, lnode)) || (rc = up(&c->mfbLpos, lpos)) ||
        (rc = up(&c->mfbLincp, lincp)) || (rc = up(&c->mfbPptr, pptr)) || (rc = up(&c->mfbLinc, linc)) ||
        (rc = up(&c->mfbLidx, lidx)) || (rc = up(&c->mfbE0, e0s)))
        return rc;
    EFE_HIP(hipMalloc(&c->mfbF, (size_t)L * 3 * sizeof(float)));
    c->mfbG = G;
    c->mfb = true;
    return EFE_OK;
}
}  // namespace

int efe_mf_alloc(efe_ctx* c) {
    const int NC = c->mfNC;
    if (c->mfC) return EFE_OK;
    const size_t EQ = (size_t)c->E * c->Q;
    const int64_t L = (int64_t)c->E * c->NN;
    EFE_HIP(hipMalloc(&c->mfC, EQ * NC * sizeof(float)));
    EFE_HIP(hipMalloc(&c->mfF, (size_t)L * 3 * sizeof(float)));
    if (!c->Minv32) EFE_HIP(hipMalloc(&c->Minv32, (size_t)c->N * 9 * sizeof(float)));
    // FP32 copies of the natural shape-function derivatives at the evaluation points and of
    // the table J is formed with, one after the other
    const size_t nd = (size_t)c->Q * c->NN * 3;
    std::vector<double> h(2 * nd);
    std::vector<float> hf(2 * nd);
    EFE_HIP(hipMemcpy(h.data(), c->dNdxi, nd * sizeof(double), hipMemcpyDeviceToHost));
    EFE_HIP(hipMemcpy(h.data() + nd, c->dJdxi, nd * sizeof(double), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < 2 * nd; ++i) hf[i] = (float)h[i];
    EFE_HIP(hipMalloc(&c->dnxi32, 2 * nd * sizeof(float)));
    EFE_HIP(hipMemcpy(c->dnxi32, hf.data(), 2 * nd * sizeof(float), hipMemcpyHostToDevice));
    return mfb_build(c);
}

// Diagonal blocks, Minv, P, F and (optionally) the right-hand-side correction K[:, dir] dir_values.
int efe_launch_mf_assembly(efe_ctx* c, double* P, double* F, const double* dir_values, double* dir_correction) {
    const int grid = (c->N + 255) / 256;
    mf_diag_kernel<<<round8_mf(grid), 256, 0, c->stream>>>(c->N, c->nptr, c->nlist, (const float*)c->Kscr, c->Pscr,
                                                           c->dirflag, c->ndir, c->Minv32, P, F,
                                                           c->useActive ? c->specActive : nullptr, c->specOfNode);
    EFE_HIP(hipGetLastError());
    if (dir_values) {
        launch_apply<double>(c, c->stream, dir_values, 0, nullptr);
        if (c->mfb)
            mf_gather_kernel<256, 1, double, double, true><<<round8_mf(grid), 256, 0, c->stream>>>(
                c->N, gather_ptr(c), c->nlist, gather_forces(c), dir_values, c->dirflag, dir_correction,
                nullptr, nullptr, 1);
        else
            mf_gather_kernel<256, 1, double, double><<<round8_mf(grid), 256, 0, c->stream>>>(
                c->N, gather_ptr(c), c->nlist, gather_forces(c), dir_values, c->dirflag, dir_correction,
                nullptr, nullptr, 1);
        EFE_HIP(hipGetLastError());
    }
    c->Kprec = EFE_FP32_MATFREE;
    c->Kversion++;
    return EFE_OK;
}

// q = K p (FP32, Dirichlet eliminated) with per-workgroup partials of p.q, for the PCG:
// nwg workgroups of bs threads, the layout of the assembled SpMV's partials.
void efe_mf_launch_operator(efe_ctx* c, hipStream_t st, const float* p, float* q, double* part, const int* flags,
                            int nwg, int bs, int kr) {
    launch_apply<float>(c, st, p, 1, flags);
    switch (bs) {
        case 64: launch_gather_op<64>(c, st, p, q, part, flags, nwg, kr); break;
        case 128: launch_gather_op<128>(c, st, p, q, part, flags, nwg, kr); break;
        case 512: launch_gather_op<512>(c, st, p, q, part, flags, nwg, kr); break;
        default: launch_gather_op<256>(c, st, p, q, part, flags, nwg, kr); break;
    }
}

// The brick product alone (hexa8 bricks only): the node partials of K p for the caller's own
// gather, and per brick the partial of p . K p (Dirichlet DOFs masked) at part[g], g below
// round8(mfbG).
void efe_mf_launch_apply_partials(efe_ctx* c, hipStream_t st, const float* p, double* part, const int* flags) {
    launch_apply<float>(c, st, p, 1, flags, nullptr, part);
}

// y = K x for FP64 vectors with the matrix-free FP32 operator (Dirichlet eliminated).
int efe_launch_spmv_mf(efe_ctx* c, const double* x, double* y) {
    launch_apply<double>(c, c->stream, x, 1, nullptr);
    if (c->mfb)
        mf_gather_kernel<256, 0, double, double, true><<<round8_mf((c->N + 255) / 256), 256, 0, c->stream>>>(
            c->N, gather_ptr(c), c->nlist, gather_forces(c), x, c->dirflag, y, nullptr, nullptr, 1);
    else
        mf_gather_kernel<256, 0, double, double><<<round8_mf((c->N + 255) / 256), 256, 0, c->stream>>>(
            c->N, gather_ptr(c), c->nlist, gather_forces(c), x, c->dirflag, y, nullptr, nullptr, 1);
    EFE_HIP(hipGetLastError());
    return EFE_OK;
}

// Batched PCG: q = K p for the active specimens, per-node p.q into npq.
void efe_mf_launch_operator_batched(efe_ctx* c, hipStream_t st, const float* p, float* q, double* npq,
                                    const int* active) {
    launch_apply<float>(c, st, p, 1, nullptr, active);
    if (c->mfb)
        mf_gather_batched_kernel<true><<<(c->N + 255) / 256, 256, 0, st>>>(
            c->N, gather_ptr(c), gather_forces(c), p, c->dirflag, c->specOfNode, active, q, npq);
    else
        mf_gather_batched_kernel<false><<<(c->N + 255) / 256, 256, 0, st>>>(
            c->N, gather_ptr(c), gather_forces(c), p, c->dirflag, c->specOfNode, active, q, npq);
}

void efe_mf_free(efe_ctx* c) {
    hipFree(c->mfC);
    hipFree(c->mfF);
    hipFree(c->dnxi32);
    c->mfC = c->mfF = c->dnxi32 = nullptr;
    void* b[] = {c->mfbLptr, c->mfbLnode, c->mfbLpos, c->mfbLincp, c->mfbPptr, c->mfbLinc, c->mfbLidx, c->mfbF,
                 c->mfbE0};
    for (void* q : b)
        if (q) hipFree(q);
    c->mfbLptr = c->mfbLnode = c->mfbLpos = c->mfbLincp = c->mfbPptr = c->mfbE0 = nullptr;
    c->mfbLinc = c->mfbLidx = nullptr;
    c->mfbF = nullptr;
    c->mfb = false;
}
