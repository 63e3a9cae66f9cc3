// Element, material and assembly kernels for gfx950.
//
// element_kernel: one 64-lane wavefront owns EPW elements.
//   phase 1, one lane per (element, QP): Jacobian from the gathered node
//            coordinates and the shape-function derivatives at the reference's
//            evaluation point, dN/dx, strain increment from the gathered dU,
//            FP64 material update of the reference state into the temporary
//            state, and (w detJ) C and (w detJ) sigma staged in LDS;
//   phase 2, one lane per (element, node a, node b), b >= a for symmetric
//            tangents: the 3x3 block
//            K_ab = sum_q B_a^T (w detJ C_q) B_b in KT precision (and K_ba = K_ab^T), written to the
//            element scratch at position cpos[e][a][b] of the contribution order
//            (block slot, then element), so that a row's contributions are
//            contiguous for the assembly;
//   phase 3, one lane per (element, node a): Pe_a = -sum_q B_a^T (w detJ sigma_q).
//   Follows DisplacementElement.computeYourself
//   (edelweissfe/elements/displacementelement/element.py:290-346).
//
// assemble_kernel: one thread per node row-block.  For every block slot of its
//   row it sums the element blocks stored for that slot, in element order (the
//   order of updateCSR, csrgenerator.pyx:112-113), eliminates Dirichlet rows and
//   columns symmetrically, writes the sliced-ELL value layout, inverts its
//   diagonal block for block Jacobi, and gathers P = sum Pe and F = sum |Pe|
//   (nonlinearimplicitstatic.py:843-844) in element order.  No atomics: the
//   result is bitwise reproducible.
#include <cstdlib>

#include "efe_internal.h"
#include "materials.h"

namespace {

template <int MAT>
struct MatState {
    static constexpr int n = (MAT == EFE_LINEAR_ELASTIC) ? 0 : 1;
};

// MF (matrix-free form, KT = float): phase 1 also stores the compact tangent
// w detJ [A, B, cnn, (cn1, c1n,)] and the deviatoric n (five components) per QP for the
// matrix-free operator, and phase 2 forms
// only the diagonal blocks K_aa (for block Jacobi), stored element-major [e][a][9] at the
// start of the element scratch.
template <int NN, int Q, int MAT, typename KT, int EPW, bool MF>
__global__ __launch_bounds__(64) void element_kernel(int E, int N, const double* __restrict__ coords,
                                                     const int32_t* __restrict__ conn,
                                                     const double* __restrict__ props, int propStride,
                                                     const double* __restrict__ dNdxi,
                                                     const double* __restrict__ dJdxi,
                                                     const double* __restrict__ qpw,
                                                     const double* __restrict__ dU,
                                                     const double* __restrict__ sRef, double* __restrict__ sTmp,
                                                     KT* __restrict__ Kscr, double* __restrict__ Pscr,
                                                     int* __restrict__ status, float* __restrict__ mfC,
                                                     int* __restrict__ specStatus,
                                                     const int32_t* __restrict__ specOfElem,
                                                     const int32_t* __restrict__ cpos,
                                                     const int* __restrict__ specActive, int formK,
                                                     const int32_t* __restrict__ elist, int ecount, int ncw) {
    constexpr int NS = 12 + MatState<MAT>::n;
    constexpr int NP = (MAT == EFE_LINEAR_ELASTIC) ? 2 : 6;
    // the tangent of every QP in compact form, w detJ [A, B, c_nn, (c_n1, c_1n)] then n:
    // C = A 1(x)1 + B Idev + c_nn n(x)n (+ c_n1 n(x)1 + c_1n 1(x)n for Drucker-Prager)
    constexpr int NCC = (MAT == EFE_DRUCKER_PRAGER) ? 11 : 9;
    __shared__ double g_s[EPW][Q * NN * 3];
    __shared__ KT fC_s[EPW][Q * NCC];
    __shared__ double fs_s[EPW][Q][6];
    // assembled passes: blocks staged for word-by-word stores, and their positions (and those
    // of their transposes, -1: none)
    constexpr int kStage = 32;
    __shared__ KT stg_s[MF ? 1 : kStage][9];
    __shared__ int spos_s[MF ? 1 : 2 * kStage];
    const int lane = threadIdx.x;
    const int64_t EQ = (int64_t)E * Q;
    // Grid-stride over the element list: a grid smaller than the list (a persistent launch)
    // takes EPW elements per workgroup and round.
    for (int e0 = blockIdx.x * EPW; e0 < ecount; e0 += gridDim.x * EPW) {
    // Element el of this round: entry e0 + el of the group's element list (elist, or all
    // elements in order when it is null); -1 past the end and, in batched contexts, for
    // specimens the pass skips.
    auto elem = [&](int el) {
        const int i = e0 + el;
        if (i >= ecount) return -1;
        const int e = elist ? elist[i] : i;
        return (specActive && !specActive[specOfElem[e]]) ? -1 : e;
    };

    for (int t = lane; t < EPW * Q; t += 64) {
        const int el = t / Q, q = t - (t / Q) * Q;
        const int e = elem(el);
        if (e < 0) continue;
        const double* dn = dNdxi + (size_t)q * NN * 3;
        const double* dj = dJdxi + (size_t)q * NN * 3;
        double J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll 4
        for (int a = 0; a < NN; ++a) {
            const int nd = conn[(int64_t)a * E + e];
            const double x = coords[nd], y = coords[N + nd], z = coords[2 * N + nd];
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                const double d = dj[a * 3 + r];
                J[r * 3 + 0] += d * x;
                J[r * 3 + 1] += d * y;
                J[r * 3 + 2] += d * z;
            }
        }
        const double det = J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) +
                           J[2] * (J[3] * J[7] - J[4] * J[6]);
        const double id = 1.0 / det;
        double iJ[9];
        iJ[0] = (J[4] * J[8] - J[5] * J[7]) * id;
        iJ[1] = (J[2] * J[7] - J[1] * J[8]) * id;
        iJ[2] = (J[1] * J[5] - J[2] * J[4]) * id;
        iJ[3] = (J[5] * J[6] - J[3] * J[8]) * id;
        iJ[4] = (J[0] * J[8] - J[2] * J[6]) * id;
        iJ[5] = (J[2] * J[3] - J[0] * J[5]) * id;
        iJ[6] = (J[3] * J[7] - J[4] * J[6]) * id;
        iJ[7] = (J[1] * J[6] - J[0] * J[7]) * id;
        iJ[8] = (J[0] * J[4] - J[1] * J[3]) * id;
        double gr[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // gr[i*3+j] = d u_i / d x_j
#pragma unroll 4
        for (int a = 0; a < NN; ++a) {
            const double d0 = dn[a * 3 + 0], d1 = dn[a * 3 + 1], d2 = dn[a * 3 + 2];
            double g[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) g[c] = iJ[c * 3 + 0] * d0 + iJ[c * 3 + 1] * d1 + iJ[c * 3 + 2] * d2;
            g_s[el][(q * NN + a) * 3 + 0] = g[0];
            g_s[el][(q * NN + a) * 3 + 1] = g[1];
            g_s[el][(q * NN + a) * 3 + 2] = g[2];
            const int nd = conn[(int64_t)a * E + e];
            const double u0 = dU[3 * nd], u1 = dU[3 * nd + 1], u2 = dU[3 * nd + 2];
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                gr[0 + j] += u0 * g[j];
                gr[3 + j] += u1 * g[j];
                gr[6 + j] += u2 * g[j];
            }
        }
        double de[6] = {gr[0], gr[4], gr[8], gr[1] + gr[3], gr[2] + gr[6], gr[5] + gr[7]};
        const int64_t sidx = (int64_t)e * Q + q;
        double s[6], ms = 0.0;
#pragma unroll
        for (int c = 0; c < 6; ++c) s[c] = sRef[c * EQ + sidx];
        if (MatState<MAT>::n) ms = sRef[12 * EQ + sidx];
        double p[NP];
#pragma unroll
        for (int k = 0; k < NP; ++k) p[k] = propStride ? props[(int64_t)k * propStride + e] : props[k];
        double C[36], cc[11];
        int st = 0;
        if (MAT == EFE_LINEAR_ELASTIC) st = efe::le_update(p, s, de, C, cc);
        else if (MAT == EFE_VONMISES) st = efe::vm_update(p, s, ms, de, C, cc);
        else st = efe::dp_update(p, s, ms, de, C, cc);
        if (st) {
            *status = 1;
            if (specStatus) specStatus[specOfElem[e]] = 1;  // which specimen asked (batched contexts)
        }
        if (MF && formK) {  // ncw = 8 [A, B, cnn, n] or 10 [A, B, cnn, cn1, c1n, n] (cn1 = c1n = 0 but DP);
            // n is deviatoric (n_33 = -n_11 - n_22): its five other components are stored
            const double fq = det * qpw[q];
            float* dc = mfC + (int64_t)q * ncw * E + e;
            for (int k = 0; k < ncw - 5; ++k) dc[(int64_t)k * E] = (float)(fq * cc[k]);
            const int nk[5] = {0, 1, 3, 4, 5};
#pragma unroll
            for (int k = 0; k < 5; ++k) dc[(int64_t)(ncw - 5 + k) * E] = (float)cc[5 + nk[k]];
        }
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            sTmp[c * EQ + sidx] = s[c];
            sTmp[(6 + c) * EQ + sidx] = sRef[(6 + c) * EQ + sidx] + de[c];
        }
        if (MatState<MAT>::n) sTmp[12 * EQ + sidx] = ms;
        (void)NS;
        const double f = det * qpw[q];
        if (formK) {
#pragma unroll
            for (int k = 0; k < NCC - 6; ++k) fC_s[el][q * NCC + k] = (KT)(f * cc[k]);
#pragma unroll
            for (int k = 0; k < 6; ++k) fC_s[el][q * NCC + NCC - 6 + k] = (KT)cc[5 + k];
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) fs_s[el][q][k] = f * s[k];
    }
    __syncthreads();

    // Symmetric tangents (linear elastic, J2): only the blocks a <= b are formed; K_ba is
    // stored as K_ab^T.  Drucker-Prager (non-associative in general) forms all NN^2 blocks.
    constexpr bool SYM = !MF && MAT != EFE_DRUCKER_PRAGER;
    constexpr int NB = MF ? NN : (SYM ? NN * (NN + 1) / 2 : NN * NN);  // blocks per element formed here
    // a residual-only pass (formK = 0) leaves the tangent blocks alone
    // Lanes take an element's blocks one after the other.  (Taking the EPW elements of one
    // block position side by side, so that the stores form runs of adjacent blocks, was
    // measured 8 % slower at 64^3 with per-lane stores, as the tangent reads lose their
    // broadcast, and no faster with the staged stores below.)  The assembled
    // passes write their blocks through LDS, kStage blocks (and their transposes) at a time:
    // the wavefront then stores consecutive words, about seven 72-byte blocks per store
    // instruction instead of one scattered block per lane.  Measured on the 64^3 hexa8 J2
    // tangent pass: 1.00 ms with per-lane stores, 0.95 / 0.91 / 0.88 / 0.95 ms staging 8 / 16 /
    // 32 / 64 blocks (0.53 ms with no block stores at all); 48^3 hexa20R DP: 2.34 -> 2.27 ms.
    const int ntask = formK ? EPW * NB : 0;
    for (int t0 = 0; t0 < ntask; t0 += 64) {
        const int t = t0 + lane;
        const int el = t / NB, r0 = t - (t / NB) * NB;
        int a = 0, b = 0;
        if (MF) {
            a = b = r0;
        } else if (SYM) {  // r0 -> (a, b), upper triangle row by row
            int r = r0;
            while (r >= NN - a) {
                r -= NN - a;
                ++a;
            }
            b = a + r;
        } else {
            a = r0 / NN;
            b = r0 - a * NN;
        }
        const int e = t < ntask ? elem(el) : -1;
        KT acc[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) acc[k] = (KT)0;
        if (e >= 0) {
            for (int q = 0; q < Q; ++q) {
                const KT ax = (KT)g_s[el][(q * NN + a) * 3 + 0], ay = (KT)g_s[el][(q * NN + a) * 3 + 1], az = (KT)g_s[el][(q * NN + a) * 3 + 2];
                const KT bx = (KT)g_s[el][(q * NN + b) * 3 + 0], by = (KT)g_s[el][(q * NN + b) * 3 + 1], bz = (KT)g_s[el][(q * NN + b) * 3 + 2];
                // CB[:, j] = C B_b e_j from the compact tangent: the strain of column j has the
                // normal component b_j and two engineering shears (12/13, 12/23, 13/23)
                const KT* cq = fC_s[el] + q * NCC;
                const KT cA = cq[0], cB = cq[1], cnn = cq[2];
                const KT cn1 = NCC == 11 ? cq[3] : (KT)0, c1n = NCC == 11 ? cq[4] : (KT)0;
                const KT* n = cq + NCC - 6;
                const KT bv[3] = {bx, by, bz};
                KT CB[6][3];
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    KT v[6] = {0, 0, 0, 0, 0, 0};
                    v[j] = bv[j];
                    if (j == 0) { v[3] = by; v[4] = bz; }
                    if (j == 1) { v[3] = bx; v[5] = bz; }
                    if (j == 2) { v[4] = bx; v[5] = by; }
                    const KT tr = bv[j];
                    KT ne = n[j] * bv[j];
#pragma unroll
                    for (int k = 3; k < 6; ++k) ne += n[k] * v[k];
                    const KT sn = cnn * ne + cn1 * tr, s1 = cA * tr + c1n * ne, t3 = tr * (KT)(1.0 / 3.0);
#pragma unroll
                    for (int i = 0; i < 3; ++i) CB[i][j] = s1 + cB * (v[i] - t3) + sn * n[i];
#pragma unroll
                    for (int k = 3; k < 6; ++k) CB[k][j] = (KT)0.5 * cB * v[k] + sn * n[k];
                }
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    acc[0 + j] += ax * CB[0][j] + ay * CB[3][j] + az * CB[4][j];
                    acc[3 + j] += ay * CB[1][j] + ax * CB[3][j] + az * CB[5][j];
                    acc[6 + j] += az * CB[2][j] + ax * CB[4][j] + ay * CB[5][j];
                }
            }
        }
        if (MF) {  // K_aa element-major at the start of the scratch: [e][a][9]
            if (e >= 0) store_block(Kscr + ((int64_t)e * NN + a) * kBlockPitch, acc);
            continue;
        }
        // the block at position cp[a][b] and, for a symmetric tangent with a != b, its
        // transpose at cp[b][a]; staged kStage lanes at a time, then stored word by word
        const int32_t* cp = cpos + (int64_t)(e >= 0 ? e : 0) * NN * NN;
        const int p0 = e >= 0 ? cp[a * NN + b] : -1;
        const int p1 = (e >= 0 && SYM && a != b) ? cp[b * NN + a] : -1;
        const int nlive = ntask - t0 < 64 ? ntask - t0 : 64;
        for (int s0 = 0; s0 < nlive; s0 += kStage) {
            const int sl = lane - s0;
            if (sl >= 0 && sl < kStage) {
#pragma unroll
                for (int k = 0; k < 9; ++k) stg_s[sl][k] = acc[k];
                spos_s[sl] = p0;
                spos_s[kStage + sl] = p1;
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            const int nb = (nlive - s0 < kStage ? nlive - s0 : kStage);
            for (int w = lane; w < (SYM ? 2 : 1) * kStage * 9; w += 64) {
                const int bi = w / 9, v = w - bi * 9;
                const int src = bi < kStage ? bi : bi - kStage;
                const int pos = src < nb ? spos_s[bi] : -1;
                if (pos >= 0) {
                    const int vv = bi < kStage ? v : (v % 3) * 3 + v / 3;  // K_ba = K_ab^T
                    Kscr[(int64_t)pos * kBlockPitch + v] = stg_s[src][vv];
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }

    for (int t = lane; t < EPW * NN; t += 64) {
        const int el = t / NN, a = t - (t / NN) * NN;
        const int e = elem(el);
        if (e < 0) continue;
        double p0 = 0, p1 = 0, p2 = 0;
        for (int q = 0; q < Q; ++q) {
            const double ax = g_s[el][(q * NN + a) * 3 + 0], ay = g_s[el][(q * NN + a) * 3 + 1], az = g_s[el][(q * NN + a) * 3 + 2];
            const double* s = fs_s[el][q];
            p0 += ax * s[0] + ay * s[3] + az * s[4];
            p1 += ay * s[1] + ax * s[3] + az * s[5];
            p2 += az * s[2] + ax * s[4] + ay * s[5];
        }
        double* dst = Pscr + ((int64_t)e * NN + a) * 3;
        dst[0] = -p0;
        dst[1] = -p1;
        dst[2] = -p2;
    }
    __syncthreads();  // the LDS tiles are refilled by the next round
    }
}

__device__ __forceinline__ void inv3(const double m[9], double o[9]) {
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

template <typename KT>
__global__ __launch_bounds__(256) void assemble_kernel(
    int N, const int32_t* __restrict__ bptr, const int32_t* __restrict__ bcol, const int32_t* __restrict__ coff,
    const int32_t* __restrict__ nptr, const int32_t* __restrict__ nlist,
    const int32_t* __restrict__ slice_off, const KT* __restrict__ Kscr, const double* __restrict__ Pscr,
    const uint8_t* __restrict__ dirflag, int ndir, const double* __restrict__ dirv, double* __restrict__ dircorr,
    KT* __restrict__ Kval, KT* __restrict__ Minv, double* __restrict__ P, double* __restrict__ F,
    const int32_t* __restrict__ slotNode, const int32_t* __restrict__ cslice, const int* __restrict__ specActive,
    const int32_t* __restrict__ specOfNode) {
    // Threads walk the row slots of the sliced layout: the value stores of a wavefront are
    // contiguous, and so are its reads of the element scratch, where contribution j of the
    // 64 rows of a slice is one run of 64 blocks.
    const int per = gridDim.x >> 3;
    const int blk = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    const int ps = blk * blockDim.x + threadIdx.x;
    if (ps >= N) return;
    const int n = slotNode[ps];
    if (specActive && !specActive[specOfNode[n]]) {
        // a skipped specimen keeps its rows; its increment is not solved, so its right-hand
        // side correction is zero
        if (dirv) dircorr[3 * n] = dircorr[3 * n + 1] = dircorr[3 * n + 2] = 0.0;
        return;
    }
    const int s = ps / kSlice, l = ps - (ps / kSlice) * kSlice;
    const int64_t vbase = (int64_t)slice_off[s] * 9 + l;
    bool fr[3] = {false, false, false};
    if (ndir) {
        fr[0] = dirflag[3 * n];
        fr[1] = dirflag[3 * n + 1];
        fr[2] = dirflag[3 * n + 2];
    }
    double t0 = 0, t1 = 0, t2 = 0;
    double dg[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    const int b0 = bptr[n], b1 = bptr[n + 1];
    // block position of this row's contribution c: cbase + (c - coff[b0]) * kSlice
    const int64_t cbase = (int64_t)cslice[s] * kSlice + l - (int64_t)coff[b0] * kSlice;
    for (int b = b0; b < b1; ++b) {
        const int m = bcol[b];
        KT acc[9];
#pragma unroll
        for (int v = 0; v < 9; ++v) acc[v] = (KT)0;
        // the slot's contributions are contiguous and in element order; unrolled so that
        // the loads of several contributions are in flight together
#pragma unroll 4
        for (int c = coff[b]; c < coff[b + 1]; ++c) add_block(Kscr + (cbase + (int64_t)c * kSlice) * kBlockPitch, acc);
        if (dirv) {
            const double d0 = dirv[3 * m], d1 = dirv[3 * m + 1], d2 = dirv[3 * m + 2];
            t0 += (double)acc[0] * d0 + (double)acc[1] * d1 + (double)acc[2] * d2;
            t1 += (double)acc[3] * d0 + (double)acc[4] * d1 + (double)acc[5] * d2;
            t2 += (double)acc[6] * d0 + (double)acc[7] * d1 + (double)acc[8] * d2;
        }
        if (ndir) {
            const bool fc[3] = {(bool)dirflag[3 * m], (bool)dirflag[3 * m + 1], (bool)dirflag[3 * m + 2]};
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j)
                    if (fr[i] || fc[j]) acc[i * 3 + j] = (m == n && i == j) ? (KT)1 : (KT)0;
        }
        const int k = b - b0;
#pragma unroll
        for (int v = 0; v < 9; ++v) Kval[vbase + (int64_t)(k * 9 + v) * kSlice] = acc[v];
        if (m == n) {
#pragma unroll
            for (int v = 0; v < 9; ++v) dg[v] = (double)acc[v];
        }
    }
    double iv[9];
    inv3(dg, iv);
#pragma unroll
    for (int v = 0; v < 9; ++v) Minv[(int64_t)n * 9 + v] = (KT)iv[v];
    if (P) {
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
    if (dirv) {
        dircorr[3 * n] = fr[0] ? 0.0 : t0;
        dircorr[3 * n + 1] = fr[1] ? 0.0 : t1;
        dircorr[3 * n + 2] = fr[2] ? 0.0 : t2;
    }
}

// assemble_kernel with the element scratch read run by run: the wavefront of a slice loads
// contribution run j (the j-th contribution of its 64 rows, 64 contiguous blocks) with nine
// coalesced 512-byte loads (FP64), stages it in LDS and each lane takes its own block from
// there, loading run j + 1 meanwhile.  Every line of the scratch is fetched once (the per-lane
// nine-load form re-fetched lines its wavefronts had in flight when they were evicted).  The
// lanes walk their rows' contributions in the same order, so K, Minv, P, F and the correction are
// bitwise those of assemble_kernel.
template <typename KT>
__global__ __launch_bounds__(256) void assemble_runs_kernel(
    int N, const int32_t* __restrict__ bptr, const int32_t* __restrict__ bcol, const int32_t* __restrict__ coff,
    const int32_t* __restrict__ nptr, const int32_t* __restrict__ nlist,
    const int32_t* __restrict__ slice_off, const KT* __restrict__ Kscr, const double* __restrict__ Pscr,
    const uint8_t* __restrict__ dirflag, int ndir, const double* __restrict__ dirv, double* __restrict__ dircorr,
    KT* __restrict__ Kval, KT* __restrict__ Minv, double* __restrict__ P, double* __restrict__ F,
    const int32_t* __restrict__ slotNode, const int32_t* __restrict__ cslice, const int* __restrict__ specActive,
    const int32_t* __restrict__ specOfNode) {
    constexpr int RUN = 9 * kSlice;  // values of one contribution run
    __shared__ KT runbuf[256 / kSlice][RUN];
    const int per = gridDim.x >> 3;
    const int blk = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    const int ps = blk * blockDim.x + threadIdx.x;
    const int s = ps / kSlice, l = ps - (ps / kSlice) * kSlice;  // one slice per wavefront: l is the lane
    if (s * kSlice >= N) return;                                  // the whole wavefront
    KT* buf = runbuf[threadIdx.x / kSlice];
    const bool valid = ps < N;
    const int n = valid ? slotNode[ps] : 0;
    const bool work = valid && !(specActive && !specActive[specOfNode[n]]);
    if (valid && !work && dirv) dircorr[3 * n] = dircorr[3 * n + 1] = dircorr[3 * n + 2] = 0.0;
    bool fr[3] = {false, false, false};
    if (work && ndir) {
        fr[0] = dirflag[3 * n];
        fr[1] = dirflag[3 * n + 1];
        fr[2] = dirflag[3 * n + 2];
    }
    double t0 = 0, t1 = 0, t2 = 0;
    double dg[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    const int64_t vbase = (int64_t)slice_off[s] * 9 + l;
    int b0 = 0, b = 0, b1 = 0, c0 = 0, cnt = 0, cend = 0;
    if (work) {
        b0 = b = bptr[n];
        b1 = bptr[n + 1];
        c0 = coff[b0];
        cnt = coff[b1] - c0;
        cend = coff[b + 1] - c0;
    }
    KT acc[9];
#pragma unroll
    for (int v = 0; v < 9; ++v) acc[v] = (KT)0;
    const int J = cslice[s + 1] - cslice[s];
    if (__any(work)) {
        const KT* runs = Kscr + (int64_t)cslice[s] * RUN;
        KT nxt[9];
#pragma unroll
        for (int t = 0; t < 9; ++t) nxt[t] = J > 0 ? runs[t * kSlice + l] : (KT)0;
        for (int j = 0; j < J; ++j) {
#pragma unroll
            for (int t = 0; t < 9; ++t) buf[t * kSlice + l] = nxt[t];
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (j + 1 < J) {
                const KT* nr = runs + (int64_t)(j + 1) * RUN;
#pragma unroll
                for (int t = 0; t < 9; ++t) nxt[t] = __builtin_nontemporal_load(nr + t * kSlice + l);
            }
            if (work && j < cnt) {
#pragma unroll
                for (int v = 0; v < 9; ++v) acc[v] += buf[l * 9 + v];
                if (j + 1 == cend) {  // the last contribution of block b: finish it
                    const int m = bcol[b];
                    if (dirv) {
                        const double d0 = dirv[3 * m], d1 = dirv[3 * m + 1], d2 = dirv[3 * m + 2];
                        t0 += (double)acc[0] * d0 + (double)acc[1] * d1 + (double)acc[2] * d2;
                        t1 += (double)acc[3] * d0 + (double)acc[4] * d1 + (double)acc[5] * d2;
                        t2 += (double)acc[6] * d0 + (double)acc[7] * d1 + (double)acc[8] * d2;
                    }
                    if (ndir) {
                        const bool fc[3] = {(bool)dirflag[3 * m], (bool)dirflag[3 * m + 1], (bool)dirflag[3 * m + 2]};
#pragma unroll
                        for (int i = 0; i < 3; ++i)
#pragma unroll
                            for (int jj = 0; jj < 3; ++jj)
                                if (fr[i] || fc[jj]) acc[i * 3 + jj] = (m == n && i == jj) ? (KT)1 : (KT)0;
                    }
                    const int k = b - b0;
#pragma unroll
                    for (int v = 0; v < 9; ++v) Kval[vbase + (int64_t)(k * 9 + v) * kSlice] = acc[v];
                    if (m == n) {
#pragma unroll
                        for (int v = 0; v < 9; ++v) dg[v] = (double)acc[v];
                    }
#pragma unroll
                    for (int v = 0; v < 9; ++v) acc[v] = (KT)0;
                    ++b;
                    if (b < b1) cend = coff[b + 1] - c0;
                }
            }
            // every lane has read its block before the next run overwrites the buffer
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
    if (!work) return;
    double iv[9];
    inv3(dg, iv);
#pragma unroll
    for (int v = 0; v < 9; ++v) Minv[(int64_t)n * 9 + v] = (KT)iv[v];
    if (P) {
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
    if (dirv) {
        dircorr[3 * n] = fr[0] ? 0.0 : t0;
        dircorr[3 * n + 1] = fr[1] ? 0.0 : t1;
        dircorr[3 * n + 2] = fr[2] ? 0.0 : t2;
    }
}

// Body force (element.py:348-371): Pe = sum_q N_q (x) f detJ_q w_q with N at the table
// point and detJ of the element's J at the evaluation point (dJdxi); written to the
// element scratch, then gathered.
// One thread per listed element (elems[k], or element k when elems is NULL).
template <int NN, int Q>
__global__ __launch_bounds__(64) void body_force_kernel(int n, int E, int N, const int32_t* __restrict__ elems,
                                                        const double* __restrict__ coords,
                                                        const int32_t* __restrict__ conn,
                                                        const double* __restrict__ dNdxi,
                                                        const double* __restrict__ shapeN,
                                                        const double* __restrict__ qpw, double fx, double fy,
                                                        double fz, double* __restrict__ Pscr) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int e = elems ? elems[k] : k;
    double acc[NN];
#pragma unroll
    for (int a = 0; a < NN; ++a) acc[a] = 0.0;
    for (int q = 0; q < Q; ++q) {
        const double* dn = dNdxi + (size_t)q * NN * 3;
        double J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int a = 0; a < NN; ++a) {
            const int nd = conn[(int64_t)a * E + e];
            const double x = coords[nd], y = coords[N + nd], z = coords[2 * N + nd];
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                J[r * 3 + 0] += dn[a * 3 + r] * x;
                J[r * 3 + 1] += dn[a * 3 + r] * y;
                J[r * 3 + 2] += dn[a * 3 + r] * z;
            }
        }
        const double det = J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) +
                           J[2] * (J[3] * J[7] - J[4] * J[6]);
#pragma unroll
        for (int a = 0; a < NN; ++a) acc[a] += shapeN[q * NN + a] * det * qpw[q];
    }
#pragma unroll
    for (int a = 0; a < NN; ++a) {
        double* dst = Pscr + ((int64_t)e * NN + a) * 3;
        dst[0] = acc[a] * fx;
        dst[1] = acc[a] * fy;
        dst[2] = acc[a] * fz;
    }
}

// Surface traction (the distributed load of nonlinearimplicitstatic.py:460-508 with type
// "surface traction"): the consistent nodal force f_a = t * integral N_a dA of a uniform
// traction t on element faces.  Faces are numbered by local node numbers as in Abaqus
// (1: 1-2-3-4, 2: 5-8-7-6, 3: 1-5-6-2, 4: 2-6-7-3, 5: 3-7-8-4, 6: 4-8-5-1; hexa20 adds the
// mid-side nodes between consecutive corners), the numbering of the box generator's
// surfaces (boxgen.py:438-446).  2x2 Gauss points on 4-node faces, 3x3 on 8-node
// (serendipity) faces.  One thread per loaded element: its faces (bit f-1 of fmask) are
// summed in face order into the element's force slots.
__constant__ int8_t kFace8[6][4] = {{0, 1, 2, 3}, {4, 7, 6, 5}, {0, 4, 5, 1},
                                    {1, 5, 6, 2}, {2, 6, 7, 3}, {3, 7, 4, 0}};
__constant__ int8_t kFace20[6][8] = {{0, 1, 2, 3, 8, 9, 10, 11},   {4, 7, 6, 5, 15, 14, 13, 12},
                                     {0, 4, 5, 1, 16, 12, 17, 8},  {1, 5, 6, 2, 17, 13, 18, 9},
                                     {2, 6, 7, 3, 18, 14, 19, 10}, {3, 7, 4, 0, 19, 15, 16, 11}};

// N, dN/ds, dN/dt of the NF-node face (4: bilinear; 8: serendipity) at (s, t); node order:
// corners (-1,-1), (1,-1), (1,1), (-1,1), then mid-sides (0,-1), (1,0), (0,1), (-1,0)
template <int NF>
__device__ __forceinline__ void face_shape(double s, double t, double* N, double* Ns, double* Nt) {
    const double cs[4] = {-1, 1, 1, -1}, ct[4] = {-1, -1, 1, 1};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const double a = 1 + s * cs[i], b = 1 + t * ct[i];
        if (NF == 4) {
            N[i] = a * b / 4;
            Ns[i] = cs[i] * b / 4;
            Nt[i] = ct[i] * a / 4;
        } else {
            N[i] = a * b * (s * cs[i] + t * ct[i] - 1) / 4;
            Ns[i] = cs[i] * b * (2 * s * cs[i] + t * ct[i]) / 4;
            Nt[i] = ct[i] * a * (s * cs[i] + 2 * t * ct[i]) / 4;
        }
    }
    if (NF == 8) {
        N[4] = (1 - s * s) * (1 - t) / 2;
        Ns[4] = -s * (1 - t);
        Nt[4] = -(1 - s * s) / 2;
        N[5] = (1 + s) * (1 - t * t) / 2;
        Ns[5] = (1 - t * t) / 2;
        Nt[5] = -t * (1 + s);
        N[6] = (1 - s * s) * (1 + t) / 2;
        Ns[6] = -s * (1 + t);
        Nt[6] = (1 - s * s) / 2;
        N[7] = (1 - s) * (1 - t * t) / 2;
        Ns[7] = -(1 - t * t) / 2;
        Nt[7] = -t * (1 - s);
    }
}

template <int NN>
__global__ __launch_bounds__(256) void surface_traction_kernel(int n, int E, int N, const int32_t* __restrict__ elems,
                                                               const uint8_t* __restrict__ fmask,
                                                               const double* __restrict__ coords,
                                                               const int32_t* __restrict__ conn, double tx, double ty,
                                                               double tz, double* __restrict__ Pscr) {
    constexpr int NF = NN == 8 ? 4 : 8, NG = NN == 8 ? 2 : 3;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int e = elems[k], mask = fmask[k];
    const double gp2[2] = {-0.57735026918962576451, 0.57735026918962576451}, gw2[2] = {1.0, 1.0};
    const double gp3[3] = {-0.77459666924148337704, 0.0, 0.77459666924148337704};
    const double gw3[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
    double acc[NN];
#pragma unroll
    for (int a = 0; a < NN; ++a) acc[a] = 0.0;
    for (int f = 0; f < 6; ++f) {
        if (!((mask >> f) & 1)) continue;
        int loc[NF];
        double X[NF][3];
#pragma unroll
        for (int j = 0; j < NF; ++j) {
            loc[j] = NN == 8 ? kFace8[f][j] : kFace20[f][j];
            const int nd = conn[(int64_t)loc[j] * E + e];
            X[j][0] = coords[nd];
            X[j][1] = coords[N + nd];
            X[j][2] = coords[2 * N + nd];
        }
        for (int i = 0; i < NG; ++i)
            for (int j2 = 0; j2 < NG; ++j2) {
                const double s = NG == 2 ? gp2[i] : gp3[i], t = NG == 2 ? gp2[j2] : gp3[j2];
                const double w = NG == 2 ? gw2[i] * gw2[j2] : gw3[i] * gw3[j2];
                double Nf[NF], Ns[NF], Nt[NF];
                face_shape<NF>(s, t, Nf, Ns, Nt);
                double xs[3] = {0, 0, 0}, xt[3] = {0, 0, 0};
#pragma unroll
                for (int j = 0; j < NF; ++j)
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        xs[c] += Ns[j] * X[j][c];
                        xt[c] += Nt[j] * X[j][c];
                    }
                const double cx = xs[1] * xt[2] - xs[2] * xt[1], cy = xs[2] * xt[0] - xs[0] * xt[2],
                             cz = xs[0] * xt[1] - xs[1] * xt[0];
                const double dA = sqrt(cx * cx + cy * cy + cz * cz) * w;
#pragma unroll
                for (int j = 0; j < NF; ++j) acc[loc[j]] += Nf[j] * dA;
            }
    }
#pragma unroll
    for (int a = 0; a < NN; ++a) {
        double* dst = Pscr + ((int64_t)e * NN + a) * 3;
        dst[0] = acc[a] * tx;
        dst[1] = acc[a] * ty;
        dst[2] = acc[a] * tz;
    }
}

// P and F of a residual-only pass: the element residuals of every node summed in element
// order, as the assembly's gather sums them.  Nodes of specimens the pass skips (batched
// contexts) keep P and F, as in the assembly.
__global__ void pf_gather_kernel(int N, const int32_t* __restrict__ nptr, const int32_t* __restrict__ nlist,
                                 const double* __restrict__ Pscr, double* __restrict__ P, double* __restrict__ F,
                                 const int* __restrict__ specActive, const int32_t* __restrict__ specOfNode) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    if (specActive && !specActive[specOfNode[n]]) return;
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

__global__ void gather_add_kernel(int N, const int32_t* __restrict__ nptr, const int32_t* __restrict__ nlist,
                                  const double* __restrict__ Pscr, double* __restrict__ out) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    double p0 = 0, p1 = 0, p2 = 0;
    for (int c = nptr[n]; c < nptr[n + 1]; ++c) {
        const double* src = Pscr + (int64_t)nlist[c] * 3;
        p0 += src[0];
        p1 += src[1];
        p2 += src[2];
    }
    out[3 * n] += p0;
    out[3 * n + 1] += p1;
    out[3 * n + 2] += p2;
}

template <typename KT>
__global__ void export_csr_kernel(int N, const int32_t* __restrict__ bptr, const int32_t* __restrict__ slice_off,
                                  const int32_t* __restrict__ nodeSlot, const KT* __restrict__ Kval,
                                  double* __restrict__ out) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const int ps = nodeSlot[n];
    const int s = ps / kSlice, l = ps - (ps / kSlice) * kSlice;
    const int64_t vbase = (int64_t)slice_off[s] * 9 + l;
    const int deg = bptr[n + 1] - bptr[n];
    const int64_t r0 = (int64_t)9 * bptr[n];
    for (int k = 0; k < deg; ++k)
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                out[r0 + (int64_t)3 * i * deg + 3 * k + j] = (double)Kval[vbase + (int64_t)(k * 9 + i * 3 + j) * kSlice];
}

template <int NN, int Q, int MAT, typename KT, int EPW>
void launch_elem_t(efe_ctx* c, const double* dU) {
    static const int cap = std::getenv("EFE_ELEMENT_GRID_CAP") ? std::atoi(std::getenv("EFE_ELEMENT_GRID_CAP")) : 0;
    int grid = (c->curCount + EPW - 1) / EPW;
    if (grid == 0) return;
    if (cap > 0 && grid > cap) grid = cap;  // persistent: the workgroups loop over the list
    const int propStride = c->props_per_element ? c->E : 0;
    const int* act = c->useActive ? c->specActive : nullptr;
    if (c->mfPass)
        element_kernel<NN, Q, MAT, float, EPW, true><<<grid, 64, 0, c->stream>>>(
            c->E, c->N, c->coords, c->conn, c->props, propStride, c->dNdxi, c->dJdxi, c->qpw, dU, c->stateRef, c->stateTmp,
            (float*)c->Kscr, c->Pscr, c->status, c->mfC, c->specStatus, c->specOfElem, c->cpos, act, c->formK,
            c->curList, c->curCount, c->mfNC);
    else
        element_kernel<NN, Q, MAT, KT, EPW, false><<<grid, 64, 0, c->stream>>>(
            c->E, c->N, c->coords, c->conn, c->props, propStride, c->dNdxi, c->dJdxi, c->qpw, dU, c->stateRef, c->stateTmp,
            (KT*)c->Kscr, c->Pscr, c->status, nullptr, c->specStatus, c->specOfElem, c->cpos, act, c->formK,
            c->curList, c->curCount, c->mfNC);
}

template <int NN, int Q, int MAT, typename KT>
int launch_elem_epw(efe_ctx* c, const double* dU) {
    // default: hexa8 4 (assembled passes) or 8 (matrix-free and residual-only passes: phase 1,
    // one lane per element and QP, then fills the wavefront; 64^3: matrix-free 0.34 against
    // 0.51 ms, residual-only 0.27 against 0.34 ms), hexa20R 2, hexa20 1
    int epw = c->epw;
    if (epw <= 0) epw = (NN == 8) ? ((c->mfPass || !c->formK) ? 8 : 4) : (Q == 8 ? 2 : 1);
    switch (epw) {
        case 1: launch_elem_t<NN, Q, MAT, KT, 1>(c, dU); break;
        case 2: launch_elem_t<NN, Q, MAT, KT, 2>(c, dU); break;
        case 4:
            if (NN == 8) { launch_elem_t<NN, Q, MAT, KT, (NN == 8 ? 4 : 1)>(c, dU); break; }
            [[fallthrough]];
        case 8:
            if (NN == 8 && Q == 8) { launch_elem_t<NN, Q, MAT, KT, (NN == 8 && Q == 8 ? 8 : 1)>(c, dU); break; }
            [[fallthrough]];
        default:
            efe_set_error("elements per workgroup must be 1, 2, 4 or 8 (hexa20: 1 or 2)");
            return EFE_ERR_ARG;
    }
    return EFE_OK;
}

template <int NN, int Q, typename KT>
int launch_elem_mat(efe_ctx* c, const double* dU) {
    switch (c->curMat) {
        case EFE_LINEAR_ELASTIC: return launch_elem_epw<NN, Q, EFE_LINEAR_ELASTIC, KT>(c, dU);
        case EFE_VONMISES: return launch_elem_epw<NN, Q, EFE_VONMISES, KT>(c, dU);
        case EFE_DRUCKER_PRAGER: return launch_elem_epw<NN, Q, EFE_DRUCKER_PRAGER, KT>(c, dU);
    }
    efe_set_error("unknown material");
    return EFE_ERR_ARG;
}

template <typename KT>
int launch_elem_type(efe_ctx* c, const double* dU) {
    switch (c->eltype) {
        case EFE_C3D8: return launch_elem_mat<8, 8, KT>(c, dU);
        case EFE_C3D20: return launch_elem_mat<20, 27, KT>(c, dU);
        case EFE_C3D20R: return launch_elem_mat<20, 8, KT>(c, dU);
    }
    efe_set_error("unknown element type");
    return EFE_ERR_ARG;
}

}  // namespace

int efe_launch_elements(efe_ctx* c, const double* dU, int precision) {
    c->mfPass = precision == EFE_FP32_MATFREE;
    for (int g = 0; g < c->ngroups; ++g) {  // one launch per material group (section)
        c->curMat = c->groupMat[g];
        c->curList = c->groupElems[g];
        c->curCount = c->groupCount[g];
        int rc = (precision == EFE_FP64) ? launch_elem_type<double>(c, dU) : launch_elem_type<float>(c, dU);
        if (rc != EFE_OK) return rc;
        EFE_HIP(hipGetLastError());
    }
    return EFE_OK;
}

int efe_launch_residual(efe_ctx* c, const double* dU, double* P, double* F, int precision) {
    c->formK = 0;
    const int rc = efe_launch_elements(c, dU, precision);  // the precision's launch shape; same residual
    c->formK = 1;
    if (rc != EFE_OK) return rc;
    pf_gather_kernel<<<(c->N + 255) / 256, 256, 0, c->stream>>>(c->N, c->nptr, c->nlist, c->Pscr, P, F,
                                                                 c->useActive ? c->specActive : nullptr,
                                                                 c->specOfNode);
    EFE_HIP(hipGetLastError());
    return EFE_OK;
}

int efe_launch_assembly(efe_ctx* c, double* P, double* F, const double* dir_values, double* dir_correction,
                        int precision) {
    const int bs = 256, grid = (((c->N + bs - 1) / bs) + 7) & ~7;  // a multiple of 8 (XCD mapping)
    static const bool perLane = std::getenv("EFE_ASSEMBLY_PER_LANE") != nullptr;  // the per-lane form, for comparison
    if (!perLane) {
#define ASM_ARGS(KT, K, M)                                                                                           \
    c->N, c->bptr, c->bcol, c->coff, c->nptr, c->nlist, c->slice_off, (const KT*)c->Kscr, c->Pscr, c->dirflag,       \
        c->ndir, dir_values, dir_correction, K, M, P, F, c->slotNode, c->cslice, c->useActive ? c->specActive : nullptr, \
        c->specOfNode
        if (precision == EFE_FP32)
            assemble_runs_kernel<float><<<grid, bs, 0, c->stream>>>(ASM_ARGS(float, c->K32, c->Minv32));
        else
            assemble_runs_kernel<double><<<grid, bs, 0, c->stream>>>(ASM_ARGS(double, c->K64, c->Minv64));
#undef ASM_ARGS
        EFE_HIP(hipGetLastError());
        c->Kprec = precision;
        c->Kversion++;
        return EFE_OK;
    }
    if (precision == EFE_FP32) {
        assemble_kernel<float><<<grid, bs, 0, c->stream>>>(
            c->N, c->bptr, c->bcol, c->coff, c->nptr, c->nlist, c->slice_off, (const float*)c->Kscr,
            c->Pscr, c->dirflag, c->ndir, dir_values, dir_correction, c->K32, c->Minv32, P, F, c->slotNode, c->cslice,
            c->useActive ? c->specActive : nullptr, c->specOfNode);
    } else {
        assemble_kernel<double><<<grid, bs, 0, c->stream>>>(
            c->N, c->bptr, c->bcol, c->coff, c->nptr, c->nlist, c->slice_off, (const double*)c->Kscr,
            c->Pscr, c->dirflag, c->ndir, dir_values, dir_correction, c->K64, c->Minv64, P, F, c->slotNode, c->cslice,
            c->useActive ? c->specActive : nullptr, c->specOfNode);
    }
    EFE_HIP(hipGetLastError());
    c->Kprec = precision;
    c->Kversion++;
    return EFE_OK;
}

int efe_launch_body_force(efe_ctx* c, const int32_t* elems, int n, const double* f3, double* PExt) {
    // a subset leaves the other elements' slots at zero
    if (elems) EFE_HIP(hipMemsetAsync(c->Pscr, 0, (size_t)c->E * c->NN * 3 * sizeof(double), c->stream));
    c->pscrValid = false;
    const int grid = (n + 63) / 64;
    if (n > 0) switch (c->eltype) {
        case EFE_C3D8:
            body_force_kernel<8, 8><<<grid, 64, 0, c->stream>>>(n, c->E, c->N, elems, c->coords, c->conn, c->dJdxi,
                                                               c->shapeN, c->qpw, f3[0], f3[1], f3[2], c->Pscr);
            break;
        case EFE_C3D20:
            body_force_kernel<20, 27><<<grid, 64, 0, c->stream>>>(n, c->E, c->N, elems, c->coords, c->conn, c->dJdxi,
                                                                 c->shapeN, c->qpw, f3[0], f3[1], f3[2], c->Pscr);
            break;
        case EFE_C3D20R:
            body_force_kernel<20, 8><<<grid, 64, 0, c->stream>>>(n, c->E, c->N, elems, c->coords, c->conn, c->dJdxi,
                                                                c->shapeN, c->qpw, f3[0], f3[1], f3[2], c->Pscr);
            break;
        default:
            efe_set_error("unknown element type");
            return EFE_ERR_ARG;
    }
    gather_add_kernel<<<(c->N + 255) / 256, 256, 0, c->stream>>>(c->N, c->nptr, c->nlist, c->Pscr, PExt);
    EFE_HIP(hipGetLastError());
    return EFE_OK;
}

int efe_launch_surface_traction(efe_ctx* c, const int32_t* elems, const uint8_t* fmask, int n, const double* t3,
                                double* PExt) {
    EFE_HIP(hipMemsetAsync(c->Pscr, 0, (size_t)c->E * c->NN * 3 * sizeof(double), c->stream));
    c->pscrValid = false;
    const int grid = (n + 255) / 256;
    if (n > 0) {
        if (c->NN == 8)
            surface_traction_kernel<8><<<grid, 256, 0, c->stream>>>(n, c->E, c->N, elems, fmask, c->coords, c->conn,
                                                                    t3[0], t3[1], t3[2], c->Pscr);
        else
            surface_traction_kernel<20><<<grid, 256, 0, c->stream>>>(n, c->E, c->N, elems, fmask, c->coords, c->conn,
                                                                     t3[0], t3[1], t3[2], c->Pscr);
    }
    gather_add_kernel<<<(c->N + 255) / 256, 256, 0, c->stream>>>(c->N, c->nptr, c->nlist, c->Pscr, PExt);
    EFE_HIP(hipGetLastError());
    return EFE_OK;
}

// Geostatic initial stress at every QP of the listed elements (all when elems is NULL):
// sigma_yy = p1 + (p2 - p1) (y - h1) / (h2 - h1) at the QP's y coordinate, sigma_xx =
// xLateral sigma_yy, sigma_zz = zLateral sigma_yy; shear stresses and all other state stay.
template <int NN>
__global__ __launch_bounds__(256) void geostatic_kernel(int E, int Q, int n, int N, const int32_t* __restrict__ elems,
                                                        const double* __restrict__ coords,
                                                        const int32_t* __restrict__ conn,
                                                        const double* __restrict__ shapeNe, double p1, double h1,
                                                        double p2, double h2, double xl, double zl,
                                                        double* __restrict__ stateRef) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)n * Q) return;
    const int k = (int)(i / Q), q = (int)(i - (int64_t)k * Q);
    const int e = elems ? elems[k] : k;
    double y = 0.0;
#pragma unroll
    for (int a = 0; a < NN; ++a) y += shapeNe[q * NN + a] * coords[N + conn[(int64_t)a * E + e]];
    const double sy = p1 + (p2 - p1) * (y - h1) / (h2 - h1);
    const int64_t EQ = (int64_t)E * Q, idx = (int64_t)e * Q + q;
    stateRef[idx] = xl * sy;
    stateRef[EQ + idx] = sy;
    stateRef[2 * EQ + idx] = zl * sy;
}

int efe_launch_geostatic(efe_ctx* c, const double* def, const int32_t* elements, int32_t n) {
    const int32_t* de = nullptr;
    if (elements) {
        if (!c->geoElems) EFE_HIP(hipMalloc(&c->geoElems, (size_t)c->E * sizeof(int32_t)));
        EFE_HIP(hipMemcpyAsync(c->geoElems, elements, (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice,
                               c->stream));
        de = c->geoElems;
    }
    if (n == 0) return EFE_OK;
    const int64_t work = (int64_t)n * c->Q;
    const unsigned grid = (unsigned)((work + 255) / 256);
    if (c->NN == 8)
        geostatic_kernel<8><<<grid, 256, 0, c->stream>>>(c->E, c->Q, n, c->N, de, c->coords, c->conn, c->shapeNe,
                                                          def[0], def[1], def[2], def[3], def[4], def[5],
                                                          c->stateRef);
    else
        geostatic_kernel<20><<<grid, 256, 0, c->stream>>>(c->E, c->Q, n, c->N, de, c->coords, c->conn, c->shapeNe,
                                                           def[0], def[1], def[2], def[3], def[4], def[5],
                                                           c->stateRef);
    EFE_HIP(hipGetLastError());
    return EFE_OK;
}

// Element matrices in the layout of the native element ABI (MarmotElement::computeYourself,
// marmotelement/element.pxd:97-103: Ke column-major nDof x nDof, Pe nDof, DOFs node-major):
// thread per (listed element, block a b); the block is read from its contribution position.
template <typename KT>
__global__ void export_element_kernel(int n, int NN, const int32_t* __restrict__ elems,
                                      const int32_t* __restrict__ cpos, const KT* __restrict__ Kscr,
                                      const double* __restrict__ Pscr, double* __restrict__ Ke,
                                      double* __restrict__ Pe) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nb = (int64_t)NN * NN;
    if (t >= (int64_t)n * nb) return;
    const int k = (int)(t / nb), ab = (int)(t - (int64_t)k * nb);
    const int a = ab / NN, b = ab - (ab / NN) * NN;
    const int e = elems[k];
    const int nd = 3 * NN;
    const KT* src = Kscr + (int64_t)cpos[(int64_t)e * nb + ab] * kBlockPitch;
    double* dst = Ke + (int64_t)k * nd * nd;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) dst[(int64_t)(3 * b + j) * nd + 3 * a + i] = (double)src[i * 3 + j];
    if (Pe && b == 0) {
#pragma unroll
        for (int i = 0; i < 3; ++i) Pe[(int64_t)k * nd + 3 * a + i] = Pscr[((int64_t)e * NN + a) * 3 + i];
    }
}

int efe_launch_export_elements(efe_ctx* c, const int32_t* elems, int n, double* Ke, double* Pe) {
    const int64_t total = (int64_t)n * c->NN * c->NN;
    const unsigned grid = (unsigned)((total + 255) / 256);
    if (c->Kprec == EFE_FP32)
        export_element_kernel<float><<<grid, 256, 0, c->stream>>>(n, c->NN, elems, c->cpos, (const float*)c->Kscr,
                                                                   c->Pscr, Ke, Pe);
    else
        export_element_kernel<double><<<grid, 256, 0, c->stream>>>(n, c->NN, elems, c->cpos,
                                                                    (const double*)c->Kscr, c->Pscr, Ke, Pe);
    EFE_HIP(hipGetLastError());
    return EFE_OK;
}

int efe_launch_export_csr(efe_ctx* c, double* values) {
    const int grid = (c->N + 255) / 256;
    if (c->Kprec == EFE_FP32)
        export_csr_kernel<float><<<grid, 256, 0, c->stream>>>(c->N, c->bptr, c->slice_off, c->nodeSlot, c->K32,
                                                               values);
    else
        export_csr_kernel<double><<<grid, 256, 0, c->stream>>>(c->N, c->bptr, c->slice_off, c->nodeSlot, c->K64,
                                                                values);
    EFE_HIP(hipGetLastError());
    return EFE_OK;
}
