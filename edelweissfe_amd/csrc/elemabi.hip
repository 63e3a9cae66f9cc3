// The native element ABI on the GPU: MarmotElement::computeYourself
// (marmotelement/element.pxd:97-103, called per element by NISTParallelForMarmotElements,
// solvers/nonlinearimplicitstaticparallel.pyx:65-167) for a list of elements in one launch,
// with caller-owned arrays:
//   dQ        [n][nDof]            increment of the element's DOFs (node-major, 3 a + c)
//   stateRef  [n][Q][nState]       QP state at the start of the increment (stress 6, strain 6,
//                                  material state; element.py:222-236)
//   stateTemp [n][Q][nState]       the updated state (may be stateRef: in place)
//   Pe        [n][nDof]            -sum_q B^T sigma detJ w
//   Ke        [n][nDof][nDof]      column-major sum_q B^T C B detJ w, or none
// The element formulation is the context's (DisplacementElement.computeYourself,
// element.py:290-346, with the reference's Jacobian), and so are its material and properties.
// One wavefront per EPW elements; phase 1 lane per (element, QP): J, dN/dx, dEps, material update
// (FP64), w detJ C and w detJ sigma staged in LDS; phase 2 lane per (element, block a b): the 3x3
// block of Ke, stored column-major; phase 3 lane per (element, node): Pe.
#include <vector>

#include "efe_internal.h"
#include "materials.h"

namespace {

template <int NN, int Q, int MAT, int EPW>
__global__ __launch_bounds__(64) void element_abi_kernel(
    int cnt, const int32_t* __restrict__ kmap, const int32_t* __restrict__ emap, int E, int N,
    const double* __restrict__ coords, const int32_t* __restrict__ conn, const double* __restrict__ props,
    int propStride, const double* __restrict__ dNdxi, const double* __restrict__ dJdxi,
    const double* __restrict__ qpw, const double* __restrict__ dQ, const double* __restrict__ sRef,
    double* __restrict__ sTmp, int ns, double* __restrict__ Pe, double* __restrict__ Ke, int* __restrict__ status) {
    constexpr int NP = (MAT == EFE_LINEAR_ELASTIC) ? 2 : 6;
    constexpr int ND = 3 * NN;
    __shared__ double g_s[EPW][Q][NN][3];
    __shared__ double fC_s[EPW][Q][36];
    __shared__ double fs_s[EPW][Q][6];
    const int lane = threadIdx.x;
    const int i0 = blockIdx.x * EPW;

    for (int t = lane; t < EPW * Q; t += 64) {
        const int el = t / Q, q = t - (t / Q) * Q;
        const int i = i0 + el;
        if (i >= cnt) continue;
        const int k = kmap[i], e = emap[i];
        const double* dn = dNdxi + (size_t)q * NN * 3;
        const double* dj = dJdxi + (size_t)q * NN * 3;
        double J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int a = 0; a < NN; ++a) {
            const int nd = conn[(int64_t)a * E + e];
            const double x = coords[nd], y = coords[N + nd], z = coords[2 * N + nd];
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                J[r * 3 + 0] += dj[a * 3 + r] * x;
                J[r * 3 + 1] += dj[a * 3 + r] * y;
                J[r * 3 + 2] += dj[a * 3 + r] * z;
            }
        }
        const double det = J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) +
                           J[2] * (J[3] * J[7] - J[4] * J[6]);
        const double id = 1.0 / det;
        const double iJ[9] = {(J[4] * J[8] - J[5] * J[7]) * id, (J[2] * J[7] - J[1] * J[8]) * id,
                              (J[1] * J[5] - J[2] * J[4]) * id, (J[5] * J[6] - J[3] * J[8]) * id,
                              (J[0] * J[8] - J[2] * J[6]) * id, (J[2] * J[3] - J[0] * J[5]) * id,
                              (J[3] * J[7] - J[4] * J[6]) * id, (J[1] * J[6] - J[0] * J[7]) * id,
                              (J[0] * J[4] - J[1] * J[3]) * id};
        const double* du = dQ + (int64_t)k * ND;
        double gr[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int a = 0; a < NN; ++a) {
            const double d0 = dn[a * 3 + 0], d1 = dn[a * 3 + 1], d2 = dn[a * 3 + 2];
            double g[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) g[c] = iJ[c * 3 + 0] * d0 + iJ[c * 3 + 1] * d1 + iJ[c * 3 + 2] * d2;
            g_s[el][q][a][0] = g[0];
            g_s[el][q][a][1] = g[1];
            g_s[el][q][a][2] = g[2];
            const double u0 = du[3 * a], u1 = du[3 * a + 1], u2 = du[3 * a + 2];
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                gr[0 + j] += u0 * g[j];
                gr[3 + j] += u1 * g[j];
                gr[6 + j] += u2 * g[j];
            }
        }
        const double de[6] = {gr[0], gr[4], gr[8], gr[1] + gr[3], gr[2] + gr[6], gr[5] + gr[7]};
        const double* sr = sRef + ((int64_t)k * Q + q) * ns;
        double* st = sTmp + ((int64_t)k * Q + q) * ns;
        double s[6], eps[6], ms = 0.0;
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            s[c] = sr[c];
            eps[c] = sr[6 + c];
        }
        if (MAT != EFE_LINEAR_ELASTIC && ns > 12) ms = sr[12];
        double p[NP];
#pragma unroll
        for (int m = 0; m < NP; ++m) p[m] = propStride ? props[(int64_t)m * propStride + e] : props[m];
        double C[36];
        int fail = 0;
        if (MAT == EFE_LINEAR_ELASTIC) fail = efe::le_update(p, s, de, C);
        else if (MAT == EFE_VONMISES) fail = efe::vm_update(p, s, ms, de, C);
        else fail = efe::dp_update(p, s, ms, de, C);
        if (fail) *status = 1;
        // the state out: other columns (a wider state of another section) are carried over
        for (int c = 0; c < ns; ++c) st[c] = sr[c];
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            st[c] = s[c];
            st[6 + c] = eps[c] + de[c];
        }
        if (MAT != EFE_LINEAR_ELASTIC && ns > 12) st[12] = ms;
        const double f = det * qpw[q];
#pragma unroll
        for (int m = 0; m < 36; ++m) fC_s[el][q][m] = f * C[m];
#pragma unroll
        for (int m = 0; m < 6; ++m) fs_s[el][q][m] = f * s[m];
    }
    __syncthreads();

    if (Ke)
        for (int t = lane; t < EPW * NN * NN; t += 64) {
            const int el = t / (NN * NN), ab = t - el * NN * NN;
            const int a = ab / NN, b = ab - (ab / NN) * NN;
            const int i = i0 + el;
            if (i >= cnt) continue;
            const int k = kmap[i];
            double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            for (int q = 0; q < Q; ++q) {
                const double ax = g_s[el][q][a][0], ay = g_s[el][q][a][1], az = g_s[el][q][a][2];
                const double bx = g_s[el][q][b][0], by = g_s[el][q][b][1], bz = g_s[el][q][b][2];
                const double* Cq = fC_s[el][q];
                double CB[6][3];
#pragma unroll
                for (int r = 0; r < 6; ++r) {
                    CB[r][0] = Cq[r * 6 + 0] * bx + Cq[r * 6 + 3] * by + Cq[r * 6 + 4] * bz;
                    CB[r][1] = Cq[r * 6 + 1] * by + Cq[r * 6 + 3] * bx + Cq[r * 6 + 5] * bz;
                    CB[r][2] = Cq[r * 6 + 2] * bz + Cq[r * 6 + 4] * bx + Cq[r * 6 + 5] * by;
                }
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    acc[0 + j] += ax * CB[0][j] + ay * CB[3][j] + az * CB[4][j];
                    acc[3 + j] += ay * CB[1][j] + ax * CB[3][j] + az * CB[5][j];
                    acc[6 + j] += az * CB[2][j] + ax * CB[4][j] + ay * CB[5][j];
                }
            }
            double* dst = Ke + (int64_t)k * ND * ND;
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c) dst[(int64_t)(3 * b + c) * ND + 3 * a + r] = acc[r * 3 + c];
        }

    for (int t = lane; t < EPW * NN; t += 64) {
        const int el = t / NN, a = t - (t / NN) * NN;
        const int i = i0 + el;
        if (i >= cnt) continue;
        const int k = kmap[i];
        double p0 = 0, p1 = 0, p2 = 0;
        for (int q = 0; q < Q; ++q) {
            const double ax = g_s[el][q][a][0], ay = g_s[el][q][a][1], az = g_s[el][q][a][2];
            const double* s = fs_s[el][q];
            p0 += ax * s[0] + ay * s[3] + az * s[4];
            p1 += ay * s[1] + ax * s[3] + az * s[5];
            p2 += az * s[2] + ax * s[4] + ay * s[5];
        }
        double* dst = Pe + (int64_t)k * ND + 3 * a;
        dst[0] = -p0;
        dst[1] = -p1;
        dst[2] = -p2;
    }
}

template <int NN, int Q, int MAT>
void launch_abi(efe_ctx* c, int cnt, const int32_t* kmap, const int32_t* emap, const double* dQ, const double* sRef,
                double* sTmp, double* Pe, double* Ke) {
    constexpr int EPW = (NN == 8) ? 8 : (Q == 8 ? 2 : 1);
    const int grid = (cnt + EPW - 1) / EPW;
    if (grid == 0) return;
    element_abi_kernel<NN, Q, MAT, EPW><<<grid, 64, 0, c->stream>>>(
        cnt, kmap, emap, c->E, c->N, c->coords, c->conn, c->props, c->props_per_element ? c->E : 0, c->dNdxi,
        c->dJdxi, c->qpw, dQ, sRef, sTmp, c->nstate, Pe, Ke, c->status);
}

template <int NN, int Q>
void launch_abi_mat(efe_ctx* c, int mat, int cnt, const int32_t* kmap, const int32_t* emap, const double* dQ,
                    const double* sRef, double* sTmp, double* Pe, double* Ke) {
    if (mat == EFE_LINEAR_ELASTIC) launch_abi<NN, Q, EFE_LINEAR_ELASTIC>(c, cnt, kmap, emap, dQ, sRef, sTmp, Pe, Ke);
    else if (mat == EFE_VONMISES) launch_abi<NN, Q, EFE_VONMISES>(c, cnt, kmap, emap, dQ, sRef, sTmp, Pe, Ke);
    else launch_abi<NN, Q, EFE_DRUCKER_PRAGER>(c, cnt, kmap, emap, dQ, sRef, sTmp, Pe, Ke);
}

}  // namespace

extern "C" int efe_element_compute(efe_ctx* c, const int32_t* elements, int32_t n, const double* QTotal,
                                   const double* dQ, const double* stateRef, double* stateTemp, double* Pe,
                                   double* Ke, const double* time, double dT, double* pNewDT) {
    (void)QTotal;
    (void)time;
    (void)dT;
    if (!c) {
        efe_set_error("null context");
        return EFE_ERR_ARG;
    }
    if (n < 0 || (n > 0 && (!elements || !dQ || !stateRef || !stateTemp || !Pe)) || !pNewDT) {
        efe_set_error("efe_element_compute: invalid arguments");
        return EFE_ERR_ARG;
    }
    // the listed elements by material group (the order of each group kept), positions k in
    // the caller's arrays and element indices e of the context
    std::vector<int32_t> km, em;
    km.reserve(n);
    em.reserve(n);
    std::vector<int> gcount(c->ngroups, 0);
    for (int32_t i = 0; i < n; ++i)
        if (elements[i] < 0 || elements[i] >= c->E) {
            efe_set_error("efe_element_compute: element out of range");
            return EFE_ERR_ARG;
        }
    for (int g = 0; g < c->ngroups; ++g)
        for (int32_t i = 0; i < n; ++i)
            if (c->ngroups == 1 || c->matOf[elements[i]] == c->groupMat[g]) {
                km.push_back(i);
                em.push_back(elements[i]);
                ++gcount[g];
            }
    if (n == 0) return EFE_OK;
    const size_t need = 2 * (size_t)n;
    if (need > c->resCap) {
        hipFree(c->resIdx);
        c->resIdx = nullptr;
        c->resCap = 0;
        EFE_HIP(hipMalloc(&c->resIdx, need * sizeof(int32_t)));
        c->resCap = need;
    }
    std::vector<int32_t> both(km);
    both.insert(both.end(), em.begin(), em.end());
    EFE_HIP(hipMemcpyAsync(c->resIdx, both.data(), need * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    EFE_HIP(hipMemsetAsync(c->status, 0, sizeof(int), c->stream));
    int off = 0;
    for (int g = 0; g < c->ngroups; ++g) {
        const int32_t* kmap = c->resIdx + off;
        const int32_t* emap = c->resIdx + n + off;
        const int cnt = gcount[g], mat = c->groupMat[g];
        switch (c->eltype) {
            case EFE_C3D8: launch_abi_mat<8, 8>(c, mat, cnt, kmap, emap, dQ, stateRef, stateTemp, Pe, Ke); break;
            case EFE_C3D20: launch_abi_mat<20, 27>(c, mat, cnt, kmap, emap, dQ, stateRef, stateTemp, Pe, Ke); break;
            default: launch_abi_mat<20, 8>(c, mat, cnt, kmap, emap, dQ, stateRef, stateTemp, Pe, Ke); break;
        }
        EFE_HIP(hipGetLastError());
        off += cnt;
    }
    int hs = 0;
    EFE_HIP(hipMemcpyAsync(&hs, c->status, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    EFE_HIP(hipStreamSynchronize(c->stream));  // both must outlive the copy
    if (hs) {
        *pNewDT = *pNewDT < 0.5 ? *pNewDT : 0.5;
        return EFE_CUTBACK;
    }
    return EFE_OK;
}
