// C ABI entry points (include/efe_gpu.h).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "efe_internal.h"

namespace {

thread_local std::string g_err;

// Natural node coordinates (xi, eta, zeta) of the reference's hexa8 / hexa20
// (_elementcomputationmatrices.py:164-212): corners n1..n8, then mid-side nodes.
const double kNat20[20][3] = {
    {-1, -1, -1}, {-1, -1, 1}, {1, -1, 1}, {1, -1, -1}, {-1, 1, -1}, {-1, 1, 1}, {1, 1, 1}, {1, 1, -1},
    {-1, -1, 0},  {0, -1, 1},  {1, -1, 0}, {0, -1, -1}, {-1, 1, 0},  {0, 1, 1},  {1, 1, 0}, {0, 1, -1},
    {-1, 0, -1},  {-1, 0, 1},  {1, 0, 1},  {1, 0, -1}};

// Trilinear (8) / serendipity (20) shape functions and their natural derivatives.
void shape(int nn, const double p[3], double* N, double* dN) {
    const double x = p[0], y = p[1], z = p[2];
    for (int a = 0; a < nn; ++a) {
        const double ax = kNat20[a][0], ay = kNat20[a][1], az = kNat20[a][2];
        if (nn == 8) {
            N[a] = (1 + ax * x) * (1 + ay * y) * (1 + az * z) / 8;
            dN[a * 3 + 0] = ax * (1 + ay * y) * (1 + az * z) / 8;
            dN[a * 3 + 1] = ay * (1 + ax * x) * (1 + az * z) / 8;
            dN[a * 3 + 2] = az * (1 + ax * x) * (1 + ay * y) / 8;
        } else if (a < 8) {
            const double s = ax * x + ay * y + az * z - 2;
            N[a] = (1 + ax * x) * (1 + ay * y) * (1 + az * z) * s / 8;
            dN[a * 3 + 0] = ax * (1 + ay * y) * (1 + az * z) * (2 * ax * x + ay * y + az * z - 1) / 8;
            dN[a * 3 + 1] = ay * (1 + ax * x) * (1 + az * z) * (ax * x + 2 * ay * y + az * z - 1) / 8;
            dN[a * 3 + 2] = az * (1 + ax * x) * (1 + ay * y) * (ax * x + ay * y + 2 * az * z - 1) / 8;
        } else if (ax == 0) {
            N[a] = (1 - x * x) * (1 + ay * y) * (1 + az * z) / 4;
            dN[a * 3 + 0] = -2 * x * (1 + ay * y) * (1 + az * z) / 4;
            dN[a * 3 + 1] = ay * (1 - x * x) * (1 + az * z) / 4;
            dN[a * 3 + 2] = az * (1 - x * x) * (1 + ay * y) / 4;
        } else if (ay == 0) {
            N[a] = (1 + ax * x) * (1 - y * y) * (1 + az * z) / 4;
            dN[a * 3 + 0] = ax * (1 - y * y) * (1 + az * z) / 4;
            dN[a * 3 + 1] = -2 * y * (1 + ax * x) * (1 + az * z) / 4;
            dN[a * 3 + 2] = az * (1 + ax * x) * (1 - y * y) / 4;
        } else {
            N[a] = (1 + ax * x) * (1 + ay * y) * (1 - z * z) / 4;
            dN[a * 3 + 0] = ax * (1 + ay * y) * (1 - z * z) / 4;
            dN[a * 3 + 1] = ay * (1 + ax * x) * (1 - z * z) / 4;
            dN[a * 3 + 2] = -2 * z * (1 + ax * x) * (1 + ay * y) / 4;
        }
    }
}

// Quadrature tables of edelweissfe/elements/library.py:33-46,212-291.
void qp_table(int eltype, std::vector<double>& xi, std::vector<double>& eta, std::vector<double>& zeta,
              std::vector<double>& w) {
    if (eltype == EFE_C3D20) {
        const double s = std::sqrt(0.6);
        const double w1 = std::pow(5.0 / 9.0, 3.0), w2 = std::pow(5.0 / 9.0, 2.0) * (8.0 / 9.0);
        const double w3 = (5.0 / 9.0) * std::pow(8.0 / 9.0, 2.0), w4 = std::pow(8.0 / 9.0, 3.0);
        const double t20[9] = {-1, -1, -1, 0, 0, 0, 1, 1, 1};
        const double s20[3] = {-1, 0, 1};
        const double wtab[27] = {w1, w2, w1, w2, w3, w2, w1, w2, w1, w2, w3, w2, w3, w4,
                                 w3, w2, w3, w2, w1, w2, w1, w2, w3, w2, w1, w2, w1};
        for (int i = 0; i < 27; ++i) {
            xi.push_back(s * (i < 9 ? -1.0 : (i < 18 ? 0.0 : 1.0)));
            eta.push_back(s * t20[i % 9]);
            zeta.push_back(s * s20[i % 3]);
            w.push_back(wtab[i]);
        }
    } else {
        const double s = 1.0 / std::sqrt(3.0);
        const double t8[4] = {-1, -1, 1, 1}, s8[4] = {-1, 1, 1, -1};
        for (int i = 0; i < 8; ++i) {
            xi.push_back(s * (i < 4 ? -1.0 : 1.0));
            eta.push_back(s * t8[i % 4]);
            zeta.push_back(s * s8[i % 4]);
            w.push_back(1.0);
        }
    }
}

template <typename T>
int upload(T** dst, const T* src, size_t n, hipStream_t st) {
    EFE_HIP(hipMalloc(dst, (n ? n : 1) * sizeof(T)));
    if (n) EFE_HIP(hipMemcpyAsync(*dst, src, n * sizeof(T), hipMemcpyHostToDevice, st));
    return EFE_OK;
}

int check(efe_ctx* c) {
    if (!c) {
        efe_set_error("null context");
        return EFE_ERR_ARG;
    }
    return EFE_OK;
}

float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0;
    hipEventElapsedTime(&ms, a, b);
    return ms;
}

}  // namespace

namespace {
// device buffers of one call, freed on every return path
struct DeviceBuffers {
    std::vector<void*> p;
    ~DeviceBuffers() {
        for (void* q : p) hipFree(q);
    }
    template <typename T>
    hipError_t alloc(T** out, size_t n) {
        void* q = nullptr;
        const hipError_t e = hipMalloc(&q, (n ? n : 1) * sizeof(T));
        if (e == hipSuccess) {
            p.push_back(q);
            *out = (T*)q;
        }
        return e;
    }
};

int check_solve_args(int32_t n, const void* ip, const void* ix, const void* v, const void* b, const void* x,
                     const efe_solve_opts* o, const efe_solve_info* info, const char* what) {
    if (n <= 0 || !ip || !ix || !v || !b || !x || !o || !info) {
        efe_set_error(std::string(what) + ": invalid arguments");
        return EFE_ERR_ARG;
    }
    if (o->solver != EFE_SOLVER_CG && o->solver != EFE_SOLVER_BICGSTAB && o->solver != EFE_SOLVER_GMRES) {
        efe_set_error(std::string(what) + ": solver must be EFE_SOLVER_CG, EFE_SOLVER_BICGSTAB or EFE_SOLVER_GMRES");
        return EFE_ERR_ARG;
    }
    return EFE_OK;
}
}  // namespace

void efe_set_error(const std::string& msg) { g_err = msg; }

// Buffers of an element pass and its assembly in `precision` (allocated once).
int efe_prepare_elements(efe_ctx* c, int precision) {
    const size_t nblk = (size_t)c->scratchBlocks * kBlockPitch;
    if (!c->Kscr) EFE_HIP(hipMalloc(&c->Kscr, nblk * sizeof(double)));
    const size_t nval = (size_t)c->hp.padded_blocks * 9;
    if (precision == EFE_FP64 && !c->K64) {
        EFE_HIP(hipMalloc(&c->K64, nval * sizeof(double)));
        EFE_HIP(hipMemsetAsync(c->K64, 0, nval * sizeof(double), c->stream));
        EFE_HIP(hipMalloc(&c->Minv64, (size_t)c->N * 9 * sizeof(double)));
    }
    if (precision == EFE_FP32 && !c->K32) {
        EFE_HIP(hipMalloc(&c->K32, nval * sizeof(float)));
        EFE_HIP(hipMemsetAsync(c->K32, 0, nval * sizeof(float), c->stream));
        // the matrix-free operator may have allocated it already: a PCG graph captured for
        // that operator holds its address, so it must not move
        if (!c->Minv32) EFE_HIP(hipMalloc(&c->Minv32, (size_t)c->N * 9 * sizeof(float)));
    }
    if (precision == EFE_FP32_MATFREE)
        if (int rc = efe_mf_alloc(c)) return rc;
    return EFE_OK;
}

extern "C" {

const char* efe_last_error(void) { return g_err.c_str(); }

int efe_ctx_create(const efe_mesh_desc* d, efe_ctx** out) {
    if (!d || !out || !d->coords || !d->conn || d->n_nodes <= 0 || d->n_elements <= 0 || !d->props) {
        efe_set_error("invalid mesh descriptor");
        return EFE_ERR_ARG;
    }
    int NN, Q;
    switch (d->element_type) {
        case EFE_C3D8: NN = 8; Q = 8; break;
        case EFE_C3D20: NN = 20; Q = 27; break;
        case EFE_C3D20R: NN = 20; Q = 8; break;
        default: efe_set_error("unknown element type"); return EFE_ERR_ARG;
    }
    int np = (d->material == EFE_LINEAR_ELASTIC) ? 2 : 6;
    const int32_t* mat_of = d->material_of_element;
    int mat_count[3] = {0, 0, 0};
    if (mat_of) {
        if (d->n_props != 6 || !d->props_per_element) {
            efe_set_error("several sections need per-element properties with n_props = 6");
            return EFE_ERR_ARG;
        }
        for (int e = 0; e < d->n_elements; ++e) {
            if (mat_of[e] < 0 || mat_of[e] > 2) {
                efe_set_error("material_of_element: unknown material");
                return EFE_ERR_ARG;
            }
            ++mat_count[mat_of[e]];
        }
        np = 6;
    } else if (d->material < 0 || d->material > 2 || d->n_props != np) {
        efe_set_error("material / n_props mismatch (linear elastic: 2, von Mises and Drucker-Prager: 6)");
        return EFE_ERR_ARG;
    }
    if ((int64_t)d->n_nodes * 3 >= INT32_MAX || (int64_t)d->n_elements * NN * NN >= INT32_MAX) {
        efe_set_error("mesh too large for 32-bit indices");
        return EFE_ERR_ARG;
    }
    efe_ctx* c = new efe_ctx();
    c->N = d->n_nodes;
    c->E = d->n_elements;
    c->NN = NN;
    c->Q = Q;
    c->D = 3 * c->N;
    c->eltype = d->element_type;
    c->material = d->material;
    c->nprops = np;
    c->props_per_element = d->props_per_element ? 1 : 0;
    c->nstate = 12 + (d->material == EFE_LINEAR_ELASTIC ? 0 : 1);
    if (mat_of) {
        c->ngroups = 0;
        for (int m = 0; m < 3; ++m)
            if (mat_count[m]) {
                c->groupMat[c->ngroups] = m;
                c->groupCount[c->ngroups] = mat_count[m];
                ++c->ngroups;
            }
        c->material = c->groupMat[c->ngroups - 1];  // the material with the most state (LE < J2 < DP)
        c->matOf.assign(mat_of, mat_of + d->n_elements);
        c->nstate = 12 + (mat_count[EFE_VONMISES] || mat_count[EFE_DRUCKER_PRAGER] ? 1 : 0);
    } else {
        c->groupMat[0] = c->material;
        c->groupCount[0] = c->E;
    }
    c->mfNC = (mat_of ? mat_count[EFE_DRUCKER_PRAGER] > 0 : d->material == EFE_DRUCKER_PRAGER) ? 10 : 8;
    c->stream = (hipStream_t)d->stream;
    c->nspec = d->specimen_of_node ? d->n_specimens : 1;
    std::string err;
    if (!efe_build_pattern(c->hp, c->N, c->E, NN, d->conn, err)) {
        efe_set_error(err);
        delete c;
        return EFE_ERR_ARG;
    }
    HostPattern& hp = c->hp;
    c->nb = hp.bptr[c->N];
    c->nnz = c->nb * 9;
    c->nslices = (int)hp.slice_deg.size();
    hipStream_t st = c->stream;
    int rc = EFE_OK;
    auto fail = [&](int r) {
        efe_ctx_destroy(c);
        return r;
    };
    {
        std::vector<double> xyz((size_t)3 * c->N);
        for (int n = 0; n < c->N; ++n)
            for (int k = 0; k < 3; ++k) xyz[(size_t)k * c->N + n] = d->coords[(size_t)n * 3 + k];
        if ((rc = upload(&c->coords, xyz.data(), xyz.size(), st))) return fail(rc);
        std::vector<int32_t> cn((size_t)NN * c->E);
        for (int e = 0; e < c->E; ++e)
            for (int a = 0; a < NN; ++a) cn[(size_t)a * c->E + e] = d->conn[(size_t)e * NN + a];
        if ((rc = upload(&c->conn, cn.data(), cn.size(), st))) return fail(rc);
        std::vector<double> pr;
        if (c->props_per_element) {
            pr.resize((size_t)np * c->E);
            for (int e = 0; e < c->E; ++e)
                for (int k = 0; k < np; ++k) pr[(size_t)k * c->E + e] = d->props[(size_t)e * np + k];
        } else {
            pr.assign(d->props, d->props + np);
        }
        if ((rc = upload(&c->props, pr.data(), pr.size(), st))) return fail(rc);
        if (mat_of && c->ngroups > 1)
            for (int g = 0; g < c->ngroups; ++g) {
                std::vector<int32_t> el;
                el.reserve(c->groupCount[g]);
                for (int e = 0; e < c->E; ++e)
                    if (mat_of[e] == c->groupMat[g]) el.push_back(e);
                if ((rc = upload(&c->groupElems[g], el.data(), el.size(), st))) return fail(rc);
            }
        std::vector<double> xi, eta, zeta, w;
        qp_table(c->eltype, xi, eta, zeta, w);
        std::vector<double> dn((size_t)Q * NN * 3), sn((size_t)Q * NN), tmp((size_t)NN * 3), tn((size_t)Q * NN);
        for (int q = 0; q < Q; ++q) {
            // J and B: the reference evaluates QP q at (eta_q, xi_q, zeta_q)
            const double pe[3] = {eta[q], xi[q], zeta[q]};
            shape(NN, pe, tn.data() + (size_t)q * NN, dn.data() + (size_t)q * NN * 3);
            // N for body forces: at the table point (xi_q, eta_q, zeta_q)
            const double pt[3] = {xi[q], eta[q], zeta[q]};
            shape(NN, pt, sn.data() + (size_t)q * NN, tmp.data());
        }
        // The table J is formed with.  hexa20: the reference's third Jacobian row takes the
        // coefficient a5 of the monomial (eta zeta) of its polynomial fit times zeta_q where the
        // derivative has xi_q (_elementcomputationmatrices.py:502,514,526); a5 = sum_a h_a x_a with
        // h_a = d2N_a / (d eta d zeta) at the centre (+-1/4 on the mid-side nodes (0, +-1, +-1)),
        // so J's third row gains (zeta_q - xi_q) h_a per node.  Zero on undistorted elements.
        std::vector<double> dj(dn);
        if (NN == 20)
            for (int q = 0; q < Q; ++q)
                for (int a = 0; a < NN; ++a)
                    if (kNat20[a][0] == 0)
                        dj[((size_t)q * NN + a) * 3 + 2] += (zeta[q] - xi[q]) * kNat20[a][1] * kNat20[a][2] / 4.0;
        if ((rc = upload(&c->dNdxi, dn.data(), dn.size(), st))) return fail(rc);
        if ((rc = upload(&c->dJdxi, dj.data(), dj.size(), st))) return fail(rc);
        if ((rc = upload(&c->shapeN, sn.data(), sn.size(), st))) return fail(rc);
        if ((rc = upload(&c->shapeNe, tn.data(), tn.size(), st))) return fail(rc);
        if ((rc = upload(&c->qpw, w.data(), w.size(), st))) return fail(rc);
        if (d->specimen_of_node) {
            // specimens: contiguous, non-decreasing node ranges; every element inside one specimen
            const int32_t* sp = d->specimen_of_node;
            std::vector<int64_t> off(c->nspec + 1, 0);
            bool okk = sp[0] == 0;
            for (int n = 0; n < c->N && okk; ++n) {
                if (sp[n] < 0 || sp[n] >= c->nspec || (n > 0 && sp[n] != sp[n - 1] && sp[n] != sp[n - 1] + 1)) okk = false;
                else off[sp[n] + 1] = n + 1;
            }
            okk = okk && sp[c->N - 1] == c->nspec - 1;
            std::vector<int32_t> se(c->E);
            for (int e = 0; e < c->E && okk; ++e) {
                se[e] = sp[d->conn[(size_t)e * NN]];
                for (int a = 1; a < NN; ++a)
                    if (sp[d->conn[(size_t)e * NN + a]] != se[e]) okk = false;
            }
            if (!okk) {
                efe_set_error("specimens must be contiguous node ranges 0..n_specimens-1 with elements inside one specimen");
                return fail(EFE_ERR_ARG);
            }
            if ((rc = upload(&c->specOfNode, d->specimen_of_node, (size_t)c->N, st))) return fail(rc);
            if ((rc = upload(&c->specNodeOff, off.data(), off.size(), st))) return fail(rc);
            if ((rc = upload(&c->specOfElem, se.data(), se.size(), st))) return fail(rc);
            EFE_HIP(hipMalloc(&c->specStatus, (size_t)c->nspec * sizeof(int)));
            EFE_HIP(hipMemsetAsync(c->specStatus, 0, (size_t)c->nspec * sizeof(int), st));
            EFE_HIP(hipMalloc(&c->specActive, (size_t)c->nspec * sizeof(int)));
        }
    }
    if ((rc = upload(&c->bptr, hp.bptr.data(), hp.bptr.size(), st))) return fail(rc);
    if ((rc = upload(&c->bcol, hp.bcol.data(), hp.bcol.size(), st))) return fail(rc);
    if ((rc = upload(&c->coff, hp.coff.data(), hp.coff.size(), st))) return fail(rc);
    if ((rc = upload(&c->cpos, hp.cpos.data(), hp.cpos.size(), st))) return fail(rc);
    if ((rc = upload(&c->cslice, hp.cslice_off.data(), hp.cslice_off.size(), st))) return fail(rc);
    c->scratchBlocks = hp.scratch_blocks;
    if ((rc = upload(&c->elemPos, hp.epos.data(), hp.epos.size(), st))) return fail(rc);
    if ((rc = upload(&c->nptr, hp.nptr.data(), hp.nptr.size(), st))) return fail(rc);
    if ((rc = upload(&c->nlist, hp.nlist.data(), hp.nlist.size(), st))) return fail(rc);
    if ((rc = upload(&c->diagk, hp.diagk.data(), hp.diagk.size(), st))) return fail(rc);
    if ((rc = upload(&c->slice_off, hp.slice_off.data(), hp.slice_off.size(), st))) return fail(rc);
    if ((rc = upload(&c->scol, hp.scol.data(), hp.scol.size(), st))) return fail(rc);
    if ((rc = upload(&c->slotNode, hp.slot_node.data(), hp.slot_node.size(), st))) return fail(rc);
    if ((rc = upload(&c->nodeSlot, hp.node_slot.data(), hp.node_slot.size(), st))) return fail(rc);
    // no host copies of the large gather lists are needed after upload
    std::vector<int32_t>().swap(hp.cpos);
    std::vector<int32_t>().swap(hp.epos);
    std::vector<int32_t>().swap(hp.coff);
    std::vector<int32_t>().swap(hp.scol);
    std::vector<int32_t>().swap(hp.nlist);
    const size_t sbytes = (size_t)c->nstate * c->E * Q * sizeof(double);
    auto dalloc = [&](void** p, size_t bytes) -> int {
        EFE_HIP(hipMalloc(p, bytes ? bytes : 1));
        EFE_HIP(hipMemsetAsync(*p, 0, bytes ? bytes : 1, st));
        return EFE_OK;
    };
    if ((rc = dalloc((void**)&c->stateRef, sbytes))) return fail(rc);
    if ((rc = dalloc((void**)&c->stateTmp, sbytes))) return fail(rc);
    if ((rc = dalloc((void**)&c->Pscr, (size_t)c->E * NN * 3 * sizeof(double)))) return fail(rc);
    if ((rc = dalloc((void**)&c->dirflag, (size_t)c->D))) return fail(rc);
    if ((rc = dalloc((void**)&c->status, 4 * sizeof(int)))) return fail(rc);
    c->redn = 8 + 5 * 1024;
    if ((rc = dalloc((void**)&c->red, (size_t)c->redn * sizeof(double)))) return fail(rc);
    for (auto& ev : c->ev) EFE_HIP(hipEventCreate(&ev));
    EFE_HIP(hipStreamSynchronize(st));
    *out = c;
    return EFE_OK;
}

void efe_ctx_destroy(efe_ctx* c) {
    if (!c) return;
    hipDeviceSynchronize();
    efe_krylov_free(c);
    efe_gmres_free(c);
    efe_batched_free(c);
    efe_mf_free(c);
    void* bufs[] = {c->shapeNe, c->dJdxi, c->geoElems, c->surfElems, c->surfMask, c->resIdx,
                    c->coords, c->conn,  c->props,  c->dNdxi,  c->shapeN, c->qpw,    c->stateRef, c->stateTmp,
                    c->Kscr,   c->Pscr,  c->bptr,   c->bcol,   c->coff,   c->cpos,   c->cslice, c->nptr,     c->nlist,
                    c->diagk,  c->slice_off, c->scol, c->K64,  c->K32,    c->Minv64, c->Minv32,   c->dirflag,
                    c->slotNode, c->nodeSlot, c->specStatus, c->specActive, c->elemPos,
                    c->groupElems[0], c->groupElems[1], c->groupElems[2],
                    c->status, c->red,   c->specOfNode, c->specNodeOff, c->specOfElem};
    for (void* p : bufs)
        if (p) hipFree(p);
    for (auto& ev : c->ev)
        if (ev) hipEventDestroy(ev);
    delete c;
}

int efe_get_sizes(efe_ctx* c, int64_t out[4]) {
    if (int rc = check(c)) return rc;
    out[0] = c->D;
    out[1] = c->nnz;
    out[2] = c->Q;
    out[3] = c->nstate;
    return EFE_OK;
}

int efe_get_csr_pattern(efe_ctx* c, int32_t* indptr, int32_t* indices) {
    if (int rc = check(c)) return rc;
    const HostPattern& hp = c->hp;
    indptr[0] = 0;
    for (int n = 0; n < c->N; ++n) {
        const int deg = hp.bptr[n + 1] - hp.bptr[n];
        for (int i = 0; i < 3; ++i) {
            const int64_t row = 3 * (int64_t)n + i;
            const int64_t r0 = 9 * (int64_t)hp.bptr[n] + 3 * (int64_t)i * deg;
            indptr[row + 1] = (int32_t)(r0 + 3 * deg);
            for (int k = 0; k < deg; ++k)
                for (int j = 0; j < 3; ++j) indices[r0 + 3 * k + j] = 3 * hp.bcol[hp.bptr[n] + k] + j;
        }
    }
    return EFE_OK;
}

int efe_set_dirichlet(efe_ctx* c, const int32_t* dofs, int32_t n) {
    if (int rc = check(c)) return rc;
    std::vector<uint8_t> f((size_t)c->D, 0);
    for (int32_t i = 0; i < n; ++i) {
        if (dofs[i] < 0 || dofs[i] >= c->D) {
            efe_set_error("Dirichlet DOF out of range");
            return EFE_ERR_ARG;
        }
        f[dofs[i]] = 1;
    }
    EFE_HIP(hipMemcpyAsync(c->dirflag, f.data(), f.size(), hipMemcpyHostToDevice, c->stream));
    EFE_HIP(hipStreamSynchronize(c->stream));
    c->ndir = n;
    return EFE_OK;
}

int efe_compute_elements(efe_ctx* c, const double* U, const double* dU, const double* time, double dT, double* P,
                         double* F, const double* dir_values, double* dir_correction, int32_t precision,
                         double* pNewDT) {
    (void)U;
    (void)time;
    (void)dT;
    if (int rc = check(c)) return rc;
    if (!dU || !P || !F || (dir_values && !dir_correction) ||
        (precision != EFE_FP64 && precision != EFE_FP32 && precision != EFE_FP32_MATFREE)) {
        efe_set_error("efe_compute_elements: invalid arguments");
        return EFE_ERR_ARG;
    }
    if (int rc = efe_prepare_elements(c, precision)) return rc;
    EFE_HIP(hipMemsetAsync(c->status, 0, sizeof(int), c->stream));
    if (c->specStatus) EFE_HIP(hipMemsetAsync(c->specStatus, 0, (size_t)c->nspec * sizeof(int), c->stream));
    EFE_HIP(hipEventRecord(c->ev[0], c->stream));
    if (int rc = efe_launch_elements(c, dU, precision)) return rc;
    EFE_HIP(hipEventRecord(c->ev[1], c->stream));
    if (precision == EFE_FP32_MATFREE) {
        if (int rc = efe_launch_mf_assembly(c, P, F, dir_values, dir_correction)) return rc;
    } else if (int rc = efe_launch_assembly(c, P, F, dir_values, dir_correction, precision)) {
        return rc;
    }
    EFE_HIP(hipEventRecord(c->ev[2], c->stream));
    c->kscrValid = precision != EFE_FP32_MATFREE;
    c->pscrValid = true;
    int hs = 0;
    EFE_HIP(hipMemcpyAsync(&hs, c->status, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    EFE_HIP(hipStreamSynchronize(c->stream));
    c->t_elem = elapsed(c->ev[0], c->ev[1]);
    c->t_asm = elapsed(c->ev[1], c->ev[2]);
    c->tot[EFE_T_ELEMENTS] += c->t_elem;
    c->tot[EFE_T_ELEMENT_PASSES] += 1;
    c->tot[EFE_T_ASSEMBLY] += c->t_asm;
    if (pNewDT) *pNewDT = hs ? 0.5 : 1.0;
    return hs ? EFE_CUTBACK : EFE_OK;
}

int efe_compute_residual(efe_ctx* c, const double* U, const double* dU, const double time[2], double dT, double* P,
                         double* F, int precision, double* pNewDT) {
    (void)U;
    (void)time;
    (void)dT;
    if (int rc = check(c)) return rc;
    if (!dU || !P || !F || (precision != EFE_FP64 && precision != EFE_FP32 && precision != EFE_FP32_MATFREE)) {
        efe_set_error("efe_compute_residual: invalid arguments");
        return EFE_ERR_ARG;
    }
    if (!c->Kscr) EFE_HIP(hipMalloc(&c->Kscr, (size_t)c->scratchBlocks * kBlockPitch * sizeof(double)));
    EFE_HIP(hipMemsetAsync(c->status, 0, sizeof(int), c->stream));
    if (c->specStatus) EFE_HIP(hipMemsetAsync(c->specStatus, 0, (size_t)c->nspec * sizeof(int), c->stream));
    EFE_HIP(hipEventRecord(c->ev[5], c->stream));
    if (int rc = efe_launch_residual(c, dU, P, F, precision)) return rc;
    EFE_HIP(hipEventRecord(c->ev[6], c->stream));
    c->kscrValid = false;  // the blocks are those of an earlier state now
    c->pscrValid = true;
    int hs = 0;
    EFE_HIP(hipMemcpyAsync(&hs, c->status, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    EFE_HIP(hipStreamSynchronize(c->stream));
    c->tot[EFE_T_RESIDUAL] += elapsed(c->ev[5], c->ev[6]);
    c->tot[EFE_T_RESIDUAL_PASSES] += 1;
    if (pNewDT) *pNewDT = hs ? 0.5 : 1.0;
    return hs ? EFE_CUTBACK : EFE_OK;
}

int efe_body_force(efe_ctx* c, const int32_t* elements, int32_t n, const double* force3, double* PExt) {
    if (int rc = check(c)) return rc;
    if (!force3 || !PExt || (elements && n < 0)) {
        efe_set_error("efe_body_force: invalid arguments");
        return EFE_ERR_ARG;
    }
    if (!elements) return efe_launch_body_force(c, nullptr, c->E, force3, PExt);
    // the element set, each element once, ascending (an element set holds an element once)
    std::vector<uint8_t> in((size_t)c->E, 0);
    for (int32_t i = 0; i < n; ++i) {
        if (elements[i] < 0 || elements[i] >= c->E) {
            efe_set_error("efe_body_force: element out of range");
            return EFE_ERR_ARG;
        }
        in[elements[i]] = 1;
    }
    std::vector<int32_t> el;
    for (int e = 0; e < c->E; ++e)
        if (in[e]) el.push_back(e);
    if (el.size() > c->resCap) {
        hipFree(c->resIdx);
        c->resIdx = nullptr;
        c->resCap = 0;
        EFE_HIP(hipMalloc(&c->resIdx, el.size() * sizeof(int32_t)));
        c->resCap = el.size();
    }
    if (!el.empty())
        EFE_HIP(hipMemcpyAsync(c->resIdx, el.data(), el.size() * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    if (int rc = efe_launch_body_force(c, c->resIdx, (int)el.size(), force3, PExt)) return rc;
    EFE_HIP(hipStreamSynchronize(c->stream));  // el must outlive the copy
    return EFE_OK;
}

int efe_surface_traction(efe_ctx* c, const int32_t* elements, const int32_t* faces, int32_t n,
                         const double* traction3, double* PExt) {
    if (int rc = check(c)) return rc;
    if (n < 0 || (n > 0 && (!elements || !faces)) || !traction3 || !PExt) {
        efe_set_error("efe_surface_traction: invalid arguments");
        return EFE_ERR_ARG;
    }
    // faces of each loaded element as a bit mask, elements in ascending order
    std::vector<uint8_t> mask((size_t)c->E, 0);
    for (int32_t i = 0; i < n; ++i) {
        if (elements[i] < 0 || elements[i] >= c->E || faces[i] < 1 || faces[i] > 6) {
            efe_set_error("efe_surface_traction: element out of range or face not in 1..6");
            return EFE_ERR_ARG;
        }
        mask[elements[i]] |= (uint8_t)(1u << (faces[i] - 1));
    }
    std::vector<int32_t> el;
    std::vector<uint8_t> fm;
    for (int e = 0; e < c->E; ++e)
        if (mask[e]) {
            el.push_back(e);
            fm.push_back(mask[e]);
        }
    const size_t m = el.size();
    if (m > c->surfCap) {
        hipFree(c->surfElems);
        hipFree(c->surfMask);
        c->surfElems = nullptr;
        c->surfMask = nullptr;
        c->surfCap = 0;
        EFE_HIP(hipMalloc(&c->surfElems, m * sizeof(int32_t)));
        EFE_HIP(hipMalloc(&c->surfMask, m));
        c->surfCap = m;
    }
    if (m) {
        EFE_HIP(hipMemcpyAsync(c->surfElems, el.data(), m * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
        EFE_HIP(hipMemcpyAsync(c->surfMask, fm.data(), m, hipMemcpyHostToDevice, c->stream));
    }
    if (int rc = efe_launch_surface_traction(c, c->surfElems, c->surfMask, (int)m, traction3, PExt)) return rc;
    // the host vectors above must outlive the copies
    EFE_HIP(hipStreamSynchronize(c->stream));
    return EFE_OK;
}

int efe_set_geostatic(efe_ctx* c, const double def[6], const int32_t* elements, int32_t n) {
    if (int rc = check(c)) return rc;
    if (!def || (elements && (n < 0 || n > c->E)) || def[3] == def[1]) {
        efe_set_error("efe_set_geostatic: invalid arguments (h1 == h2 or element list)");
        return EFE_ERR_ARG;
    }
    if (elements)
        for (int32_t i = 0; i < n; ++i)
            if (elements[i] < 0 || elements[i] >= c->E) {
                efe_set_error("efe_set_geostatic: element index out of range");
                return EFE_ERR_ARG;
            }
    int rc = efe_launch_geostatic(c, def, elements, elements ? n : c->E);
    if (rc) return rc;
    EFE_HIP(hipStreamSynchronize(c->stream));
    return EFE_OK;
}

int efe_export_csr_values(efe_ctx* c, double* values) {
    if (int rc = check(c)) return rc;
    if (c->Kprec < 0 || c->Kprec == EFE_FP32_MATFREE) {
        efe_set_error("no assembled operator");
        return EFE_ERR_STATE;
    }
    return efe_launch_export_csr(c, values);
}

int efe_get_element_matrices(efe_ctx* c, const int32_t* elements, int32_t n, double* Ke, double* Pe) {
    if (int rc = check(c)) return rc;
    if (n < 0 || (n > 0 && (!elements || !Ke))) {
        efe_set_error("efe_get_element_matrices: invalid arguments");
        return EFE_ERR_ARG;
    }
    if (c->Kprec < 0 || c->Kprec == EFE_FP32_MATFREE || !c->kscrValid) {
        efe_set_error("efe_get_element_matrices: the last element pass formed no element matrices "
                      "(matrix-free or residual-only)");
        return EFE_ERR_STATE;
    }
    if (Pe && !c->pscrValid) {
        efe_set_error("efe_get_element_matrices: a load computation has overwritten the element residuals "
                      "since the last element pass");
        return EFE_ERR_STATE;
    }
    for (int32_t i = 0; i < n; ++i)
        if (elements[i] < 0 || elements[i] >= c->E) {
            efe_set_error("efe_get_element_matrices: element out of range");
            return EFE_ERR_ARG;
        }
    if (n == 0) return EFE_OK;
    if ((size_t)n > c->resCap) {
        hipFree(c->resIdx);
        c->resIdx = nullptr;
        c->resCap = 0;
        EFE_HIP(hipMalloc(&c->resIdx, (size_t)n * sizeof(int32_t)));
        c->resCap = (size_t)n;
    }
    EFE_HIP(hipMemcpyAsync(c->resIdx, elements, (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    if (int rc = efe_launch_export_elements(c, c->resIdx, n, Ke, Pe)) return rc;
    EFE_HIP(hipStreamSynchronize(c->stream));
    return EFE_OK;
}

int efe_spmv(efe_ctx* c, const double* x, double* y) {
    if (int rc = check(c)) return rc;
    if (!x || !y) {
        efe_set_error("efe_spmv: invalid arguments");
        return EFE_ERR_ARG;
    }
    if (c->Kprec == EFE_FP32_MATFREE) return efe_launch_spmv_mf(c, x, y);
    return efe_launch_spmv64(c, x, y);
}

int efe_linsolve(efe_ctx* c, const double* b, double* x, const efe_solve_opts* o, efe_solve_info* info) {
    if (int rc = check(c)) return rc;
    if (!b || !x || !o || !info) {
        efe_set_error("efe_linsolve: invalid arguments");
        return EFE_ERR_ARG;
    }
    return efe_krylov_solve(c, b, x, o, info);
}

int efe_convergence_norms(efe_ctx* c, const double* R, const double* ddU, const double* F, double out[5]) {
    if (int rc = check(c)) return rc;
    EFE_HIP(hipEventRecord(c->ev[3], c->stream));
    int rc = efe_launch_norms(c, R, ddU, F, out);
    EFE_HIP(hipEventRecord(c->ev[4], c->stream));
    EFE_HIP(hipEventSynchronize(c->ev[4]));
    c->t_norm = elapsed(c->ev[3], c->ev[4]);
    c->tot[EFE_T_NORMS] += c->t_norm;
    c->tot[EFE_T_NORM_CALLS] += 1;
    return rc;
}

int efe_convergence_norms_batched(efe_ctx* c, const double* R, const double* ddU, const double* F, double* out) {
    if (int rc = check(c)) return rc;
    return efe_launch_norms_batched(c, R, ddU, F, out);
}

int efe_linsolve_batched(efe_ctx* c, const double* b, double* x, const efe_solve_opts* o, const double* spec_atol,
                         const double* spec_rtol, const uint8_t* spec_active, int32_t* spec_iterations,
                         efe_solve_info* info) {
    if (int rc = check(c)) return rc;
    if (!b || !x || !o || !info || !spec_atol) {
        efe_set_error("efe_linsolve_batched: invalid arguments");
        return EFE_ERR_ARG;
    }
    return efe_batched_linsolve(c, b, x, o, spec_atol, spec_rtol, spec_active, spec_iterations, info);
}

int efe_accept_state_batched(efe_ctx* c, const uint8_t* accept) {
    if (int rc = check(c)) return rc;
    if (!c->specOfElem || !accept) {
        efe_set_error("efe_accept_state_batched needs a context created with specimens");
        return EFE_ERR_STATE;
    }
    return efe_launch_accept_batched(c, accept);
}

int efe_newton_round_batched(efe_ctx* c, double* dU, double* P, double* F, const double* PExt, const double* deltaT,
                             const uint8_t* live, const uint8_t* extrapolated, const uint8_t* has_dirichlet,
                             const int32_t* max_iter, const int32_t* max_grow, const double* n_acc,
                             const efe_newton_opts* o, int32_t* state, int32_t* iterations, int32_t* reason,
                             int32_t* work) {
    if (int rc = check(c)) return rc;
    if (!c->specOfNode || !c->specActive) {
        efe_set_error("efe_newton_round_batched needs a context created with specimens");
        return EFE_ERR_STATE;
    }
    if (!dU || !deltaT || !live || !extrapolated || !has_dirichlet || !max_iter || !max_grow || !n_acc || !o ||
        !state || !iterations || !reason || !work ||
        (o->precision != EFE_FP64 && o->precision != EFE_FP32 && o->precision != EFE_FP32_MATFREE)) {
        efe_set_error("efe_newton_round_batched: invalid arguments");
        return EFE_ERR_ARG;
    }
    for (int s = 0; s < c->nspec; ++s)
        if (max_iter[s] < 0 || max_grow[s] < 0 || !(n_acc[s] > 0.0)) {
            efe_set_error("efe_newton_round_batched: max_iter and max_grow must be >= 0, n_acc > 0");
            return EFE_ERR_ARG;
        }
    return efe_newton_round_batched_impl(c, dU, P, F, PExt, deltaT, live, extrapolated, has_dirichlet, max_iter,
                                         max_grow, n_acc, o, state, iterations, reason, work);
}

int efe_get_specimen_status(efe_ctx* c, uint8_t* out) {
    if (int rc = check(c)) return rc;
    if (!out) {
        efe_set_error("efe_get_specimen_status: out is NULL");
        return EFE_ERR_ARG;
    }
    if (!c->specStatus) {
        efe_set_error("efe_get_specimen_status needs a context created with specimens");
        return EFE_ERR_STATE;
    }
    std::vector<int> h(c->nspec);
    EFE_HIP(hipMemcpyAsync(h.data(), c->specStatus, h.size() * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    EFE_HIP(hipStreamSynchronize(c->stream));
    for (int s = 0; s < c->nspec; ++s) out[s] = h[s] ? 1 : 0;
    return EFE_OK;
}

int efe_set_active_specimens(efe_ctx* c, const uint8_t* active) {
    if (int rc = check(c)) return rc;
    if (!c->specActive) {
        efe_set_error("efe_set_active_specimens needs a context created with specimens");
        return EFE_ERR_STATE;
    }
    if (!active) {
        c->useActive = false;
        return EFE_OK;
    }
    std::vector<int> a(c->nspec);
    for (int s = 0; s < c->nspec; ++s) a[s] = active[s] ? 1 : 0;
    EFE_HIP(hipMemcpyAsync(c->specActive, a.data(), a.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    EFE_HIP(hipStreamSynchronize(c->stream));  // a must outlive the copy
    c->useActive = true;
    return EFE_OK;
}

int efe_get_specimen_count(efe_ctx* c) {
    if (int rc = check(c)) return rc;
    return c->nspec;
}

int efe_accept_state(efe_ctx* c) {
    if (int rc = check(c)) return rc;
    std::swap(c->stateRef, c->stateTmp);
    return EFE_OK;
}

namespace {
// out[i][j][k] = state component first + k of element elems[i] (i when elems is null) at QP
// qps[j]; state layout [component][element][QP]
__global__ void collect_results_kernel(int64_t total, int nq, int cnt, int64_t EQ, int Q, const int32_t* __restrict__ elems,
                                       const int32_t* __restrict__ qps, int first, const double* __restrict__ st,
                                       double* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int k = (int)(t % cnt);
    const int j = (int)((t / cnt) % nq);
    const int64_t i = t / ((int64_t)cnt * nq);
    const int64_t e = elems ? elems[i] : i;
    out[t] = st[(int64_t)(first + k) * EQ + e * Q + qps[j]];
}
}  // namespace

int efe_collect_results(efe_ctx* c, int32_t which, const int32_t* elements, int32_t n, const int32_t* qps, int32_t nq,
                        int32_t first, int32_t count, double* out) {
    if (int rc = check(c)) return rc;
    if (!out || !qps || nq <= 0 || count <= 0 || first < 0 || first + count > c->nstate || (elements && n < 0)) {
        efe_set_error("efe_collect_results: invalid arguments");
        return EFE_ERR_ARG;
    }
    const int64_t ne = elements ? n : c->E;
    for (int32_t j = 0; j < nq; ++j)
        if (qps[j] < 0 || qps[j] >= c->Q) {
            efe_set_error("efe_collect_results: quadrature point out of range");
            return EFE_ERR_ARG;
        }
    for (int64_t i = 0; elements && i < ne; ++i)
        if (elements[i] < 0 || elements[i] >= c->E) {
            efe_set_error("efe_collect_results: element out of range");
            return EFE_ERR_ARG;
        }
    if (ne == 0) return EFE_OK;
    // index lists on the device (kept between calls, grown on demand)
    const size_t need = (size_t)(elements ? ne : 0) + (size_t)nq;
    if (need > c->resCap) {
        hipFree(c->resIdx);
        c->resIdx = nullptr;
        c->resCap = 0;
        EFE_HIP(hipMalloc(&c->resIdx, need * sizeof(int32_t)));
        c->resCap = need;
    }
    std::vector<int32_t> idx(need);
    if (elements) std::copy(elements, elements + ne, idx.begin());
    std::copy(qps, qps + nq, idx.begin() + (elements ? ne : 0));
    EFE_HIP(hipMemcpyAsync(c->resIdx, idx.data(), need * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    const int64_t total = ne * nq * count;
    collect_results_kernel<<<(unsigned)((total + 255) / 256), 256, 0, c->stream>>>(
        total, nq, count, (int64_t)c->E * c->Q, c->Q, elements ? c->resIdx : nullptr, c->resIdx + (elements ? ne : 0),
        first, which ? c->stateTmp : c->stateRef, out);
    EFE_HIP(hipGetLastError());
    EFE_HIP(hipStreamSynchronize(c->stream));  // idx must outlive the copy
    return EFE_OK;
}


int efe_csr_linsolve(int32_t n, const int32_t* indptr, const int32_t* indices, const double* data, const double* b,
                     double* x, const efe_solve_opts* o, efe_solve_info* info) {
    if (int rc = check_solve_args(n, indptr, indices, data, b, x, o, info, "efe_csr_linsolve")) return rc;
    if (indptr[0] != 0) {
        efe_set_error("efe_csr_linsolve: indptr[0] must be 0");
        return EFE_ERR_ARG;
    }
    for (int32_t i = 0; i < n; ++i)
        if (indptr[i + 1] < indptr[i]) {
            efe_set_error("efe_csr_linsolve: indptr must be non-decreasing");
            return EFE_ERR_ARG;
        }
    const int64_t nnz = indptr[n];
    for (int64_t k = 0; k < nnz; ++k)
        if (indices[k] < 0 || indices[k] >= n) {
            efe_set_error("efe_csr_linsolve: column index out of range");
            return EFE_ERR_ARG;
        }
    DeviceBuffers db;
    int32_t *dip, *dix;
    double *dv, *dbv, *dx, *ws;
    EFE_HIP(db.alloc(&dip, (size_t)n + 1));
    EFE_HIP(db.alloc(&dix, (size_t)nnz));
    EFE_HIP(db.alloc(&dv, (size_t)nnz));
    EFE_HIP(db.alloc(&dbv, (size_t)n));
    EFE_HIP(db.alloc(&dx, (size_t)n));
    EFE_HIP(db.alloc(&ws, efe_csr_workspace_doubles(n, o->solver, o->restart)));
    hipStream_t st = nullptr;
    EFE_HIP(hipMemcpyAsync(dip, indptr, ((size_t)n + 1) * sizeof(int32_t), hipMemcpyHostToDevice, st));
    EFE_HIP(hipMemcpyAsync(dix, indices, (size_t)nnz * sizeof(int32_t), hipMemcpyHostToDevice, st));
    EFE_HIP(hipMemcpyAsync(dv, data, (size_t)nnz * sizeof(double), hipMemcpyHostToDevice, st));
    EFE_HIP(hipMemcpyAsync(dbv, b, (size_t)n * sizeof(double), hipMemcpyHostToDevice, st));
    if (int rc = efe_csr_solve_ws(n, dip, dix, dv, dbv, dx, o, info, ws, st)) return rc;
    EFE_HIP(hipMemcpyAsync(x, dx, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, st));
    EFE_HIP(hipStreamSynchronize(st));
    return EFE_OK;
}

int efe_csr_linsolve_device(int32_t n, const int32_t* indptr, const int32_t* indices, const double* data,
                            const double* b, double* x, void* stream, const efe_solve_opts* o, efe_solve_info* info) {
    if (int rc = check_solve_args(n, indptr, indices, data, b, x, o, info, "efe_csr_linsolve_device")) return rc;
    hipStream_t st = (hipStream_t)stream;
    DeviceBuffers db;
    double* ws;
    int* flag;
    EFE_HIP(db.alloc(&ws, efe_csr_workspace_doubles(n, o->solver, o->restart)));
    EFE_HIP(db.alloc(&flag, 1));
    int bad = 0;
    if (int rc = efe_csr_check_pattern(n, indptr, indices, flag, st, &bad)) return rc;
    if (bad) {
        efe_set_error("efe_csr_linsolve_device: invalid CSR pattern (indptr or column indices)");
        return EFE_ERR_ARG;
    }
    return efe_csr_solve_ws(n, indptr, indices, data, b, x, o, info, ws, st);
}

int efe_get_qp_state(efe_ctx* c, int32_t which, double* out) {
    if (int rc = check(c)) return rc;
    const size_t EQ = (size_t)c->E * c->Q, ns = c->nstate;
    if (!out) {
        efe_set_error("efe_get_qp_state: out is NULL");
        return EFE_ERR_ARG;
    }
    std::vector<double> buf(EQ * ns);
    EFE_HIP(hipMemcpyAsync(buf.data(), which ? c->stateTmp : c->stateRef, buf.size() * sizeof(double),
                           hipMemcpyDeviceToHost, c->stream));
    EFE_HIP(hipStreamSynchronize(c->stream));
    for (size_t k = 0; k < ns; ++k)
        for (size_t i = 0; i < EQ; ++i) out[i * ns + k] = buf[k * EQ + i];
    return EFE_OK;
}

int efe_set_qp_state(efe_ctx* c, const double* in) {
    if (int rc = check(c)) return rc;
    const size_t EQ = (size_t)c->E * c->Q, ns = c->nstate;
    if (!in) {  // zero state of a fresh model, cleared on the device
        EFE_HIP(hipMemsetAsync(c->stateRef, 0, EQ * ns * sizeof(double), c->stream));
        EFE_HIP(hipStreamSynchronize(c->stream));
        return EFE_OK;
    }
    std::vector<double> buf(EQ * ns);
    for (size_t k = 0; k < ns; ++k)
        for (size_t i = 0; i < EQ; ++i) buf[k * EQ + i] = in[i * ns + k];
    EFE_HIP(hipMemcpyAsync(c->stateRef, buf.data(), buf.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    EFE_HIP(hipStreamSynchronize(c->stream));
    return EFE_OK;
}

int efe_set_launch(efe_ctx* c, int32_t epw, int32_t kblock) {
    if (int rc = check(c)) return rc;
    if (epw > 0) c->epw = epw;
    if (kblock > 0) c->kblock = kblock;
    return EFE_OK;
}

int efe_get_phase_times(efe_ctx* c, double out[16]) {
    if (int rc = check(c)) return rc;
    out[0] = c->t_elem;
    out[1] = c->t_asm;
    out[2] = c->t_kry;
    out[3] = c->t_norm;
    for (int k = 0; k < 12; ++k) out[4 + k] = k < EFE_T_COUNT ? c->tot[k] : 0.0;
    return EFE_OK;
}

}  // extern "C"
