// Internal declarations of the gfx950 NIST-step library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/efe_gpu.h"

#define EFE_HIP(call)                                                                   \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            efe_set_error(std::string(#call) + ": " + hipGetErrorString(e_));           \
            return EFE_ERR_HIP;                                                         \
        }                                                                               \
    } while (0)

void efe_set_error(const std::string& msg);

// Slice height of the sliced-ELL 3x3 block layout: one wavefront of row blocks.
constexpr int kSlice = 64;
// Rows of the sliced operator are ordered by decreasing length inside windows of this many
// consecutive nodes, so that slices hold rows of equal length (hexa20 grids mix 81-block
// corner rows with 51-block mid-side rows) while rows stay near their mesh neighbours.
constexpr int kSortWindow = 1024;

// Element scratch: one 3x3 block (9 values) per contribution.  (Padding the blocks to 16-byte
// vectors was measured slower on MI355X: the extra bytes cost more than the wider accesses
// save; at an 80-byte pitch with five 16-byte stores the 64^3 hexa8 pass took 1.01 against
// 0.98 ms.)
constexpr int kBlockPitch = 9;

template <typename KT>
__device__ __forceinline__ void store_block(KT* dst, const KT (&a)[9]) {
#pragma unroll
    for (int v = 0; v < 9; ++v) dst[v] = a[v];
}

template <typename KT>
__device__ __forceinline__ void add_block(const KT* src, KT (&acc)[9]) {
#pragma unroll
    for (int v = 0; v < 9; ++v) acc[v] += src[v];
}

struct HostPattern {
    int N = 0, E = 0, NN = 0;
    std::vector<int32_t> bptr, bcol;      // node adjacency (block CSR), columns ascending
    std::vector<int32_t> diagk;           // position of the diagonal block in each row
    std::vector<int32_t> coff, cpos;      // per block slot: its contribution range; per element block: its position
    std::vector<int32_t> cslice_off;      // [nslices+1] contribution rows of the element scratch per slice
    int64_t scratch_blocks = 0;           // element scratch size in 3x3 blocks (slices padded)
    std::vector<int32_t> epos;            // [a][e]: index of (e, a) in nlist
    std::vector<int32_t> nptr, nlist;     // per node: (e*NN + a) incidences, element order
    std::vector<int32_t> slice_off;       // [nslices+1] first padded block of each slice
    std::vector<int32_t> slice_deg;       // [nslices] padded row length of each slice
    std::vector<int32_t> scol;            // [padded blocks] block column (padding: the row itself)
    std::vector<int32_t> slot_node;       // [N] node of operator row slot p (p = 64 s + l)
    std::vector<int32_t> node_slot;       // [N] inverse
    int64_t padded_blocks = 0;
};

// Multicolor block ILU(0) structure (rows in elimination order = (color, node)).
struct IluPattern {
    int ncolors = 0;
    std::vector<int32_t> color_start;  // [ncolors+1] first factor row of each color
    std::vector<int32_t> perm, iperm;  // factor row -> node, node -> factor row
    std::vector<int32_t> slice_off;    // sliced layout of the factor
    std::vector<int32_t> scol;         // factor block columns (new numbering), ascending per row
    std::vector<int32_t> kpos;         // per factor block: value base of the same block in the operator
    std::vector<int32_t> rowlen, dslot;
};

bool efe_build_ilu_pattern(const HostPattern& hp, IluPattern& ip, std::string& err);

// Builds the adjacency, gather maps and sliced layout from the element connectivity
// (host, conn [E][NN] node positions).  Returns false on inconsistent input.
bool efe_build_pattern(HostPattern& hp, int N, int E, int NN, const int32_t* conn, std::string& err);

// running totals of efe_get_phase_times (ms, or counts)
enum {
    EFE_T_ELEMENTS = 0,   // element kernel of tangent passes
    EFE_T_ELEMENT_PASSES, // count
    EFE_T_ASSEMBLY,       // assembly (or matrix-free diagonal gather)
    EFE_T_RESIDUAL,       // residual-only passes (element kernel + P/F gather)
    EFE_T_RESIDUAL_PASSES,
    EFE_T_KRYLOV,         // Krylov iterations (without preconditioner set-up)
    EFE_T_KRYLOV_ITERS,
    EFE_T_SOLVES,
    EFE_T_PRECOND,        // ILU(0) factorisation
    EFE_T_NORMS,
    EFE_T_NORM_CALLS,
    EFE_T_COUNT
};

struct KrylovWork;
struct GmresWork;
struct BatchWork;

struct efe_ctx {
    int N = 0, E = 0, NN = 0, Q = 0, D = 0;
    int eltype = 0, material = 0, nprops = 0, props_per_element = 0;
    int nstate = 0;  // per QP: 12 + material state
    int64_t nnz = 0; // CSR entries
    int64_t nb = 0;  // 3x3 blocks
    int nspec = 1;
    hipStream_t stream = nullptr;
    HostPattern hp;

    // device mesh and element data
    double* coords = nullptr;    // [3][N]
    int32_t* conn = nullptr;     // [NN][E]
    double* props = nullptr;     // [nprops][E] or [nprops]
    double* dNdxi = nullptr;     // [Q][NN][3] at the evaluation points
    double* dJdxi = nullptr;     // [Q][NN][3] the table J is formed with (hexa20: the reference's third row)
    double* shapeN = nullptr;    // [Q][NN] at the table points (body force)
    double* shapeNe = nullptr;   // [Q][NN] at the evaluation points (QP coordinates)
    // matrix-free FP32 operator: [Q][NC][E] compact tangent, [E*NN][3] node-sorted force slots,
    // dnxi32 = FP32 dNdxi then dJdxi ([2][Q][NN][3])
    float *mfC = nullptr, *mfF = nullptr, *dnxi32 = nullptr;
    // hexa8 matrix-free operator by workgroup bricks (256 consecutive elements): per workgroup its
    // local nodes (lptr [G+1] into lnode / lpos / lincp), per local node its partial-force slot in
    // node order (lpos) and its element-node incidences t*8+a (linc, element order, from lincp);
    // per element its 8 local node indices (lidx); per node its partial range (pptr [N+1]) in
    // the partial forces mfbF [L][3]
    bool mfb = false;
    int mfbG = 0;                // workgroups (bricks); in batched contexts a brick holds one specimen's elements
    int32_t* mfbE0 = nullptr;    // [G+1] first element of each brick
    int32_t *mfbLptr = nullptr, *mfbLnode = nullptr, *mfbLpos = nullptr, *mfbLincp = nullptr, *mfbPptr = nullptr;
    uint16_t *mfbLinc = nullptr, *mfbLidx = nullptr;
    float* mfbF = nullptr;
    int32_t* elemPos = nullptr;  // [NN][E] index of (e, a) in the node-sorted incidence list nlist
                                 // (force slot of the matrix-free operator; block-row (e, a) of the
                                 // element scratch starts at block elemPos * NN)
    bool mfPass = false;             // the current element launch is a matrix-free pass
    int mfNC = 8;                    // compact tangent values per QP of the matrix-free operator
                                     // (10 when a Drucker-Prager section is present)
    // material groups (sections): the element kernel runs once per group over its element list
    int ngroups = 1;
    int groupMat[3] = {0, 0, 0};
    int groupCount[3] = {0, 0, 0};
    int32_t* groupElems[3] = {nullptr, nullptr, nullptr};  // device, ascending; nullptr: all elements
    std::vector<int8_t> matOf;       // host material of every element (several sections only)
    int curMat = 0;                  // the group of the current element launch
    const int32_t* curList = nullptr;
    int curCount = 0;
    int32_t* geoElems = nullptr; // [E] element subset of a geostatic initial state
    int32_t* surfElems = nullptr;  // loaded elements of the last surface traction call
    uint8_t* surfMask = nullptr;   // their loaded faces (bit f-1)
    size_t surfCap = 0;
    int32_t* resIdx = nullptr;     // element and QP lists of the last result collection
    size_t resCap = 0;
    double* qpw = nullptr;       // [Q]
    double* stateRef = nullptr;  // [nstate][E][Q]
    double* stateTmp = nullptr;
    void* Kscr = nullptr;        // [E][NN][NN][9] in the assembly precision
    int KscrPrec = -1;
    double* Pscr = nullptr;      // [E][NN][3]
    int32_t* specOfNode = nullptr;   // [N] specimen of each node (batched contexts)
    int64_t* specNodeOff = nullptr;  // [nspec+1] first node of each specimen
    int32_t* specOfElem = nullptr;   // [E]

    // device pattern
    int32_t *bptr = nullptr, *bcol = nullptr, *coff = nullptr, *cpos = nullptr, *cslice = nullptr;
    int64_t scratchBlocks = 0;
    int32_t *nptr = nullptr, *nlist = nullptr, *diagk = nullptr;
    int32_t *slice_off = nullptr, *scol = nullptr;
    int32_t *slotNode = nullptr, *nodeSlot = nullptr;  // operator row slot <-> node
    int nslices = 0;

    // assembled operator
    double* K64 = nullptr;       // sliced layout, [padded*9]
    float* K32 = nullptr;
    double* Minv64 = nullptr;    // [N][9] inverse diagonal blocks
    float* Minv32 = nullptr;
    int Kprec = -1;              // precision of the current K
    uint8_t* dirflag = nullptr;  // [D]
    int ndir = 0;
    int* status = nullptr;       // device status word (cutback / NaN)
    int* specStatus = nullptr;   // [nspec] cutback request per specimen (batched contexts)
    int* specActive = nullptr;   // [nspec] element passes skip specimens with 0 (batched contexts)
    bool useActive = false;      // whether the element pass reads specActive
    int formK = 1;               // 0: the element pass forms residuals only (efe_compute_residual)
    bool kscrValid = false;      // Kscr holds the element blocks of the last element pass (assembled)
    bool pscrValid = false;      // Pscr holds the element residuals of the last element pass (no load
                                 // computation has used it as scratch since)
    double* red = nullptr;       // reduction scratch
    int redn = 0;

    KrylovWork* kw = nullptr;
    GmresWork* gw = nullptr;
    BatchWork* bw = nullptr;
    int Kversion = 0;            // incremented by every assembly

    // launch configuration
    int epw = 0;       // elements per workgroup of the element kernel (0 = default)
    int kblock = 256;  // threads per workgroup of Krylov kernels

    // phase timing: the last call of each phase, and totals over the context's life
    hipEvent_t ev[8] = {};
    double t_elem = 0, t_asm = 0, t_kry = 0, t_norm = 0;
    double tot[EFE_T_COUNT] = {};
};

// buffers of element passes in a precision (capi.hip)
int efe_prepare_elements(efe_ctx* c, int precision);
// element kernels (elements.hip)
int efe_launch_elements(efe_ctx* c, const double* dU, int precision);
int efe_launch_residual(efe_ctx* c, const double* dU, double* P, double* F, int precision);
int efe_launch_assembly(efe_ctx* c, double* P, double* F, const double* dir_values, double* dir_correction,
                        int precision);
int efe_launch_body_force(efe_ctx* c, const int32_t* elems, int n, const double* f3, double* PExt);
int efe_launch_surface_traction(efe_ctx* c, const int32_t* elems, const uint8_t* fmask, int n, const double* t3,
                                double* PExt);
int efe_launch_geostatic(efe_ctx* c, const double* def, const int32_t* elements, int32_t n);
// matrix-free FP32 operator (matfree.hip)
int efe_mf_alloc(efe_ctx* c);
int efe_launch_mf_assembly(efe_ctx* c, double* P, double* F, const double* dir_values, double* dir_correction);
// hexa8 bricks: the product's node partials (mfbF over mfbPptr) and per-brick partials of p.Kp
void efe_mf_launch_apply_partials(efe_ctx* c, hipStream_t st, const float* p, double* part, const int* flags);
void efe_mf_launch_operator(efe_ctx* c, hipStream_t st, const float* p, float* q, double* part, const int* flags,
                            int nwg, int bs, int kr);
void efe_mf_free(efe_ctx* c);
int efe_launch_spmv_mf(efe_ctx* c, const double* x, double* y);
void efe_mf_launch_operator_batched(efe_ctx* c, hipStream_t st, const float* p, float* q, double* npq,
                                    const int* active);
int efe_launch_export_csr(efe_ctx* c, double* values);
int efe_launch_export_elements(efe_ctx* c, const int32_t* elems, int n, double* Ke, double* Pe);

// Krylov (krylov.hip)
int efe_krylov_solve(efe_ctx* c, const double* b, double* x, const efe_solve_opts* o, efe_solve_info* info);
int efe_launch_spmv64(efe_ctx* c, const double* x, double* y);
int efe_launch_norms(efe_ctx* c, const double* R, const double* ddU, const double* F, double* out);
int efe_launch_norms_batched(efe_ctx* c, const double* R, const double* ddU, const double* F, double* out);
void efe_krylov_free(efe_ctx* c);

// batched specimens (batched.hip)
int efe_batched_linsolve(efe_ctx* c, const double* b, double* x, const efe_solve_opts* o, const double* spec_tol,
                         const double* spec_rtol, const uint8_t* spec_active, int32_t* spec_iters,
                         efe_solve_info* info);
int efe_launch_accept_batched(efe_ctx* c, const uint8_t* accept);
int efe_newton_round_batched_impl(efe_ctx* c, double* dU, double* P, double* F, const double* PExt,
                                  const double* deltaT, const uint8_t* live, const uint8_t* extrap,
                                  const uint8_t* hasDir, const int32_t* maxIter, const int32_t* maxGrow,
                                  const double* nAcc, const efe_newton_opts* o, int32_t* state, int32_t* iterations,
                                  int32_t* reason, int32_t* work);
void efe_batched_free(efe_ctx* c);

// scalar CSR Krylov solves (csrsolve.hip)
constexpr int kCsrPartials = 1024;  // workgroup partials per dot product
size_t efe_csr_workspace_doubles(int32_t n, int32_t solver, int32_t restart);
int efe_csr_check_pattern(int32_t n, const int32_t* indptr, const int32_t* indices, int* flag, hipStream_t st,
                          int* bad);
int efe_csr_solve_ws(int32_t n, const int32_t* ip, const int32_t* ix, const double* v, const double* b, double* xout,
                     const efe_solve_opts* o, efe_solve_info* info, double* ws, hipStream_t st);

// GMRES (gmres.hip)
int efe_gmres_solve(efe_ctx* c, const double* b, double* x, const efe_solve_opts* o, efe_solve_info* info);
void efe_gmres_free(efe_ctx* c);
