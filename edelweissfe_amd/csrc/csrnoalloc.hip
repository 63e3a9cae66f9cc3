// efe_csr_linsolve_device_noalloc: the CSR Krylov solve (csrsolve.hip) with its workspace in a
// static device array instead of an allocation, for callers that cannot allocate device memory.
#include <mutex>

#include "efe_internal.h"

// ---- the CSR solve without allocation -------------------------------------------------------
// Scratch of efe_csr_linsolve_device_noalloc: the CG workspace of a system of up to
// kNoAllocMaxDofs unknowns (BiCGStab: up to the size whose workspace fits), plus one int of
// pattern-check flag.  One copy per GPU, held from the library's load until the process ends.
namespace {
constexpr int32_t kNoAllocMaxDofs = 1 << 20;
constexpr size_t kScratchDoubles = 6 * (size_t)kNoAllocMaxDofs + kNoAllocMaxDofs / 8 + 4 * kCsrPartials + 16 + 8;
}  // namespace

static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per device with that
// device current (each GPU holds its own copy at its own address).
int scratch_address(double** out) {
    int dev = 0;
    EFE_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        efe_set_error("efe_csr_linsolve_device_noalloc: device index beyond the supported 64");
        return EFE_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        EFE_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = (double*)p;
    }
    *out = g_scratch_ptr[dev];
    return EFE_OK;
}
}  // namespace

extern "C" int efe_csr_linsolve_device_noalloc(int32_t n, const int32_t* indptr, const int32_t* indices,
                                               const double* data, const double* b, double* x, void* stream,
                                               const efe_solve_opts* o, efe_solve_info* info) {
    if (n <= 0 || !indptr || !indices || !data || !b || !x || !o || !info) {
        efe_set_error("efe_csr_linsolve_device_noalloc: invalid arguments");
        return EFE_ERR_ARG;
    }
    if (o->solver != EFE_SOLVER_CG && o->solver != EFE_SOLVER_BICGSTAB) {
        efe_set_error("efe_csr_linsolve_device_noalloc: solver must be EFE_SOLVER_CG or EFE_SOLVER_BICGSTAB");
        return EFE_ERR_ARG;
    }
    const size_t need = efe_csr_workspace_doubles(n, o->solver, 0);
    if (n > kNoAllocMaxDofs || need + 1 > kScratchDoubles) {
        efe_set_error("efe_csr_linsolve_device_noalloc: the system exceeds the static scratch (CG up to 1048576 "
                      "unknowns, BiCGStab up to about 700000); use efe_csr_linsolve_device");
        return EFE_ERR_ARG;
    }
    double* ws = nullptr;
    if (int rc = scratch_address(&ws)) return rc;
    hipStream_t st = (hipStream_t)stream;
    int* flag = (int*)(ws + need);  // after the workspace; set before it is read
    int bad = 0;
    if (int rc = efe_csr_check_pattern(n, indptr, indices, flag, st, &bad)) return rc;
    if (bad) {
        efe_set_error("efe_csr_linsolve_device_noalloc: invalid CSR pattern (indptr or column indices)");
        return EFE_ERR_ARG;
    }
    return efe_csr_solve_ws(n, indptr, indices, data, b, x, o, info, ws, st);
}
