// C ABI: the solve of the assembled system without any device allocation, for callers that
// cannot allocate (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc).  Its Krylov workspace (basis, residual and scaled-matrix copies, Hessenberg
// state, reduction partials) lives in a file-scope device array of fixed size: each GPU the
// library runs on holds its own copy until the process ends.  The host never uses the array's
// name as an address: the device address is looked up once per device, at context creation.
#include <hip/hip_runtime.h>

#include <array>
#include <mutex>
#include <string>

#include "ctx.hpp"

using namespace vfvm;

namespace {

// 16 Mi doubles (128 MiB) per GPU
constexpr int64_t NOALLOC_ENTRIES = (int64_t)1 << 24;
static __device__ __align__(256) double scratch[NOALLOC_ENTRIES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mutex;
std::array<double*, MAX_DEVICES> g_scratch{};  // device address of `scratch` per device

}  // namespace

// With `device` current: its copy of the scratch array's address.
int vfvm::noalloc_setup(int device) {
    if (device < 0 || device >= MAX_DEVICES) return set_error(VFVM_ERR_INVALID_ARGUMENT, "device index above the supported range");
    std::lock_guard<std::mutex> lock(g_mutex);
    if (g_scratch[device]) return VFVM_OK;
    double* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return set_error(VFVM_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[device] = ptr;
    return VFVM_OK;
}

extern "C" int vfvm_solve_device_noalloc(vfvm_ctx* ctx, const vfvm_solve_opts* opts, double* x_dev, vfvm_solve_stats* stats) {
    CTX_GUARD(ctx);
    if (!opts || !x_dev) return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_solve_device_noalloc: NULL argument");
    if (!c.has_matrix) return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_solve_device_noalloc: nothing assembled");
    double* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_mutex);
        base = g_scratch[c.device];
    }
    if (!base) return set_error(VFVM_ERR_HIP, "vfvm_solve_device_noalloc: the scratch array of this device was not set up");
    vfvm_solve_stats st;
    const int r = gmres_solve_noalloc(c, primal_view(c), *opts, x_dev, st, base, NOALLOC_ENTRIES);
    if (stats) *stats = st;
    if (r != VFVM_OK) return r;
    return clear_error();
}

