// cf_cg_solve_static: the CG solve of cf_cg_solve on work buffers that the
// library holds in a static device array, for callers that cannot allocate
// device memory. The array (`scratch`) is part of the library's code object:
// each GPU the library runs on holds its own copy from the first HIP call
// until the process ends. Host code never passes the array's name to a kernel
// (in host code it is a stand-in, not a device address): the device address is
// looked up once per device with hipGetSymbolAddress and kept.
#include "cf_internal.h"

#include <mutex>

namespace cf {
namespace {

constexpr size_t align256(size_t b) { return (b + 255) & ~size_t(255); }
constexpr size_t kStaticN = CF_CG_STATIC_MAX_N;
constexpr size_t kOffR = 0;
constexpr size_t kOffD = align256(kOffR + kStaticN * sizeof(float));
constexpr size_t kOffQ = align256(kOffD + kStaticN * sizeof(float));
constexpr size_t kOffSt = align256(kOffQ + kStaticN * sizeof(float));
constexpr size_t kOffPa = align256(kOffSt + sizeof(CgScalars));
constexpr size_t kOffPb = align256(kOffPa + kMaxPartials * sizeof(double));
constexpr size_t kScratchBytes = align256(kOffPb + kMaxPartials * sizeof(double));
static_assert(kScratchBytes == CF_CG_STATIC_BYTES, "the header states the static scratch's size");

// r, d, q, the scalar record and the two partial arrays of one solve
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char* g_ptr[kMaxDevices] = {};

// The current device's address of `scratch`, looked up on its first use.
int scratch_base(unsigned char** out) {
    int dev = 0;
    CF_CHECK_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("cf_cg_solve_static: device %d beyond the %d the library tracks", dev, kMaxDevices);
        return CF_ERR_UNSUPPORTED;
    }
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error("cf_cg_solve_static: hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                      hipGetErrorString(e));
            return CF_ERR_HIP;
        }
        g_ptr[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_ptr[dev];
    return CF_OK;
}

}  // namespace
}  // namespace cf

using namespace cf;

extern "C" int cf_cg_solve_static(void* stream, uint32_t n, cf_av_fn av, cf_constraints_fn apply_constraints,
                                  void* user, float* d_x, const float* d_b, float rel_epsilon, float abs_epsilon,
                                  uint32_t max_iterations, cf_cg_result* result) {
    if (!av || !d_x || !d_b || !result || n == 0 || max_iterations == 0) {
        set_error("cf_cg_solve_static: invalid arguments");
        return CF_ERR_INVALID_ARG;
    }
    if (n > CF_CG_STATIC_MAX_N) {
        set_error("cf_cg_solve_static: n = %u exceeds the static scratch's %u unknowns (use cf_cg_create/cf_cg_solve)",
                  n, (unsigned)CF_CG_STATIC_MAX_N);
        return CF_ERR_CAPACITY;
    }
    unsigned char* base = nullptr;
    CF_TRY(scratch_base(&base));
    const CgWork w{reinterpret_cast<float*>(base + kOffR), reinterpret_cast<float*>(base + kOffD),
                   reinterpret_cast<float*>(base + kOffQ), reinterpret_cast<CgScalars*>(base + kOffSt),
                   reinterpret_cast<double*>(base + kOffPa), reinterpret_cast<double*>(base + kOffPb)};
    int32_t flag = 0;  // host memory of this call for the stop-flag read-back
    return cg_solve_on(w, n, static_cast<hipStream_t>(stream), av, apply_constraints, user, d_x, d_b, rel_epsilon,
                       abs_epsilon, max_iterations, &flag, result);
}
