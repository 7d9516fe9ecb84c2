// mgs_noalloc.hip -- mgs_solve_noalloc: one whole solve that allocates no
// device memory, for callers that cannot call hipMalloc (code under HIP graph
// capture, tight serving loops, harnesses that forbid allocation).
//
// The solve's multigrid hierarchy, norm partials and every other working
// array live in `scratch`, a file-scope device array of MGS_NOALLOC_SCRATCH_BYTES
// that each GPU the library runs on holds until the process ends.  Host code
// never passes the array's name to a kernel (on the host it names a stand-in,
// not the device copy): the first call on a device looks up that device's
// address with hipGetSymbolAddress and keeps it, and every later use goes
// through the kept pointer.
//
// The body is mgs_solve's (same kernels, same launch order, the coarse levels
// launched directly instead of from a captured graph), so the field it
// returns is bitwise the one mgs_create + mgs_set_density + mgs_solve +
// mgs_get_field return for the same inputs.  Every call on a device shares that
// device's array and leaves its contents behind; everything a solve reads is
// written by that solve first.  Calls that overlap on one device race.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>
#include <mutex>
#include <string>

#include "../../include/mgsolve.h"
#include "mgs_internal.h"
#include "mgs_kernels.h"

static __device__ __align__(256) unsigned char scratch[MGS_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
unsigned char* g_scratch_ptr[MAX_DEVICES] = {};

// the current device's address of `scratch` (looked up once per device)
int scratch_of_current_device(unsigned char** out)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return set_err(nullptr, MGS_E_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= MAX_DEVICES) return set_err(nullptr, MGS_E_HIP, "device index out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return set_err(nullptr, MGS_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = (unsigned char*)ptr;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

// the handle's error becomes the thread's creation-time error (mgs_last_error(NULL))
int fail(mgs_handle* h, int rc) { return set_err(nullptr, rc, h->err.empty() ? "mgs_solve_noalloc failed" : h->err); }

}  // namespace

extern "C" int mgs_solve_noalloc(const mgs_params* params, const mgs_solve_opts* opts, const void* delta,
                                 int32_t delta_dtype, void* u, int32_t u_dtype, mgs_solve_report* report)
{
    if (!params || !opts || !delta || !u || !report) return set_err(nullptr, MGS_E_ARG, "NULL argument");
    if (delta_dtype != MGS_F32 && delta_dtype != MGS_F64) return set_err(nullptr, MGS_E_ARG, "delta_dtype");
    if (u_dtype != MGS_F32 && u_dtype != MGS_F64) return set_err(nullptr, MGS_E_ARG, "u_dtype");
    std::string why;
    int rc = mgs::validate_params(params, &why);
    if (rc) return set_err(nullptr, rc, why);
    if (opts->max_cycles < 1) return set_err(nullptr, MGS_E_ARG, "max_cycles must be >= 1");
    if (opts->cycle < 0 || opts->cycle > 2) return set_err(nullptr, MGS_E_ARG, "cycle must be 0 (V), 1 (F) or 2 (W)");
    std::memset(report, 0, sizeof(*report));

    // bytes this hierarchy needs (a counting pass touches no device memory)
    size_t need = 0;
    {
        mgs_handle c;
        rc = mgs::init_handle(&c, params, true, nullptr, 0);
        if (rc) return fail(&c, rc);
        need = c.arena_used;
    }
    if (need > (size_t)MGS_NOALLOC_SCRATCH_BYTES)
        return set_err(nullptr, MGS_E_ARG,
                       "mgs_solve_noalloc: this grid needs " + std::to_string(need) + " bytes of scratch, more than its " +
                           std::to_string((size_t)MGS_NOALLOC_SCRATCH_BYTES) +
                           " (MGS_NOALLOC_SCRATCH_BYTES); use mgs_create/mgs_solve");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return set_err(nullptr, MGS_E_HIP, "no HIP device available: libmgsolve has no CPU fallback");
    unsigned char* base = nullptr;
    if ((rc = scratch_of_current_device(&base))) return rc;

    // a handle whose arrays are pieces of this device's scratch, on the null
    // stream, with no captured graphs, no profiling events and no staging
    mgs_handle h;
    h.s = nullptr;
    h.use_graphs = false;
    rc = mgs::init_handle(&h, params, true, base, (size_t)MGS_NOALLOC_SCRATCH_BYTES);
    if (rc) return fail(&h, rc);
    const int N = params->ngrid;
    if (delta_dtype == MGS_F32)
        mgs::launch_pack_f64_from_f32((const float*)delta, h.L[0].rho, N, h.s);
    else
        mgs::launch_pack_f64_from_f64((const double*)delta, h.L[0].rho, N, h.s);
    if (!opts->init_guess) {
        if (u_dtype == MGS_F32)
            mgs::launch_pack_f64_from_f32((const float*)u, h.L[0].u, N, h.s);
        else
            mgs::launch_pack_f64_from_f64((const double*)u, h.L[0].u, N, h.s);
    }
    h.has_density = true;
    h.rho_cached = false;
    const auto t0 = std::chrono::steady_clock::now();
    rc = mgs::solve_on_handle(&h, *opts, report);
    if (rc) {
        (void)hipStreamSynchronize(h.s);
        return fail(&h, rc);
    }
    if (u_dtype == MGS_F32)
        mgs::launch_unpack_f64_to_f32(h.L[0].u, (float*)u, N, 1.0, h.s);
    else
        mgs::launch_unpack_f64_to_f64(h.L[0].u, (double*)u, N, 1.0, h.s);
    hipError_t e = hipStreamSynchronize(h.s);
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) return set_err(nullptr, MGS_E_HIP, std::string("mgs_solve_noalloc: ") + hipGetErrorString(e));
    report->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return 0;
}
