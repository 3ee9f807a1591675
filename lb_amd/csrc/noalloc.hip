// noalloc.hip -- lbm_step_diagnostics_noalloc: the diagnostics step for
// callers that cannot allocate device memory (code under stream capture,
// tight stepping loops, harnesses that forbid hipMalloc).
//
// lbm_step_diagnostics reduces mass, momentum and the wall force of one step
// into per-tile FP64 records held by the context and sums them on the host.
// This variant writes the records into a file-scope device array of fixed
// size, sums them on the device in a fixed order, and copies the 7 results
// to the caller.  It allocates nothing: no hipMalloc, no hipFree, no
// allocating library call.
//
// The array exists once per GPU that the library runs on and lives until the
// process ends.  Host code never passes the array's name to a kernel or a
// copy: its device address is looked up once per device with
// hipGetSymbolAddress (the name is only a host-side stand-in), and every use
// goes through that address.  Every call on one GPU shares the array; calls
// that overlap on one GPU (two host threads, or two contexts whose streams
// run concurrently) race on it.  Each call rewrites every record it reads,
// so nothing needs clearing between calls.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>

#include "../../include/lbm_core.h"

lbm_status lbm_internal_diag_info(lbm_ctx* c, int* ntiles, int* device, hipStream_t* stream);
lbm_status lbm_internal_diag_step(lbm_ctx* c, double* dst);
lbm_status lbm_internal_reduce(const double* parts, int nb, int stride, int ncomp, double* out,
                               hipStream_t stream);
void lbm_internal_mark_nan(lbm_ctx* c);

// doubles per tile record (7 used)
#define NOALLOC_REC 8

// per-tile records plus the 8 reduced values: 16 MiB + 64 B per GPU
static __device__ __align__(256) double scratch[LBM_NOALLOC_MAX_TILES * NOALLOC_REC + NOALLOC_REC];

namespace {

constexpr int REC = NOALLOC_REC;
constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
double* g_scratch[MAX_DEVICES] = {};   // device address of `scratch`, per device

// Device address of the array on `device`, looked up the first time with that
// device current.
lbm_status scratch_on(int device, double** out)
{
    if (device < 0 || device >= MAX_DEVICES) return LBM_E_UNSUPPORTED;
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[device]) {
        int cur = -1;
        if (hipGetDevice(&cur) != hipSuccess) return LBM_E_HIP;
        if (cur != device && hipSetDevice(device) != hipSuccess) return LBM_E_HIP;
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return LBM_E_HIP;
        }
        g_scratch[device] = (double*)p;
    }
    *out = g_scratch[device];
    return LBM_OK;
}

}  // namespace

extern "C" lbm_status lbm_step_diagnostics_noalloc(lbm_ctx* c, double out[7])
{
    if (!c || !out) return LBM_E_ARG;
    int ntiles = 0, device = 0;
    hipStream_t stream = nullptr;
    lbm_status st = lbm_internal_diag_info(c, &ntiles, &device, &stream);
    if (st != LBM_OK) return st;
    if (ntiles > LBM_NOALLOC_MAX_TILES) return LBM_E_UNSUPPORTED;   // before any step is taken
    double* buf = nullptr;
    st = scratch_on(device, &buf);
    if (st != LBM_OK) return st;
    double* res = buf + (size_t)LBM_NOALLOC_MAX_TILES * REC;
    st = lbm_internal_diag_step(c, buf);
    if (st != LBM_OK) return st;
    st = lbm_internal_reduce(buf, ntiles, REC, 7, res, stream);
    if (st != LBM_OK) return st;
    double s[7];
    if (hipMemcpyAsync(s, res, sizeof(s), hipMemcpyDeviceToHost, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess) {
        (void)hipGetLastError();
        return LBM_E_HIP;
    }
    st = lbm_sync(c);   // resolves the step's timing; reports an earlier NaN
    bool finite = true;
    for (int k = 0; k < 7; ++k) {
        out[k] = s[k];
        finite = finite && std::isfinite(s[k]);
    }
    if (!finite) {
        lbm_internal_mark_nan(c);   // as lbm_step_diagnostics: lbm_sync reports it from now on
        return LBM_E_NAN;
    }
    return st;
}
