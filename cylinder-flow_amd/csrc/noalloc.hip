// noalloc.hip -- lbm_get_fields_noalloc: lbm_get_fields for callers that cannot allocate.
//
// lbm_get_fields stages each field in a device buffer it allocates per call (hipMalloc/hipFree
// of N doubles) and, in reference mode, in a host vector.  This variant stages through a
// fixed device array of this file, `scratch` (NOALLOC_MAX_NODES doubles, 32 MiB, held by every
// GPU the library runs on until the process ends), and copies each field straight into the
// caller's arrays, so it calls no allocating routine.  Its address on each device is looked up
// once, at set-up (lbm_create), with that device current; host code uses only that address.
// Every call on one GPU shares the array: calls that overlap on one GPU (two host threads, two
// contexts) race and must be serialised by the caller.  Nothing in the array has to start at
// zero: each field is written in full before it is copied out.
#include <mutex>

#include "internal.hpp"

using namespace lbm;

namespace {

constexpr size_t NOALLOC_MAX_NODES = (size_t)1 << 22;   // 4 194 304 doubles = 32 MiB
static __device__ __align__(256) double scratch[NOALLOC_MAX_NODES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
double *g_ptr[MAX_DEVICES] = {};

// interior (ghost layers stripped) of one reference-mode moment plane, x fastest
__global__ void __launch_bounds__(256) k_ref_strip(const double *__restrict__ mac, RefGeo r, double *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
    if (i >= r.glx) return;
    const int gz = r.D == 3 ? r.g : 0;
    out[(size_t)i + (size_t)r.glx * ((size_t)j + (size_t)r.gly * k)] =
        mac[(size_t)(i + r.g) + (size_t)r.px * ((size_t)(j + r.g) + (size_t)r.ly * (k + gz))];
}

}  // namespace

// the scratch array's address on the current device, looked up once per device
int noalloc_setup(double **out, std::string &err) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        err = std::string("hipGetDevice: ") + hipGetErrorString(e);
        return LBM_ERR_HIP;
    }
    if (dev < 0 || dev >= MAX_DEVICES) {
        err = "device index beyond the no-alloc table";
        return LBM_ERR_INVALID;
    }
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return LBM_ERR_HIP;
        }
        g_ptr[dev] = (double *)p;
    }
    *out = g_ptr[dev];
    return LBM_OK;
}

extern "C" int lbm_get_fields_noalloc(lbm_ctx *c, double *p, double *u, double *v, double *w) {
    if (!c) return LBM_ERR_INVALID;
    if (!c->initialized) {
        c->err = "lbm_get_fields_noalloc before lbm_initialize";
        return LBM_ERR_STATE;
    }
    if (c->N > NOALLOC_MAX_NODES) {
        c->err = "lbm_get_fields_noalloc: " + std::to_string(c->N) + " nodes exceed its staging maximum of " +
                 std::to_string(NOALLOC_MAX_NODES) + " (32 MiB); use lbm_get_fields";
        return LBM_ERR_INVALID;
    }
    if (!c->noalloc_scratch) {
        c->err = "lbm_get_fields_noalloc: no staging array on this device (" + c->noalloc_err + ")";
        return LBM_ERR_HIP;
    }
    if (c->stage == 0) {
        const int rc = ensure_fields(c);
        if (rc != LBM_OK) return rc;
    }
    double *const tmp = c->noalloc_scratch;
    double *outs[4] = {p, u, v, w};
    const bool ref = c->cfg.bc_mode != LBM_BC_PLAIN;
    for (int k = 0; k < 4; k++) {
        if (!outs[k]) continue;
        hipError_t e;
        if (ref) {
            const RefGeo &r = c->rg;
            hipLaunchKernelGGL(k_ref_strip, dim3((unsigned)((r.glx + 255) / 256), (unsigned)r.gly, (unsigned)r.glz),
                               dim3(256), 0, c->stream, c->rmac + (size_t)k * r.plane, r, tmp);
            e = hipGetLastError();
        } else {
            e = macros_extract(c->macros, c->flags, tmp, k, c->geo, c->stream);
        }
        if (e == hipSuccess) e = hipMemcpyAsync(outs[k], tmp, sizeof(double) * c->N, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) {
            c->err = std::string("lbm_get_fields_noalloc: ") + hipGetErrorString(e);
            return LBM_ERR_HIP;
        }
    }
    return LBM_OK;
}
