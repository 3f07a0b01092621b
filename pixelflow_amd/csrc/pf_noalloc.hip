// pf_noalloc.hip -- pf_force_noalloc: the body-force reduction of pf_force without any
// device allocation, for callers that cannot allocate (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc).
//
// pf_force allocates its reduction buffer on its first call.  This variant keeps the
// per-block partial sums and the six totals in a file-scope device array instead: one copy
// per GPU, held until the process ends, of a fixed size (PF_NOALLOC_FORCE_BLOCKS partial
// blocks; the force reduction never uses more).  Host code never passes the array's name to
// a kernel -- in host code the name is a host-side stand-in, not a device address -- so the
// address of each GPU's copy is looked up once per device, with that device current, by
// hipGetSymbolAddress, and that address is what the kernels receive.  Every call on one GPU
// shares its copy, which keeps the last call's contents; every entry the reduction reads is
// written earlier in the same call, so nothing needs clearing.  Calls that overlap on one
// GPU (two host threads, or two contexts on different streams) race on the array.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/pixelflow.h"
#include "pf_kernels.h"

namespace pf {
int force_into(pf_ctx *c, double *work, double *out);
bool ctx_ready(const pf_ctx *c);
int ctx_device(const pf_ctx *c);
const PfGeom &ctx_geom(const pf_ctx *c);
int ctx_fail(pf_ctx *c, int code, const std::string &m);
}  // namespace pf

#define PF_NOALLOC_FORCE_BLOCKS 1024
#define PF_NOALLOC_SCRATCH_DOUBLES (PF_NOALLOC_FORCE_BLOCKS * 6 + 6)
#define PF_NOALLOC_MAX_DEVICES 64

static __device__ __align__(256) double scratch[PF_NOALLOC_SCRATCH_DOUBLES];

static std::mutex g_scratch_mu;
static double *g_scratch_ptr[PF_NOALLOC_MAX_DEVICES] = {};

// Device address of `scratch` on device `dev`, looked up with dev current on the first call
// for that device and kept.  Returns nullptr (and the HIP error in *e) on failure.
static double *scratch_on_device(int dev, hipError_t *e) {
    *e = hipSuccess;
    if (dev < 0 || dev >= PF_NOALLOC_MAX_DEVICES) {
        *e = hipErrorInvalidDevice;
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[dev]) return g_scratch_ptr[dev];
    int cur = 0;
    if ((*e = hipGetDevice(&cur)) != hipSuccess) return nullptr;
    if (cur != dev && (*e = hipSetDevice(dev)) != hipSuccess) return nullptr;
    void *p = nullptr;
    *e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    if (*e != hipSuccess) return nullptr;
    g_scratch_ptr[dev] = (double *)p;
    return g_scratch_ptr[dev];
}

extern "C" int pf_force_noalloc(pf_ctx *c, double *out) {
    if (!c || !out) return PF_EINVAL;
    if (!pf::ctx_ready(c)) return pf::ctx_fail(c, PF_ESTATE, "pf_force_noalloc before pf_initial");
    const int nb = pf::force_blocks(pf::ctx_geom(c));
    if ((int64_t)nb * 6 + 6 > PF_NOALLOC_SCRATCH_DOUBLES)
        return pf::ctx_fail(c, PF_EINVAL,
                            "pf_force_noalloc: the reduction needs " + std::to_string(nb) +
                                " partial blocks, more than its fixed scratch holds (" +
                                std::to_string(PF_NOALLOC_FORCE_BLOCKS) + ")");
    hipError_t e;
    double *sp = scratch_on_device(pf::ctx_device(c), &e);
    if (!sp) return pf::ctx_fail(c, PF_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    return pf::force_into(c, sp, out);
}
