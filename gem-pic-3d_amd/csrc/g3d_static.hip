// g3d_stream_spectra_static (include/g3d.h): g3d_stream_spectra with its workspace (6 nk complex values: the x
// transforms of the half-cell current and the column buffer) in a file-scope device array instead of a caller's
// buffer, for callers that cannot allocate device memory (code under HIP graph capture, serving loops).  It calls no
// allocating routine.
//
// In host code the name `scratch` is a host-side stand-in, not a device address: the device address of each GPU's
// copy comes from hipGetSymbolAddress, taken on the entry point's first call on that device with the device current,
// and only that address reaches the kernels.  Every call on a GPU shares its copy: calls that overlap on one GPU race
// (the kernels write the whole workspace before reading it, so nothing needs clearing between calls).
#include <mutex>
#include <stdexcept>
#include <string>
#include "../../include/g3d.h"
#include "g3d_runtime.hpp"

namespace g3d {

constexpr long kStaticModes = 65536;                     // the workspace serves nk up to 256 x 256 modes
__device__ __align__(256) g3d_cpx scratch[6 * kStaticModes];   // 6 MiB per device, held until the process ends

namespace {
constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
g3d_cpx* scratch_dev[kMaxDevices] = {};

// the device address of this device's copy of `scratch` (looked up once per device)
g3d_cpx* static_scratch() {
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("g3d_stream_spectra_static: device index out of range");
    std::lock_guard<std::mutex> lk(scratch_mu);
    if (!scratch_dev[dev]) {
        void* p = nullptr;
        hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress");
        if (!p) throw std::runtime_error("g3d_stream_spectra_static: no device address for the workspace");
        scratch_dev[dev] = static_cast<g3d_cpx*>(p);
    }
    return scratch_dev[dev];
}
}  // namespace
}  // namespace g3d

extern "C" int g3d_stream_spectra_static(const g3d_spec_ctx* c, double* jsr, g3d_cpx* jsn) {
    int rc = 0;
    const int g = g3d::guard([&] {
        if (!c) throw std::invalid_argument("g3d: missing spectral context");
        if (c->nk > g3d::kStaticModes || c->nk < 1)
            throw std::invalid_argument("g3d_stream_spectra_static: nk = " + std::to_string(c->nk) +
                                        " is outside 1 ... " + std::to_string(g3d::kStaticModes) +
                                        " (the static workspace holds 6 x " + std::to_string(g3d::kStaticModes) +
                                        " complex values); use g3d_stream_spectra with a workspace");
        rc = g3d_stream_spectra(c, jsr, jsn, g3d::static_scratch());
    });
    return g ? g : rc;
}
