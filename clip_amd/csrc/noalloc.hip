// noalloc.hip — clip_sum_noalloc: the allocation-free variant of clip_sum
// (Reporter::sum, src/utils/Reporter.cu:95-114) for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving
// loops).  The caller passes the field and the output in device memory; the
// reduction partials and the non-finite counter live in `scratch`, a
// file-scope device array of fixed size that each GPU holds until the
// process ends.  No hipMalloc, hipFree, memset or host copy is issued: a call
// is four kernel launches on the caller's stream.
#include "clip_amd.h"
#include "common.hpp"
#include "kernels.hpp"

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

namespace {

// Reduction partials (one per block of the first launch; the launch strides
// over any grid), then one slot for the non-finite counter.
constexpr int kParts = 1024;
constexpr int kMaxDevices = 64;
constexpr unsigned kMaxExtent = 1u << 20;  // per axis, ghost ring included

static __device__ __align__(256) double scratch[kParts + 1];

// Device address of `scratch` on each GPU (each holds its own copy), taken
// once per device with that device current; host code uses only these.
double* g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mu;

__global__ void k_noalloc_clear(unsigned* bad) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *bad = 0u;
}

__global__ void k_noalloc_count(const unsigned* bad, unsigned* out) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *out = *bad;
}

int scratch_of_current_device(double** out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        (void)hipGetLastError();
        return clipamd::set_error(CLIP_ERR_HIP, "clip_sum_noalloc: no current HIP device (or device id >= 64)");
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            (void)hipGetLastError();
            return clipamd::set_error(CLIP_ERR_HIP, std::string("clip_sum_noalloc: hipGetSymbolAddress: ") +
                                                        hipGetErrorString(e));
        }
        g_scratch[dev] = (double*)ptr;
    }
    *out = g_scratch[dev];
    return CLIP_OK;
}

}  // namespace

extern "C" int clip_sum_noalloc(const void* field, int precision, int dim, const unsigned ext[3], double* out,
                                unsigned* nonfinite, void* hip_stream) {
    using namespace clipamd;
    if (!field || !out || !ext) return set_error(CLIP_ERR_ARG, "clip_sum_noalloc: null argument");
    if (dim != 2 && dim != 3) return set_error(CLIP_ERR_ARG, "clip_sum_noalloc: dim must be 2 or 3");
    if (precision < CLIP_PREC_F64 || precision > CLIP_PREC_MIXED_BF16)
        return set_error(CLIP_ERR_ARG, "clip_sum_noalloc: unknown precision");
    for (int a = 0; a < 3; ++a) {
        const bool used = a < dim;
        if (used && (ext[a] < 3 || ext[a] > kMaxExtent))
            return set_error(CLIP_ERR_ARG, "clip_sum_noalloc: extent " + std::to_string(ext[a]) + " of axis " +
                                               std::to_string(a) + " outside [3, 2^20]");
        if (!used && ext[a] != 1) return set_error(CLIP_ERR_ARG, "clip_sum_noalloc: a 2D grid has ext[2] = 1");
    }
    double* sc = nullptr;
    const int rc = scratch_of_current_device(&sc);
    if (rc != CLIP_OK) return rc;
    DevGrid g{};
    for (int a = 0; a < 3; ++a) {
        g.ext[a] = (int)ext[a];
        g.dmin[a] = a < dim ? 1 : 0;
        g.dmax[a] = a < dim ? (int)ext[a] - 2 : 0;
    }
    g.sy = g.ext[0];
    g.sz = (long long)g.ext[0] * g.ext[1];
    g.nodes = g.sz * g.ext[2];
    hipStream_t st = (hipStream_t)hip_stream;
    double* part = sc;
    unsigned* bad = (unsigned*)(sc + kParts);
    k_noalloc_clear<<<1, 64, 0, st>>>(bad);
    if (precision == CLIP_PREC_F64)
        k_reduce_interior<double><<<kParts, 256, 0, st>>>((const double*)field, g, part, bad);
    else
        k_reduce_interior<float><<<kParts, 256, 0, st>>>((const float*)field, g, part, bad);
    k_reduce_final<double><<<1, 256, 0, st>>>(part, kParts, out);
    if (nonfinite) k_noalloc_count<<<1, 64, 0, st>>>(bad, nonfinite);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(CLIP_ERR_HIP, std::string("clip_sum_noalloc: launch: ") + hipGetErrorString(e));
    return CLIP_OK;
}
