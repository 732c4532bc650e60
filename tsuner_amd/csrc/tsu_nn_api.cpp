// tsu_nn_api.cpp -- C ABI of the correction-network kernels (device pointers
// in, work enqueued on the caller's stream).
#define TSU_BUILDING_LIBRARY 1
#include "../../include/tsuner.h"

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "tsu_host.h"
#include "tsu_nn.h"

namespace {

int hip_fail(const char* where, hipError_t e)
{
    return tsu::set_error(TSU_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

bool aligned16(const void* p) { return ((uintptr_t)p & 15) == 0; }

// rows (MY) and columns (KX) of the weight gradient per kind
bool wgrad_shape(int kind, int* my, int* kx)
{
    if (kind & ~(TSU_NN_W64 | 7)) return false;
    const int width = (kind & TSU_NN_W64) ? 64 : 32;
    switch (kind & 7) {
    case TSU_NN_WGRAD_IN: *my = width; *kx = 8; return true;
    case TSU_NN_WGRAD_MID: *my = width; *kx = width; return true;
    case TSU_NN_WGRAD_OUT: *my = 8; *kx = width; return true;
    default: return false;
    }
}

// Workspace of tsu_nn_wgrad_noalloc: the largest weight-gradient workspace
// (hidden width 64, the 64 -> 64 layer: per-wave partials of 9 x 64 x 64
// weights and 64 biases), 9,453,568 floats = 37.8 MB.  Every device the
// library's code object is loaded on holds one copy until the process ends.
constexpr long long kScratchFloats = 9LL * tsu::kWgradWaves * 64 * 64 + (long long)tsu::kWgradWaves * 64;

}  // namespace

static __device__ __align__(256) float scratch[kScratchFloats];

namespace {

// Device address of `scratch` per device, looked up once with that device
// current (the host symbol is not a device address).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch_ptr[kMaxDevices] = {};

int scratch_of_device(int dev, float** out)
{
    if (dev < 0 || dev >= kMaxDevices) return tsu::set_error(TSU_ERR_ARG, "tsu_nn_wgrad_noalloc: device id out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        int prev = -1;
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev && hipSetDevice(dev) != hipSuccess)
            return tsu::set_error(TSU_ERR_HIP, "tsu_nn_wgrad_noalloc: cannot select the stream's device");
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (prev >= 0 && prev != dev) (void)hipSetDevice(prev);
        if (e != hipSuccess || !p)
            return hip_fail("tsu_nn_wgrad_noalloc: hipGetSymbolAddress(scratch)", e != hipSuccess ? e : hipErrorUnknown);
        g_scratch_ptr[dev] = (float*)p;
    }
    *out = g_scratch_ptr[dev];
    return TSU_OK;
}

}  // namespace

extern "C" {

TSU_API int tsu_nn_conv(int kind, const void* in, const void* w, const float* bias, const void* act, void* out, int n,
                        int h, int w_, void* stream)
{
    const int base = kind & ~TSU_NN_W64;
    if (base < TSU_NN_FWD_IN || base > TSU_NN_DGRAD_MID) return tsu::set_error(TSU_ERR_ARG, "tsu_nn_conv: bad kind");
    if (!in || !w || !out || n < 1 || h < 1 || w_ < 1) return tsu::set_error(TSU_ERR_ARG, "tsu_nn_conv: bad argument");
    const bool needs_bias = base <= TSU_NN_FWD_OUT, needs_act = base >= TSU_NN_DGRAD_OUT;
    if ((needs_bias && !bias) || (needs_act && !act)) return tsu::set_error(TSU_ERR_ARG, "tsu_nn_conv: missing operand");
    if (!aligned16(in) || !aligned16(w) || !aligned16(out) || (act && !aligned16(act)))
        return tsu::set_error(TSU_ERR_ARG, "tsu_nn_conv: operands must be 16-byte aligned");
    hipError_t e = tsu::nn_launch_conv(kind, (const uint16_t*)in, (const uint16_t*)w, bias, (const uint16_t*)act, out,
                                       n, h, w_, (hipStream_t)stream);
    return e == hipSuccess ? TSU_OK : hip_fail("tsu_nn_conv", e);
}

TSU_API long long tsu_nn_wgrad_workspace(int kind)
{
    int my, kx;
    if (!wgrad_shape(kind, &my, &kx)) return -1;
    return 9LL * tsu::kWgradWaves * my * kx + (long long)tsu::kWgradWaves * my;
}

TSU_API int tsu_nn_wgrad(int kind, const void* x, const void* dy, float* dw, float* db, int mrows, int n, int h, int w_,
                         float* workspace, long long workspace_floats, void* stream)
{
    int my, kx;
    if (!wgrad_shape(kind, &my, &kx)) return tsu::set_error(TSU_ERR_ARG, "tsu_nn_wgrad: bad kind");
    if (!x || !dy || !dw || !db || !workspace || n < 1 || h < 1 || w_ < 1 || mrows < my)
        return tsu::set_error(TSU_ERR_ARG, "tsu_nn_wgrad: bad argument");
    if (workspace_floats < tsu_nn_wgrad_workspace(kind))
        return tsu::set_error(TSU_ERR_ARG, "tsu_nn_wgrad: workspace too small");
    float* part = workspace;
    float* bpart = workspace + 9LL * tsu::kWgradWaves * my * kx;
    hipError_t e = tsu::nn_launch_wgrad(kind, (const uint16_t*)x, (const uint16_t*)dy, part, bpart, dw, db, mrows, n,
                                        h, w_, (hipStream_t)stream);
    return e == hipSuccess ? TSU_OK : hip_fail("tsu_nn_wgrad", e);
}

TSU_API int tsu_nn_wgrad_noalloc(int kind, const void* x, const void* dy, float* dw, float* db, int mrows, int n,
                                 int h, int w_, void* stream)
{
    int my, kx;
    if (!wgrad_shape(kind, &my, &kx)) return tsu::set_error(TSU_ERR_ARG, "tsu_nn_wgrad_noalloc: bad kind");
    if (!x || !dy || !dw || !db || n < 1 || h < 1 || w_ < 1 || mrows < my)
        return tsu::set_error(TSU_ERR_ARG, "tsu_nn_wgrad_noalloc: bad argument");
    const long long need = tsu_nn_wgrad_workspace(kind);
    if (need > kScratchFloats)
        return tsu::set_error(TSU_ERR_ARG, "tsu_nn_wgrad_noalloc: this kind needs " + std::to_string(need) +
                                               " workspace floats, more than the " + std::to_string(kScratchFloats) +
                                               "-float device scratch");
    int dev = -1;
    const hipError_t ed = stream ? hipStreamGetDevice((hipStream_t)stream, &dev) : hipGetDevice(&dev);
    if (ed != hipSuccess) return hip_fail("tsu_nn_wgrad_noalloc: device of the stream", ed);
    float* ws = nullptr;
    int rc = scratch_of_device(dev, &ws);
    if (rc) return rc;
    // every wave writes all of its partial sums each call: nothing in the
    // scratch has to start at zero
    float* part = ws;
    float* bpart = ws + 9LL * tsu::kWgradWaves * my * kx;
    hipError_t e = tsu::nn_launch_wgrad(kind, (const uint16_t*)x, (const uint16_t*)dy, part, bpart, dw, db, mrows, n,
                                        h, w_, (hipStream_t)stream);
    return e == hipSuccess ? TSU_OK : hip_fail("tsu_nn_wgrad_noalloc", e);
}

TSU_API int tsu_nn_mse(const float* out, const float* target, void* dy, double* loss_part, int nblocks, long long npix,
                       float scale, void* stream)
{
    if (!out || !target || !dy || !loss_part || nblocks < 1 || npix < 1)
        return tsu::set_error(TSU_ERR_ARG, "tsu_nn_mse: bad argument");
    if (!aligned16(out) || !aligned16(target) || !aligned16(dy))
        return tsu::set_error(TSU_ERR_ARG, "tsu_nn_mse: operands must be 16-byte aligned");
    hipError_t e = tsu::nn_launch_mse(out, target, (uint16_t*)dy, loss_part, nblocks, npix, scale, (hipStream_t)stream);
    return e == hipSuccess ? TSU_OK : hip_fail("tsu_nn_mse", e);
}

TSU_API int tsu_nn_adam(float* p, const float* g, float* m, float* v, long long n, float lr_t, float beta1,
                        float beta2, float eps, void* stream)
{
    if (!p || !g || !m || !v || n < 1) return tsu::set_error(TSU_ERR_ARG, "tsu_nn_adam: bad argument");
    hipError_t e = tsu::nn_launch_adam(p, g, m, v, n, lr_t, beta1, beta2, eps, (hipStream_t)stream);
    return e == hipSuccess ? TSU_OK : hip_fail("tsu_nn_adam", e);
}

TSU_API int tsu_nn_pack(const float* w, void* wb, void* wt, int m, int k, int mt, void* stream)
{
    if (!w || !wb || m < 1 || k < 1 || (wt && (mt < 1 || mt > m)))
        return tsu::set_error(TSU_ERR_ARG, "tsu_nn_pack: bad argument");
    hipError_t e = tsu::nn_launch_pack(w, (uint16_t*)wb, (uint16_t*)wt, m, k, mt, (hipStream_t)stream);
    return e == hipSuccess ? TSU_OK : hip_fail("tsu_nn_pack", e);
}

}  // extern "C"
