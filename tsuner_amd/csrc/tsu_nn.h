// tsu_nn.h -- launchers of the correction-network kernels (internal).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tsu_kernels.h"

namespace tsu {

enum NnKind {
    NN_FWD_IN = 0,      // 8 -> 32 channels, bias + ReLU
    NN_FWD_MID = 1,     // 32 -> 32, bias + ReLU
    NN_FWD_OUT = 2,     // 32 -> 4, bias, FP32 output
    NN_DGRAD_OUT = 3,   // gradient through the output layer: dy 8 (4 used) -> 32, ReLU mask
    NN_DGRAD_MID = 4,   // 32 -> 32, ReLU mask
    NN_WGRAD_IN = 5,    // weight gradient of the input layer
    NN_WGRAD_MID = 6,
    NN_WGRAD_OUT = 7,
    NN_W64 = 16,        // flag: hidden width 64 (8 -> 64 -> ... -> 64 -> 4)
};

// Waves of the weight-gradient contraction (partial sums per wave).
constexpr int kWgradWaves = 256;

// Fine cells deeper than this count as wet when a 4x4 block is restricted
// onto a coarse cell [m].
constexpr double kRestrictWetH = 1e-3;

// Feature and correction scales (normalisation of the network's inputs and
// outputs).  Input channels: h, qx, qy, z = -H, dz/dx, dz/dy, eta (wet
// cells, else 0), wet flag.  Outputs: corrections of h, qx, qy, H.
struct NnScales {
    float in[8];
    float out[4];
    float wet_h;        // wet threshold of the eta and flag channels [m]
    float shore_h;      // least depth a correction may give a dry shoreline cell [m]
    float pad[2];
};

hipError_t nn_launch_conv(int kind, const uint16_t* in, const uint16_t* w, const float* bias, const uint16_t* act,
                          void* out, int N, int H, int W, hipStream_t s);
hipError_t nn_launch_wgrad(int kind, const uint16_t* x, const uint16_t* dy, float* part, float* bpart, float* dw,
                           float* db, int mrows, int N, int H, int W, hipStream_t s);
hipError_t nn_launch_mse(const float* out, const float* tgt, uint16_t* dy, double* loss_part, int nblocks,
                         long long npix, float scale, hipStream_t s);
hipError_t nn_launch_adam(float* p, const float* g, float* m, float* v, long long n, float lr_t, float b1, float b2,
                          float eps, hipStream_t s);
hipError_t nn_launch_pack(const float* w, uint16_t* wb, uint16_t* wt, int m, int k, int mt, hipStream_t s);

// Ensemble-side kernels.  args: the members' last-stage arguments (their
// .out is the current state); all members share npx x npy.
hipError_t nn_launch_features(const StageArgs* args, int n, int npx, int npy, int precision, const NnScales& sc,
                              uint16_t* x, hipStream_t s);
// Adds out * scale to members whose step count equals expect_step (wet
// cells; with x, the network inputs, also dry shoreline cells); if resid is
// not null, resid[i] = max(resid[i], largest |correction of h - H|).  The
// corrected members' dt for the next step is recomputed from the corrected
// state.
hipError_t nn_launch_apply(const StageArgs* args, int n, int npx, int npy, int precision, const NnScales& sc,
                           const float* out, long long expect_step, float* resid, const uint16_t* x,
                           hipStream_t s);
// Coarse state = 4x4 block means of the fine state (fine members are 4x the
// coarse size in both directions); also resets the coarse dt slot.
hipError_t nn_launch_restrict(const StageArgs* fine, const StageArgs* coarse, int n, int npx, int npy,
                              int precision, hipStream_t s);
// target = (restrict(fine) - coarse) / scale, FP32 [n][npy][npx][4].
hipError_t nn_launch_target(const StageArgs* fine, const StageArgs* coarse, int n, int npx, int npy, int precision,
                            const NnScales& sc, float* target, hipStream_t s);

}  // namespace tsu
