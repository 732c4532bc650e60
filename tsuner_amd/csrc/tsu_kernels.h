// tsu_kernels.h -- host-side launchers of the gfx950 kernels (internal).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tsu_device.h"

namespace tsu {

// Columns of output per wave: a 64-lane wave covers 64 consecutive columns,
// the two outer lanes on each side are halo.
constexpr int kStripOut = 60;

struct StageArgs {
    const void* in;       // stage input planes (h, qx, qy, H, hc), stride nvol
    const void* Hin;      // bed plane read by the stencil when sediment is off
    const void* v0;       // U^n planes, read at own cells only (may alias out)
    void* out;            // output planes
    int pingpong;         // 1: on odd steps in and out swap roles (single-stage scheme)
    const uint8_t* fix;   // 1 = erodible bed, 0 = fixed
    Ctrl* ctrl;
    double* ext;          // hmax, max H, eta_max planes (nvol each) or null
    int k;                // RK stage index
    int last;             // 1 on the last stage of a step (dt epilogue)
    int tile_rows;
    int nstrips;
    int ntiles;
    const DevParams* P;   // run constants in device memory
};

// precision: 0 = FP64, 1 = FP32 planes and arithmetic (compensated h, H, hc).  fast_bed:
// the run constants satisfy fast_bedload (tsu_device.h); with sediment the
// launch then takes the kernel that evaluates bedload among the HLL
// arithmetic.
hipError_t launch_stage(const StageArgs& a, bool sed, int precision, hipStream_t s, bool fast_bed = false);

// Ensemble member as seen by the batched stage kernel: byte offsets into
// the ensemble's pools instead of pointers, so that every address the kernel
// forms derives from a kernel-argument base pointer (global address space:
// global_load with scalar bases, not flat loads with 64-bit VGPR addresses).
struct BatchMember {
    long long off;        // state offset in pools A and B (bytes)
    long long plane;      // bytes per plane (nvol * element size)
    long long foff;       // mask offset (bytes)
    long long xoff;       // extrema offset (doubles), -1: none
    int tile_rows, nstrips, ntiles, order;
};
struct BatchPools {
    char* A;              // U^n (stage 2 writes it in place)
    char* B;              // stage-1 result
    const uint8_t* fix;
    Ctrl* ctrl;           // one per member
    double* ext;
    const DevParams* P;   // one per member
};
// Ensemble: one launch over all tiles of all members (tiles[i] = member, tile).
// fast_bed: every member's run constants satisfy fast_bedload.
hipError_t launch_stage_batch(const BatchPools& pools, const BatchMember* mem, const int2* tiles, int ntiles, int k,
                              int last, bool sed, int precision, hipStream_t s, bool fast_bed = false);
hipError_t launch_advance_batch(Ctrl* ctrl, const DevParams* P, int n, hipStream_t s);
hipError_t launch_init_dt(const void* V, const DevParams& P, Ctrl* ctrl, int precision, hipStream_t s);
hipError_t launch_warmup(hipStream_t s);
hipError_t launch_copy(const void* src, void* dst, long long bytes, hipStream_t s);
hipError_t launch_advance(Ctrl* ctrl, const DevParams& P, hipStream_t s);
hipError_t launch_fill(double* p, long long n, double v, hipStream_t s);
// Sums of h and of (H + hc) in FP64 into out[0], out[1] (device pointer),
// using partial (>= 2 * nblocks doubles) as scratch.
hipError_t launch_mass(const void* V, long long nvol, int precision, double* partial, int nblocks, double* out,
                       hipStream_t s);
int mass_blocks(long long nvol);

// Known-answer kernels running the device functions on arrays of inputs.
hipError_t launch_kat_edge(const DevParams& P, int n, const double* left, const double* right, const int* kind,
                           const int* fixed, double* Dm, double* Dp, hipStream_t s);
// Edges 2k and 2k + 1 of the same arrays solved as a pair (the stage
// kernels' edge_flux_pair), npairs pairs.
hipError_t launch_kat_edge_pairs(const DevParams& P, int npairs, const double* left, const double* right,
                                 const int* kind, const int* fixed, double* Dm, double* Dp, hipStream_t s);
hipError_t launch_kat_point(int what, int law, int n, const double* a, const double* b, const double* c,
                            const double* d, double* o1, double* o2, double* o3, double* o4, const DevParams& P,
                            hipStream_t s);

}  // namespace tsu
