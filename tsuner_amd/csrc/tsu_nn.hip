// tsu_nn.hip -- BF16 MFMA kernels of the learned coarse-grid correction.
//
// Network: four 3x3 'same' convolutions (8 -> 32 -> 32 -> 32 -> 4 channels),
// ReLU after the first three.  Activations are NHWC BF16, weights BF16 with
// FP32 accumulation on v_mfma_f32_16x16x32_bf16, master weights and
// optimiser state FP32.
//
// Every convolution is computed transposed, D^T[m][p] = sum_k Wt[m][k] X^T[k][p]
// with m = output channel and p = pixel, so that both MFMA operands are
// 16-byte vectors along the channel axis of an NHWC row (weights [tap][m][k],
// activations [pixel][k]) and each lane's four results are four consecutive
// output channels of one pixel (one 8-byte store).  The K axis of an MFMA
// (32 values) is four "octets" of 8 channels; octet o of the flattened
// (tap, channel) axis is tap o / (K/8), channels 8*(o % (K/8)) + 0..7, so an
// 8-channel layer packs four taps into one MFMA and a 32-channel layer one
// tap.  The data-gradient is the same kernel with the tap offsets negated and
// the per-tap transposed weights; the weight gradient contracts over pixels.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tsu_nn.h"

namespace tsu {
namespace nn {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t to_bf16(float f)
{
    // round to nearest, ties to even (finite inputs)
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float from_bf16(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

__device__ __forceinline__ bf16x8 as_bf16x8(uint4 v) { return __builtin_bit_cast(bf16x8, v); }

__device__ __forceinline__ int tap_dy(int t) { return t / 3 - 1; }
__device__ __forceinline__ int tap_dx(int t) { return t % 3 - 1; }

// out[p][m] = post( sum_t sum_k in[p + SIGN*off(t)][k] * Wt[t][m][k] )
//   K: input channels (8 or 32), M: output channels held by the weights
//   (multiple of 16), MOUT: channels stored per output pixel.
//   POST 0: + bias, ReLU, BF16 out.  1: + bias, FP32 out.  2: zero where
//   act[p][m] == 0 (ReLU derivative), BF16 out.
// A workgroup of 4 waves covers 4 consecutive rows x 64 columns of one
// image, one row per wave.  The input rows r0-1 .. r0+4 and columns
// c0-1 .. c0+64 are staged in LDS once (zeros outside the image: the 'same'
// padding), so every tap's B operand is one 16-byte LDS read instead of a
// global load re-fetched by nine taps and three waves.
constexpr int kConvRows = 4;
constexpr int kTileRows = kConvRows + 2, kTileCols = 64 + 2;

// LDS slot of octet q of tile pixel p: the 16 lanes of a ds_read_b128 read
// the same octet of 16 consecutive pixels, which with KO octets per pixel
// are KO*16 bytes apart and would share banks; XOR-ing q with the pixel's
// bits above log2(16/KO) spreads them over all 64 banks.
template <int KO>
__device__ __forceinline__ int swz(int p, int q)
{
    return q ^ ((p / (16 / KO)) & (KO - 1));
}

template <int K, int M, int MOUT, int SIGN, int POST>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void conv_kernel(const uint16_t* __restrict__ in, const uint16_t* __restrict__ Wt,
                                                   const float* __restrict__ bias, const uint16_t* __restrict__ act,
                                                   void* __restrict__ out, int N, int H, int W)
{
    constexpr int KO = K / 8;                  // octets per tap
    constexpr int STEPS = (9 * KO + 3) / 4;    // MFMAs per (m-tile, n-tile)
    constexpr int MT = M / 16;
    constexpr int TILE = kTileRows * kTileCols * KO;
    constexpr int PER_THREAD = (TILE + 255) / 256;
    constexpr int AHEAD = 2;                   // weight operands prefetched this many steps ahead
    __shared__ uint4 tile[TILE];
    const int segs = (W + 63) / 64, rblocks = (H + kConvRows - 1) / kConvRows;
    long long blk = blockIdx.x;
    const int seg = (int)(blk % segs);
    blk /= segs;
    const int r0 = (int)(blk % rblocks) * kConvRows;
    const long long n = blk / rblocks;
    if (n >= N) return;
    const int c0 = seg * 64;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int g = lane >> 4, li = lane & 15;

    // weight operands of step s: A[a] = Wt[t][a*16 + li][8q .. 8q+7] for this
    // lane's octet o = 4s + g (t = o / KO, q = o % KO; zero past the 9 taps)
    auto load_a = [&](int s, bf16x8* A) {
        const int o = 4 * s + g;
        const int t = o / KO, q = o % KO;
#pragma unroll
        for (int a = 0; a < MT; a++) {
            uint4 v = make_uint4(0, 0, 0, 0);
            if (t < 9) v = *(const uint4*)(Wt + ((size_t)(t * M + a * 16 + li) * K + 8 * q));
            A[a] = as_bf16x8(v);
        }
    };
    // the first steps' weights are in flight while the input tile is staged
    bf16x8 Abuf[AHEAD + 1][MT];
#pragma unroll
    for (int s = 0; s < AHEAD && s < STEPS; s++) load_a(s, Abuf[s]);

    // input tile, in batches of up to 7 loads per thread, each batch's loads
    // issued before its LDS stores
    constexpr int BATCH = PER_THREAD < 7 ? PER_THREAD : 7;
#pragma unroll
    for (int j0 = 0; j0 < PER_THREAD; j0 += BATCH) {
        uint4 stage[BATCH];
#pragma unroll
        for (int j = 0; j < BATCH; j++) {
            const int idx = threadIdx.x + 256 * (j0 + j);
            const int q = idx % KO, rest = idx / KO;
            const int gc = c0 - 1 + rest % kTileCols, gr = r0 - 1 + rest / kTileCols;
            stage[j] = make_uint4(0, 0, 0, 0);
            if (j0 + j < PER_THREAD && idx < TILE && gr >= 0 && gr < H && gc >= 0 && gc < W)
                stage[j] = *(const uint4*)(in + (((size_t)n * H + gr) * W + gc) * K + 8 * q);
        }
#pragma unroll
        for (int j = 0; j < BATCH; j++) {
            const int idx = threadIdx.x + 256 * (j0 + j);
            const int q = idx % KO, rest = idx / KO;
            if (j0 + j < PER_THREAD && idx < TILE) tile[rest * KO + swz<KO>(rest, q)] = stage[j];
        }
    }
    __syncthreads();
    const int r = r0 + wv;
    if (r >= H) return;                        // no barrier follows
    const long long row = n * H + r;

    f32x4 acc[MT][4];
#pragma unroll
    for (int a = 0; a < MT; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};

#pragma unroll
    for (int s = 0; s < STEPS; s++) {
        if (s + AHEAD < STEPS) load_a(s + AHEAD, Abuf[(s + AHEAD) % (AHEAD + 1)]);
        const bf16x8* A = Abuf[s % (AHEAD + 1)];
        const int o = 4 * s + g;               // this lane's octet
        const int t = o / KO, q = o % KO;
        const bool tv = t < 9;
        const int dy = SIGN * tap_dy(t), dx = SIGN * tap_dx(t);
        const int lrow = (wv + 1 + dy) * kTileCols + 1 + dx + li;
#pragma unroll
        for (int b = 0; b < 4; b++) {
            uint4 v = make_uint4(0, 0, 0, 0);
            if (tv) v = tile[(lrow + b * 16) * KO + swz<KO>(lrow + b * 16, q)];
            const bf16x8 B = as_bf16x8(v);
#pragma unroll
            for (int a = 0; a < MT; a++)
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[a], B, acc[a][b], 0, 0, 0);
        }
    }

#pragma unroll
    for (int b = 0; b < 4; b++) {
        const int c = c0 + b * 16 + li;
        if (c >= W) continue;
        const size_t p = (size_t)row * W + c;
#pragma unroll
        for (int a = 0; a < MT; a++) {
            const int m0 = a * 16 + 4 * g;
            if (m0 >= MOUT) continue;
            float v[4];
#pragma unroll
            for (int i = 0; i < 4; i++) v[i] = acc[a][b][i];
            if (POST == 0 || POST == 1) {
#pragma unroll
                for (int i = 0; i < 4; i++) v[i] += bias[m0 + i];
            }
            if (POST == 0) {
#pragma unroll
                for (int i = 0; i < 4; i++) v[i] = fmaxf(v[i], 0.f);
            }
            if (POST == 1) {
                *(float4*)((float*)out + p * MOUT + m0) = make_float4(v[0], v[1], v[2], v[3]);
            } else {
                if (POST == 2) {
                    const uint2 av = *(const uint2*)(act + p * MOUT + m0);
                    // ReLU outputs are >= 0: only +0 / -0 bit patterns mean inactive
                    if ((av.x & 0x7FFFu) == 0) v[0] = 0.f;
                    if ((av.x & 0x7FFF0000u) == 0) v[1] = 0.f;
                    if ((av.y & 0x7FFFu) == 0) v[2] = 0.f;
                    if ((av.y & 0x7FFF0000u) == 0) v[3] = 0.f;
                }
                uint2 st;
                st.x = (uint32_t)to_bf16(v[0]) | ((uint32_t)to_bf16(v[1]) << 16);
                st.y = (uint32_t)to_bf16(v[2]) | ((uint32_t)to_bf16(v[3]) << 16);
                *(uint2*)((uint16_t*)out + p * MOUT + m0) = st;
            }
        }
    }
}

// Convolutions of the wide (64-channel) network: 64 -> 64 (middle layers
// and their data gradients), 8 -> 64 (input layer) and 64 -> 4 (output
// layer).  Same arithmetic and operand layout as conv_kernel, arranged so
// that neither memory latency nor address arithmetic bounds it:
//   * the layer's weights are staged in LDS once per workgroup, as
//     [tap][octet][output channel] 16-byte rows, so a step's A operands are
//     LDS reads (16 lanes read 16 consecutive rows);
//   * the input tile is planar by octet, [octet][pixel]: the 16 lanes of a
//     read take 16 consecutive pixels (256 contiguous bytes, no bank
//     conflicts) and a tap's pixel offset is a constant, so with the steps
//     unrolled every LDS address is a per-lane base plus an immediate;
//   * a workgroup of 8 waves covers 8 rows x 64 columns, one row per wave
//     (two waves per SIMD hide each other's LDS latency);
//   * workgroups are persistent (one per CU: for 64 -> 64 the weights and
//     the 10 x 66 pixel input tile fill 155 KB of the 160 KB LDS) and load
//     the next tile's input into registers while computing the current one.
constexpr int kWideRows = 8;
constexpr int kWideTileRows = kWideRows + 2;
constexpr int kWidePix = kWideTileRows * kTileCols;          // 660 pixels per tile

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// two floats -> packed BF16 pair, round to nearest even (v_cvt_pk_bf16_f32)
__device__ __forceinline__ uint32_t pack_bf16x2(float a, float b)
{
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){a, b}, bf16x2));
}

template <int K, int M, int MOUT, int SIGN, int POST>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2, 2))) void conv_wide_kernel(
    const uint16_t* __restrict__ in, const uint16_t* __restrict__ Wt, const float* __restrict__ bias,
    const uint16_t* __restrict__ act, void* __restrict__ out, int N, int H, int W)
{
    constexpr int KO = K / 8, MT = M / 16;
    constexpr int STEPS = (9 * KO + 3) / 4;
    constexpr bool TAP_UNIFORM = KO >= 4;                    // a step's 4 octets lie in one tap
    constexpr int TILE = KO * kWidePix;
    constexpr int WSLOTS = 9 * KO * M;
    constexpr int PER_THREAD = (TILE + 511) / 512;
    constexpr int WPER = (WSLOTS + 511) / 512;
    constexpr int SLOTS = PER_THREAD > WPER ? PER_THREAD : WPER;
    __shared__ uint4 wl[WSLOTS];
    __shared__ uint4 tile[TILE];
    const int segs = (W + 63) / 64, rblocks = (H + kWideRows - 1) / kWideRows;
    const long long nblk = (long long)N * rblocks * segs;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int g = lane >> 4, li = lane & 15;
    if ((long long)blockIdx.x >= nblk) return;              // whole workgroup, before any barrier

    uint4 stage[SLOTS];
    // weights: global [tap][m][k] -> LDS [tap][octet][m]
#pragma unroll
    for (int j = 0; j < WPER; j++) {
        const int idx = threadIdx.x + 512 * j;              // over (tap, m, octet), octet fastest
        if (idx < WSLOTS) stage[j] = *(const uint4*)(Wt + (size_t)idx * 8);
    }
#pragma unroll
    for (int j = 0; j < WPER; j++) {
        const int idx = threadIdx.x + 512 * j;
        const int q = idx % KO, m = (idx / KO) % M, t = idx / (KO * M);
        if (idx < WSLOTS) wl[(t * KO + q) * M + m] = stage[j];
    }

    auto origin = [&](long long b, long long& n, int& r0, int& c0) {
        c0 = (int)(b % segs) * 64;
        b /= segs;
        r0 = (int)(b % rblocks) * kWideRows;
        n = b / rblocks;
    };
    auto fetch = [&](long long b) {
        long long n;
        int r0, c0;
        origin(b, n, r0, c0);
#pragma unroll
        for (int j = 0; j < PER_THREAD; j++) {
            const int idx = threadIdx.x + 512 * j;          // over (pixel, octet), octet fastest
            const int q = idx % KO, pix = idx / KO;
            const int gc = c0 - 1 + pix % kTileCols, gr = r0 - 1 + pix / kTileCols;
            stage[j] = make_uint4(0, 0, 0, 0);
            if (idx < TILE && gr >= 0 && gr < H && gc >= 0 && gc < W)
                stage[j] = *(const uint4*)(in + (((size_t)n * H + gr) * W + gc) * K + 8 * q);
        }
    };
    auto put = [&]() {
#pragma unroll
        for (int j = 0; j < PER_THREAD; j++) {
            const int idx = threadIdx.x + 512 * j;
            const int q = idx % KO, pix = idx / KO;
            if (idx < TILE) tile[q * kWidePix + pix] = stage[j];
        }
    };

    long long blk = blockIdx.x;
    fetch(blk);
    put();
    __syncthreads();
    // per-lane bases: this lane's pixel (row wv + 1, column 1 + li) and octet g
    const uint4* tbase = tile + g * kWidePix + (wv + 1) * kTileCols + 1 + li;
    const uint4* wbase = wl + g * M + li;
    for (;;) {
        const long long next = blk + gridDim.x;
        const bool more = next < nblk;                       // uniform over the workgroup
        if (more) fetch(next);                               // in flight during the MFMAs
        long long n;
        int r0, c0;
        origin(blk, n, r0, c0);
        f32x4 acc[MT][4];
#pragma unroll
        for (int a = 0; a < MT; a++)
#pragma unroll
            for (int b = 0; b < 4; b++) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < STEPS; s++) {
            // octet o = 4s + g: tap t = o / KO, octet q = o % KO; with
            // TAP_UNIFORM, t and q - g are constants of the step
            bf16x8 A[MT];
            const int t = TAP_UNIFORM ? (4 * s) / KO : (4 * s + g) / KO;
            const int q = TAP_UNIFORM ? (4 * s) % KO + g : (4 * s + g) % KO;
            const bool tv = t < 9;
            const int d = SIGN * (tap_dy(t) * kTileCols + tap_dx(t));   // tap's pixel offset
#pragma unroll
            for (int a = 0; a < MT; a++) {
                uint4 v = make_uint4(0, 0, 0, 0);
                if (TAP_UNIFORM)
                    v = wbase[((4 * s) / KO * KO + (4 * s) % KO) * M + a * 16];
                else if (tv)
                    v = wl[(t * KO + q) * M + a * 16 + li];
                A[a] = as_bf16x8(v);
            }
#pragma unroll
            for (int b = 0; b < 4; b++) {
                uint4 v = make_uint4(0, 0, 0, 0);
                if (TAP_UNIFORM)
                    v = tbase[((4 * s) % KO) * kWidePix + d + b * 16];
                else if (tv)
                    v = tile[q * kWidePix + (wv + 1) * kTileCols + 1 + li + d + b * 16];
                const bf16x8 B = as_bf16x8(v);
#pragma unroll
                for (int a = 0; a < MT; a++)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[a], B, acc[a][b], 0, 0, 0);
            }
        }
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const int r = r0 + wv, c = c0 + b * 16 + li;
            if (r >= H || c >= W) continue;
            const size_t p = ((size_t)n * H + r) * W + c;
#pragma unroll
            for (int a = 0; a < MT; a++) {
                const int m0 = a * 16 + 4 * g;
                if (m0 >= MOUT) continue;
                float v[4];
#pragma unroll
                for (int i = 0; i < 4; i++) v[i] = acc[a][b][i];
                if (POST == 0 || POST == 1) {
#pragma unroll
                    for (int i = 0; i < 4; i++) v[i] += bias[m0 + i];
                }
                if (POST == 0) {
#pragma unroll
                    for (int i = 0; i < 4; i++) v[i] = fmaxf(v[i], 0.f);
                }
                if (POST == 1) {
                    *(float4*)((float*)out + p * MOUT + m0) = make_float4(v[0], v[1], v[2], v[3]);
                } else {
                    if (POST == 2) {
                        const uint2 av = *(const uint2*)(act + p * MOUT + m0);
                        // ReLU outputs are >= 0: only +0 / -0 bit patterns mean inactive
                        if ((av.x & 0x7FFFu) == 0) v[0] = 0.f;
                        if ((av.x & 0x7FFF0000u) == 0) v[1] = 0.f;
                        if ((av.y & 0x7FFFu) == 0) v[2] = 0.f;
                        if ((av.y & 0x7FFF0000u) == 0) v[3] = 0.f;
                    }
                    *(uint2*)((uint16_t*)out + p * MOUT + m0) = make_uint2(pack_bf16x2(v[0], v[1]),
                                                                           pack_bf16x2(v[2], v[3]));
                }
            }
        }
        __syncthreads();                                     // every wave is done reading the tile
        if (!more) break;
        put();
        __syncthreads();
        blk = next;
    }
}

// Weight gradient of out = conv(x, w): dW[t][m][k] = sum_p dy[p][m] x[p + off(t)][k],
// contracted over pixels on the MFMA K axis (32 pixels per MFMA: four runs of
// 8 consecutive columns of one row).  Grid: (WAVES/4, 9 taps); wave w of tap t
// takes rows w, w + WAVES, ... and writes its partial sums to
// part[t][w][MY][KX]; the centre tap also writes column sums of dy (the bias
// gradient) to bpart[w][MY].  The reduction kernel adds the partials in a
// fixed order, so the result is deterministic.
template <int KX, int MY>
__global__ __launch_bounds__(256) void wgrad_kernel(const uint16_t* __restrict__ x, const uint16_t* __restrict__ dy,
                                                    float* __restrict__ part, float* __restrict__ bpart, int N, int H,
                                                    int W, int waves)
{
    constexpr int MT = (MY + 15) / 16, NT = (KX + 15) / 16;
    const int lane = threadIdx.x & 63;
    const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int t = blockIdx.y;
    if (w >= waves) return;
    const int g = lane >> 4, li = lane & 15;
    const int ty = tap_dy(t), tx = tap_dx(t);
    f32x4 acc[MT][NT];
#pragma unroll
    for (int a = 0; a < MT; a++)
#pragma unroll
        for (int b = 0; b < NT; b++) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    float bsum[MT];
#pragma unroll
    for (int a = 0; a < MT; a++) bsum[a] = 0.f;

    const long long rows = (long long)N * H;
    for (long long row = w; row < rows; row += waves) {
        const int r = (int)(row % H);
        const int rr = r + ty;
        const bool rv = rr >= 0 && rr < H;
        for (int c0 = 0; c0 < W; c0 += 32) {
            const int cb = c0 + 8 * g;          // this lane's 8 pixels: columns cb .. cb + 7
            bf16x8 A[MT], B[NT];
#pragma unroll
            for (int a = 0; a < MT; a++) {
                const int m = a * 16 + li;
                uint16_t v[8];
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    const int c = cb + j;
                    v[j] = (m < MY && c < W) ? dy[((size_t)row * W + c) * MY + m] : (uint16_t)0;
                }
                uint4 u;
                u.x = v[0] | ((uint32_t)v[1] << 16);
                u.y = v[2] | ((uint32_t)v[3] << 16);
                u.z = v[4] | ((uint32_t)v[5] << 16);
                u.w = v[6] | ((uint32_t)v[7] << 16);
                A[a] = as_bf16x8(u);
                if (t == 4) {
#pragma unroll
                    for (int j = 0; j < 8; j++) bsum[a] += from_bf16(v[j]);
                }
            }
#pragma unroll
            for (int b = 0; b < NT; b++) {
                const int k = b * 16 + li;
                uint16_t v[8];
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    const int c = cb + j, cs = c + tx;
                    v[j] = (k < KX && c < W && rv && cs >= 0 && cs < W)
                               ? x[((size_t)(row + ty) * W + cs) * KX + k] : (uint16_t)0;
                }
                uint4 u;
                u.x = v[0] | ((uint32_t)v[1] << 16);
                u.y = v[2] | ((uint32_t)v[3] << 16);
                u.z = v[4] | ((uint32_t)v[5] << 16);
                u.w = v[6] | ((uint32_t)v[7] << 16);
                B[b] = as_bf16x8(u);
            }
#pragma unroll
            for (int a = 0; a < MT; a++)
#pragma unroll
                for (int b = 0; b < NT; b++)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[a], B[b], acc[a][b], 0, 0, 0);
        }
    }
    // D[m][k]: lane holds m = a*16 + 4*g + i, k = b*16 + li
    float* dst = part + ((size_t)t * waves + w) * MY * KX;
#pragma unroll
    for (int a = 0; a < MT; a++)
#pragma unroll
        for (int b = 0; b < NT; b++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int m = a * 16 + 4 * g + i, k = b * 16 + li;
                if (m < MY && k < KX) dst[m * KX + k] = acc[a][b][i];
            }
    if (t == 4) {
#pragma unroll
        for (int a = 0; a < MT; a++) {
            float s = bsum[a];
            s += __shfl_xor(s, 16);
            s += __shfl_xor(s, 32);
            const int m = a * 16 + li;
            if (g == 0 && m < MY) bpart[(size_t)w * MY + m] = s;
        }
    }
}

// dw[t][m][k] = sum_w part[t][w][m][k] in order w = 0, 1, ...; dw has
// mrows >= my rows per tap (rows my .. mrows-1 are left untouched).
__global__ void wgrad_reduce_kernel(const float* __restrict__ part, const float* __restrict__ bpart,
                                    float* __restrict__ dw, float* __restrict__ db, int my, int kx, int mrows,
                                    int waves)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int t = blockIdx.y;
    const int n = my * kx;
    if (i < n) {
        float s = 0.f;
        for (int w = 0; w < waves; w++) s += part[((size_t)t * waves + w) * n + i];
        dw[((size_t)t * mrows + i / kx) * kx + i % kx] = s;
    }
    if (t == 4 && i < my) {
        float s = 0.f;
        for (int w = 0; w < waves; w++) s += bpart[(size_t)w * my + i];
        db[i] = s;
    }
}

// Mean-squared-error gradient of the 4-channel FP32 output against the
// target: dy[p][0..3] = BF16(2 (out - target) / count), dy[p][4..7] = 0; the
// per-block partial sums of squared error go to loss_part (FP64).
__global__ __launch_bounds__(256) void mse_kernel(const float* __restrict__ out, const float* __restrict__ tgt,
                                                  uint16_t* __restrict__ dy, double* __restrict__ loss_part,
                                                  long long npix, float scale)
{
    __shared__ double sh[4];
    double acc = 0.0;
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < npix;
         p += (long long)gridDim.x * blockDim.x) {
        const float4 o = *(const float4*)(out + 4 * p);
        const float4 y = *(const float4*)(tgt + 4 * p);
        const float d0 = o.x - y.x, d1 = o.y - y.y, d2 = o.z - y.z, d3 = o.w - y.w;
        acc += (double)d0 * d0 + (double)d1 * d1 + (double)d2 * d2 + (double)d3 * d3;
        uint4 st;
        st.x = (uint32_t)to_bf16(scale * d0) | ((uint32_t)to_bf16(scale * d1) << 16);
        st.y = (uint32_t)to_bf16(scale * d2) | ((uint32_t)to_bf16(scale * d3) << 16);
        st.z = 0;
        st.w = 0;
        *(uint4*)(dy + 8 * p) = st;
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) loss_part[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

// Adam (Kingma & Ba 2015) with bias correction folded into the step size.
__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ gr, float* __restrict__ m,
                            float* __restrict__ v, long long n, float lr_t, float b1, float b2, float eps)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float gi = gr[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr_t * mi / (sqrtf(vi) + eps);
    }
}

// FP32 master weights [9][m][k] -> BF16 forward weights (same layout) and,
// if wt != null, BF16 per-tap transposes wt[9][k][mt] keeping m < mt.
__global__ void pack_kernel(const float* __restrict__ w, uint16_t* __restrict__ wb, uint16_t* __restrict__ wt, int m,
                            int k, int mt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 9 * m * k) return;
    const int t = i / (m * k), rem = i % (m * k), mi = rem / k, ki = rem % k;
    const uint16_t b = to_bf16(w[i]);
    wb[i] = b;
    if (wt && mi < mt) wt[((size_t)t * k + ki) * mt + mi] = b;
}

// ---- ensemble-side kernels -------------------------------------------------
// State planes: h, qx, qy, H, hc (stride nvol); FP32 states add the low
// words of h, H, hc in planes 5, 6, 7.

template <typename S, bool COMP>
__device__ __forceinline__ double plane(const S* V, int f, long long nv, long long p)
{
    double v = (double)V[f * nv + p];
    if (COMP && (f == 0 || f == 3 || f == 4)) v += (double)V[(f == 0 ? 5 : f == 3 ? 6 : 7) * nv + p];
    return v;
}

template <typename S, bool COMP>
__device__ __forceinline__ void store_plane(S* V, int f, long long nv, long long p, double v)
{
    if (COMP && (f == 0 || f == 3 || f == 4)) {
        const float hi = (float)v;
        V[f * nv + p] = (S)hi;
        V[(f == 0 ? 5 : f == 3 ? 6 : 7) * nv + p] = (S)(float)(v - (double)hi);
    } else {
        V[f * nv + p] = (S)v;
    }
}

template <typename S, bool COMP>
__global__ __launch_bounds__(256) void features_kernel(const StageArgs* __restrict__ args, int npx, int npy,
                                                       NnScales sc, uint16_t* __restrict__ x)
{
    const int i = blockIdx.y;
    const S* V = (const S*)args[i].out;
    const double dx = args[i].P->dx, dy = args[i].P->dy;
    const long long nv = (long long)npx * npy;
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < nv; p += (long long)gridDim.x * blockDim.x) {
        const int r = (int)(p / npx), c = (int)(p % npx);
        const double h = plane<S, COMP>(V, 0, nv, p), H = plane<S, COMP>(V, 3, nv, p);
        const long long pl = c > 0 ? p - 1 : p, pr = c < npx - 1 ? p + 1 : p;
        const long long pd = r > 0 ? p - npx : p, pu = r < npy - 1 ? p + npx : p;
        const double sx = (double)((c > 0) + (c < npx - 1)) * dx, sy = (double)((r > 0) + (r < npy - 1)) * dy;
        const double dzdx = sx > 0 ? (plane<S, COMP>(V, 3, nv, pl) - plane<S, COMP>(V, 3, nv, pr)) / sx : 0.0;
        const double dzdy = sy > 0 ? (plane<S, COMP>(V, 3, nv, pd) - plane<S, COMP>(V, 3, nv, pu)) / sy : 0.0;
        const bool wet = h > sc.wet_h;
        float f[8];
        f[0] = (float)h * sc.in[0];
        f[1] = (float)V[nv + p] * sc.in[1];
        f[2] = (float)V[2 * nv + p] * sc.in[2];
        f[3] = (float)(-H) * sc.in[3];
        f[4] = (float)dzdx * sc.in[4];
        f[5] = (float)dzdy * sc.in[5];
        f[6] = wet ? (float)(h - H) * sc.in[6] : 0.f;
        f[7] = wet ? sc.in[7] : 0.f;
        uint4 st;
        st.x = (uint32_t)to_bf16(f[0]) | ((uint32_t)to_bf16(f[1]) << 16);
        st.y = (uint32_t)to_bf16(f[2]) | ((uint32_t)to_bf16(f[3]) << 16);
        st.z = (uint32_t)to_bf16(f[4]) | ((uint32_t)to_bf16(f[5]) << 16);
        st.w = (uint32_t)to_bf16(f[6]) | ((uint32_t)to_bf16(f[7]) << 16);
        *(uint4*)(x + ((size_t)i * nv + p) * 8) = st;
    }
}

template <typename S, bool COMP>
__global__ __launch_bounds__(256) void apply_kernel(const StageArgs* __restrict__ args, int npx, int npy, NnScales sc,
                                                    const float* __restrict__ out, long long expect,
                                                    float* __restrict__ resid, const uint16_t* __restrict__ x)
{
    const int i = blockIdx.y;
    const StageArgs& a = args[i];
    if (a.ctrl->step != expect) return;
    S* V = (S*)a.out;
    const double eps_h = a.P->epsilon_h;
    const long long nv = (long long)npx * npy;
    // wet flag (input channel 7) of cell q before this correction
    auto wet_in = [&](long long q) -> bool { return x[((size_t)i * nv + q) * 8 + 7] != 0; };
    float rmax = 0.f;
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < nv; p += (long long)gridDim.x * blockDim.x) {
        const float4 o = *(const float4*)(out + ((size_t)i * nv + p) * 4);
        const double h = plane<S, COMP>(V, 0, nv, p), H = plane<S, COMP>(V, 3, nv, p);
        double h2 = h + (double)(o.x * sc.out[0]);
        double H2 = H;
        if (h > sc.wet_h) {
            // residual indicator: largest correction of the surface (h - H)
            rmax = fmaxf(rmax, fabsf(o.x * sc.out[0] - o.w * sc.out[3]));
            if (!(h2 > 0.0)) h2 = 0.0;
            H2 = H + (double)(o.w * sc.out[3]);
        } else {
            // a dry cell is wetted only on the shoreline (a wet neighbour in
            // the network's inputs) and only deeper than shore_h; the bed of
            // dry cells is left alone
            if (!x || !(h2 > sc.shore_h)) continue;
            const int r = (int)(p / npx), c = (int)(p % npx);
            const bool shore = (c > 0 && wet_in(p - 1)) || (c < npx - 1 && wet_in(p + 1)) ||
                               (r > 0 && wet_in(p - npx)) || (r < npy - 1 && wet_in(p + npx));
            if (!shore) continue;
        }
        double qx = 0.0, qy = 0.0;
        if (h2 > 0.0) {
            qx = (double)V[nv + p] + (double)(o.y * sc.out[1]);
            qy = (double)V[2 * nv + p] + (double)(o.z * sc.out[2]);
        }
        store_plane<S, COMP>(V, 0, nv, p, h2);
        store_plane<S, COMP>(V, 1, nv, p, qx);
        store_plane<S, COMP>(V, 2, nv, p, qy);
        store_plane<S, COMP>(V, 3, nv, p, H2);
        if (h2 == 0.0) store_plane<S, COMP>(V, 4, nv, p, 0.0);
        if (a.ext) {
            // the stage kernels fold the stored leading words (FP32: hi)
            const double hx = (double)V[p], Hx = (double)V[3 * nv + p];
            double* hmax = a.ext;
            double* bmax = a.ext + nv;
            double* emax = a.ext + 2 * nv;
            if (hx > eps_h) {
                hmax[p] = fmax(hmax[p], hx);
                emax[p] = fmax(emax[p], hx - Hx);
            }
            bmax[p] = fmax(bmax[p], Hx);
        }
    }
    if (resid) {
        // one atomic per workgroup (the early return above is uniform over
        // it: one member per grid row); non-negative floats order like their
        // bit patterns
        __shared__ float wmax[4];
        for (int off = 32; off > 0; off >>= 1) rmax = fmaxf(rmax, __shfl_xor(rmax, off));
        if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = rmax;
        __syncthreads();
        if (threadIdx.x == 0) {
            const float m = fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]));
            if (m > 0.f) atomicMax((unsigned int*)(resid + i), __float_as_uint(m));
        }
    }
}

// The correction changes depths and momenta, so the dt the last stage left
// for the next step (from the uncorrected state) no longer bounds it.  The
// corrected members' dt slot is re-armed and refilled from the corrected
// state, as the initial dt is computed (GPU/main_TSUNER.cu:147-148), on the
// stream and without a host round trip.
__global__ void apply_dt_reset_kernel(const StageArgs* __restrict__ args, int n, long long expect)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Ctrl* C = args[i].ctrl;
    if (C->step == expect) C->dt_bits[expect & 1] = 0x7FF0000000000000ull;
}

template <typename S>
__global__ __launch_bounds__(256) void apply_dt_kernel(const StageArgs* __restrict__ args, int npx, int npy,
                                                       long long expect)
{
    const StageArgs& a = args[blockIdx.y];
    Ctrl* C = a.ctrl;
    if (C->step != expect) return;
    const S* V = (const S*)a.out;
    const DevParams& P = *a.P;
    const long long nv = (long long)npx * npy;
    const double inf = __longlong_as_double(0x7FF0000000000000ll);
    double m = inf;
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < nv; p += (long long)gridDim.x * blockDim.x)
        m = fmin(m, (double)cell_dt(P, V[p], V[nv + p], V[2 * nv + p]));
    for (int o = 32; o > 0; o >>= 1) m = fmin(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0 && m < inf)
        atomicMin(&C->dt_bits[expect & 1], (unsigned long long)__double_as_longlong(m));
}

// Free-surface restriction of the 4x4 block of fine cells under coarse cell
// (r, c): bed = block mean of H; surface = mean of h - H over the block's wet
// fine cells (h > kRestrictWetH), depth = max(surface + bed, 0) (0 when no
// fine cell is wet); discharge and suspended load = the block totals of q and
// hc scaled by depth / block total of h (mass-weighted velocity and
// concentration times the new depth).  A lake at rest restricts to a lake at
// rest: a block mean of h would lift the surface of every block that
// straddles the shoreline.
template <typename S, bool COMP>
__device__ __forceinline__ void restrict_block(const S* F, long long fnv, int fnx, int r, int c, double v[5])
{
    double sh = 0.0, sqx = 0.0, sqy = 0.0, sH = 0.0, shc = 0.0, seta = 0.0;
    int nw = 0;
    for (int dr = 0; dr < 4; dr++)
        for (int dc = 0; dc < 4; dc++) {
            const long long q = (long long)(4 * r + dr) * fnx + 4 * c + dc;
            const double h = plane<S, COMP>(F, 0, fnv, q), H = plane<S, COMP>(F, 3, fnv, q);
            sh += h;
            sqx += plane<S, COMP>(F, 1, fnv, q);
            sqy += plane<S, COMP>(F, 2, fnv, q);
            sH += H;
            shc += plane<S, COMP>(F, 4, fnv, q);
            if (h > kRestrictWetH) {
                seta += h - H;
                nw++;
            }
        }
    const double Hc = sH * 0.0625;
    const double hc = nw ? fmax(seta / (double)nw + Hc, 0.0) : 0.0;
    const double k = (hc > 0.0 && sh > 0.0) ? hc / sh : 0.0;
    v[0] = hc;
    v[1] = sqx * k;
    v[2] = sqy * k;
    v[3] = Hc;
    v[4] = shc * k;
}

template <typename S, bool COMP>
__global__ __launch_bounds__(256) void restrict_kernel(const StageArgs* __restrict__ fine,
                                                       const StageArgs* __restrict__ coarse, int npx, int npy)
{
    const int i = blockIdx.y;
    const S* F = (const S*)fine[i].out;
    S* Cs = (S*)coarse[i].out;
    const long long nv = (long long)npx * npy, fnv = 16 * nv;
    const int fnx = 4 * npx;
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < nv; p += (long long)gridDim.x * blockDim.x) {
        double v[5];
        restrict_block<S, COMP>(F, fnv, fnx, (int)(p / npx), (int)(p % npx), v);
        for (int f = 0; f < 5; f++) store_plane<S, COMP>(Cs, f, nv, p, v[f]);
    }
}

template <typename S, bool COMP>
__global__ __launch_bounds__(256) void target_kernel(const StageArgs* __restrict__ fine,
                                                     const StageArgs* __restrict__ coarse, int npx, int npy,
                                                     NnScales sc, float* __restrict__ target)
{
    const int i = blockIdx.y;
    const S* F = (const S*)fine[i].out;
    const S* Cs = (const S*)coarse[i].out;
    const long long nv = (long long)npx * npy, fnv = 16 * nv;
    const int fnx = 4 * npx;
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < nv; p += (long long)gridDim.x * blockDim.x) {
        double v[5];
        restrict_block<S, COMP>(F, fnv, fnx, (int)(p / npx), (int)(p % npx), v);
        float o[4];
        for (int j = 0; j < 4; j++) o[j] = (float)((v[j] - plane<S, COMP>(Cs, j, nv, p)) / (double)sc.out[j]);
        *(float4*)(target + ((size_t)i * nv + p) * 4) = make_float4(o[0], o[1], o[2], o[3]);
    }
}

}  // namespace nn

using namespace nn;

namespace {
dim3 member_grid(long long nv, int n)
{
    long long b = (nv + 255) / 256;
    if (b > 1024) b = 1024;
    return dim3((unsigned)b, (unsigned)n);
}
}  // namespace

hipError_t nn_launch_features(const StageArgs* args, int n, int npx, int npy, int precision, const NnScales& sc,
                              uint16_t* x, hipStream_t s)
{
    const dim3 g = member_grid((long long)npx * npy, n);
    if (precision == 0)
        features_kernel<double, false><<<g, 256, 0, s>>>(args, npx, npy, sc, x);
    else
        features_kernel<float, true><<<g, 256, 0, s>>>(args, npx, npy, sc, x);
    return hipGetLastError();
}

hipError_t nn_launch_apply(const StageArgs* args, int n, int npx, int npy, int precision, const NnScales& sc,
                           const float* out, long long expect_step, float* resid, const uint16_t* x,
                           hipStream_t s)
{
    const dim3 g = member_grid((long long)npx * npy, n);
    if (precision == 0)
        apply_kernel<double, false><<<g, 256, 0, s>>>(args, npx, npy, sc, out, expect_step, resid, x);
    else
        apply_kernel<float, true><<<g, 256, 0, s>>>(args, npx, npy, sc, out, expect_step, resid, x);
    apply_dt_reset_kernel<<<(n + 63) / 64, 64, 0, s>>>(args, n, expect_step);
    if (precision == 0)
        apply_dt_kernel<double><<<g, 256, 0, s>>>(args, npx, npy, expect_step);
    else
        apply_dt_kernel<float><<<g, 256, 0, s>>>(args, npx, npy, expect_step);
    return hipGetLastError();
}

hipError_t nn_launch_restrict(const StageArgs* fine, const StageArgs* coarse, int n, int npx, int npy,
                              int precision, hipStream_t s)
{
    const dim3 g = member_grid((long long)npx * npy, n);
    if (precision == 0)
        restrict_kernel<double, false><<<g, 256, 0, s>>>(fine, coarse, npx, npy);
    else
        restrict_kernel<float, true><<<g, 256, 0, s>>>(fine, coarse, npx, npy);
    return hipGetLastError();
}

hipError_t nn_launch_target(const StageArgs* fine, const StageArgs* coarse, int n, int npx, int npy, int precision,
                            const NnScales& sc, float* target, hipStream_t s)
{
    const dim3 g = member_grid((long long)npx * npy, n);
    if (precision == 0)
        target_kernel<double, false><<<g, 256, 0, s>>>(fine, coarse, npx, npy, sc, target);
    else
        target_kernel<float, true><<<g, 256, 0, s>>>(fine, coarse, npx, npy, sc, target);
    return hipGetLastError();
}

namespace {
// conv_wide_kernel's persistent grid: one workgroup per CU (its LDS holds one),
// at most one per tile.
dim3 wide_grid(int N, int H, int W)
{
    static int cus = 0;
    if (cus == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
            cus = 256;
    }
    const long long tiles = (long long)N * ((H + kWideRows - 1) / kWideRows) * ((W + 63) / 64);
    return dim3((unsigned)(tiles < cus ? tiles : cus));
}
}  // namespace

hipError_t nn_launch_conv(int kind, const uint16_t* in, const uint16_t* w, const float* bias, const uint16_t* act,
                          void* out, int N, int H, int W, hipStream_t s)
{
    const long long blocks = (long long)N * ((H + kConvRows - 1) / kConvRows) * ((W + 63) / 64);
    const dim3 grid((unsigned)blocks), block(256);
    switch (kind) {
    case NN_FWD_IN:      // 8 -> 32, ReLU
        conv_kernel<8, 32, 32, 1, 0><<<grid, block, 0, s>>>(in, w, bias, act, out, N, H, W);
        break;
    case NN_FWD_MID:     // 32 -> 32, ReLU
        conv_kernel<32, 32, 32, 1, 0><<<grid, block, 0, s>>>(in, w, bias, act, out, N, H, W);
        break;
    case NN_FWD_OUT:     // 32 -> 4 (weights padded to 16 rows), FP32 out
        conv_kernel<32, 16, 4, 1, 1><<<grid, block, 0, s>>>(in, w, bias, act, out, N, H, W);
        break;
    case NN_DGRAD_OUT:   // dy 8 channels (4 used) -> dx 32, ReLU mask
        conv_kernel<8, 32, 32, -1, 2><<<grid, block, 0, s>>>(in, w, bias, act, out, N, H, W);
        break;
    case NN_DGRAD_MID:   // dy 32 -> dx 32, ReLU mask
        conv_kernel<32, 32, 32, -1, 2><<<grid, block, 0, s>>>(in, w, bias, act, out, N, H, W);
        break;
    case NN_FWD_IN | NN_W64:
        conv_wide_kernel<8, 64, 64, 1, 0><<<wide_grid(N, H, W), 512, 0, s>>>(in, w, bias, act, out, N, H, W);
        break;
    case NN_FWD_MID | NN_W64:
        conv_wide_kernel<64, 64, 64, 1, 0><<<wide_grid(N, H, W), 512, 0, s>>>(in, w, bias, act, out, N, H, W);
        break;
    case NN_FWD_OUT | NN_W64:
        conv_wide_kernel<64, 16, 4, 1, 1><<<wide_grid(N, H, W), 512, 0, s>>>(in, w, bias, act, out, N, H, W);
        break;
    case NN_DGRAD_OUT | NN_W64:
        conv_wide_kernel<8, 64, 64, -1, 2><<<wide_grid(N, H, W), 512, 0, s>>>(in, w, bias, act, out, N, H, W);
        break;
    case NN_DGRAD_MID | NN_W64:
        conv_wide_kernel<64, 64, 64, -1, 2><<<wide_grid(N, H, W), 512, 0, s>>>(in, w, bias, act, out, N, H, W);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t nn_launch_wgrad(int kind, const uint16_t* x, const uint16_t* dy, float* part, float* bpart, float* dw,
                           float* db, int mrows, int N, int H, int W, hipStream_t s)
{
    const int waves = kWgradWaves;
    const dim3 grid((unsigned)((waves + 3) / 4), 9), block(256);
    int my, kx;
    switch (kind) {
    case NN_WGRAD_IN:
        wgrad_kernel<8, 32><<<grid, block, 0, s>>>(x, dy, part, bpart, N, H, W, waves);
        my = 32;
        kx = 8;
        break;
    case NN_WGRAD_MID:
        wgrad_kernel<32, 32><<<grid, block, 0, s>>>(x, dy, part, bpart, N, H, W, waves);
        my = 32;
        kx = 32;
        break;
    case NN_WGRAD_OUT:
        wgrad_kernel<32, 8><<<grid, block, 0, s>>>(x, dy, part, bpart, N, H, W, waves);
        my = 8;
        kx = 32;
        break;
    case NN_WGRAD_IN | NN_W64:
        wgrad_kernel<8, 64><<<grid, block, 0, s>>>(x, dy, part, bpart, N, H, W, waves);
        my = 64;
        kx = 8;
        break;
    case NN_WGRAD_MID | NN_W64:
        wgrad_kernel<64, 64><<<grid, block, 0, s>>>(x, dy, part, bpart, N, H, W, waves);
        my = 64;
        kx = 64;
        break;
    case NN_WGRAD_OUT | NN_W64:
        wgrad_kernel<64, 8><<<grid, block, 0, s>>>(x, dy, part, bpart, N, H, W, waves);
        my = 8;
        kx = 64;
        break;
    default:
        return hipErrorInvalidValue;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int n = my * kx;
    wgrad_reduce_kernel<<<dim3((n + 255) / 256, 9), 256, 0, s>>>(part, bpart, dw, db, my, kx, mrows, waves);
    return hipGetLastError();
}

hipError_t nn_launch_mse(const float* out, const float* tgt, uint16_t* dy, double* loss_part, int nblocks,
                         long long npix, float scale, hipStream_t s)
{
    mse_kernel<<<nblocks, 256, 0, s>>>(out, tgt, dy, loss_part, npix, scale);
    return hipGetLastError();
}

hipError_t nn_launch_adam(float* p, const float* g, float* m, float* v, long long n, float lr_t, float b1, float b2,
                          float eps, hipStream_t s)
{
    adam_kernel<<<(unsigned)((n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024), 256, 0, s>>>(p, g, m, v, n, lr_t, b1,
                                                                                           b2, eps);
    return hipGetLastError();
}

hipError_t nn_launch_pack(const float* w, uint16_t* wb, uint16_t* wt, int m, int k, int mt, hipStream_t s)
{
    const int n = 9 * m * k;
    pack_kernel<<<(n + 255) / 256, 256, 0, s>>>(w, wb, wt, m, k, mt);
    return hipGetLastError();
}

}  // namespace tsu
