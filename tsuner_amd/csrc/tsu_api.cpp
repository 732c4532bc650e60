// tsu_api.cpp -- host runtime behind the libtsuner C ABI (include/tsuner.h).
//
// A tsu_state owns everything one simulation needs on one device: two
// planar SoA state buffers (U^n and the first RK stage), the bed mask as
// bytes, the inflow series, the step-control block and one stream.  Steps are
// enqueued without host synchronisation: dt and the simulation time live on
// the device (GPU/main_TSUNER.cu:147-197 keeps them on the host and blocks on
// a device-to-host copy every step).
#define TSU_BUILDING_LIBRARY 1
#include "../../include/tsuner.h"

#include <hip/hip_runtime.h>

#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "tsu_host.h"
#include "tsu_kernels.h"

using tsu::Ctrl;
using tsu::DevParams;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess) return fail(TSU_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr unsigned long long kInfBits = 0x7FF0000000000000ull;

unsigned long long dbits(double v)
{
    unsigned long long b;
    std::memcpy(&b, &v, 8);
    return b;
}

double bitsd(unsigned long long b)
{
    double v;
    std::memcpy(&v, &b, 8);
    return v;
}

// Host twin of tsu::finalize_dt.
double finalize_host(const DevParams& P, unsigned long long bits)
{
    double m = bitsd(bits);
    if (P.dt_narrow && P.nvol > 4096) m = (double)(float)m;
    return (m < 1e30) ? m : 1e30;
}

int check_scheme(const tsu_scheme* s)
{
    if (s->order != 1 && s->order != 2) return fail(TSU_ERR_UNSUPPORTED, "order must be 1 or 2");
    for (int i = 0; i < 4; i++)
        if (s->bc[i] < TSU_BC_FREE || s->bc[i] > TSU_BC_INFLOW) return fail(TSU_ERR_ARG, "bad boundary type");
    if ((s->bc[0] == TSU_BC_PERIODIC) != (s->bc[1] == TSU_BC_PERIODIC) ||
        (s->bc[2] == TSU_BC_PERIODIC) != (s->bc[3] == TSU_BC_PERIODIC))
        return fail(TSU_ERR_UNSUPPORTED, "periodic boundaries must be paired on an axis");
    if (s->limiter < 0 || s->limiter > 1 || s->bedload_law < 0 || s->bedload_law > 1 || s->friction_law < 0 ||
        s->friction_law > 1)
        return fail(TSU_ERR_ARG, "bad law selector");
    if ((s->bedload != 0 && s->bedload != 1) || (s->suspension != 0 && s->suspension != 1))
        return fail(TSU_ERR_ARG, "bedload and suspension are 0/1 switches");
    return TSU_OK;
}

}  // namespace

namespace tsu {

// Run constants from the reference parameters, evaluated in the reference's
// operation order (Reconstructor.cu:64-77, SedFlux.cuh:56-83).
void make_dev_params(const tsu_sim_params* p, const tsu_scheme* s, DevParams* P)
{
    std::memset(P, 0, sizeof(*P));
    const double G = 9.81;
    P->npx = p->npx;
    P->npy = p->npy;
    P->nvol = (long long)p->npx * p->npy;
    P->dx = p->dx;
    P->dy = p->dy;
    P->inv_dx = 1.0 / p->dx;
    P->inv_dy = 1.0 / p->dy;
    P->cfl_half = p->CFL * 0.5;
    P->FR = p->FR;
    P->epsilon_h = p->epsilon_h;
    P->Q = p->Q;
    P->Opar = p->Opar;
    P->OC = p->OC;
    P->sqrtOC = std::sqrt(p->OC);
    P->HQB = p->HQB;
    P->fiveDS = 5 * p->DS;
    const double nu = 1.002e-6;
    const double sd = p->R - 1.0;
    P->vs = std::sqrt((13.95 * nu / p->DS) * (13.95 * nu / p->DS) + 1.09 * sd * G * p->DS) - 13.95 * nu / p->DS;
    const double Rp = (p->DS * std::sqrt(sd * G * p->DS)) / nu;
    P->ka = (Rp > 2.36) ? 1.0 * std::pow(Rp, 0.6) : 0.586 * std::pow(Rp, 1.23);
    P->sqrtFR = std::sqrt(p->FR);
    P->cbk = 1.64 + 0.4 * std::pow(p->DS / p->Dsg, 1.64);
    P->vs1p = P->vs * (1 - p->porosity);
    P->inv_vs = 1.0 / P->vs;
    P->grass_k = s->grass_ag / (1.0 - p->porosity);
    P->manning_n = s->manning_n;
    P->bedload = s->bedload;
    P->suspension = s->suspension;
    P->limiter = s->limiter;
    P->bedload_law = s->bedload_law;
    P->friction_law = s->friction_law;
    P->order = s->order;
    P->dt_narrow = s->dt_narrow;
    P->extrema = s->extrema;
    // bedload square roots: |u|^2 > OC (1 - 2^-40) / Opar and O > OC there
    const double tiny = std::ldexp(1.0, -700);
    P->bed_sqrt_unscaled = (P->OC > tiny && P->Opar > 0.0 && std::isfinite(P->Opar) &&
                            P->OC * (1.0 - std::ldexp(1.0, -40)) / P->Opar > tiny) ? 1 : 0;
    for (int i = 0; i < 4; i++) P->bc[i] = s->bc[i];
    P->ring = 2LL * p->npx + 2LL * p->npy;
}

int validate_scheme(const tsu_scheme* s) { return check_scheme(s); }

int set_error(int code, const std::string& msg) { return fail(code, msg); }

}  // namespace tsu

struct tsu_state {
    int device = 0;
    int precision = TSU_FP64;
    bool sed = true;
    bool uploaded = false;
    tsu_sim_params p{};
    tsu_scheme s{};
    DevParams P{};
    size_t esize = 8;
    long long nvol = 0;
    void* A = nullptr;
    void* B = nullptr;
    uint8_t* fix = nullptr;
    Ctrl* ctrl = nullptr;
    double* ext = nullptr;
    double* CC = nullptr;
    double* tCC = nullptr;
    unsigned long timeCC = 0;
    double* mass_buf = nullptr;
    int mass_nb = 0;
    DevParams* dP = nullptr;   // device copy of P read by the stage kernels
    hipStream_t stream = nullptr;
    int tile_rows = 0, nstrips = 0, nchunks = 0, ntiles = 0;
    int graph_chunk = 0;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    bool capturing = false;
    // stage-kernel event timing
    bool timing = false;
    std::vector<hipEvent_t> ev;        // start/end pairs, reused
    std::vector<int> ev_stage;         // stage index of each pair
    size_t ev_used = 0;                // pairs recorded since the last read
    double stage_ms[2] = {0.0, 0.0};
    long long launches[2] = {0, 0};
};

namespace {

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev)
    {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        (void)hipSetDevice(dev);
    }
    ~DeviceGuard()
    {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

// FP64: h, qx, qy, H, hc.  FP32: the same five planes plus the low-order
// parts of h, H and hc (compensated accumulation).
size_t planes_of(const tsu_state* st) { return st->precision == TSU_FP64 ? 5 : 8; }

int default_tile_rows(const tsu_state* st)
{
    // Enough waves to cover the 1024 SIMDs a few times, tall enough to keep
    // the vertical halo overhead small.  Measured at 8192^2: 32 rows beat 16,
    // 24, 48, 64, 96 and 128 in both precisions (smaller last-round tail);
    // at 2048^2, 8 rows (8960 tiles) beat 4 and 16.  Small grids that cannot
    // fill 8192 tiles at 8 rows go down to 2 rows while under 2048 tiles:
    // at 512^2, 2 rows (2304 tiles) run 1.37x faster than 8 and beat 1 and 4.
    const long long strips = (st->p.npx + tsu::kStripOut - 1) / tsu::kStripOut;
    auto tiles = [&](int r) { return strips * ((st->p.npy + r - 1) / r); };
    int rows = 32;
    while (rows > 8 && tiles(rows) < 8192) rows >>= 1;
    while (rows > 2 && tiles(rows) < 2048) rows >>= 1;
    return rows;
}

void set_tiling(tsu_state* st, int rows)
{
    st->tile_rows = rows;
    st->nstrips = (st->p.npx + tsu::kStripOut - 1) / tsu::kStripOut;
    st->nchunks = (st->p.npy + rows - 1) / rows;
    st->ntiles = st->nstrips * st->nchunks;
}

void drop_graph(tsu_state* st)
{
    if (st->gexec) (void)hipGraphExecDestroy(st->gexec);
    if (st->graph) (void)hipGraphDestroy(st->graph);
    st->gexec = nullptr;
    st->graph = nullptr;
}

void free_state(tsu_state* st)
{
    if (!st) return;
    DeviceGuard g(st->device);
    drop_graph(st);
    if (st->stream) (void)hipStreamSynchronize(st->stream);
    for (hipEvent_t e : st->ev) (void)hipEventDestroy(e);
    (void)hipFree(st->A);
    (void)hipFree(st->B);
    (void)hipFree(st->fix);
    (void)hipFree(st->ctrl);
    (void)hipFree(st->ext);
    (void)hipFree(st->CC);
    (void)hipFree(st->tCC);
    (void)hipFree(st->mass_buf);
    (void)hipFree(st->dP);
    if (st->stream) (void)hipStreamDestroy(st->stream);
    delete st;
}

tsu::StageArgs stage_args(const tsu_state* st, int k)
{
    tsu::StageArgs a{};
    const size_t plane = (size_t)st->nvol * st->esize;
    a.Hin = (const char*)st->A + 3 * plane;
    a.fix = st->fix;
    a.ctrl = st->ctrl;
    a.ext = st->s.extrema ? st->ext : nullptr;
    a.tile_rows = st->tile_rows;
    a.nstrips = st->nstrips;
    a.ntiles = st->ntiles;
    a.P = st->dP;
    a.k = k;
    if (st->s.order == 1) {
        a.in = st->A;
        a.out = st->B;
        a.v0 = st->A;
        a.pingpong = 1;
        a.last = 1;
    } else if (k == 0) {
        a.in = st->A;
        a.out = st->B;
        a.v0 = st->A;
        a.last = 0;
    } else {
        a.in = st->B;
        a.out = st->A;
        a.v0 = st->A;
        a.last = 1;
    }
    return a;
}

int enqueue_step(tsu_state* st)
{
    const bool timed = st->timing && !st->capturing;
    for (int k = 0; k < st->s.order; k++) {
        tsu::StageArgs a = stage_args(st, k);
        if (timed) {
            if (st->ev_used * 2 + 2 > st->ev.size()) {
                for (int i = 0; i < 2; i++) {
                    hipEvent_t e;
                    HIPCHK(hipEventCreate(&e));
                    st->ev.push_back(e);
                }
                st->ev_stage.push_back(0);
            }
            st->ev_stage[st->ev_used] = k;
            HIPCHK(hipEventRecord(st->ev[2 * st->ev_used], st->stream));
        }
        HIPCHK(tsu::launch_stage(a, st->sed, st->precision, st->stream, tsu::fast_bedload(st->P)));
        if (timed) {
            HIPCHK(hipEventRecord(st->ev[2 * st->ev_used + 1], st->stream));
            st->ev_used++;
        }
    }
    HIPCHK(tsu::launch_advance(st->ctrl, st->P, st->stream));
    return TSU_OK;
}

int read_ctrl(tsu_state* st, Ctrl* c)
{
    HIPCHK(hipMemcpyAsync(c, st->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, st->stream));
    HIPCHK(hipStreamSynchronize(st->stream));
    return TSU_OK;
}

// Buffer holding U^n (the single-stage scheme alternates by step parity).
int current_buffer(tsu_state* st, void** buf)
{
    if (st->s.order == 2) {
        *buf = st->A;
        return TSU_OK;
    }
    Ctrl c;
    int rc = read_ctrl(st, &c);
    if (rc) return rc;
    *buf = (c.step & 1) ? st->B : st->A;
    return TSU_OK;
}

}  // namespace

extern "C" {

TSU_API const char* tsu_last_error(void) { return g_err.c_str(); }

TSU_API const char* tsu_version(void) { return "tsuner-amd 0.1 (gfx950)"; }

TSU_API int checkCUDAsupport(int device)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return 1;
    if (device < 0 || device >= n) return 3;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return 2;
    return std::strncmp(prop.gcnArchName, "gfx950", 6) == 0 ? 0 : 2;
}

TSU_API int tsu_device_check(int device) { return checkCUDAsupport(device); }

TSU_API void initializeGPU(void)
{
    (void)tsu::launch_warmup(nullptr);
    (void)hipDeviceSynchronize();
}

TSU_API int tsu_state_create(tsu_state** out, const tsu_sim_params* p, const tsu_scheme* s, int precision,
                             int device)
{
    if (!out || !p || !s) return fail(TSU_ERR_ARG, "null argument");
    *out = nullptr;
    if (p->npx < 1 || p->npy < 1) return fail(TSU_ERR_ARG, "npx and npy must be positive");
    if (!(p->dx > 0) || !(p->dy > 0)) return fail(TSU_ERR_ARG, "dx and dy must be positive");
    if (precision != TSU_FP64 && precision != TSU_FP32) return fail(TSU_ERR_ARG, "precision must be FP64 or FP32");
    int rc = check_scheme(s);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(TSU_ERR_HIP, "no HIP device");
    if (device < 0 || device >= ndev) return fail(TSU_ERR_ARG, "bad device index");

    tsu_state* st = new tsu_state();
    st->device = device;
    st->precision = precision;
    st->p = *p;
    st->p.nvol = p->npx * p->npy;
    st->s = *s;
    st->esize = precision == TSU_FP64 ? 8 : 4;
    st->nvol = (long long)p->npx * p->npy;
    tsu::make_dev_params(&st->p, &st->s, &st->P);
    set_tiling(st, default_tile_rows(st));

    DeviceGuard g(device);
    const size_t state_bytes = planes_of(st) * (size_t)st->nvol * st->esize;
    st->mass_nb = tsu::mass_blocks(st->nvol);
    hipError_t e = hipStreamCreateWithFlags(&st->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&st->A, state_bytes);
    if (e == hipSuccess) e = hipMalloc(&st->B, state_bytes);
    if (e == hipSuccess) e = hipMalloc((void**)&st->fix, (size_t)st->nvol);
    if (e == hipSuccess) e = hipMalloc((void**)&st->ctrl, sizeof(Ctrl));
    if (e == hipSuccess) e = hipMalloc((void**)&st->mass_buf, (2 * (size_t)st->mass_nb + 2) * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&st->dP, sizeof(DevParams));
    if (e == hipSuccess) e = hipMemcpy(st->dP, &st->P, sizeof(DevParams), hipMemcpyHostToDevice);
    if (e == hipSuccess && s->extrema) e = hipMalloc((void**)&st->ext, 3 * (size_t)st->nvol * sizeof(double));
    if (e != hipSuccess) {
        free_state(st);
        (void)hipGetLastError();
        return fail(TSU_ERR_DEVICE_MEMORY, std::string("device allocation failed: ") + hipGetErrorString(e));
    }
    Ctrl c{};
    c.t = 0.0;
    c.stop_time = INFINITY;
    c.dt_bits[0] = c.dt_bits[1] = kInfBits;
    c.step = 0;
    c.stop_step = LLONG_MAX;
    e = hipMemcpy(st->ctrl, &c, sizeof(Ctrl), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        free_state(st);
        return fail(TSU_ERR_HIP, std::string("ctrl init: ") + hipGetErrorString(e));
    }
    *out = st;
    return TSU_OK;
}

TSU_API int tsu_state_destroy(tsu_state* st)
{
    free_state(st);
    return TSU_OK;
}

TSU_API int tsu_state_upload(tsu_state* st, const double* V, const int* fixH)
{
    if (!st || !V || !fixH) return fail(TSU_ERR_ARG, "null argument");
    const long long nv = st->nvol;
    std::vector<uint8_t> fx((size_t)nv);
    for (long long i = 0; i < nv; i++) {
        if (fixH[i] != 0 && fixH[i] != 1) return fail(TSU_ERR_ARG, "fixH values must be 0 or 1");
        fx[(size_t)i] = (uint8_t)fixH[i];
    }
    bool hc_zero = true;
    for (long long i = 0; i < nv && hc_zero; i++) hc_zero = V[4 * nv + i] == 0.0;
    st->sed = st->s.bedload || st->s.suspension || !hc_zero;
    DeviceGuard g(st->device);
    drop_graph(st);
    const size_t bytes = planes_of(st) * (size_t)nv * st->esize;
    if (st->precision == TSU_FP64) {
        HIPCHK(hipMemcpyAsync(st->A, V, bytes, hipMemcpyHostToDevice, st->stream));
    } else {
        // hi planes 0-4, then the lo parts of h, H, hc (planes 5-7)
        std::vector<float> f((size_t)8 * nv);
        for (size_t i = 0; i < (size_t)5 * nv; i++) f[i] = (float)V[i];
        const int src[3] = {0, 3, 4};
        for (int j = 0; j < 3; j++)
            for (size_t i = 0; i < (size_t)nv; i++) {
                const size_t k = (size_t)src[j] * nv + i;
                f[(size_t)(5 + j) * nv + i] = (float)(V[k] - (double)f[k]);
            }
        HIPCHK(hipMemcpyAsync(st->A, f.data(), bytes, hipMemcpyHostToDevice, st->stream));
        HIPCHK(hipStreamSynchronize(st->stream));
    }
    HIPCHK(hipMemcpyAsync(st->B, st->A, bytes, hipMemcpyDeviceToDevice, st->stream));
    HIPCHK(hipMemcpyAsync(st->fix, fx.data(), (size_t)nv, hipMemcpyHostToDevice, st->stream));
    Ctrl c{};
    c.t = 0.0;
    c.stop_time = INFINITY;
    c.dt_bits[0] = c.dt_bits[1] = kInfBits;
    c.step = 0;
    c.stop_step = LLONG_MAX;
    HIPCHK(hipMemcpyAsync(st->ctrl, &c, sizeof(Ctrl), hipMemcpyHostToDevice, st->stream));
    if (st->ext) HIPCHK(tsu::launch_fill(st->ext, 3 * nv, -1e30, st->stream));
    HIPCHK(hipStreamSynchronize(st->stream));
    st->uploaded = true;
    return TSU_OK;
}

TSU_API int tsu_state_download(tsu_state* st, double* V)
{
    if (!st || !V) return fail(TSU_ERR_ARG, "null argument");
    DeviceGuard g(st->device);
    void* cur;
    int rc = current_buffer(st, &cur);
    if (rc) return rc;
    const long long nv = st->nvol;
    const size_t bytes = planes_of(st) * (size_t)nv * st->esize;
    if (st->precision == TSU_FP64) {
        HIPCHK(hipMemcpyAsync(V, cur, bytes, hipMemcpyDeviceToHost, st->stream));
        HIPCHK(hipStreamSynchronize(st->stream));
    } else {
        std::vector<float> f((size_t)8 * nv);
        HIPCHK(hipMemcpyAsync(f.data(), cur, bytes, hipMemcpyDeviceToHost, st->stream));
        HIPCHK(hipStreamSynchronize(st->stream));
        for (size_t i = 0; i < (size_t)5 * nv; i++) V[i] = (double)f[i];
        const int dst[3] = {0, 3, 4};
        for (int j = 0; j < 3; j++)
            for (size_t i = 0; i < (size_t)nv; i++) V[(size_t)dst[j] * nv + i] += (double)f[(size_t)(5 + j) * nv + i];
    }
    return TSU_OK;
}

TSU_API int tsu_state_set_inflow(tsu_state* st, const double* CC, const double* tCC, unsigned long timeCC)
{
    if (!st || !CC || !tCC) return fail(TSU_ERR_ARG, "null argument");
    if (timeCC < 2) return fail(TSU_ERR_ARG, "the inflow series needs at least two time levels");
    // interpolate_time divides by tCC[1] - tCC[0] and truncates the quotient
    // to an int level index (VolumeSolver.cu:174-176): the times must be
    // finite and increasing
    for (unsigned long k = 0; k < timeCC; k++)
        if (!std::isfinite(tCC[k]) || (k > 0 && !(tCC[k] > tCC[k - 1])))
            return fail(TSU_ERR_ARG, "inflow times must be finite and strictly increasing");
    DeviceGuard g(st->device);
    // steps already enqueued may still read the old series and run constants
    HIPCHK(hipStreamSynchronize(st->stream));
    drop_graph(st);
    (void)hipFree(st->CC);
    (void)hipFree(st->tCC);
    st->CC = st->tCC = nullptr;
    const size_t ncc = (size_t)st->P.ring * 4 * timeCC;
    hipError_t e = hipMalloc((void**)&st->CC, ncc * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&st->tCC, timeCC * sizeof(double));
    if (e != hipSuccess) return fail(TSU_ERR_DEVICE_MEMORY, "inflow series allocation failed");
    HIPCHK(hipMemcpy(st->CC, CC, ncc * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(st->tCC, tCC, timeCC * sizeof(double), hipMemcpyHostToDevice));
    st->timeCC = timeCC;
    st->P.CC = st->CC;
    st->P.tCC = st->tCC;
    st->P.timeCC = (long long)timeCC;
    st->P.tcc0 = tCC[0];
    st->P.dtcc = tCC[1] - tCC[0];
    st->P.tcc_end = tCC[timeCC - 2];
    HIPCHK(hipMemcpy(st->dP, &st->P, sizeof(DevParams), hipMemcpyHostToDevice));
    return TSU_OK;
}

static int check_ready(tsu_state* st)
{
    if (!st) return fail(TSU_ERR_ARG, "null state");
    if (!st->uploaded) return fail(TSU_ERR_ARG, "state not uploaded");
    for (int i = 0; i < 4; i++)
        if (st->s.bc[i] == TSU_BC_INFLOW && !st->CC)
            return fail(TSU_ERR_ARG, "an inflow side needs tsu_state_set_inflow first");
    return TSU_OK;
}

TSU_API int tsu_init_dt(tsu_state* st)
{
    int rc = check_ready(st);
    if (rc) return rc;
    DeviceGuard g(st->device);
    Ctrl c;
    rc = read_ctrl(st, &c);
    if (rc) return rc;
    c.dt_bits[0] = c.dt_bits[1] = kInfBits;
    HIPCHK(hipMemcpyAsync(st->ctrl, &c, sizeof(Ctrl), hipMemcpyHostToDevice, st->stream));
    void* cur = (st->s.order == 1 && (c.step & 1)) ? st->B : st->A;
    HIPCHK(tsu::launch_init_dt(cur, st->P, st->ctrl, st->precision, st->stream));
    HIPCHK(hipStreamSynchronize(st->stream));
    return TSU_OK;
}

TSU_API int tsu_step(tsu_state* st, int nsteps)
{
    int rc = check_ready(st);
    if (rc) return rc;
    if (nsteps < 0) return fail(TSU_ERR_ARG, "negative step count");
    DeviceGuard g(st->device);
    int done = 0;
    if (st->graph_chunk > 0 && nsteps >= st->graph_chunk) {
        if (!st->gexec) {
            HIPCHK(hipStreamBeginCapture(st->stream, hipStreamCaptureModeThreadLocal));
            st->capturing = true;
            for (int i = 0; i < st->graph_chunk; i++) {
                rc = enqueue_step(st);
                if (rc) {
                    hipGraph_t gph = nullptr;
                    st->capturing = false;
                    (void)hipStreamEndCapture(st->stream, &gph);
                    if (gph) (void)hipGraphDestroy(gph);
                    return rc;
                }
            }
            st->capturing = false;
            HIPCHK(hipStreamEndCapture(st->stream, &st->graph));
            HIPCHK(hipGraphInstantiate(&st->gexec, st->graph, nullptr, nullptr, 0));
        }
        while (nsteps - done >= st->graph_chunk) {
            HIPCHK(hipGraphLaunch(st->gexec, st->stream));
            done += st->graph_chunk;
        }
    }
    for (; done < nsteps; done++) {
        rc = enqueue_step(st);
        if (rc) return rc;
    }
    return TSU_OK;
}

TSU_API int tsu_synchronize(tsu_state* st)
{
    if (!st) return fail(TSU_ERR_ARG, "null state");
    DeviceGuard g(st->device);
    HIPCHK(hipStreamSynchronize(st->stream));
    return TSU_OK;
}

TSU_API int tsu_get_time(tsu_state* st, double* t, double* dt, long long* steps)
{
    if (!st) return fail(TSU_ERR_ARG, "null state");
    DeviceGuard g(st->device);
    Ctrl c;
    int rc = read_ctrl(st, &c);
    if (rc) return rc;
    if (t) *t = c.t;
    if (dt) *dt = finalize_host(st->P, c.dt_bits[c.step & 1]);
    if (steps) *steps = c.step;
    return TSU_OK;
}

TSU_API int tsu_set_time(tsu_state* st, double t, double dt)
{
    if (!st) return fail(TSU_ERR_ARG, "null state");
    DeviceGuard g(st->device);
    Ctrl c;
    int rc = read_ctrl(st, &c);
    if (rc) return rc;
    c.t = t;
    if (dt > 0) {
        c.dt_bits[c.step & 1] = dbits(dt);
        c.dt_bits[(c.step + 1) & 1] = kInfBits;
    }
    HIPCHK(hipMemcpyAsync(st->ctrl, &c, sizeof(Ctrl), hipMemcpyHostToDevice, st->stream));
    HIPCHK(hipStreamSynchronize(st->stream));
    return TSU_OK;
}

TSU_API int tsu_set_stop_time(tsu_state* st, double stop_time)
{
    if (!st) return fail(TSU_ERR_ARG, "null state");
    DeviceGuard g(st->device);
    static_assert(offsetof(Ctrl, stop_time) == 8, "ctrl layout");
    HIPCHK(hipMemcpyAsync((char*)st->ctrl + offsetof(Ctrl, stop_time), &stop_time, sizeof(double),
                          hipMemcpyHostToDevice, st->stream));
    HIPCHK(hipStreamSynchronize(st->stream));
    return TSU_OK;
}

TSU_API int tsu_set_clamp(tsu_state* st, int on)
{
    if (!st) return fail(TSU_ERR_ARG, "null state");
    DeviceGuard g(st->device);
    const long long v = on ? 1 : 0;
    HIPCHK(hipMemcpyAsync((char*)st->ctrl + offsetof(Ctrl, clamp), &v, sizeof(long long), hipMemcpyHostToDevice,
                          st->stream));
    HIPCHK(hipStreamSynchronize(st->stream));
    return TSU_OK;
}

TSU_API int tsu_mass(tsu_state* st, double out[2])
{
    if (!st || !out) return fail(TSU_ERR_ARG, "null argument");
    DeviceGuard g(st->device);
    void* cur;
    int rc = current_buffer(st, &cur);
    if (rc) return rc;
    double* res = st->mass_buf + 2 * st->mass_nb;
    HIPCHK(tsu::launch_mass(cur, st->nvol, st->precision, st->mass_buf, st->mass_nb, res, st->stream));
    HIPCHK(hipMemcpyAsync(out, res, 2 * sizeof(double), hipMemcpyDeviceToHost, st->stream));
    HIPCHK(hipStreamSynchronize(st->stream));
    return TSU_OK;
}

TSU_API int tsu_get_extrema(tsu_state* st, double* hmax, double* bmax, double* eta_max)
{
    if (!st) return fail(TSU_ERR_ARG, "null state");
    if (!st->ext) return fail(TSU_ERR_ARG, "extrema are not enabled in the scheme");
    DeviceGuard g(st->device);
    const size_t b = (size_t)st->nvol * sizeof(double);
    if (hmax) HIPCHK(hipMemcpyAsync(hmax, st->ext, b, hipMemcpyDeviceToHost, st->stream));
    if (bmax) HIPCHK(hipMemcpyAsync(bmax, st->ext + st->nvol, b, hipMemcpyDeviceToHost, st->stream));
    if (eta_max) HIPCHK(hipMemcpyAsync(eta_max, st->ext + 2 * st->nvol, b, hipMemcpyDeviceToHost, st->stream));
    HIPCHK(hipStreamSynchronize(st->stream));
    return TSU_OK;
}

TSU_API int tsu_set_tile_rows(tsu_state* st, int rows)
{
    if (!st) return fail(TSU_ERR_ARG, "null state");
    if (rows < 0 || rows > 4096) return fail(TSU_ERR_ARG, "tile rows out of range");
    drop_graph(st);
    set_tiling(st, rows == 0 ? default_tile_rows(st) : rows);
    return TSU_OK;
}

TSU_API int tsu_get_tile_rows(tsu_state* st) { return st ? st->tile_rows : -1; }

TSU_API int tsu_set_graph_chunk(tsu_state* st, int chunk)
{
    if (!st) return fail(TSU_ERR_ARG, "null state");
    if (chunk < 0) return fail(TSU_ERR_ARG, "negative graph chunk");
    if (st->s.order == 1 && (chunk & 1)) return fail(TSU_ERR_ARG, "single-stage graphs need an even chunk");
    drop_graph(st);
    st->graph_chunk = chunk;
    return TSU_OK;
}

TSU_API int tsu_set_timing(tsu_state* st, int on)
{
    if (!st) return fail(TSU_ERR_ARG, "null state");
    st->timing = on != 0;
    return TSU_OK;
}

TSU_API int tsu_get_timing(tsu_state* st, double stage_ms[2], long long launches[2], int reset)
{
    if (!st) return fail(TSU_ERR_ARG, "null state");
    DeviceGuard g(st->device);
    HIPCHK(hipStreamSynchronize(st->stream));
    for (size_t i = 0; i < st->ev_used; i++) {
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, st->ev[2 * i], st->ev[2 * i + 1]));
        st->stage_ms[st->ev_stage[i]] += ms;
        st->launches[st->ev_stage[i]]++;
    }
    st->ev_used = 0;
    for (int k = 0; k < 2; k++) {
        if (stage_ms) stage_ms[k] = st->stage_ms[k];
        if (launches) launches[k] = st->launches[k];
        if (reset) {
            st->stage_ms[k] = 0.0;
            st->launches[k] = 0;
        }
    }
    return TSU_OK;
}

TSU_API void* tsu_stream(tsu_state* st) { return st ? (void*)st->stream : nullptr; }

TSU_API void* tsu_device_state(tsu_state* st) { return st ? st->A : nullptr; }

TSU_API int tsu_bandwidth_probe(int device, long long bytes, int reps, double* gbs)
{
    if (bytes < 16 || reps < 1 || !gbs) return fail(TSU_ERR_ARG, "bad argument");
    bytes &= ~15LL;
    DeviceGuard g(device);
    void *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&a, (size_t)bytes);
    if (e == hipSuccess) e = hipMalloc(&b, (size_t)bytes);
    if (e == hipSuccess) e = hipMemset(a, 0, (size_t)bytes);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = tsu::launch_copy(a, b, bytes, nullptr);   // warm-up
    if (e == hipSuccess) e = hipEventRecord(e0, nullptr);
    for (int i = 0; i < reps && e == hipSuccess; i++) e = tsu::launch_copy(a, b, bytes, nullptr);
    if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    if (e != hipSuccess) return fail(TSU_ERR_HIP, std::string("bandwidth probe: ") + hipGetErrorString(e));
    *gbs = 2.0 * (double)bytes * reps / (ms * 1e-3) / 1e9;
    return TSU_OK;
}

namespace {

// Shared host side of the edge known-answer entries: pairs = false solves
// every edge alone (edge_flux), true solves edges 2k and 2k + 1 together
// (edge_flux_pair, n even).
int kat_edges(const tsu_sim_params* p, const tsu_scheme* s, int n, const double* left, const double* right,
              const int* normal, const int* fixed, double* Dm, double* Dp, int device, bool pairs)
{
    if (!p || !s || n < 0 || (n > 0 && (!left || !right || !normal || !fixed || !Dm || !Dp)))
        return fail(TSU_ERR_ARG, "bad argument");
    if (pairs && n % 2 != 0) return fail(TSU_ERR_ARG, "edge pairs need an even number of edges");
    if (n == 0) return TSU_OK;
    DevParams P;
    tsu::make_dev_params(p, s, &P);
    std::vector<int> kind((size_t)n);
    for (int i = 0; i < n; i++) {
        int nx = normal[2 * i], ny = normal[2 * i + 1];
        if (nx == 1 && ny == 0) kind[i] = tsu::N_PX;
        else if (nx == -1 && ny == 0) kind[i] = tsu::N_MX;
        else if (nx == 0 && ny == 1) kind[i] = tsu::N_PY;
        else if (nx == 0 && ny == -1) kind[i] = tsu::N_MY;
        else return fail(TSU_ERR_ARG, "normals must be axis-aligned unit vectors");
    }
    DeviceGuard g(device);
    double *dl = nullptr, *dr = nullptr, *dm = nullptr, *dp = nullptr;
    int *dk = nullptr, *df = nullptr;
    const size_t b5 = 5 * (size_t)n * sizeof(double), bi = (size_t)n * sizeof(int);
    hipError_t e = hipMalloc((void**)&dl, b5);
    if (e == hipSuccess) e = hipMalloc((void**)&dr, b5);
    if (e == hipSuccess) e = hipMalloc((void**)&dm, b5);
    if (e == hipSuccess) e = hipMalloc((void**)&dp, b5);
    if (e == hipSuccess) e = hipMalloc((void**)&dk, bi);
    if (e == hipSuccess) e = hipMalloc((void**)&df, bi);
    if (e == hipSuccess) e = hipMemcpy(dl, left, b5, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dr, right, b5, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dk, kind.data(), bi, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(df, fixed, bi, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = pairs ? tsu::launch_kat_edge_pairs(P, n / 2, dl, dr, dk, df, dm, dp, nullptr)
                  : tsu::launch_kat_edge(P, n, dl, dr, dk, df, dm, dp, nullptr);
    if (e == hipSuccess) e = hipMemcpy(Dm, dm, b5, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(Dp, dp, b5, hipMemcpyDeviceToHost);
    (void)hipFree(dl);
    (void)hipFree(dr);
    (void)hipFree(dm);
    (void)hipFree(dp);
    (void)hipFree(dk);
    (void)hipFree(df);
    if (e != hipSuccess) return fail(TSU_ERR_HIP, std::string("kat edge: ") + hipGetErrorString(e));
    return TSU_OK;
}

}  // namespace

TSU_API int tsu_kat_edge_flux(const tsu_sim_params* p, const tsu_scheme* s, int n, const double* left,
                              const double* right, const int* normal, const int* fixed, double* Dm, double* Dp,
                              int device)
{
    return kat_edges(p, s, n, left, right, normal, fixed, Dm, Dp, device, false);
}

TSU_API int tsu_kat_edge_flux_pairs(const tsu_sim_params* p, const tsu_scheme* s, int n, const double* left,
                                    const double* right, const int* normal, const int* fixed, double* Dm,
                                    double* Dp, int device)
{
    return kat_edges(p, s, n, left, right, normal, fixed, Dm, Dp, device, true);
}

TSU_API int tsu_kat_point(const tsu_sim_params* p, const tsu_scheme* s, int what, int n, const double* a,
                          const double* b, const double* c, const double* d, double* o1, double* o2, double* o3,
                          double* o4, int device)
{
    if (!p || !s || n < 0 || what < 0 || what > 7) return fail(TSU_ERR_ARG, "bad argument");
    if (n == 0) return TSU_OK;
    DevParams P;
    tsu::make_dev_params(p, s, &P);
    DeviceGuard g(device);
    const size_t bytes = (size_t)n * sizeof(double);
    const double* in[4] = {a, b, c, d};
    double* outs[4] = {o1, o2, o3, o4};
    double* din[4] = {nullptr, nullptr, nullptr, nullptr};
    double* dout[4] = {nullptr, nullptr, nullptr, nullptr};
    hipError_t e = hipSuccess;
    for (int i = 0; i < 4 && e == hipSuccess; i++) {
        e = hipMalloc((void**)&din[i], bytes);
        if (e == hipSuccess && in[i]) e = hipMemcpy(din[i], in[i], bytes, hipMemcpyHostToDevice);
        if (e == hipSuccess && !in[i]) e = hipMemset(din[i], 0, bytes);
        if (e == hipSuccess) e = hipMalloc((void**)&dout[i], bytes);
    }
    if (e == hipSuccess)
        e = tsu::launch_kat_point(what, s->limiter, n, din[0], din[1], din[2], din[3], dout[0], dout[1], dout[2],
                                  dout[3], P, nullptr);
    for (int i = 0; i < 4 && e == hipSuccess; i++)
        if (outs[i]) e = hipMemcpy(outs[i], dout[i], bytes, hipMemcpyDeviceToHost);
    for (int i = 0; i < 4; i++) {
        (void)hipFree(din[i]);
        (void)hipFree(dout[i]);
    }
    if (e != hipSuccess) return fail(TSU_ERR_HIP, std::string("kat point: ") + hipGetErrorString(e));
    return TSU_OK;
}

}  // extern "C"
