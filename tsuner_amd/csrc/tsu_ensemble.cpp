// tsu_ensemble.cpp -- many independent scenarios on grids of uneven size,
// advanced together: one stage launch covers every tile of every member
// (stage_kernel_batch), each member keeps its own buffers, run constants and
// step-control block (so its own dt and time), and whole-ensemble steps are
// replayed from a captured HIP graph.  The reference runs one scenario per
// process (GPU/main_TSUNER.cu:21-222); this is the batched form of that loop.
#define TSU_BUILDING_LIBRARY 1
#include "../../include/tsuner.h"

#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "tsu_host.h"
#include "tsu_kernels.h"
#include "tsu_nn.h"

using tsu::Ctrl;
using tsu::DevParams;

namespace {

constexpr unsigned long long kInfBits = 0x7FF0000000000000ull;

#define ECHK(expr)                                                                                    \
    do {                                                                                              \
        hipError_t e_ = (expr);                                                                       \
        if (e_ != hipSuccess)                                                                         \
            return tsu::set_error(TSU_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));    \
    } while (0)

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

struct tsu_ensemble {
    int device = 0, precision = TSU_FP64, n = 0;
    bool sed = false;
    std::vector<tsu_sim_params> p;
    std::vector<tsu_scheme> s;
    std::vector<DevParams> P;
    std::vector<size_t> off;        // member offset in the state pools (bytes)
    std::vector<size_t> foff;       // member offset in the mask pool
    std::vector<size_t> xoff;       // member offset in the extrema pool (doubles), if extrema on
    std::vector<int> rows;          // tile rows per member
    std::vector<bool> uploaded;
    std::vector<double*> cc, tcc;   // per-member inflow series (device), or null
    size_t esize = 8, planes = 5, pool_bytes = 0, fix_bytes = 0;
    char* A = nullptr;
    char* B = nullptr;
    char* snapA = nullptr;          // snapshot for tuning
    uint8_t* fix = nullptr;
    double* ext = nullptr;          // hmax, max H, eta_max planes per member with extrema on
    size_t ext_doubles = 0;
    Ctrl* ctrl = nullptr;
    Ctrl* snapC = nullptr;
    DevParams* dP = nullptr;
    tsu::StageArgs* dargs = nullptr;    // [2][n]: stage 0 and stage 1 arguments (hooks)
    tsu::BatchMember* dmem = nullptr;   // [n]: members as the batched stage kernel sees them
    int2* tiles = nullptr;
    int ntiles = 0;
    double* mass_buf = nullptr;
    int mass_nb = 0;
    hipStream_t stream = nullptr;
    int graph_chunk = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
};

namespace {

struct Guard {
    int prev = -1;
    explicit Guard(int dev)
    {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        (void)hipSetDevice(dev);
    }
    ~Guard()
    {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

void drop_graph(tsu_ensemble* e)
{
    if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    e->gexec = nullptr;
    e->graph = nullptr;
}

void destroy(tsu_ensemble* e)
{
    if (!e) return;
    Guard g(e->device);
    drop_graph(e);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    (void)hipFree(e->A);
    (void)hipFree(e->B);
    (void)hipFree(e->snapA);
    (void)hipFree(e->fix);
    (void)hipFree(e->ext);
    (void)hipFree(e->ctrl);
    (void)hipFree(e->snapC);
    (void)hipFree(e->dP);
    (void)hipFree(e->dargs);
    (void)hipFree(e->dmem);
    (void)hipFree(e->tiles);
    (void)hipFree(e->mass_buf);
    for (double* p : e->cc) (void)hipFree(p);
    for (double* p : e->tcc) (void)hipFree(p);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

int nstrips(const tsu_ensemble* e, int i) { return (e->p[i].npx + tsu::kStripOut - 1) / tsu::kStripOut; }
int nchunks(const tsu_ensemble* e, int i) { return (e->p[i].npy + e->rows[i] - 1) / e->rows[i]; }

// Per-member stage arguments and the tile table (member, tile) for the
// current tile heights.
int build_launch_tables(tsu_ensemble* e)
{
    std::vector<tsu::StageArgs> args(2 * (size_t)e->n);
    std::vector<tsu::BatchMember> mem((size_t)e->n);
    std::vector<int2> tiles;
    for (int i = 0; i < e->n; i++) {
        const size_t plane = (size_t)e->P[i].nvol * e->esize;
        tsu::BatchMember& bm = mem[(size_t)i];
        bm.off = (long long)e->off[i];
        bm.plane = (long long)plane;
        bm.foff = (long long)e->foff[i];
        bm.xoff = e->s[i].extrema ? (long long)e->xoff[i] : -1;
        bm.tile_rows = e->rows[i];
        bm.nstrips = nstrips(e, i);
        bm.ntiles = bm.nstrips * nchunks(e, i);
        bm.order = e->s[i].order;
        for (int k = 0; k < 2; k++) {
            tsu::StageArgs a{};
            char* Ai = e->A + e->off[i];
            char* Bi = e->B + e->off[i];
            a.Hin = Ai + 3 * plane;
            a.fix = e->fix + e->foff[i];
            a.ctrl = e->ctrl + i;
            a.ext = e->s[i].extrema ? e->ext + e->xoff[i] : nullptr;
            a.tile_rows = e->rows[i];
            a.nstrips = nstrips(e, i);
            a.ntiles = a.nstrips * nchunks(e, i);
            a.P = e->dP + i;
            if (e->s[i].order == 1) {
                a.in = Ai;
                a.out = Bi;
                a.v0 = Ai;
                a.pingpong = 1;
            } else if (k == 0) {
                a.in = Ai;
                a.out = Bi;
                a.v0 = Ai;
            } else {
                a.in = Bi;
                a.out = Ai;
                a.v0 = Ai;
            }
            args[(size_t)k * e->n + i] = a;
        }
        const int nt = args[i].ntiles;
        for (int t = 0; t < nt; t++) tiles.push_back(make_int2(i, t));
    }
    drop_graph(e);
    (void)hipFree(e->tiles);
    e->tiles = nullptr;
    e->ntiles = (int)tiles.size();
    ECHK(hipMalloc((void**)&e->tiles, tiles.size() * sizeof(int2)));
    ECHK(hipMemcpy(e->tiles, tiles.data(), tiles.size() * sizeof(int2), hipMemcpyHostToDevice));
    ECHK(hipMemcpy(e->dargs, args.data(), args.size() * sizeof(tsu::StageArgs), hipMemcpyHostToDevice));
    ECHK(hipMemcpy(e->dmem, mem.data(), mem.size() * sizeof(tsu::BatchMember), hipMemcpyHostToDevice));
    return TSU_OK;
}

int enqueue_step(tsu_ensemble* e)
{
    const int order = e->s[0].order;
    const tsu::BatchPools pools{e->A, e->B, e->fix, e->ctrl, e->ext, e->dP};
    bool fast_bed = true;
    for (int i = 0; i < e->n; i++) fast_bed = fast_bed && tsu::fast_bedload(e->P[i]);
    for (int k = 0; k < order; k++)
        ECHK(tsu::launch_stage_batch(pools, e->dmem, e->tiles, e->ntiles, k, k == order - 1, e->sed, e->precision,
                                     e->stream, fast_bed));
    ECHK(tsu::launch_advance_batch(e->ctrl, e->dP, e->n, e->stream));
    return TSU_OK;
}

int member_ok(const tsu_ensemble* e, int i)
{
    if (!e) return tsu::set_error(TSU_ERR_ARG, "null ensemble");
    if (i < 0 || i >= e->n) return tsu::set_error(TSU_ERR_ARG, "member index out of range");
    return TSU_OK;
}

// Every member uploaded, and every member with an inflow side given its
// series.
int members_ready(const tsu_ensemble* e)
{
    for (int i = 0; i < e->n; i++) {
        if (!e->uploaded[i]) return tsu::set_error(TSU_ERR_ARG, "member " + std::to_string(i) + " not uploaded");
        for (int sd = 0; sd < 4; sd++)
            if (e->s[i].bc[sd] == TSU_BC_INFLOW && !e->cc[i])
                return tsu::set_error(TSU_ERR_ARG, "member " + std::to_string(i) +
                                                       " has an inflow side: tsu_ensemble_set_inflow first");
    }
    return TSU_OK;
}

int current_buffer(tsu_ensemble* e, int i, char** buf)
{
    if (e->s[i].order == 2) {
        *buf = e->A + e->off[i];
        return TSU_OK;
    }
    Ctrl c;
    ECHK(hipMemcpyAsync(&c, e->ctrl + i, sizeof(Ctrl), hipMemcpyDeviceToHost, e->stream));
    ECHK(hipStreamSynchronize(e->stream));
    *buf = ((c.step & 1) ? e->B : e->A) + e->off[i];
    return TSU_OK;
}

}  // namespace

extern "C" {

TSU_API int tsu_ensemble_create(tsu_ensemble** out, int n, const tsu_sim_params* params, const tsu_scheme* scheme,
                                int precision, int device)
{
    if (!out || !params || !scheme || n < 1) return tsu::set_error(TSU_ERR_ARG, "bad argument");
    *out = nullptr;
    if (precision != TSU_FP64 && precision != TSU_FP32) return tsu::set_error(TSU_ERR_ARG, "bad precision");
    for (int i = 0; i < n; i++) {
        int rc = tsu::validate_scheme(&scheme[i]);
        if (rc) return rc;
        if (params[i].npx < 1 || params[i].npy < 1 || !(params[i].dx > 0) || !(params[i].dy > 0))
            return tsu::set_error(TSU_ERR_ARG, "bad grid of member " + std::to_string(i));
        if (scheme[i].order != scheme[0].order) return tsu::set_error(TSU_ERR_UNSUPPORTED, "members must share order");
    }
    tsu_ensemble* e = new tsu_ensemble();
    e->device = device;
    e->precision = precision;
    e->n = n;
    e->esize = precision == TSU_FP64 ? 8 : 4;
    e->planes = precision == TSU_FP64 ? 5 : 8;
    e->p.assign(params, params + n);
    e->s.assign(scheme, scheme + n);
    e->P.resize(n);
    e->off.resize(n);
    e->foff.resize(n);
    e->xoff.assign(n, 0);
    e->rows.assign(n, 16);
    e->uploaded.assign(n, false);
    e->cc.assign(n, nullptr);
    e->tcc.assign(n, nullptr);
    long long maxnv = 0;
    for (int i = 0; i < n; i++) {
        e->p[i].nvol = e->p[i].npx * e->p[i].npy;
        tsu::make_dev_params(&e->p[i], &e->s[i], &e->P[i]);
        e->off[i] = e->pool_bytes;
        e->pool_bytes += align256(e->planes * (size_t)e->P[i].nvol * e->esize);
        e->foff[i] = e->fix_bytes;
        e->fix_bytes += align256((size_t)e->P[i].nvol);
        if (e->P[i].nvol > maxnv) maxnv = e->P[i].nvol;
        e->sed = e->sed || scheme[i].bedload || scheme[i].suspension;
        if (scheme[i].extrema) {
            e->xoff[i] = e->ext_doubles;
            e->ext_doubles += 3 * (size_t)((e->P[i].nvol + 31) / 32 * 32);
        }
    }
    e->mass_nb = tsu::mass_blocks(maxnv);
    Guard g(device);
    hipError_t err = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
    if (err == hipSuccess) err = hipMalloc((void**)&e->A, e->pool_bytes);
    if (err == hipSuccess) err = hipMalloc((void**)&e->B, e->pool_bytes);
    if (err == hipSuccess) err = hipMalloc((void**)&e->fix, e->fix_bytes);
    if (err == hipSuccess && e->ext_doubles) err = hipMalloc((void**)&e->ext, e->ext_doubles * sizeof(double));
    if (err == hipSuccess) err = hipMalloc((void**)&e->ctrl, n * sizeof(Ctrl));
    if (err == hipSuccess) err = hipMalloc((void**)&e->dP, n * sizeof(DevParams));
    if (err == hipSuccess) err = hipMalloc((void**)&e->dargs, 2 * (size_t)n * sizeof(tsu::StageArgs));
    if (err == hipSuccess) err = hipMalloc((void**)&e->dmem, (size_t)n * sizeof(tsu::BatchMember));
    if (err == hipSuccess) err = hipMalloc((void**)&e->mass_buf, (2 * (size_t)e->mass_nb + 2) * sizeof(double));
    if (err == hipSuccess) err = hipMemcpy(e->dP, e->P.data(), n * sizeof(DevParams), hipMemcpyHostToDevice);
    if (err != hipSuccess) {
        destroy(e);
        (void)hipGetLastError();
        return tsu::set_error(TSU_ERR_DEVICE_MEMORY, std::string("ensemble allocation: ") + hipGetErrorString(err));
    }
    std::vector<Ctrl> c(n);
    for (int i = 0; i < n; i++) {
        std::memset(&c[i], 0, sizeof(Ctrl));
        c[i].stop_time = INFINITY;
        c[i].stop_step = LLONG_MAX;
        c[i].dt_bits[0] = c[i].dt_bits[1] = kInfBits;
    }
    if (hipMemcpy(e->ctrl, c.data(), n * sizeof(Ctrl), hipMemcpyHostToDevice) != hipSuccess) {
        destroy(e);
        return tsu::set_error(TSU_ERR_HIP, "ensemble ctrl init failed");
    }
    int rc = build_launch_tables(e);
    if (rc) {
        destroy(e);
        return rc;
    }
    *out = e;
    return TSU_OK;
}

TSU_API int tsu_ensemble_destroy(tsu_ensemble* e)
{
    destroy(e);
    return TSU_OK;
}

TSU_API int tsu_ensemble_upload(tsu_ensemble* e, int i, const double* V, const int* fixH)
{
    int rc = member_ok(e, i);
    if (rc) return rc;
    if (!V || !fixH) return tsu::set_error(TSU_ERR_ARG, "null argument");
    const long long nv = e->P[i].nvol;
    std::vector<uint8_t> fx((size_t)nv);
    for (long long k = 0; k < nv; k++) {
        if (fixH[k] != 0 && fixH[k] != 1) return tsu::set_error(TSU_ERR_ARG, "fixH values must be 0 or 1");
        fx[(size_t)k] = (uint8_t)fixH[k];
    }
    bool hc_zero = true;
    for (long long k = 0; k < nv && hc_zero; k++) hc_zero = V[4 * nv + k] == 0.0;
    if (!hc_zero && !e->sed) return tsu::set_error(TSU_ERR_UNSUPPORTED, "nonzero hc needs a sediment scheme");
    Guard g(e->device);
    drop_graph(e);
    const size_t bytes = e->planes * (size_t)nv * e->esize;
    if (e->precision == TSU_FP64) {
        ECHK(hipMemcpyAsync(e->A + e->off[i], V, bytes, hipMemcpyHostToDevice, e->stream));
        ECHK(hipStreamSynchronize(e->stream));
    } else {
        std::vector<float> f((size_t)8 * nv);
        for (size_t k = 0; k < (size_t)5 * nv; k++) f[k] = (float)V[k];
        const int src[3] = {0, 3, 4};
        for (int j = 0; j < 3; j++)
            for (size_t k = 0; k < (size_t)nv; k++) {
                const size_t q = (size_t)src[j] * nv + k;
                f[(size_t)(5 + j) * nv + k] = (float)(V[q] - (double)f[q]);
            }
        ECHK(hipMemcpyAsync(e->A + e->off[i], f.data(), bytes, hipMemcpyHostToDevice, e->stream));
        ECHK(hipStreamSynchronize(e->stream));
    }
    ECHK(hipMemcpyAsync(e->B + e->off[i], e->A + e->off[i], bytes, hipMemcpyDeviceToDevice, e->stream));
    ECHK(hipMemcpyAsync(e->fix + e->foff[i], fx.data(), (size_t)nv, hipMemcpyHostToDevice, e->stream));
    if (e->s[i].extrema) ECHK(tsu::launch_fill(e->ext + e->xoff[i], 3 * nv, -1e30, e->stream));
    Ctrl c;
    std::memset(&c, 0, sizeof(c));
    c.stop_time = INFINITY;
    c.stop_step = LLONG_MAX;
    c.dt_bits[0] = c.dt_bits[1] = kInfBits;
    ECHK(hipMemcpyAsync(e->ctrl + i, &c, sizeof(Ctrl), hipMemcpyHostToDevice, e->stream));
    ECHK(hipStreamSynchronize(e->stream));
    e->uploaded[i] = true;
    return TSU_OK;
}

TSU_API int tsu_ensemble_download(tsu_ensemble* e, int i, double* V)
{
    int rc = member_ok(e, i);
    if (rc) return rc;
    if (!V) return tsu::set_error(TSU_ERR_ARG, "null argument");
    Guard g(e->device);
    char* cur;
    if ((rc = current_buffer(e, i, &cur))) return rc;
    const long long nv = e->P[i].nvol;
    const size_t bytes = e->planes * (size_t)nv * e->esize;
    if (e->precision == TSU_FP64) {
        ECHK(hipMemcpyAsync(V, cur, bytes, hipMemcpyDeviceToHost, e->stream));
        ECHK(hipStreamSynchronize(e->stream));
    } else {
        std::vector<float> f((size_t)8 * nv);
        ECHK(hipMemcpyAsync(f.data(), cur, bytes, hipMemcpyDeviceToHost, e->stream));
        ECHK(hipStreamSynchronize(e->stream));
        for (size_t k = 0; k < (size_t)5 * nv; k++) V[k] = (double)f[k];
        const int dst[3] = {0, 3, 4};
        for (int j = 0; j < 3; j++)
            for (size_t k = 0; k < (size_t)nv; k++) V[(size_t)dst[j] * nv + k] += (double)f[(size_t)(5 + j) * nv + k];
    }
    return TSU_OK;
}

TSU_API int tsu_ensemble_set_inflow(tsu_ensemble* e, int i, const double* CC, const double* tCC, unsigned long timeCC)
{
    int rc = member_ok(e, i);
    if (rc) return rc;
    if (!CC || !tCC) return tsu::set_error(TSU_ERR_ARG, "null argument");
    if (timeCC < 2) return tsu::set_error(TSU_ERR_ARG, "the inflow series needs at least two time levels");
    for (unsigned long k = 0; k < timeCC; k++)
        if (!std::isfinite(tCC[k]) || (k > 0 && !(tCC[k] > tCC[k - 1])))
            return tsu::set_error(TSU_ERR_ARG, "inflow times must be finite and strictly increasing");
    Guard g(e->device);
    // steps already enqueued may still read the member's old series
    ECHK(hipStreamSynchronize(e->stream));
    drop_graph(e);
    (void)hipFree(e->cc[i]);
    (void)hipFree(e->tcc[i]);
    e->cc[i] = e->tcc[i] = nullptr;
    tsu::DevParams& P = e->P[i];
    const size_t ncc = (size_t)P.ring * 4 * timeCC;
    hipError_t err = hipMalloc((void**)&e->cc[i], ncc * sizeof(double));
    if (err == hipSuccess) err = hipMalloc((void**)&e->tcc[i], timeCC * sizeof(double));
    if (err != hipSuccess) {
        (void)hipFree(e->cc[i]);
        e->cc[i] = nullptr;
        return tsu::set_error(TSU_ERR_DEVICE_MEMORY, "member inflow series allocation failed");
    }
    ECHK(hipMemcpy(e->cc[i], CC, ncc * sizeof(double), hipMemcpyHostToDevice));
    ECHK(hipMemcpy(e->tcc[i], tCC, timeCC * sizeof(double), hipMemcpyHostToDevice));
    P.CC = e->cc[i];
    P.tCC = e->tcc[i];
    P.timeCC = (long long)timeCC;
    P.tcc0 = tCC[0];
    P.dtcc = tCC[1] - tCC[0];
    P.tcc_end = tCC[timeCC - 2];
    ECHK(hipMemcpy(e->dP + i, &P, sizeof(tsu::DevParams), hipMemcpyHostToDevice));
    return TSU_OK;
}

TSU_API int tsu_ensemble_init_dt(tsu_ensemble* e)
{
    if (!e) return tsu::set_error(TSU_ERR_ARG, "null ensemble");
    Guard g(e->device);
    if (int rc = members_ready(e)) return rc;
    for (int i = 0; i < e->n; i++) {
        Ctrl c;
        ECHK(hipMemcpyAsync(&c, e->ctrl + i, sizeof(Ctrl), hipMemcpyDeviceToHost, e->stream));
        ECHK(hipStreamSynchronize(e->stream));
        c.dt_bits[0] = c.dt_bits[1] = kInfBits;
        ECHK(hipMemcpyAsync(e->ctrl + i, &c, sizeof(Ctrl), hipMemcpyHostToDevice, e->stream));
        char* cur = ((e->s[i].order == 1 && (c.step & 1)) ? e->B : e->A) + e->off[i];
        ECHK(tsu::launch_init_dt(cur, e->P[i], e->ctrl + i, e->precision, e->stream));
    }
    ECHK(hipStreamSynchronize(e->stream));
    return TSU_OK;
}

TSU_API int tsu_ensemble_set_stop_step(tsu_ensemble* e, long long stop_step)
{
    if (!e) return tsu::set_error(TSU_ERR_ARG, "null ensemble");
    Guard g(e->device);
    for (int i = 0; i < e->n; i++)
        ECHK(hipMemcpyAsync((char*)(e->ctrl + i) + offsetof(Ctrl, stop_step), &stop_step, sizeof(long long),
                            hipMemcpyHostToDevice, e->stream));
    ECHK(hipStreamSynchronize(e->stream));
    return TSU_OK;
}

TSU_API int tsu_ensemble_step(tsu_ensemble* e, int nsteps)
{
    if (!e) return tsu::set_error(TSU_ERR_ARG, "null ensemble");
    if (nsteps < 0) return tsu::set_error(TSU_ERR_ARG, "negative step count");
    if (int rc = members_ready(e)) return rc;
    Guard g(e->device);
    int done = 0, rc;
    if (e->graph_chunk > 0 && nsteps >= e->graph_chunk) {
        if (!e->gexec) {
            ECHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
            for (int k = 0; k < e->graph_chunk; k++) {
                if ((rc = enqueue_step(e))) {
                    hipGraph_t gph = nullptr;
                    (void)hipStreamEndCapture(e->stream, &gph);
                    if (gph) (void)hipGraphDestroy(gph);
                    return rc;
                }
            }
            ECHK(hipStreamEndCapture(e->stream, &e->graph));
            ECHK(hipGraphInstantiate(&e->gexec, e->graph, nullptr, nullptr, 0));
        }
        while (nsteps - done >= e->graph_chunk) {
            ECHK(hipGraphLaunch(e->gexec, e->stream));
            done += e->graph_chunk;
        }
    }
    for (; done < nsteps; done++)
        if ((rc = enqueue_step(e))) return rc;
    return TSU_OK;
}

TSU_API int tsu_ensemble_synchronize(tsu_ensemble* e)
{
    if (!e) return tsu::set_error(TSU_ERR_ARG, "null ensemble");
    Guard g(e->device);
    ECHK(hipStreamSynchronize(e->stream));
    return TSU_OK;
}

TSU_API int tsu_ensemble_get_time(tsu_ensemble* e, int i, double* t, double* dt, long long* steps)
{
    int rc = member_ok(e, i);
    if (rc) return rc;
    Guard g(e->device);
    Ctrl c;
    ECHK(hipMemcpyAsync(&c, e->ctrl + i, sizeof(Ctrl), hipMemcpyDeviceToHost, e->stream));
    ECHK(hipStreamSynchronize(e->stream));
    if (t) *t = c.t;
    if (dt) {
        double m;
        unsigned long long b = c.dt_bits[c.step & 1];
        std::memcpy(&m, &b, 8);
        if (e->P[i].dt_narrow && e->P[i].nvol > 4096) m = (double)(float)m;
        *dt = (m < 1e30) ? m : 1e30;
    }
    if (steps) *steps = c.step;
    return TSU_OK;
}

TSU_API int tsu_ensemble_mass(tsu_ensemble* e, int i, double out[2])
{
    int rc = member_ok(e, i);
    if (rc) return rc;
    if (!out) return tsu::set_error(TSU_ERR_ARG, "null argument");
    Guard g(e->device);
    char* cur;
    if ((rc = current_buffer(e, i, &cur))) return rc;
    double* res = e->mass_buf + 2 * e->mass_nb;
    ECHK(tsu::launch_mass(cur, e->P[i].nvol, e->precision, e->mass_buf, e->mass_nb, res, e->stream));
    ECHK(hipMemcpyAsync(out, res, 2 * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    ECHK(hipStreamSynchronize(e->stream));
    return TSU_OK;
}

TSU_API int tsu_ensemble_set_tile_rows(tsu_ensemble* e, const int* rows)
{
    if (!e || !rows) return tsu::set_error(TSU_ERR_ARG, "null argument");
    for (int i = 0; i < e->n; i++)
        if (rows[i] < 1 || rows[i] > 4096) return tsu::set_error(TSU_ERR_ARG, "tile rows out of range");
    e->rows.assign(rows, rows + e->n);
    Guard g(e->device);
    return build_launch_tables(e);
}

TSU_API int tsu_ensemble_set_graph_chunk(tsu_ensemble* e, int chunk)
{
    if (!e || chunk < 0) return tsu::set_error(TSU_ERR_ARG, "bad argument");
    if (e->s[0].order == 1 && (chunk & 1)) return tsu::set_error(TSU_ERR_ARG, "single-stage graphs need an even chunk");
    Guard g(e->device);
    drop_graph(e);
    e->graph_chunk = chunk;
    return TSU_OK;
}

TSU_API int tsu_ensemble_snapshot(tsu_ensemble* e, int restore)
{
    if (!e) return tsu::set_error(TSU_ERR_ARG, "null ensemble");
    if (e->s[0].order != 2) return tsu::set_error(TSU_ERR_UNSUPPORTED, "snapshots need the two-stage scheme");
    Guard g(e->device);
    if (!e->snapA) {
        if (restore) return tsu::set_error(TSU_ERR_ARG, "no snapshot taken");
        ECHK(hipMalloc((void**)&e->snapA, e->pool_bytes));
        ECHK(hipMalloc((void**)&e->snapC, e->n * sizeof(Ctrl)));
    }
    if (restore) {
        ECHK(hipMemcpyAsync(e->A, e->snapA, e->pool_bytes, hipMemcpyDeviceToDevice, e->stream));
        ECHK(hipMemcpyAsync(e->B, e->snapA, e->pool_bytes, hipMemcpyDeviceToDevice, e->stream));
        ECHK(hipMemcpyAsync(e->ctrl, e->snapC, e->n * sizeof(Ctrl), hipMemcpyDeviceToDevice, e->stream));
    } else {
        ECHK(hipMemcpyAsync(e->snapA, e->A, e->pool_bytes, hipMemcpyDeviceToDevice, e->stream));
        ECHK(hipMemcpyAsync(e->snapC, e->ctrl, e->n * sizeof(Ctrl), hipMemcpyDeviceToDevice, e->stream));
    }
    ECHK(hipStreamSynchronize(e->stream));
    return TSU_OK;
}

TSU_API int tsu_ensemble_time_steps(tsu_ensemble* e, int nsteps, double* ms)
{
    if (!e || !ms || nsteps < 1) return tsu::set_error(TSU_ERR_ARG, "bad argument");
    Guard g(e->device);
    hipEvent_t a = nullptr, b = nullptr;
    ECHK(hipEventCreate(&a));
    ECHK(hipEventCreate(&b));
    int rc = TSU_OK;
    hipError_t er = hipEventRecord(a, e->stream);
    for (int k = 0; k < nsteps && er == hipSuccess && rc == TSU_OK; k++) rc = enqueue_step(e);
    if (er == hipSuccess) er = hipEventRecord(b, e->stream);
    if (er == hipSuccess) er = hipEventSynchronize(b);
    float f = 0.0f;
    if (er == hipSuccess) er = hipEventElapsedTime(&f, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (rc) return rc;
    if (er != hipSuccess) return tsu::set_error(TSU_ERR_HIP, std::string("timing: ") + hipGetErrorString(er));
    *ms = f;
    return TSU_OK;
}

TSU_API int tsu_ensemble_set_stop_time(tsu_ensemble* e, const double* stop_time, int clamp)
{
    if (!e || !stop_time) return tsu::set_error(TSU_ERR_ARG, "null argument");
    Guard g(e->device);
    const long long cl = clamp ? 1 : 0;
    for (int i = 0; i < e->n; i++) {
        ECHK(hipMemcpyAsync((char*)(e->ctrl + i) + offsetof(Ctrl, stop_time), stop_time + i, sizeof(double),
                            hipMemcpyHostToDevice, e->stream));
        ECHK(hipMemcpyAsync((char*)(e->ctrl + i) + offsetof(Ctrl, clamp), &cl, sizeof(long long),
                            hipMemcpyHostToDevice, e->stream));
    }
    ECHK(hipStreamSynchronize(e->stream));
    return TSU_OK;
}

TSU_API int tsu_ensemble_get_extrema(tsu_ensemble* e, int i, double* hmax, double* bmax, double* eta_max)
{
    int rc = member_ok(e, i);
    if (rc) return rc;
    if (!e->s[i].extrema) return tsu::set_error(TSU_ERR_ARG, "extrema are not enabled for this member");
    Guard g(e->device);
    const size_t nv = (size_t)e->P[i].nvol, b = nv * sizeof(double);
    const double* x = e->ext + e->xoff[i];
    if (hmax) ECHK(hipMemcpyAsync(hmax, x, b, hipMemcpyDeviceToHost, e->stream));
    if (bmax) ECHK(hipMemcpyAsync(bmax, x + nv, b, hipMemcpyDeviceToHost, e->stream));
    if (eta_max) ECHK(hipMemcpyAsync(eta_max, x + 2 * nv, b, hipMemcpyDeviceToHost, e->stream));
    ECHK(hipStreamSynchronize(e->stream));
    return TSU_OK;
}

TSU_API void* tsu_ensemble_stream(tsu_ensemble* e) { return e ? (void*)e->stream : nullptr; }

TSU_API int tsu_ensemble_get_times(tsu_ensemble* e, double* t, long long* steps)
{
    if (!e || (!t && !steps)) return tsu::set_error(TSU_ERR_ARG, "bad argument");
    Guard g(e->device);
    std::vector<Ctrl> c((size_t)e->n);
    ECHK(hipMemcpyAsync(c.data(), e->ctrl, e->n * sizeof(Ctrl), hipMemcpyDeviceToHost, e->stream));
    ECHK(hipStreamSynchronize(e->stream));
    for (int i = 0; i < e->n; i++) {
        if (t) t[i] = c[(size_t)i].t;
        if (steps) steps[i] = c[(size_t)i].step;
    }
    return TSU_OK;
}

}  // extern "C"

// ---- hooks of the learned coarse-grid correction ---------------------------

namespace {

// Members of one grid and a two-stage scheme: the stage-2 arguments' .out is
// the current state.
int nn_uniform(const tsu_ensemble* e, const char* what)
{
    if (!e) return tsu::set_error(TSU_ERR_ARG, std::string(what) + ": null ensemble");
    if (e->s[0].order != 2) return tsu::set_error(TSU_ERR_UNSUPPORTED, std::string(what) + ": needs the two-stage scheme");
    for (int i = 1; i < e->n; i++)
        if (e->p[i].npx != e->p[0].npx || e->p[i].npy != e->p[0].npy)
            return tsu::set_error(TSU_ERR_UNSUPPORTED, std::string(what) + ": members must share one grid");
    for (int i = 0; i < e->n; i++)
        if (!e->uploaded[i]) return tsu::set_error(TSU_ERR_ARG, std::string(what) + ": member not uploaded");
    return TSU_OK;
}

tsu::NnScales make_scales(const float* in8, const float* out4, float wet_h)
{
    tsu::NnScales s{};
    for (int j = 0; j < 8; j++) s.in[j] = in8 ? in8[j] : 0.f;
    for (int j = 0; j < 4; j++) s.out[j] = out4 ? out4[j] : 1.f;
    s.wet_h = wet_h;
    return s;
}

int nn_pair(const tsu_ensemble* f, const tsu_ensemble* c, const char* what)
{
    int rc = nn_uniform(f, what);
    if (rc) return rc;
    if ((rc = nn_uniform(c, what))) return rc;
    if (f->n != c->n || f->precision != c->precision)
        return tsu::set_error(TSU_ERR_ARG, std::string(what) + ": ensembles differ in size or precision");
    if (f->p[0].npx != 4 * c->p[0].npx || f->p[0].npy != 4 * c->p[0].npy)
        return tsu::set_error(TSU_ERR_ARG, std::string(what) + ": fine grid must be 4x the coarse grid");
    if (f->device != c->device) return tsu::set_error(TSU_ERR_ARG, std::string(what) + ": ensembles on different devices");
    return TSU_OK;
}

}  // namespace

extern "C" {

TSU_API int tsu_ensemble_nn_features(tsu_ensemble* e, const float scale_in[8], float wet_h, void* x)
{
    int rc = nn_uniform(e, "tsu_ensemble_nn_features");
    if (rc) return rc;
    if (!scale_in || !x) return tsu::set_error(TSU_ERR_ARG, "null argument");
    Guard g(e->device);
    ECHK(tsu::nn_launch_features(e->dargs + e->n, e->n, e->p[0].npx, e->p[0].npy, e->precision,
                                 make_scales(scale_in, nullptr, wet_h), (uint16_t*)x, e->stream));
    return TSU_OK;
}

TSU_API int tsu_ensemble_nn_apply(tsu_ensemble* e, const float scale_out[4], float wet_h, const float* out,
                                  long long expect_step, float* resid_max, const void* x, float shore_h)
{
    int rc = nn_uniform(e, "tsu_ensemble_nn_apply");
    if (rc) return rc;
    if (!scale_out || !out) return tsu::set_error(TSU_ERR_ARG, "null argument");
    Guard g(e->device);
    tsu::NnScales sc = make_scales(nullptr, scale_out, wet_h);
    sc.shore_h = shore_h;
    ECHK(tsu::nn_launch_apply(e->dargs + e->n, e->n, e->p[0].npx, e->p[0].npy, e->precision, sc, out, expect_step,
                              resid_max, (const uint16_t*)x, e->stream));
    return TSU_OK;
}

TSU_API int tsu_ensemble_restrict(tsu_ensemble* fine, tsu_ensemble* coarse)
{
    int rc = nn_pair(fine, coarse, "tsu_ensemble_restrict");
    if (rc) return rc;
    Guard g(coarse->device);
    ECHK(hipStreamSynchronize(fine->stream));
    ECHK(tsu::nn_launch_restrict(fine->dargs + fine->n, coarse->dargs + coarse->n, coarse->n, coarse->p[0].npx,
                                 coarse->p[0].npy, coarse->precision, coarse->stream));
    // keep the stage-1 buffer equal to the state, as after an upload
    for (int i = 0; i < coarse->n; i++)
        ECHK(hipMemcpyAsync(coarse->B + coarse->off[i], coarse->A + coarse->off[i],
                            coarse->planes * (size_t)coarse->P[i].nvol * coarse->esize, hipMemcpyDeviceToDevice,
                            coarse->stream));
    ECHK(hipStreamSynchronize(coarse->stream));
    return tsu_ensemble_init_dt(coarse);
}

TSU_API int tsu_ensemble_nn_target(tsu_ensemble* fine, tsu_ensemble* coarse, const float scale_out[4], float* target)
{
    int rc = nn_pair(fine, coarse, "tsu_ensemble_nn_target");
    if (rc) return rc;
    if (!scale_out || !target) return tsu::set_error(TSU_ERR_ARG, "null argument");
    Guard g(coarse->device);
    ECHK(hipStreamSynchronize(fine->stream));
    ECHK(tsu::nn_launch_target(fine->dargs + fine->n, coarse->dargs + coarse->n, coarse->n, coarse->p[0].npx,
                               coarse->p[0].npy, coarse->precision, make_scales(nullptr, scale_out, 0.f), target,
                               coarse->stream));
    return TSU_OK;
}

}  // extern "C"
