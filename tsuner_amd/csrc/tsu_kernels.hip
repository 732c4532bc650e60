// tsu_kernels.hip -- gfx950 kernels of the TSUNER per-step solver.
//
// One launch of stage_kernel performs a whole RK stage of the reference
// pipeline (reconstructor -> VolumeSolver -> states_updater,
// GPU/main_TSUNER.cu:170-183) for a tile of cells, with the reference's
// face arrays reconsL/R/D/U and edge slots FL/FR/FD/FU kept in registers:
//
//   * a 64-lane wave owns 64 consecutive columns (lanes 2..61 are outputs,
//     lanes 0,1,62,63 are halo) and marches over tile_rows rows;
//   * x neighbours are exchanged with lane shuffles, y neighbours stay in a
//     three-row register window, so each cell's velocity, each face and each
//     edge flux is evaluated exactly once per stage;
//   * the last stage of a step also computes the per-cell CFL dt
//     (Volume_kernel.cu:19-43), reduces it over the wave with shuffles and
//     folds it into the device-resident step control block with one 64-bit
//     atomicMin per wave (Reduction_kernel_min.cu:155-196), so dt never
//     leaves the device.
//
// Two arithmetic variants share the code: R = double (reference arithmetic,
// FP64 planes) and R = float (FP32 planes and fluxes; h, H and hc carried as
// hi + lo float pairs updated with TwoSum, the lo planes stored after the
// five state planes).
//
// Workgroups are dealt round-robin over the 8 XCDs; tile ids are remapped so
// every XCD receives a contiguous band of tiles (neighbouring tiles share
// halo rows and partial cache lines through that XCD's L2).
#include <hip/hip_runtime.h>

#include "tsu_device.h"
#include "tsu_kernels.h"

namespace tsu {

namespace {

constexpr unsigned long long kInfBits = 0x7FF0000000000000ull;

__device__ __forceinline__ int map_index(int i, int n, bool periodic)
{
    if (periodic) {
        i %= n;
        return i < 0 ? i + n : i;
    }
    return min(max(i, 0), n - 1);
}

__device__ __forceinline__ int xcd_tile(int b, int nb)
{
    const int q = nb >> 3, r = nb & 7, x = b & 7;
    return x * q + min(x, r) + (b >> 3);
}

// Whole-wave shifts by one lane as DPP moves (wave_shr:1 gives lane i the
// value of lane i - 1, wave_shl:1 that of lane i + 1): one VALU move per
// dword instead of an LDS permute and its wait.  Lanes shifted in from
// outside the wave read 0; they are halo lanes, whose results are not stored.
__device__ __forceinline__ int dpp_up1(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xF, 0xF, true); }
__device__ __forceinline__ int dpp_dn1(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x130, 0xF, 0xF, true); }
__device__ __forceinline__ double up1(double v)
{
    const long long x = __double_as_longlong(v);
    const int lo = dpp_up1((int)x), hi = dpp_up1((int)(x >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double dn1(double v)
{
    const long long x = __double_as_longlong(v);
    const int lo = dpp_dn1((int)x), hi = dpp_dn1((int)(x >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ float up1(float v) { return __int_as_float(dpp_up1(__float_as_int(v))); }
__device__ __forceinline__ float dn1(float v) { return __int_as_float(dpp_dn1(__float_as_int(v))); }
__device__ __forceinline__ int up1i(int v) { return dpp_up1(v); }

__device__ __forceinline__ double wave_min(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}
template <typename R>
__device__ __forceinline__ R wave_max(R v)
{
    for (int o = 32; o > 0; o >>= 1) v = rmax(v, __shfl_xor(v, o, 64));
    return v;
}

// Knuth's TwoSum: s + e == a + b exactly.
__device__ __forceinline__ void two_sum(float a, float b, float& s, float& e)
{
    s = a + b;
    float bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

// Compensated value hi + lo.
struct Pair {
    float hi, lo;
};

// Adds x into p exactly (the rounding error joins lo).
__device__ __forceinline__ void pacc(Pair& p, float x)
{
    float s, e;
    two_sum(p.hi, x, s, e);
    p.hi = s;
    p.lo += e;
}

__device__ __forceinline__ Pair pnorm(Pair p)
{
    Pair r;
    two_sum(p.hi, p.lo, r.hi, r.lo);
    return r;
}

// (a + b) / 2 in pair arithmetic (Volume_kernel.cu:244-253 with coef 1/2).
__device__ __forceinline__ Pair pavg(Pair a, Pair b)
{
    Pair r;
    two_sum(a.hi, b.hi, r.hi, r.lo);
    r.lo += a.lo + b.lo;
    r = pnorm(r);
    r.hi *= 0.5f;
    r.lo *= 0.5f;
    return r;
}

template <typename R>
struct Cell {
    R h, qx, qy, H, hc;   // state (hi part in the FP32 solver)
    R ux, uy, us, c;      // velocity, |u|^2 and concentration
};

template <typename R, typename S, bool SED>
__device__ __forceinline__ void load_cell(const S* in, const S* Hin, long long nv, long long off, Cell<R>& v)
{
    v.h = (R)in[off];
    v.qx = (R)in[nv + off];
    v.qy = (R)in[2 * nv + off];
    v.H = (R)Hin[off];
    v.hc = SED ? (R)in[4 * nv + off] : (R)0;
}

template <typename R, bool SED>
__device__ __forceinline__ void derive_cell(Cell<R>& v)
{
    Vel<R> u;
    velocity_conc<SED>(v.h, v.qx, v.qy, v.hc, u, v.c);
    v.ux = u.x;
    v.uy = u.y;
    v.us = u.s;
}

// Reconstructor.cu:146-194 along one axis: faces on the minus side (Fm) and
// plus side (Fp) of the centre cell from its minus (l) and plus (r)
// neighbours; returns the limited free-surface slope used by the
// non-conservative source (Reconstructor.cu:197,274).
template <typename R, bool SED>
__device__ __forceinline__ R recons_axis(int lim, const Cell<R>& ci, R hl, R Hl, R cl, R ulx, R uly, R hr, R Hr, R cr,
                                         R urx, R ury, R* Fm, R* Fp)
{
    const R half = (R)0.5, zero = (R)0;
    const R hi = ci.h, Hi = ci.H;
    R dl = hi - hl, dr = hr - hi;
    R dh = limiter(lim, dl, dr);
    if (half * rabs(dh) > hi) dh = (R)2.0 * hi * (R)sgn(dh);
    Fm[0] = hi - half * dh;
    Fp[0] = hi + half * dh;

    R zm = rmin(Hl, Hi);
    const R dle = rmax(hi - Hi + zm, zero) - rmax(hl - Hl + zm, zero);
    zm = rmin(Hr, Hi);
    const R dre = rmax(hr - Hr + zm, zero) - rmax(hi - Hi + zm, zero);
    const R dlu = ci.ux - ulx, dru = urx - ci.ux;
    const R dlv = ci.uy - uly, drv = ury - ci.uy;
    const R dlc = SED ? ci.c - cl : zero, drc = SED ? cr - ci.c : zero;
    R deta, dux, duy, dc = zero;
    if (lim == LIM_MINMOD) {
        deta = lim_minmod(dle, dre);
        dux = lim_minmod(dlu, dru);
        duy = lim_minmod(dlv, drv);
        if (SED) dc = lim_minmod(dlc, drc);
    } else {
        // Free-surface, velocity and concentration slopes all vanish in
        // still water: one wave-wide vote skips their quotients when no lane
        // takes one (the limited slopes are then the limiter's zeros).
        const bool ue = vl_use(dle, dre), uu = vl_use(dlu, dru), uv = vl_use(dlv, drv);
        const bool uc = SED && vl_use(dlc, drc);
        deta = dux = duy = zero;
        if (__builtin_amdgcn_ballot_w64(ue || uu || uv || uc) != 0) {
            deta = vl_val(dle, dre, ue);
            dux = vl_val(dlu, dru, uu);
            duy = vl_val(dlv, drv, uv);
            if (SED) dc = vl_val(dlc, drc, uc);
        }
    }
    R dz = dh - deta;
    Fm[3] = Hi - half * dz;
    Fp[3] = Hi + half * dz;

    R dq = rmad(hi, dux, ci.ux * dh);
    Fm[1] = rmad(-half, dq, ci.qx);
    Fp[1] = rmad(half, dq, ci.qx);

    dq = rmad(hi, duy, ci.uy * dh);
    Fm[2] = rmad(-half, dq, ci.qy);
    Fp[2] = rmad(half, dq, ci.qy);

    if (SED) {
        Fm[4] = rmad(-half, dc, ci.c) * Fm[0];
        Fp[4] = rmad(half, dc, ci.c) * Fp[0];
    } else {
        Fm[4] = Fp[4] = zero;
    }
    return deta;
}

// Faces after the state filter (SedFlux.cuh:118-132).  FP64 keeps them as
// (h, qx, qy, H, hc) = (h, h u, h u, H, h c); the FP32 solver keeps the
// filter's (h, ux, uy, H, c), which is what the Riemann solver reads.
template <typename R, bool SED>
__device__ __forceinline__ void filter_face(R* F)
{
    if constexpr (sizeof(R) == 4) {
        const float h = (F[0] > 0.f) ? F[0] : 0.f;
        Vel<float> u;
        float c;
        velocity_conc<SED>(h, F[1], F[2], F[4], u, c);
        F[0] = h;
        F[1] = u.x;
        F[2] = u.y;
        F[4] = (SED && c > 0.f) ? c : 0.f;
    } else {
        state_filter<R, SED>(F[0], F[1], F[2], F[4]);
    }
}

// The two faces of a cell filtered together (values of filter_face on each);
// in FP64 their velocity quotients share one wave-wide test and basic block.
template <typename R, bool SED>
__device__ __forceinline__ void filter_faces(R* A, R* B)
{
    if constexpr (sizeof(R) == 8) {
        A[0] = (A[0] > 0.0) ? A[0] : 0.0;
        B[0] = (B[0] > 0.0) ? B[0] : 0.0;
        Vel<double> ua, ub;
        double ca, cb;
        velocity_conc_pair<SED>(A[0], A[1], A[2], A[4], B[0], B[1], B[2], B[4], ua, ca, ub, cb);
        A[1] = A[0] * ua.x;
        A[2] = A[0] * ua.y;
        B[1] = B[0] * ub.x;
        B[2] = B[0] * ub.y;
        if (SED) {
            ca = (ca > 0.0) ? ca : 0.0;
            cb = (cb > 0.0) ? cb : 0.0;
            A[4] = A[0] * ca;
            B[4] = B[0] * cb;
        }
    } else {
        filter_face<R, SED>(A);
        filter_face<R, SED>(B);
    }
}

// First-order faces: the cell state, as filtered faces are kept.
template <typename R>
__device__ __forceinline__ void cell_face(const Cell<R>& c, R* F)
{
    F[0] = c.h;
    F[1] = (sizeof(R) == 4) ? c.ux : c.qx;
    F[2] = (sizeof(R) == 4) ? c.uy : c.qy;
    F[3] = c.H;
    F[4] = (sizeof(R) == 4) ? c.c : c.hc;
}

// Riemann problem on one edge with left/right faces in the unrotated frame;
// Dm, Dp come back in the unrotated frame (VolumeSolver.cu:19-170).
template <typename R, bool SED, bool FB = false>
__device__ __forceinline__ void edge_flux(const DevParams& P, int kind, bool mobile, const R* Lf, const R* Rf, R* Dm,
                                          R* Dp)
{
    R qnl, qtl, qnr, qtr, dm[5], dp[5];
    to_normal(kind, Lf[1], Lf[2], qnl, qtl);
    to_normal(kind, Rf[1], Rf[2], qnr, qtr);
    riemann<R, SED, FB>(P, mobile, Lf[0], Lf[3], qnl, qtl, Lf[4], Rf[0], Rf[3], qnr, qtr, Rf[4], dm, dp);
    Dm[0] = dm[0];
    Dp[0] = dp[0];
    from_normal(kind, dm[1], dm[2], Dm[1], Dm[2]);
    from_normal(kind, dp[1], dp[2], Dp[1], Dp[2]);
    if (SED) {
        Dm[3] = dm[3];
        Dm[4] = dm[4];
        Dp[3] = dp[3];
        Dp[4] = dp[4];
    } else {
        Dm[3] = Dm[4] = Dp[3] = Dp[4] = (R)0;
    }
}

// Two edges' Riemann problems side by side (riemann_pair), each edge as
// edge_flux: edge a (kind ka, bed mobility ma, faces La/Ra, fluctuations
// Dma/Dpa) and edge b likewise.
template <typename R, bool SED, bool FB>
__device__ __forceinline__ void edge_flux_pair(const DevParams& P, int ka, bool ma, const R* La, const R* Ra, R* Dma,
                                               R* Dpa, int kb, bool mb, const R* Lb, const R* Rb, R* Dmb, R* Dpb)
{
    const int kind[2] = {ka, kb};
    const bool mobile[2] = {ma, mb};
    const R* const L[2] = {La, Lb};
    const R* const Rt[2] = {Ra, Rb};
    R* const Dm[2] = {Dma, Dmb};
    R* const Dp[2] = {Dpa, Dpb};
    R side[2][10], dm[2][5], dp[2][5];
    for (int i = 0; i < 2; i++) {
        side[i][0] = L[i][0];
        side[i][1] = L[i][3];
        to_normal(kind[i], L[i][1], L[i][2], side[i][2], side[i][3]);
        side[i][4] = L[i][4];
        side[i][5] = Rt[i][0];
        side[i][6] = Rt[i][3];
        to_normal(kind[i], Rt[i][1], Rt[i][2], side[i][7], side[i][8]);
        side[i][9] = Rt[i][4];
    }
    riemann_pair<R, SED, FB>(P, mobile, side, dm, dp);
    for (int i = 0; i < 2; i++) {
        Dm[i][0] = dm[i][0];
        Dp[i][0] = dp[i][0];
        from_normal(kind[i], dm[i][1], dm[i][2], Dm[i][1], Dm[i][2]);
        from_normal(kind[i], dp[i][1], dp[i][2], Dp[i][1], Dp[i][2]);
        for (int f = 3; f < 5; f++) {
            Dm[i][f] = SED ? dm[i][f] : (R)0;
            Dp[i][f] = SED ? dp[i][f] : (R)0;
        }
    }
}

// FB: the run constants satisfy fast_bedload (kernels with sediment only).
template <typename R, typename S, bool SED, bool FB>
__device__ __forceinline__ void stage_body(const StageArgs& a, const int tile)
{
    constexpr bool COMP = sizeof(R) == 4;   // FP32 solver: compensated h, H, hc
    // Run constants, copied into LDS once per workgroup (one wave).  Read
    // through the global pointer, every use in the row loop was a vector
    // load, since the kernel's own stores might alias the constants; in LDS
    // they cannot, and a use is a short LDS read.  The loop re-reads them
    // per row (see there) rather than holding them in registers.
    __shared__ DevParams Ps;
    static_assert(sizeof(DevParams) % 8 == 0 && sizeof(DevParams) <= 64 * 8, "one 8-byte word per lane");
    if (threadIdx.x < sizeof(DevParams) / 8)
        reinterpret_cast<unsigned long long*>(&Ps)[threadIdx.x] =
            reinterpret_cast<const unsigned long long*>(a.P)[threadIdx.x];
    __syncthreads();
    const DevParams& P = Ps;
    const int lane = threadIdx.x;
    const int strip = tile % a.nstrips, chunk = tile / a.nstrips;

    Ctrl* C = a.ctrl;
    const double ct = C->t;
    if (!(ct < C->stop_time) || C->step >= C->stop_step) return;
    const long long step = C->step;
    const double dtd = clamp_dt(C, ct, finalize_dt(P, C->dt_bits[step & 1]));

    const int npx = P.npx, npy = P.npy;
    const long long nv = P.nvol;
    const S* in = (const S*)a.in;
    S* out = (S*)a.out;
    if (a.pingpong && (step & 1)) {
        in = (const S*)a.out;
        out = (S*)a.in;
    }
    const S* Hin = SED ? in + 3 * nv : (const S*)a.Hin;
    const S* v0p = (a.k == 0) ? in : (const S*)a.v0;
    const uint8_t* fix = a.fix;

    const bool perx = P.bc[0] == BC_PERIODIC, pery = P.bc[2] == BC_PERIODIC;
    const int x = strip * kStripOut + lane - 2;
    const int xc = map_index(x, npx, perx);
    const int y0 = chunk * a.tile_rows;
    const int y1 = min(y0 + a.tile_rows, npy);
    const bool out_lane = lane >= 2 && lane < 62 && x < npx;
    const int lim = P.limiter;
    const bool second = P.order == 2;
    const bool need_fix = SED && P.bedload;

    // Per-step constants, in the reference's evaluation order.
    const double invxd = dtd * P.inv_dx, invyd = dtd * P.inv_dy;  // main_TSUNER.cu:182
    const R dt = (R)dtd;
    const R invx = (R)invxd, invy = (R)invyd;
    const R dtaxG = (R)(invxd * kG), dtayG = (R)(invyd * kG);   // Reconstructor.cu:197,274
    const R dtFR = (R)(dtd * P.FR);                               // Reconstructor.cu:45

    auto row_off = [&](int r) -> long long { return (long long)map_index(r, npy, pery) * npx + xc; };
    auto fix_at = [&](int r) -> int { return need_fix ? (int)fix[row_off(r)] : 0; };

    // ---- prologue: faces of row y0 and the edge y0 - 1/2 -----------------
    Cell<R> Rm, R0, R1;
    load_cell<R, S, SED>(in, Hin, nv, row_off(y0 - 1), Rm);
    load_cell<R, S, SED>(in, Hin, nv, row_off(y0), R0);
    load_cell<R, S, SED>(in, Hin, nv, row_off(y0 + 1), R1);
    derive_cell<R, SED>(Rm);
    derive_cell<R, SED>(R0);
    derive_cell<R, SED>(R1);
    int fx0 = fix_at(y0), fx1;

    R D0[5], U0[5], FD0[5];
    R detay0 = (R)0;
    if (second) {
        detay0 = recons_axis<R, SED>(lim, R0, Rm.h, Rm.H, Rm.c, Rm.ux, Rm.uy, R1.h, R1.H, R1.c, R1.ux, R1.uy, D0, U0);
        filter_faces<R, SED>(D0, U0);
    } else {
        cell_face(R0, D0);
        cell_face(R0, U0);
    }
    if (y0 == 0 && !pery) {
        R G5[5], dp[5];
        ghost_state(P, 2, D0, G5, (long long)xc, 2, ct);
        edge_flux<R, SED, FB>(P, N_MY, fx0 != 0, D0, G5, FD0, dp);
    } else {
        R Um[5], Dd[5], dm[5];
        if (second) {
            Cell<R> Rmm;
            load_cell<R, S, SED>(in, Hin, nv, row_off(y0 - 2), Rmm);
            derive_cell<R, SED>(Rmm);
            recons_axis<R, SED>(lim, Rm, Rmm.h, Rmm.H, Rmm.c, Rmm.ux, Rmm.uy, R0.h, R0.H, R0.c, R0.ux, R0.uy, Dd,
                                Um);
            filter_face<R, SED>(Um);
        } else {
            cell_face(Rm, Um);
        }
        int fxm = fix_at(y0 - 1);
        int fixed = (y0 == 0) ? fxm : fxm * fx0;   // periodic wrap edge: VolumeSolver.cu:386
        edge_flux<R, SED, FB>(P, N_PY, fixed != 0, Um, D0, dm, FD0);
    }

    // largest wave speeds of the tile's updated cells (last stage: the next dt)
    R lxmax = (R)0, lymax = (R)0;

    // ---- march over the rows of the tile -----------------------------------
    for (int r = y0; r < y1; ++r) {
        // the run constants' LDS address passed through an empty asm per
        // row, so that the compiler reads them where they are used instead
        // of hoisting all of them into registers ahead of the loop
        typedef const __attribute__((address_space(3))) DevParams* LdsParams;
        LdsParams Pl = (LdsParams)&Ps;
        asm volatile("" : "+s"(Pl));
        const DevParams& P = *(const DevParams*)Pl;
        // next row (r + 2): issued first, consumed after the x work below
        Cell<R> R2;
        load_cell<R, S, SED>(in, Hin, nv, row_off(r + 2), R2);
        fx1 = fix_at(r + 1);
        R v0[5];
        R lo_in[3] = {(R)0, (R)0, (R)0}, lo_v0[3] = {(R)0, (R)0, (R)0};
        const long long pos = (long long)r * npx + xc;
        if (out_lane) {
            if (a.k != 0) {
                v0[0] = (R)v0p[pos];
                v0[1] = (R)v0p[nv + pos];
                v0[2] = (R)v0p[2 * nv + pos];
                v0[3] = SED ? (R)v0p[3 * nv + pos] : R0.H;
                v0[4] = SED ? (R)v0p[4 * nv + pos] : (R)0;
            }
            if (COMP) {
                lo_in[0] = (R)in[5 * nv + pos];
                if (SED) {
                    lo_in[1] = (R)in[6 * nv + pos];
                    lo_in[2] = (R)in[7 * nv + pos];
                }
                if (a.k != 0) {
                    lo_v0[0] = (R)v0p[5 * nv + pos];
                    if (SED) {
                        lo_v0[1] = (R)v0p[6 * nv + pos];
                        lo_v0[2] = (R)v0p[7 * nv + pos];
                    }
                }
            }
        }

        // x faces of row r (Reconstructor.cu:120-197)
        R hl = up1(R0.h), Hl = up1(R0.H), ulx = up1(R0.ux), uly = up1(R0.uy);
        R hr = dn1(R0.h), Hr = dn1(R0.H), urx = dn1(R0.ux), ury = dn1(R0.uy);
        R cl = (R)0, cr = (R)0;
        if (SED) {
            cl = up1(R0.c);
            cr = dn1(R0.c);
        }
        R Lf[5], Rf[5], detax = (R)0;
        if (second) {
            detax = recons_axis<R, SED>(lim, R0, hl, Hl, cl, ulx, uly, hr, Hr, cr, urx, ury, Lf, Rf);
            filter_faces<R, SED>(Lf, Rf);
        } else {
            cell_face(R0, Lf);
            cell_face(R0, Rf);
        }

        // x edge at x - 1/2 (VolumeSolver.cu:227-317)
        R Rl[5];
        for (int f = 0; f < 5; f++) Rl[f] = up1(Rf[f]);
        int fxl = need_fix ? up1i(fx0) : 0;
        R El[5], Er[5];
        int kind = N_PX;
        int fixed;
        if (x == 0 && !perx) {
            for (int f = 0; f < 5; f++) El[f] = Lf[f];
            ghost_state(P, 0, Lf, Er, 2LL * npx + npy + r, 1, ct);
            kind = N_MX;
            fixed = fx0;
        } else if (x == npx && !perx) {
            for (int f = 0; f < 5; f++) El[f] = Rl[f];
            ghost_state(P, 1, Rl, Er, (long long)npx + r, 1, ct);
            fixed = fxl;
        } else {
            for (int f = 0; f < 5; f++) {
                El[f] = Rl[f];
                Er[f] = Lf[f];
            }
            fixed = (perx && (x == 0 || x == npx)) ? fxl : fxl * fx0;
        }

        // y faces of row r + 1 and the edge r + 1/2
        derive_cell<R, SED>(R2);
        R D1[5], U1[5], detay1 = (R)0;
        if (second) {
            detay1 = recons_axis<R, SED>(lim, R1, R0.h, R0.H, R0.c, R0.ux, R0.uy, R2.h, R2.H, R2.c, R2.ux, R2.uy, D1,
                                         U1);
            filter_faces<R, SED>(D1, U1);
        } else {
            cell_face(R1, D1);
            cell_face(R1, U1);
        }
        // upper side of the edge r + 1/2: row r + 1's lower faces, or the
        // ghost state above the domain's last row
        R Yr[5];
        if (r + 1 == npy && !pery)
            ghost_state(P, 3, U0, Yr, (long long)npx + npy + xc, 2, ct);
        else
            for (int f = 0; f < 5; f++) Yr[f] = D1[f];
        const int fy = (r + 1 == npy) ? fx0 : fx0 * fx1;

        // the x edge x - 1/2 and the y edge r + 1/2, solved side by side
        R Dmx[5], Dpx[5], FU0[5], FD1[5];   // FD1: not used after the domain's last row
        edge_flux_pair<R, SED, FB>(P, kind, fixed != 0, El, Er, Dmx, Dpx, N_PY, fy != 0, U0, Yr, FU0, FD1);
        R FL[5], FR[5];
        const bool left_wall = (x == 0 && !perx);
        for (int f = 0; f < 5; f++) {
            FL[f] = left_wall ? Dmx[f] : Dpx[f];
            FR[f] = dn1(Dmx[f]);
        }

        // cell update of row r (Reconstructor.cu:197,274,353-354;
        // Volume_kernel.cu:191-195,244-282)
        if (out_lane) {
            if (a.k == 0) {
                v0[0] = R0.h;
                v0[1] = R0.qx;
                v0[2] = R0.qy;
                v0[3] = R0.H;
                v0[4] = R0.hc;
                if (COMP)
                    for (int j = 0; j < 3; j++) lo_v0[j] = lo_in[j];
            }
            R k1 = R0.qx, k2 = R0.qy;
            if (second) {
                k1 -= dtaxG * R0.h * detax;
                k2 -= dtayG * R0.h * detay0;
            }
            Vel<R> un;
            // the quadratic law divides by h^n + DBL_EPSILON, the velocity's
            // denominator: one shared reciprocal, computed where a quotient
            // needs it (have_dh is uniform over the wave)
            Den<R> Dh = {(R)1, (R)1};
            bool have_dh = false;
            if constexpr (COMP) {
                Dh = den(v0[0] + (R)kEpsilon);
                have_dh = true;
            }
            if (a.k == 0) {   // U^n is this stage's input: its velocity is known
                un.x = R0.ux;
                un.y = R0.uy;
                un.s = R0.us;
            } else if constexpr (COMP) {
                un = velocity(v0[0], v0[1], v0[2], &Dh.y);
            } else if (__builtin_amdgcn_ballot_w64(v0[0] >= (R)kEps && (v0[1] != (R)0 || v0[2] != (R)0)) != 0) {
                Dh = den(v0[0] + (R)kEpsilon);
                have_dh = true;
                un = velocity64(v0[0], v0[1], v0[2], &Dh);
            } else {
                un = velocity64(v0[0], v0[1], v0[2]);   // no lane divides
            }
            // |u| of U^n (friction, exchange).  In FP64 a wave at rest (|u| =
            // 0 on every lane) has c = 0 and divides both momenta by 1 + 0 =
            // 1, which leaves them unchanged: it skips the friction step (the
            // FP32 solver, where the step is a few instructions, measured
            // slower with the vote).
            R modu = (R)0;
            const bool still = !COMP && __builtin_amdgcn_ballot_w64(un.s != (R)0) == 0;
            if (!still) {
                modu = sqrt_voted(un.s);
                if (!have_dh) Dh = den(v0[0] + (R)kEpsilon);
                R cf;
                if (P.friction_law == FRIC_MANNING) {
                    R hp = v0[0] + (R)kEpsilon;
                    R dtGnn = (R)(dtd * kG * P.manning_n * P.manning_n);
                    if constexpr (COMP)
                        cf = dtGnn * modu / (hp * cbrtf(hp));
                    else
                        cf = dtGnn * modu / (hp * cbrt(hp));
                } else {
                    cf = qdiv(dtFR * modu, Dh);
                }
                // Reconstructor.cu:52-53: both momenta divided by 1 + c
                const Den<R> Dc = den((R)1 + cf);
                k1 = qdiv(k1, Dc);
                k2 = qdiv(k2, Dc);
            }
            R phi = (R)0;
            if (SED && P.suspension) phi = erosion_rate(P, v0[0], modu, v0[4], still);

            R W[5];
            W[1] = k1 - (invx * (FL[1] + FR[1]) + invy * (FD0[1] + FU0[1]));
            W[2] = k2 - (invx * (FL[2] + FR[2]) + invy * (FD0[2] + FU0[2]));
            if constexpr (!COMP) {
                R k3 = R0.H, k4 = R0.hc;
                if (SED && P.suspension) {
                    k3 -= dt * phi;
                    k4 += dt * phi;
                }
                W[0] = R0.h - (invx * (FL[0] + FR[0]) + invy * (FD0[0] + FU0[0]));
                if (SED) {
                    W[3] = k3 - (invx * (FL[3] + FR[3]) + invy * (FD0[3] + FU0[3]));
                    W[4] = k4 - (invx * (FL[4] + FR[4]) + invy * (FD0[4] + FU0[4]));
                } else {
                    W[3] = R0.H;
                    W[4] = (R)0;
                }
                if (a.k == 1) {
                    const int nf = SED ? 5 : 3;
                    for (int f = 0; f < nf; f++) W[f] = 0.5 * v0[f] + 0.5 * W[f];
                }
                state_filter<R, SED>(W[0], W[1], W[2], W[4]);
                out[pos] = (S)W[0];
                out[nv + pos] = (S)W[1];
                out[2 * nv + pos] = (S)W[2];
                if (SED) {
                    out[3 * nv + pos] = (S)W[3];
                    out[4 * nv + pos] = (S)W[4];
                }
            } else {
                // Compensated accumulation in FP64 adds (issued at the FP32
                // rate): h, H and hc enter as the exact double hi + lo, each
                // flux term (a float product that leaves the neighbour across
                // the edge with the opposite sign, Dp = -Dm) is added with one
                // 53-bit rounding, and the result is stored back as a hi + lo
                // float pair (48 significant bits), so exchanges cancel to
                // FP64-level round-off.
                const int nm = SED ? 3 : 1;
                const int plane[3] = {0, 3, 4};
                double acc[3];
                for (int j = 0; j < nm; j++) {
                    const int f = plane[j];
                    const R hi = (j == 0) ? R0.h : (j == 1 ? R0.H : R0.hc);
                    double v = (double)hi + (double)lo_in[j];
                    v -= (double)(invx * FL[f]);
                    v -= (double)(invx * FR[f]);
                    v -= (double)(invy * FD0[f]);
                    v -= (double)(invy * FU0[f]);
                    acc[j] = v;
                }
                if (SED && P.suspension) {
                    const double ex = (double)(dt * phi);
                    acc[1] -= ex;
                    acc[2] += ex;
                }
                if (a.k == 1) {
                    for (int j = 0; j < nm; j++)
                        acc[j] = 0.5 * (((double)v0[plane[j]] + (double)lo_v0[j]) + acc[j]);
                    W[1] = (R)0.5 * v0[1] + (R)0.5 * W[1];
                    W[2] = (R)0.5 * v0[2] + (R)0.5 * W[2];
                }
                // state filter (SedFlux.cuh:118-132) on the accumulated values
                if (!(acc[0] > 0.0)) acc[0] = 0.0;
                Pair m[3];
                for (int j = 0; j < nm; j++) {
                    m[j].hi = (R)acc[j];
                    m[j].lo = (R)(acc[j] - (double)m[j].hi);
                }
                const R h = m[0].hi;
                if (h < kNearDry) {
                    R hh = h, hcv = SED ? m[2].hi : (R)0;
                    state_filter<R, SED>(hh, W[1], W[2], hcv);
                    if (SED && hcv == (R)0) m[2].hi = m[2].lo = (R)0;
                } else {
                    // wet: q = h * (q / h) is q to float rounding; only the cap acts
                    R rh = __builtin_amdgcn_rcpf(h);
                    R ux = W[1] * rh, uy = W[2] * rh;
                    R s = ux * ux + uy * uy;
                    if (s > (R)100.0f) {
                        R kk = (R)kVmax * __builtin_amdgcn_rsqf(s);
                        W[1] = h * (ux * kk);
                        W[2] = h * (uy * kk);
                    }
                }
                if (SED && m[2].hi < (R)0) m[2].hi = m[2].lo = (R)0;
                W[0] = m[0].hi;
                out[pos] = (S)m[0].hi;
                out[5 * nv + pos] = (S)m[0].lo;
                out[nv + pos] = (S)W[1];
                out[2 * nv + pos] = (S)W[2];
                if (SED) {
                    W[3] = m[1].hi;
                    W[4] = m[2].hi;
                    out[3 * nv + pos] = (S)m[1].hi;
                    out[6 * nv + pos] = (S)m[1].lo;
                    out[4 * nv + pos] = (S)m[2].hi;
                    out[7 * nv + pos] = (S)m[2].lo;
                } else {
                    W[3] = R0.H;
                    W[4] = (R)0;
                }
            }
            if (a.last) {
                // the next dt is evaluated on the stored state
                const R h = W[0];
                R lx, ly;
                cell_speeds(h, W[1], W[2], lx, ly);
                lxmax = rmax(lxmax, lx);
                lymax = rmax(lymax, ly);
                if (a.ext) {
                    double hd = (double)h, Hs = (double)W[3];
                    double* hmax = a.ext;
                    double* bmax = a.ext + nv;
                    double* emax = a.ext + 2 * nv;
                    if (hd > P.epsilon_h) {
                        hmax[pos] = fmax(hmax[pos], hd);
                        emax[pos] = fmax(emax[pos], hd - Hs);
                    }
                    bmax[pos] = fmax(bmax[pos], Hs);
                }
            }
        }

        // slide the window
        R0 = R1;
        R1 = R2;
        for (int f = 0; f < 5; f++) {
            U0[f] = U1[f];
            FD0[f] = FD1[f];
        }
        detay0 = detay1;
        fx0 = fx1;
    }

    if (a.last) {
        lxmax = wave_max(lxmax);
        lymax = wave_max(lymax);
        const double m = (double)dt_of_speeds(P, lxmax, lymax);
        if (lane == 0 && m < __longlong_as_double((long long)kInfBits))
            atomicMin(&C->dt_bits[(step + 1) & 1], (unsigned long long)__double_as_longlong(m));
    }
}

// Waves per SIMD the register allocation targets.  Measured on the 8192^2
// bench: FP32 at 3 waves (168 VGPRs, a few spilled) runs 24% faster than at
// 2; 4 waves spill 39 VGPRs and lose.  FP64 with sediment needs 245-255
// VGPRs; at 3 waves it spills 79 or more and runs at half speed, so it stays
// at 2.  FP64 without sediment fits 168 VGPRs, i.e. 3 waves, unspilled.
#ifndef TSU_WAVES_FP64
#define TSU_WAVES_FP64 2
#endif
#ifndef TSU_WAVES_FP64_NOSED
#define TSU_WAVES_FP64_NOSED 3
#endif
#ifndef TSU_WAVES_FP32
#define TSU_WAVES_FP32 3
#endif
template <typename R, bool SED>
constexpr int stage_waves()
{
    return sizeof(R) == 4 ? TSU_WAVES_FP32 : (SED ? TSU_WAVES_FP64 : TSU_WAVES_FP64_NOSED);
}

template <typename R, typename S, bool SED, bool FB = false>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(stage_waves<R, SED>(), stage_waves<R, SED>())))
void stage_kernel(StageArgs a)
{
    stage_body<R, S, SED, FB>(a, xcd_tile(blockIdx.x, gridDim.x));
}

// Ensemble: one launch covers every tile of every scenario.  tiles[b] =
// (scenario, tile within the scenario); scen[s] holds that scenario's
// buffers, constants and step-control block.
template <typename R, typename S, bool SED, bool FB = false>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(stage_waves<R, SED>(), stage_waves<R, SED>())))
void stage_kernel_batch(BatchPools pools, const BatchMember* __restrict__ mem, const int2* __restrict__ tiles, int k,
                        int last)
{
    const int2 e = tiles[xcd_tile(blockIdx.x, gridDim.x)];
    const BatchMember m = mem[e.x];
    char* Ai = pools.A + m.off;
    char* Bi = pools.B + m.off;
    StageArgs a;
    a.pingpong = 0;
    if (m.order == 1) {
        a.in = Ai;
        a.out = Bi;
        a.pingpong = 1;
    } else if (k == 0) {
        a.in = Ai;
        a.out = Bi;
    } else {
        a.in = Bi;
        a.out = Ai;
    }
    a.v0 = Ai;
    a.Hin = Ai + 3 * m.plane;
    a.fix = pools.fix + m.foff;
    a.ctrl = pools.ctrl + e.x;
    a.ext = m.xoff >= 0 ? pools.ext + m.xoff : nullptr;
    a.k = k;
    a.last = last;
    a.tile_rows = m.tile_rows;
    a.nstrips = m.nstrips;
    a.ntiles = m.ntiles;
    a.P = pools.P + e.x;
    stage_body<R, S, SED, FB>(a, e.y);
}

// GPU/main_TSUNER.cu:147-148: dt of the initial state into dt_bits[step & 1].
template <typename R, typename S>
__global__ __launch_bounds__(256) void init_dt_kernel(const S* V, DevParams P, Ctrl* C)
{
    const long long nv = P.nvol;
    double m = __longlong_as_double((long long)kInfBits);
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nv; i += (long long)gridDim.x * blockDim.x)
        m = fmin(m, (double)cell_dt(P, (R)V[i], (R)V[nv + i], (R)V[2 * nv + i]));
    m = wave_min(m);
    if ((threadIdx.x & 63) == 0 && m < __longlong_as_double((long long)kInfBits))
        atomicMin(&C->dt_bits[C->step & 1], (unsigned long long)__double_as_longlong(m));
}

// GPU/main_TSUNER.cu:185: t += dt after the stages; the dt slot of the
// finished step is re-armed for step + 2.
__device__ __forceinline__ void advance_one(Ctrl* C, const DevParams& P)
{
    const double t = C->t;
    const long long s = C->step;
    if (!(t < C->stop_time) || s >= C->stop_step) return;
    const double dt = finalize_dt(P, C->dt_bits[s & 1]);
    // a clamped step (see clamp_dt) ends exactly on stop_time
    C->t = (C->clamp && !(dt < C->stop_time - t)) ? C->stop_time : t + dt;
    C->dt_bits[s & 1] = kInfBits;
    C->step = s + 1;
}

__global__ void advance_kernel(Ctrl* C, DevParams P)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    advance_one(C, P);
}

// One thread per ensemble member.
__global__ void advance_batch_kernel(Ctrl* C, const DevParams* __restrict__ P, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) advance_one(C + i, P[i]);
}

__global__ void warmup_kernel() {}

// Plain streaming copy (16 B per lane per access), the reference point for
// achievable HBM bandwidth and for calibrating byte counters.
// Plain streaming copy of n 16-byte words, the achievable-bandwidth
// reference: each thread keeps four independent 16-byte loads in flight
// (a workgroup moves 16 KiB per pass, coalesced 4 KiB at a time), and
// loads and stores carry the nontemporal hint (the data is touched once).
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void copy_kernel(const u32x4* __restrict__ src, u32x4* __restrict__ dst,
                                                   long long n)
{
    constexpr int kU = 4;
    const long long stride = (long long)gridDim.x * blockDim.x * kU;
    long long i = (long long)blockIdx.x * blockDim.x * kU + threadIdx.x;
    for (; i + (kU - 1) * (long long)blockDim.x < n; i += stride) {
        u32x4 v[kU];
#pragma unroll
        for (int u = 0; u < kU; u++) v[u] = __builtin_nontemporal_load(src + i + u * (long long)blockDim.x);
#pragma unroll
        for (int u = 0; u < kU; u++) __builtin_nontemporal_store(v[u], dst + i + u * (long long)blockDim.x);
    }
    for (; i < n; i += blockDim.x) dst[i] = src[i];   // the last partial pass
}

__global__ void fill_kernel(double* p, long long n, double v)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

// Deterministic two-pass FP64 sums of h and of (H + hc); the FP32 solver
// adds the lo planes (planes 5, 6, 7).
template <typename S, bool COMP>
__global__ __launch_bounds__(256) void mass_partial_kernel(const S* V, long long nv, double* partial)
{
    __shared__ double sh[2][4];
    double a = 0.0, b = 0.0;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nv; i += (long long)gridDim.x * blockDim.x) {
        a += (double)V[i];
        b += (double)V[3 * nv + i] + (double)V[4 * nv + i];
        if (COMP) {
            a += (double)V[5 * nv + i];
            b += (double)V[6 * nv + i] + (double)V[7 * nv + i];
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o, 64);
        b += __shfl_xor(b, o, 64);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sh[0][w] = a;
        sh[1][w] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = ((sh[0][0] + sh[0][1]) + sh[0][2]) + sh[0][3];
        partial[2 * blockIdx.x + 1] = ((sh[1][0] + sh[1][1]) + sh[1][2]) + sh[1][3];
    }
}

__global__ void mass_final_kernel(const double* partial, int nb, double* out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double a = 0.0, b = 0.0;
    for (int i = 0; i < nb; i++) {
        a += partial[2 * i];
        b += partial[2 * i + 1];
    }
    out[0] = a;
    out[1] = b;
}

template <bool FB>
__global__ void kat_edge_kernel(DevParams P, int n, const double* left, const double* right, const int* kind,
                                const int* fixed, double* Dm, double* Dp)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double L[5], R[5], dm[5], dp[5];
    for (int f = 0; f < 5; f++) {
        L[f] = left[5 * i + f];
        R[f] = right[5 * i + f];
    }
    edge_flux<double, true, FB>(P, kind[i], fixed[i] != 0 && P.bedload, L, R, dm, dp);
    for (int f = 0; f < 5; f++) {
        Dm[5 * i + f] = dm[f];
        Dp[5 * i + f] = dp[f];
    }
}

// Edges 2k and 2k + 1 solved together by lane k, as the stage kernels
// solve a row's x edge and y edge (edge_flux_pair).
template <bool FB>
__global__ void kat_edge_pair_kernel(DevParams P, int npairs, const double* left, const double* right,
                                     const int* kind, const int* fixed, double* Dm, double* Dp)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npairs) return;
    const int a = 2 * k, b = 2 * k + 1;
    double La[5], Ra[5], Lb[5], Rb[5], dma[5], dpa[5], dmb[5], dpb[5];
    for (int f = 0; f < 5; f++) {
        La[f] = left[5 * a + f];
        Ra[f] = right[5 * a + f];
        Lb[f] = left[5 * b + f];
        Rb[f] = right[5 * b + f];
    }
    edge_flux_pair<double, true, FB>(P, kind[a], fixed[a] != 0 && P.bedload, La, Ra, dma, dpa, kind[b],
                                     fixed[b] != 0 && P.bedload, Lb, Rb, dmb, dpb);
    for (int f = 0; f < 5; f++) {
        Dm[5 * a + f] = dma[f];
        Dp[5 * a + f] = dpa[f];
        Dm[5 * b + f] = dmb[f];
        Dp[5 * b + f] = dpb[f];
    }
}

// what: 0 limiter(a, b); 1 velocity(a=h, b=qx, c=qy); 2 state filter
// (a, b, c, d) = (h, qx, qy, hc); 3 cell dt (a, b, c); 4 erosion rate
// (h, qx, qy, hc); 5 bedload (h, un, ut).
__global__ void kat_point_kernel(int what, int law, int n, const double* a, const double* b, const double* c,
                                 const double* d, double* o1, double* o2, double* o3, double* o4, DevParams P)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    switch (what) {
    case 0:
        o1[i] = limiter(law, a[i], b[i]);
        break;
    case 1: {
        Vel<double> u = velocity(a[i], b[i], c[i]);
        o1[i] = u.x;
        o2[i] = u.y;
        o3[i] = speed(u);
        break;
    }
    case 2: {
        double h = a[i], qx = b[i], qy = c[i], hc = d[i];
        state_filter<double, true>(h, qx, qy, hc);
        o1[i] = h;
        o2[i] = qx;
        o3[i] = qy;
        o4[i] = hc;
        break;
    }
    case 3:
        o1[i] = cell_dt(P, a[i], b[i], c[i]);
        break;
    case 4: {
        Vel<double> u = velocity(a[i], b[i], c[i]);
        o1[i] = erosion_rate(P, a[i], speed(u), d[i]);
        break;
    }
    case 5: {
        Vel<double> u;
        u.x = b[i];
        u.y = c[i];
        u.s = u.x * u.x + u.y * u.y;
        o1[i] = bedload(P, a[i], u);
        break;
    }
    case 6: {
        // the shared-denominator division: 1 / a and b / a through den(a),
        // beside the IEEE divisions
        const Den<double> D = den(a[i]);
        o1[i] = D.y;
        o2[i] = qdiv(b[i], D, wave_any(qtiny(b[i])));
        o3[i] = 1.0 / a[i];
        o4[i] = b[i] / a[i];
        break;
    }
    case 7:
        // the square root without operand scaling beside sqrt()
        o1[i] = sqrt_unscaled0(a[i]);
        o2[i] = sqrt(a[i]);
        break;
    }
}

}  // namespace

hipError_t launch_stage(const StageArgs& a, bool sed, int precision, hipStream_t s, bool fast_bed)
{
    dim3 grid(a.ntiles), block(64);
    if (precision == 0) {
        if (sed && fast_bed)
            hipLaunchKernelGGL((stage_kernel<double, double, true, true>), grid, block, 0, s, a);
        else if (sed)
            hipLaunchKernelGGL((stage_kernel<double, double, true>), grid, block, 0, s, a);
        else
            hipLaunchKernelGGL((stage_kernel<double, double, false>), grid, block, 0, s, a);
    } else {
        if (sed && fast_bed)
            hipLaunchKernelGGL((stage_kernel<float, float, true, true>), grid, block, 0, s, a);
        else if (sed)
            hipLaunchKernelGGL((stage_kernel<float, float, true>), grid, block, 0, s, a);
        else
            hipLaunchKernelGGL((stage_kernel<float, float, false>), grid, block, 0, s, a);
    }
    return hipGetLastError();
}

hipError_t launch_stage_batch(const BatchPools& pools, const BatchMember* mem, const int2* tiles, int ntiles, int k,
                              int last, bool sed, int precision, hipStream_t s, bool fast_bed)
{
    dim3 grid(ntiles), block(64);
    if (precision == 0) {
        if (sed && fast_bed)
            hipLaunchKernelGGL((stage_kernel_batch<double, double, true, true>), grid, block, 0, s, pools, mem, tiles, k,
                               last);
        else if (sed)
            hipLaunchKernelGGL((stage_kernel_batch<double, double, true>), grid, block, 0, s, pools, mem, tiles, k, last);
        else
            hipLaunchKernelGGL((stage_kernel_batch<double, double, false>), grid, block, 0, s, pools, mem, tiles, k,
                               last);
    } else {
        if (sed && fast_bed)
            hipLaunchKernelGGL((stage_kernel_batch<float, float, true, true>), grid, block, 0, s, pools, mem, tiles, k,
                               last);
        else if (sed)
            hipLaunchKernelGGL((stage_kernel_batch<float, float, true>), grid, block, 0, s, pools, mem, tiles, k, last);
        else
            hipLaunchKernelGGL((stage_kernel_batch<float, float, false>), grid, block, 0, s, pools, mem, tiles, k,
                               last);
    }
    return hipGetLastError();
}

hipError_t launch_advance_batch(Ctrl* ctrl, const DevParams* P, int n, hipStream_t s)
{
    hipLaunchKernelGGL(advance_batch_kernel, dim3((n + 63) / 64), dim3(64), 0, s, ctrl, P, n);
    return hipGetLastError();
}

hipError_t launch_init_dt(const void* V, const DevParams& P, Ctrl* ctrl, int precision, hipStream_t s)
{
    long long nb = (P.nvol + 255) / 256;
    int grid = (int)(nb < 4096 ? (nb > 0 ? nb : 1) : 4096);
    if (precision == 0)
        hipLaunchKernelGGL((init_dt_kernel<double, double>), dim3(grid), dim3(256), 0, s, (const double*)V, P, ctrl);
    else
        hipLaunchKernelGGL((init_dt_kernel<float, float>), dim3(grid), dim3(256), 0, s, (const float*)V, P, ctrl);
    return hipGetLastError();
}

hipError_t launch_warmup(hipStream_t s)
{
    hipLaunchKernelGGL(warmup_kernel, dim3(2, 2), dim3(16, 16), 0, s);
    return hipGetLastError();
}

hipError_t launch_copy(const void* src, void* dst, long long bytes, hipStream_t s)
{
    long long n = bytes / 16;
    // one pass per workgroup where possible, at most 2^16 workgroups
    long long nb = (n + 4 * 256 - 1) / (4 * 256);
    const unsigned grid = (unsigned)(nb < 1 ? 1 : (nb > 65536 ? 65536 : nb));
    hipLaunchKernelGGL(copy_kernel, dim3(grid), dim3(256), 0, s, (const u32x4*)src, (u32x4*)dst, n);
    return hipGetLastError();
}

hipError_t launch_advance(Ctrl* ctrl, const DevParams& P, hipStream_t s)
{
    hipLaunchKernelGGL(advance_kernel, dim3(1), dim3(64), 0, s, ctrl, P);
    return hipGetLastError();
}

hipError_t launch_fill(double* p, long long n, double v, hipStream_t s)
{
    long long nb = (n + 255) / 256;
    int grid = (int)(nb < 4096 ? (nb > 0 ? nb : 1) : 4096);
    hipLaunchKernelGGL(fill_kernel, dim3(grid), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

int mass_blocks(long long nvol)
{
    long long nb = (nvol + 255) / 256;
    return (int)(nb < 1024 ? (nb > 0 ? nb : 1) : 1024);
}

hipError_t launch_mass(const void* V, long long nvol, int precision, double* partial, int nblocks, double* out,
                       hipStream_t s)
{
    if (precision == 0)
        hipLaunchKernelGGL((mass_partial_kernel<double, false>), dim3(nblocks), dim3(256), 0, s, (const double*)V,
                           nvol, partial);
    else
        hipLaunchKernelGGL((mass_partial_kernel<float, true>), dim3(nblocks), dim3(256), 0, s, (const float*)V, nvol,
                           partial);
    hipLaunchKernelGGL(mass_final_kernel, dim3(1), dim3(64), 0, s, partial, nblocks, out);
    return hipGetLastError();
}

hipError_t launch_kat_edge(const DevParams& P, int n, const double* left, const double* right, const int* kind,
                           const int* fixed, double* Dm, double* Dp, hipStream_t s)
{
    // the stage kernels' choice of bedload evaluation for these run constants
    if (fast_bedload(P))
        hipLaunchKernelGGL(kat_edge_kernel<true>, dim3((n + 63) / 64), dim3(64), 0, s, P, n, left, right, kind, fixed,
                           Dm, Dp);
    else
        hipLaunchKernelGGL(kat_edge_kernel<false>, dim3((n + 63) / 64), dim3(64), 0, s, P, n, left, right, kind,
                           fixed, Dm, Dp);
    return hipGetLastError();
}

hipError_t launch_kat_edge_pairs(const DevParams& P, int npairs, const double* left, const double* right,
                                 const int* kind, const int* fixed, double* Dm, double* Dp, hipStream_t s)
{
    if (fast_bedload(P))
        hipLaunchKernelGGL(kat_edge_pair_kernel<true>, dim3((npairs + 63) / 64), dim3(64), 0, s, P, npairs, left,
                           right, kind, fixed, Dm, Dp);
    else
        hipLaunchKernelGGL(kat_edge_pair_kernel<false>, dim3((npairs + 63) / 64), dim3(64), 0, s, P, npairs, left,
                           right, kind, fixed, Dm, Dp);
    return hipGetLastError();
}

hipError_t launch_kat_point(int what, int law, int n, const double* a, const double* b, const double* c,
                            const double* d, double* o1, double* o2, double* o3, double* o4, const DevParams& P,
                            hipStream_t s)
{
    hipLaunchKernelGGL(kat_point_kernel, dim3((n + 63) / 64), dim3(64), 0, s, what, law, n, a, b, c, d, o1, o2, o3,
                       o4, P);
    return hipGetLastError();
}

}  // namespace tsu
