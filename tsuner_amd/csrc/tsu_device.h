// tsu_device.h -- device-side arithmetic of the TSUNER per-step solver for
// gfx950 (CDNA4), shared by the fused stage kernels and the known-answer
// kernels.
//
// Every function is a template on the compute type R.
//
// R = double is the reference arithmetic: each function computes the same
// IEEE value as the reference expression it cites (no FMA contraction: the
// library is built with -ffp-contract=off, as the reference is built with
// -fmad=false, GPU/makefile:10).  Where this code is shaped differently from
// the reference, the change is value-preserving:
//   * products with a 0/1 gate, e.g. x*(h>=EPS), become a select: the reference
//     yields +-0 for a closed gate, this code yields +0 (equal as values);
//   * rotations by the axis-aligned edge normals are signed permutations
//     (q*1 + q'*0 equals q as a value);
//   * a quantity the reference recomputes from identical operands (the cell
//     velocity, |u|) is computed once;
//   * divisions whose numerator is exactly zero are skipped (0/d == 0).
// Results therefore equal the reference's as values (+0 == -0), which is how
// the parity tests compare them.
//
// R = float is the FP32 solver: the same scheme with single-instruction
// reciprocal and square root (v_rcp_f32, v_sqrt_f32, about 1 ulp), the
// velocity of near-dry cells (h < 1 cm) evaluated in FP64 exactly as the
// reference does, and mass, bed and suspension fluctuations made exactly
// antisymmetric across each edge (Dp = -Dm) so that compensated cell
// accumulation conserves them to FP64-level round-off.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tsu {

// Sim_parameters.hxx:28-36
constexpr double kEpsilon = 2.2204460492503131e-16;  // DBL_EPSILON
constexpr double kG = 9.81;
constexpr double kSqrtG = 3.1320919526731650;   // sqrt(9.81), FP32 solver only
constexpr double kEps = 1e-6;
constexpr double kVmax = 10.0;
// FP32 solver: cells shallower than this get their velocity in FP64.
constexpr float kNearDry = 1e-2f;

enum { BC_FREE = 0, BC_WALL = 1, BC_PERIODIC = 2, BC_INFLOW = 3 };
enum { LIM_VANLEER = 0, LIM_MINMOD = 1 };
enum { BEDLOAD_AM = 0, BEDLOAD_GRASS = 1 };
enum { FRIC_CHEZY = 0, FRIC_MANNING = 1 };

// Edge normal kinds: +x, -x, +y, -y.
enum { N_PX = 0, N_MX = 1, N_PY = 2, N_MY = 3 };

// Run constants, evaluated once on the host from the reference parameters.
struct DevParams {
    int npx, npy;
    long long nvol;
    double dx, dy, inv_dx, inv_dy;
    double cfl_half;        // CFL * 0.5            (Volume_kernel.cu:39)
    double FR, epsilon_h, Q, Opar, OC, sqrtOC, HQB, fiveDS;
    double vs;              // settling velocity    (Reconstructor.cu:76)
    double inv_vs;          // correctly rounded 1 / vs (shared-denominator quotients)
    double ka;              // alpha1 * pow(Rp, alpha2)   (SedFlux.cuh:56-65)
    double sqrtFR;
    double cbk;             // 1.64 + 0.4*pow(DS/Dsg, 1.64) (SedFlux.cuh:83)
    double vs1p;            // vs * (1 - porosity)  (SedFlux.cuh:70)
    double grass_k;         // A_g / (1 - porosity)
    double manning_n;
    int bedload, suspension, limiter, bedload_law, friction_law, order, dt_narrow, extrema;
    int bed_sqrt_unscaled;  // 1: both bedload square roots have arguments above 2^-700 (see bedload)
    int bc[4];              // left, right, bottom, top
    // boundary inflow series (VolumeSolver.cu:172-202)
    const double* CC;
    const double* tCC;
    long long timeCC;
    long long ring;         // 2*npx + 2*npy
    double tcc0, dtcc, tcc_end;   // tCC[0], tCC[1]-tCC[0], tCC[timeCC-2]
};

// Step control block, resident in device memory.  t and step change only in
// the advance kernel; dt_bits[s & 1] holds the IEEE bits of the minimum cell
// dt of step s (positive doubles order like their bit patterns).
struct Ctrl {
    double t;
    double stop_time;
    unsigned long long dt_bits[2];
    long long step;
    long long stop_step;   // steps enqueued once step >= stop_step do nothing
    long long clamp;       // 1: the step crossing stop_time ends on it (dt = stop_time - t)
    long long pad;
};

// dt of the step starting at t: the finalized minimum, shortened to land on
// stop_time when clamping is on.  Stage and advance kernels use the same rule.
__device__ __forceinline__ double clamp_dt(const Ctrl* C, double t, double dt)
{
    if (C->clamp && !(dt < C->stop_time - t)) return C->stop_time - t;
    return dt;
}

// ---- primitive operations per compute type --------------------------------

__device__ __forceinline__ double rdiv(double a, double b) { return a / b; }
__device__ __forceinline__ float rdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double rsqrt_(double x) { return sqrt(x); }
__device__ __forceinline__ float rsqrt_(float x) { return __builtin_amdgcn_sqrtf(x); }
// a*b + c: FP64 rounds the product and the sum separately, as the
// reference (built without FMA contraction); the FP32 solver fuses them.
__device__ __forceinline__ double rmad(double a, double b, double c) { return a * b + c; }
__device__ __forceinline__ float rmad(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double rmin(double a, double b) { return fmin(a, b); }
__device__ __forceinline__ float rmin(float a, float b) { return fminf(a, b); }
__device__ __forceinline__ double rmax(double a, double b) { return fmax(a, b); }
__device__ __forceinline__ float rmax(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ double rabs(double a) { return fabs(a); }
__device__ __forceinline__ float rabs(float a) { return fabsf(a); }

template <typename R>
__device__ __forceinline__ int sgn(R x)
{
    return (rabs(x) < (R)kEpsilon) ? 0 : ((x > (R)0) ? 1 : -1);
}

// Reconstructor.cu:19-31.  When a and b do not share a strict sign the
// numerator a|b| + b|a| is exactly zero (for finite products), so the
// division is only evaluated for same-sign slopes; there a|b| and b|a| are
// both +-RN(ab) (rounding is sign-symmetric), so the numerator is
// +-2 RN(ab) exactly.
// In two parts: whether the quotient is taken (vl_use) and the limited
// slope given that flag (vl_val), so that callers can skip the quotients of
// a set of slopes with one wave-wide vote.
template <typename R>
__device__ __forceinline__ bool vl_use(R a, R b)
{
    if constexpr (sizeof(R) == 4) {
        // FP32 solver: for slopes of one strict sign |a| + |b| = |a + b| and
        // the quotient is 2ab / (a + b); ab > 0 selects exactly those pairs
        // (a product that underflows gives 0 where the reference's value,
        // at most 2 min(|a|, |b|) <= 2 sqrt(ab), is below 1e-22), and then
        // |a + b| >= 2 sqrt(ab) > 0, so the reciprocal is finite.
        return a * b > 0.f;
    } else {
        // same strict sign as ab > 0, except where ab underflows to 0: there
        // the reference's numerator 2 RN(ab) is 0 and so is its quotient
        const R s = rabs(a) + rabs(b);
        return a * b > (R)0 && s > (R)kEpsilon;
    }
}

template <typename R>
__device__ __forceinline__ R vl_val(R a, R b, bool use)
{
    const R p = a * b;   // > 0 for same-sign slopes
    if constexpr (sizeof(R) == 4) {
        const float r = __builtin_amdgcn_rcpf(use ? a + b : 1.f);
        return use ? (p + p) * r : 0.f;
    } else {
        // evaluated for every lane and selected, so the limiters of a face
        // set share one basic block; the divisor is 1 where the result is
        // unused
        const R s = rabs(a) + rabs(b);
        const R d = rdiv(a > (R)0 ? p + p : -(p + p), use ? s : (R)1);
        return use ? d : (R)0;
    }
}

template <typename R>
__device__ __forceinline__ R lim_vanleer(R a, R b)
{
    return vl_val(a, b, vl_use(a, b));
}

template <typename R>
__device__ __forceinline__ R lim_minmod(R a, R b)
{
    if (a > (R)0 && b > (R)0) return rmin(a, b);
    if (a < (R)0 && b < (R)0) return rmax(a, b);
    return (R)0;
}

template <typename R>
__device__ __forceinline__ R limiter(int law, R a, R b)
{
    return law == LIM_MINMOD ? lim_minmod(a, b) : lim_vanleer(a, b);
}

// sqrt is monotone and correctly rounded, so sqrt(s) > 10 exactly when
// s > nextafter(100, +inf).
constexpr double kVmax2Cut = 100.00000000000001421085471520200371742248535156250;

template <typename R>
struct Vel {
    R x, y;
    R s;   // u.x*u.x + u.y*u.y of the returned (possibly capped) velocity
};

// The correctly rounded FP64 square root without its operand scaling:
// rsqrt estimate, one Newton step on (g, h) = (x y, y / 2) and two residual
// corrections, the compiler's own sequence for sqrt() with the scale
// factors at 1 (they act only below 2^-767).  Equal to sqrt(x) for x in
// [2^-767, DBL_MAX]; sqrt_unscaled0 also returns 0 for x = 0.
__device__ __forceinline__ double sqrt_unscaled(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = y * 0.5;
    const double r = fma(-h, g, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    double d = fma(-g, g, x);
    g = fma(d, h, g);
    d = fma(-g, g, x);
    return fma(d, h, g);
}
__device__ __forceinline__ double sqrt_unscaled0(double x) { return x == 0.0 ? x : sqrt_unscaled(x); }
// x > 0 and below the unscaled square root's range
__device__ __forceinline__ bool sqrt_needs_scaling(double x) { return x > 0.0 && x < 0x1p-767; }

// |u| as the reference evaluates it after u_calculator (Reconstructor.cu:43,74;
// SedFlux.cuh:157): sqrt(u.x*u.x + u.y*u.y).
template <typename R>
__device__ __forceinline__ R speed(const Vel<R>& u) { return rsqrt_(u.s); }

// Quotients x / d that share one denominator d.  FP64: y = 1/d is one IEEE
// division; each quotient is q = x*y corrected by one residual step,
// q + (x - d*q)*y with both products fused.  With y the correctly rounded
// reciprocal and q within an ulp of x/d, that is the correctly rounded x/d
// (Markstein's theorem), i.e. the value of x / d, provided the residual
// does not underflow (|x| above 2^-969; x == 0 gives a zero).  FP32: the
// solver's single-instruction reciprocal, as rdiv.
template <typename R>
struct Den {
    R d, y;
};

// RN(1/d) for d in [2^-60, 2^60]: the steps of the IEEE f64 division with
// numerator 1 (hardware reciprocal, two Newton refinements, one residual
// correction) without its operand scaling and special-case fixup, which
// leave operands and result unchanged in that range.
__device__ __forceinline__ double recip_normal(double d)
{
    const double r0 = __builtin_amdgcn_rcp(d);
    double r = fma(r0, fma(-d, r0, 1.0), r0);
    r = fma(r, fma(-d, r, 1.0), r);
    return fma(fma(-d, r, 1.0), r, r);
}

// Every shared denominator of the solver lies in [2^-60, 2^60]: h + DBL_EPSILON
// (>= 2^-52), S_R - S_L (> DBL_EPSILON, a few hundred m/s at most), 1 + c
// (>= 1, c <= dt FR 10 / DBL_EPSILON) and |u| > 10 under the speed cap.
__device__ __forceinline__ Den<double> den(double d) { return {d, recip_normal(d)}; }
__device__ __forceinline__ Den<float> den(float d) { return {d, __builtin_amdgcn_rcpf(d)}; }
__device__ __forceinline__ double qdiv(double x, const Den<double>& D)
{
    const double q = x * D.y;
    return fma(fma(-D.d, q, x), D.y, q);
}
__device__ __forceinline__ float qdiv(float x, const Den<float>& D) { return x * D.y; }

// A nonzero numerator below 2^-968 can make qdiv's residual underflow, and
// then the corrected product may miss x / d by an ulp (zero is exact).  The
// solver's quotients of state values (velocities, concentrations, momenta
// over 1 + c) test their numerators as a group with one wave-wide vote and
// take the IEEE division for the whole group when any lane has such a
// numerator (`slow`, uniform over the wave); every other quotient of the
// solver has a numerator that is zero or far above 2^-968 (products and
// differences of speeds at least sqrt(g * 2^-1074), or a factor of a
// quantity that is only used as 1 + c).
__device__ __forceinline__ bool qtiny(double x) { return x != 0.0 && fabs(x) < 0x1p-968; }
__device__ __forceinline__ bool wave_any(bool b) { return __builtin_expect(__builtin_amdgcn_ballot_w64(b) != 0, 0); }
__device__ __forceinline__ double qdiv(double x, const Den<double>& D, bool slow)
{
    if (slow) return x / D.d;
    return qdiv(x, D);
}
__device__ __forceinline__ float qdiv(float x, const Den<float>& D, bool) { return x * D.y; }

// SedFlux.cuh:26-49 in FP64.  Divisions with a zero numerator are skipped
// (0/d is 0); the cap test uses s instead of sqrt(s) (see kVmax2Cut).  D,
// when given, is den(h + DBL_EPSILON) computed by the caller.
__device__ __forceinline__ Vel<double> velocity64(double h, double qx, double qy, const Den<double>* D = nullptr)
{
    Vel<double> u;
    u.x = 0.0;
    u.y = 0.0;
    if (h >= kEps && (qx != 0.0 || qy != 0.0)) {
        const Den<double> Dh = D ? *D : den(h + kEpsilon);
        if (qx != 0.0) u.x = qdiv(qx, Dh);
        if (qy != 0.0) u.y = qdiv(qy, Dh);
    }
    double s = u.x * u.x + u.y * u.y;
    if (s > kVmax2Cut) {
        const Den<double> Dm = den(sqrt(s));
        u.x = qdiv(u.x, Dm) * kVmax;
        u.y = qdiv(u.y, Dm) * kVmax;
        s = u.x * u.x + u.y * u.y;
    }
    u.s = s;
    return u;
}

__device__ __forceinline__ Vel<double> velocity(double h, double qx, double qy) { return velocity64(h, qx, qy); }

// FP32 solver: near-dry cells take the reference's FP64 desingularisation,
// wet cells a single reciprocal.  rh, when given, is rcp(h + DBL_EPSILON)
// computed by the caller: for h >= 1 cm, h + DBL_EPSILON rounds to h in FP32,
// so it is the wet path's rcp(h).
__device__ __forceinline__ Vel<float> velocity(float h, float qx, float qy, const float* rh = nullptr)
{
    Vel<float> u;
    if (h < kNearDry) {
        Vel<double> d = velocity64((double)h, (double)qx, (double)qy);
        u.x = (float)d.x;
        u.y = (float)d.y;
        u.s = (float)d.s;
        return u;
    }
    const float r = rh ? *rh : __builtin_amdgcn_rcpf(h);
    u.x = qx * r;
    u.y = qy * r;
    float s = u.x * u.x + u.y * u.y;
    if (s > 100.0f) {
        float k = (float)kVmax * __builtin_amdgcn_rsqf(s);
        u.x *= k;
        u.y *= k;
        s = u.x * u.x + u.y * u.y;
    }
    u.s = s;
    return u;
}

// Velocity and concentration hc / (h + DBL_EPSILON) of one state
// (SedFlux.cuh:26-49, 103-132); in FP64 the three quotients share one
// reciprocal of h + DBL_EPSILON.
// The speed cap of velocity64 on an uncapped (u.x, u.y): |u| > 10 m/s is
// scaled to 10 m/s, and u.s set.
__device__ __forceinline__ void cap_speed(Vel<double>& u)
{
    double s = u.x * u.x + u.y * u.y;
    if (s > kVmax2Cut) {
        const Den<double> Dm = den(sqrt(s));
        u.x = qdiv(u.x, Dm) * kVmax;
        u.y = qdiv(u.y, Dm) * kVmax;
        s = u.x * u.x + u.y * u.y;
    }
    u.s = s;
}

// velocity_conc for the two sides of an edge at once (values identical):
// one wave-wide test whether any lane of either side divides, then both
// sides' reciprocals and quotients for every lane in one basic block, the
// unused ones discarded by selects (h + DBL_EPSILON is a valid shared
// denominator for any face depth).  Returns true (uniform over the wave)
// when no lane divides: every velocity and concentration is then 0.
template <bool SED>
__device__ __forceinline__ bool velocity_conc_pair(double h0, double qx0, double qy0, double hc0, double h1,
                                                   double qx1, double qy1, double hc1, Vel<double>& u0, double& c0,
                                                   Vel<double>& u1, double& c1)
{
    const bool nu0 = h0 >= kEps && (qx0 != 0.0 || qy0 != 0.0), nc0 = SED && hc0 != 0.0;
    const bool nu1 = h1 >= kEps && (qx1 != 0.0 || qy1 != 0.0), nc1 = SED && hc1 != 0.0;
    u0.x = u0.y = u0.s = 0.0;
    u1.x = u1.y = u1.s = 0.0;
    c0 = c1 = 0.0;
    if (__builtin_amdgcn_ballot_w64(nu0 || nc0 || nu1 || nc1) == 0) return true;
    const Den<double> D0 = den(h0 + kEpsilon), D1 = den(h1 + kEpsilon);
    const double x0 = qdiv(qx0, D0), y0 = qdiv(qy0, D0);
    const double x1 = qdiv(qx1, D1), y1 = qdiv(qy1, D1);
    if (SED) {
        const bool slow = wave_any(qtiny(hc0) || qtiny(hc1));
        const double k0 = qdiv(hc0, D0, slow), k1 = qdiv(hc1, D1, slow);
        c0 = nc0 ? k0 : 0.0;
        c1 = nc1 ? k1 : 0.0;
    }
    u0.x = (nu0 && qx0 != 0.0) ? x0 : 0.0;
    u0.y = (nu0 && qy0 != 0.0) ? y0 : 0.0;
    u1.x = (nu1 && qx1 != 0.0) ? x1 : 0.0;
    u1.y = (nu1 && qy1 != 0.0) ? y1 : 0.0;
    cap_speed(u0);
    cap_speed(u1);
    return false;
}

template <bool SED>
__device__ __forceinline__ void velocity_conc(double h, double qx, double qy, double hc, Vel<double>& u, double& c)
{
    c = 0.0;
    const bool need_u = h >= kEps && (qx != 0.0 || qy != 0.0);
    const bool need_c = SED && hc != 0.0;
    if (need_u || need_c) {
        const Den<double> D = den(h + kEpsilon);
        if (need_c) c = qdiv(hc, D, wave_any(qtiny(hc)));
        u = velocity64(h, qx, qy, &D);
    } else {
        u = velocity64(h, qx, qy);
    }
}

template <bool SED>
__device__ __forceinline__ void velocity_conc(float h, float qx, float qy, float hc, Vel<float>& u, float& c)
{
    // one reciprocal for the concentration and the velocity, evaluated for
    // every lane and selected (no branch)
    const float r = __builtin_amdgcn_rcpf(h + (float)kEpsilon);
    c = (SED && hc != 0.f) ? hc * r : 0.f;
    u = velocity(h, qx, qy, &r);
}

// SedFlux.cuh:118-132
template <typename R, bool SED>
__device__ __forceinline__ void state_filter(R& h, R& qx, R& qy, R& hc)
{
    h = (h > (R)0) ? h : (R)0;
    R c;
    Vel<R> u;
    velocity_conc<SED>(h, qx, qy, hc, u, c);
    if (SED) c = (c > (R)0) ? c : (R)0;
    qx = h * u.x;
    qy = h * u.y;
    if (SED) hc = h * c;
}

// Bed-plane flux before the mask: SedFlux.cuh:150-168 (Ashida-Michiue), or
// Grass q_b = A_g/(1-p) u_n |u|^2.  Sign as in the reference (returns -q_b).
template <typename R>
__device__ __forceinline__ R bedload(const DevParams& P, R h, const Vel<R>& u)
{
    if (!(h > (R)P.HQB && h > (R)P.fiveDS)) return (R)0;
    if (P.bedload_law == BEDLOAD_GRASS) {
        R m = speed(u);
        return -((R)P.grass_k * u.x * m * m);
    }
    // FP64: the Shields number O = Opar m m with m = sqrt(|u|^2) lies within
    // a few roundings of Opar |u|^2, so when Opar |u|^2 is below OC by far
    // more than that (the factor 1 - 2^-40) max(O - OC, 0) is 0 and the
    // square root is not needed.
    if constexpr (sizeof(R) == 4) {
        // FP32 solver: O from |u|^2 directly, the square root only past the
        // threshold
        const R O = (R)P.Opar * u.s;
        if (!(O > (R)P.OC)) return (R)0;
        const R m = speed(u);
        const R K1 = rdiv((R)17.0 * u.x, m + (R)kEpsilon);
        const R qb = K1 * (R)P.Q * (O - (R)P.OC) * rmax(rsqrt_(O) - (R)P.sqrtOC, (R)0);
        return -qb;
    } else {
        if (!(P.Opar * u.s > P.OC * (1.0 - 0x1p-40))) return 0.0;
        // Past that test |u|^2 > OC margin / Opar and O > OC; when the run
        // constants put both above 2^-700 (bed_sqrt_unscaled), the square
        // roots need no operand scaling.
        const R m = P.bed_sqrt_unscaled ? sqrt_unscaled(u.s) : speed(u);
        const R O = (R)P.Opar * m * m;
        if (!(O > (R)P.OC)) return (R)0;
        const R K1 = rdiv((R)17.0 * u.x, m + (R)kEpsilon);
        const R sO = P.bed_sqrt_unscaled ? sqrt_unscaled(O) : rsqrt_(O);
        const R qb = K1 * (R)P.Q * (O - (R)P.OC) * rmax(sO - (R)P.sqrtOC, (R)0);
        return -qb;
    }
}

// bedload for the two sides of an edge at once (values identical): with the
// Ashida-Michiue law in FP64, one wave-wide test whether any lane of either
// side moves its bed, then both sides' square roots and quotients for every
// lane in one basic block, so that the two dependent chains interleave; the
// unused results (lanes below a threshold, whose roots and quotients may be
// NaN) are discarded by selects.  mobile gates both sides (the edge's bed
// mask).
template <typename R>
__device__ __forceinline__ void bedload_pair(const DevParams& P, bool mobile, R h0, const Vel<R>& u0, R h1,
                                             const Vel<R>& u1, R& f0, R& f1)
{
    if constexpr (sizeof(R) == 8) {
        if (P.bedload_law == BEDLOAD_AM) {
            const double thr = P.OC * (1.0 - 0x1p-40);
            const bool a0 = mobile && h0 > P.HQB && h0 > P.fiveDS && P.Opar * u0.s > thr;
            const bool a1 = mobile && h1 > P.HQB && h1 > P.fiveDS && P.Opar * u1.s > thr;
            f0 = f1 = 0.0;
            if (__builtin_amdgcn_ballot_w64(a0 || a1) == 0) return;
            double m0, m1;
            if (P.bed_sqrt_unscaled) {
                m0 = sqrt_unscaled(u0.s);
                m1 = sqrt_unscaled(u1.s);
            } else {
                m0 = speed(u0);
                m1 = speed(u1);
            }
            const double O0 = P.Opar * m0 * m0, O1 = P.Opar * m1 * m1;
            const bool b0 = a0 && O0 > P.OC, b1 = a1 && O1 > P.OC;
            const double K0 = (17.0 * u0.x) / (m0 + kEpsilon), K1 = (17.0 * u1.x) / (m1 + kEpsilon);
            double s0, s1;
            if (P.bed_sqrt_unscaled) {
                s0 = sqrt_unscaled(O0);
                s1 = sqrt_unscaled(O1);
            } else {
                s0 = sqrt(O0);
                s1 = sqrt(O1);
            }
            const double q0 = K0 * P.Q * (O0 - P.OC) * fmax(s0 - P.sqrtOC, 0.0);
            const double q1 = K1 * P.Q * (O1 - P.OC) * fmax(s1 - P.sqrtOC, 0.0);
            f0 = b0 ? -q0 : 0.0;
            f1 = b1 ? -q1 : 0.0;
            return;
        }
    }
    f0 = f1 = (R)0;
    if (mobile) {
        f0 = bedload(P, h0, u0);
        f1 = bedload(P, h1, u1);
    }
}

// Run constants for which the stage kernels take bedload_am_pair: the
// Ashida-Michiue law with both bedload square roots in the unscaled range.
__host__ __device__ inline bool fast_bedload(const DevParams& P)
{
    return P.bedload && P.bedload_law == BEDLOAD_AM && P.bed_sqrt_unscaled;
}

// bedload_pair in FP64 for fast_bedload run constants, without the vote
// and without branches: every lane evaluates both chains and the selects keep
// the lanes past the thresholds (values identical).  Placed among the HLL
// arithmetic of the Riemann solver, it shares that basic block, so the
// scheduler interleaves its square root -> quotient -> square root chains
// with the HLL chains.
__device__ __forceinline__ void bedload_am_pair(const DevParams& P, bool mobile, double h0, const Vel<double>& u0,
                                                double h1, const Vel<double>& u1, double& f0, double& f1)
{
    const double thr = P.OC * (1.0 - 0x1p-40);
    const bool a0 = mobile && h0 > P.HQB && h0 > P.fiveDS && P.Opar * u0.s > thr;
    const bool a1 = mobile && h1 > P.HQB && h1 > P.fiveDS && P.Opar * u1.s > thr;
    const double m0 = sqrt_unscaled(u0.s), m1 = sqrt_unscaled(u1.s);
    const double O0 = P.Opar * m0 * m0, O1 = P.Opar * m1 * m1;
    const bool b0 = a0 && O0 > P.OC, b1 = a1 && O1 > P.OC;
    const double K0 = (17.0 * u0.x) / (m0 + kEpsilon), K1 = (17.0 * u1.x) / (m1 + kEpsilon);
    const double s0 = sqrt_unscaled(O0), s1 = sqrt_unscaled(O1);
    const double q0 = K0 * P.Q * (O0 - P.OC) * fmax(s0 - P.sqrtOC, 0.0);
    const double q1 = K1 * P.Q * (O1 - P.OC) * fmax(s1 - P.sqrtOC, 0.0);
    f0 = b0 ? -q0 : 0.0;
    f1 = b1 ? -q1 : 0.0;
}

// The FP32 solver's bedload pair in the same branch-free form (its square
// root and reciprocal need no range condition).
__device__ __forceinline__ void bedload_am_pair(const DevParams& P, bool mobile, float h0, const Vel<float>& u0,
                                                float h1, const Vel<float>& u1, float& f0, float& f1)
{
    const float Opar = (float)P.Opar, OC = (float)P.OC;
    const float O0 = Opar * u0.s, O1 = Opar * u1.s;
    const bool b0 = mobile && h0 > (float)P.HQB && h0 > (float)P.fiveDS && O0 > OC;
    const bool b1 = mobile && h1 > (float)P.HQB && h1 > (float)P.fiveDS && O1 > OC;
    const float K0 = 17.0f * u0.x * __builtin_amdgcn_rcpf(rsqrt_(u0.s) + (float)kEpsilon);
    const float K1 = 17.0f * u1.x * __builtin_amdgcn_rcpf(rsqrt_(u1.s) + (float)kEpsilon);
    const float sq = (float)P.sqrtOC, Q = (float)P.Q;
    const float q0 = K0 * Q * (O0 - OC) * fmaxf(rsqrt_(O0) - sq, 0.f);
    const float q1 = K1 * Q * (O1 - OC) * fmaxf(rsqrt_(O1) - sq, 0.f);
    f0 = b0 ? -q0 : 0.f;
    f1 = b1 ? -q1 : 0.f;
}

// Unrotated (qx, qy) -> normal frame (qn, qt) for the four axis normals
// (VolumeSolver.cu:28-31 with normal_x, normal_y in {-1, 0, 1}): +x (qx, qy),
// -x (-qx, -qy), +y (qy, -qx), -y (-qy, qx).  Written as selects and a
// negation (the kind may differ between lanes), no branches.
template <typename R>
__device__ __forceinline__ void to_normal(int kind, R qx, R qy, R& qn, R& qt)
{
    const bool ydir = kind == N_PY || kind == N_MY, neg = kind == N_MX || kind == N_MY;
    const R n = ydir ? qy : qx, t = ydir ? -qx : qy;
    qn = neg ? -n : n;
    qt = neg ? -t : t;
}

// Normal frame -> unrotated (VolumeSolver.cu:142-152): +x (a, b), -x (-a, -b),
// +y (-b, a), -y (b, -a).
template <typename R>
__device__ __forceinline__ void from_normal(int kind, R a, R b, R& fx, R& fy)
{
    const bool ydir = kind == N_PY || kind == N_MY, neg = kind == N_MX || kind == N_MY;
    const R x = ydir ? -b : a, y = ydir ? a : b;
    fx = neg ? -x : x;
    fy = neg ? -y : y;
}

// VolumeSolver.cu:19-126 in the normal frame: hydrostatic reconstruction and
// the HLL/PVM path-conservative fluctuations Dm (to the left cell) and Dp (to
// the right cell).  Inputs: left/right face (h, H, qn, qt, hc) in FP64; in
// the FP32 solver the faces carry (h, H, u_n, u_t, c), the velocity and
// concentration that the state filter (or the cell's own velocity, first
// order) already computed, which the FP64 path recomputes from h u and h c.
// With SED false the bed and concentration rows are identically zero (no
// bedload, no suspension, hc == 0) and are not computed.
// FB (sediment, fast_bedload(P) true): bedload by bedload_am_pair among the
// HLL arithmetic; otherwise by bedload_pair with the bed terms.
//
// The solve is written in phases (rp_*) so that riemann_pair can run two
// edges' solves side by side in one basic block.

// One edge after its hydrostatic reconstruction: bed levels, depths,
// velocities and concentrations of both sides.
template <typename R>
struct EdgeRP {
    R zl, zr, h0, h1, c0, c1;
    Vel<R> u0, u1;
};

// Hydrostatic depths (VolumeSolver.cu:37-40).
template <typename R>
__device__ __forceinline__ void rp_depths(R hl, R zl, R hr, R zr, EdgeRP<R>& e)
{
    const R zmH = rmin(zl, zr);
    e.zl = zl;
    e.zr = zr;
    e.h0 = rmax(hl - zl + zmH, (R)0);
    e.h1 = rmax(hr - zr + zmH, (R)0);
}

// Velocities and concentrations of both sides (VolumeSolver.cu:32-35).
// Returns true (uniform over the wave) when no lane has a velocity or a
// concentration.
template <typename R, bool SED>
__device__ __forceinline__ bool rp_velocities(R hl, R qnl, R qtl, R hcl, R hr, R qnr, R qtr, R hcr, EdgeRP<R>& e)
{
    if constexpr (sizeof(R) == 8) {
        return velocity_conc_pair<SED>(hl, qnl, qtl, hcl, hr, qnr, qtr, hcr, e.u0, e.c0, e.u1, e.c1);
    } else {
        const R zero = (R)0;
        e.u0 = {qnl, qtl, qnl * qnl + qtl * qtl};
        e.u1 = {qnr, qtr, qnr * qnr + qtr * qtr};
        e.c0 = SED ? hcl : zero;
        e.c1 = SED ? hcr : zero;
        bool moving = e.u0.x != zero || e.u0.y != zero || e.u1.x != zero || e.u1.y != zero;
        if (SED) moving = moving || e.c0 != zero || e.c1 != zero;
        return __builtin_amdgcn_ballot_w64(moving) == 0;
    }
}

// A lane at rest (equal hydrostatic depths, no velocity and no
// concentration on either side) has every flux, jump and bed term of the
// solve equal to +-0 (the speeds and PVM coefficients stay finite; bedload
// is 0 at u = 0), so Dm = Dp = 0: when no lane of the wave moves (quiet)
// and no lane has a depth jump, the solve is skipped.
template <typename R>
__device__ __forceinline__ bool rp_has_jump(const EdgeRP<R>& e) { return e.h0 != e.h1; }

// Square roots of the depths and the HLL speed bounds before the Roe
// average (VolumeSolver.cu:63-73, SedFlux.cuh:181-189).
template <typename R>
struct RPSpeeds {
    R sh0, sh1, SL, SR, cm;
};

// FP64: every square-root argument is 0 or at least 2^-767 unless a depth
// is below it (g hm >= 4.9 min(h0, h1) when both are wet); a wave with no
// such lane takes the square roots without their scaling.  FP32 has one
// form (UNSCALED is ignored).
template <typename R>
__device__ __forceinline__ bool rp_needs_scaling(const EdgeRP<R>& e)
{
    if constexpr (sizeof(R) == 8)
        return sqrt_needs_scaling(e.h0) || sqrt_needs_scaling(e.h1);
    else
        return false;
}

template <typename R, bool UNSCALED>
__device__ __forceinline__ void rp_speeds(const EdgeRP<R>& e, RPSpeeds<R>& s)
{
    const R g = (R)kG;
    const R hm = (R)0.5 * (e.h0 + e.h1);
    if constexpr (sizeof(R) == 8) {
        if constexpr (UNSCALED) {
            s.sh0 = sqrt_unscaled0(e.h0);
            s.sh1 = sqrt_unscaled0(e.h1);
            s.SL = e.u0.x - sqrt_unscaled0(g * e.h0);
            s.SR = e.u1.x + sqrt_unscaled0(g * e.h1);
            s.cm = sqrt_unscaled0(g * hm);
        } else {
            s.sh0 = rsqrt_(e.h0);
            s.sh1 = rsqrt_(e.h1);
            s.SL = e.u0.x - rsqrt_(g * e.h0);
            s.SR = e.u1.x + rsqrt_(g * e.h1);
            s.cm = rsqrt_(g * hm);
        }
    } else {
        s.sh0 = rsqrt_(e.h0);
        s.sh1 = rsqrt_(e.h1);
        // FP32 solver: sqrt(g h) as sqrt(g) sqrt(h), reusing the Roe weights
        const R sg = (R)kSqrtG;
        s.SL = e.u0.x - sg * s.sh0;
        s.SR = e.u1.x + sg * s.sh1;
        s.cm = sg * rsqrt_(hm);
    }
}

// The rest of the solve: Roe velocity, PVM coefficients, fluxes, jumps and
// the bed and suspended-load rows.
template <typename R, bool SED, bool FB>
__device__ __forceinline__ void rp_solve(const DevParams& P, bool bed_mobile, const EdgeRP<R>& e,
                                         const RPSpeeds<R>& sp, R* Dm, R* Dp)
{
    const R g = (R)kG, half = (R)0.5, zero = (R)0;
    const R h0 = e.h0, h1 = e.h1;
    const Vel<R>& u0 = e.u0;
    const Vel<R>& u1 = e.u1;
    R q0n = u0.x * h0, q0t = u0.y * h0;
    R q1n = u1.x * h1, q1t = u1.y * h1;
    R deta = h1 - h0;

    R F00 = h0 * u0.x, F01 = F00 * u0.x, F02 = F00 * u0.y;
    R F10 = h1 * u1.x, F11 = F10 * u1.x, F12 = F10 * u1.y;

    R hm = half * (h0 + h1);
    R F03 = zero, F13 = zero;
    if constexpr (FB && SED) bedload_am_pair(P, bed_mobile, h0, u0, h1, u1, F03, F13);
    // h0 + h1 > 0 here, so sqrt(h0) + sqrt(h1) > 0: the Roe velocity is
    // always defined (the reference's 0/0 of two dry sides exits before).
    R ssum = sp.sh0 + sp.sh1;
    R umx = rdiv(sp.sh0 * u0.x + sp.sh1 * u1.x, ssum);
    R SL = rmin(sp.SL, umx - sp.cm);
    R SR = rmax(sp.SR, umx + sp.cm);

    // PVM coefficients; both branches of VolumeSolver.cu:75-83 evaluated and
    // selected (no block split), the division on a denominator kept in range.
    R dS = SR - SL;
    const bool spread = dS > (R)kEpsilon;
    R a0, a1;
    if constexpr (sizeof(R) == 8) {
        const Den<double> D = den(spread ? dS : 1.0);
        a0 = qdiv(SR * fabs(SL) - SL * fabs(SR), D);
        a1 = qdiv(fabs(SR) - fabs(SL), D);
    } else {
        R r = __builtin_amdgcn_rcpf(spread ? dS : (R)1);
        a0 = (SR * rabs(SL) - SL * rabs(SR)) * r;
        a1 = (rabs(SR) - rabs(SL)) * r;
    }
    a0 = spread ? a0 : rmax(rabs(SL), rabs(SR));
    a1 = spread ? a1 : zero;

    R df0 = F10 - F00, df1 = F11 - F01, df2 = F12 - F02;
    df1 = rmad(g * hm, deta, df1);
    R DU0 = h1 - h0, DU1 = q1n - q0n, DU2 = q1t - q0t;
    R v0 = rmad(a0, DU0, a1 * df0);
    R v1 = rmad(a0, DU1, a1 * df1);
    R v2 = rmad(a0, DU2, a1 * df2);
    Dm[0] = rmad(half, df0 - v0, F00);
    Dm[1] = rmad(half, df1 - v1, F01);
    Dm[2] = rmad(half, df2 - v2, F02);
    if constexpr (sizeof(R) == 8)
        Dp[0] = half * (df0 + v0) - F10;
    else
        Dp[0] = -Dm[0];
    Dp[1] = rmad(half, df1 + v1, -F11);
    Dp[2] = rmad(half, df2 + v2, -F12);

    if (SED) {
        R hc0 = h0 * e.c0, hc1 = h1 * e.c1;
        if constexpr (!FB) bedload_pair(P, bed_mobile, h0, u0, h1, u1, F03, F13);
        R F04 = hc0 * u0.x, F14 = hc1 * u1.x;
        R df3 = F13 - F03, df4 = F14 - F04;
        R DU4 = hc1 - hc0;
        int s0 = sgn(e.zr - e.zl), s1 = sgn(df3), s2 = sgn(a1);
        s1 = (s0 != 0) ? s1 * s0 : 0;
        R v3 = (s2 == s1) ? a1 * df3 : -a1 * df3;
        R v4 = rmad(a0, DU4, a1 * df4);
        Dm[3] = rmad(half, df3 - v3, F03);
        Dm[4] = rmad(half, df4 - v4, F04);
        if constexpr (sizeof(R) == 8) {
            Dp[3] = half * (df3 + v3) - F13;
            Dp[4] = half * (df4 + v4) - F14;
        } else {
            Dp[3] = -Dm[3];
            Dp[4] = -Dm[4];
        }
    }
}

template <typename R, bool SED, bool FB = false>
__device__ __forceinline__ void riemann(const DevParams& P, bool bed_mobile,
                                        R hl, R zl, R qnl, R qtl, R hcl,
                                        R hr, R zr, R qnr, R qtr, R hcr,
                                        R* Dm, R* Dp)
{
    EdgeRP<R> e;
    rp_depths(hl, zl, hr, zr, e);
    if (e.h0 == (R)0 && e.h1 == (R)0) {
        // Both reconstructed depths vanish: every flux, jump and bed term
        // is a product with h0, h1, their difference or a zero jump, so
        // Dm = Dp = 0 (the speeds stay finite: |u| <= VMAX).
        for (int f = 0; f < (SED ? 5 : 3); f++) Dm[f] = Dp[f] = (R)0;
        return;
    }
    const bool quiet = rp_velocities<R, SED>(hl, qnl, qtl, hcl, hr, qnr, qtr, hcr, e);
    if (quiet && __builtin_amdgcn_ballot_w64(rp_has_jump(e)) == 0) {
        for (int f = 0; f < (SED ? 5 : 3); f++) Dm[f] = Dp[f] = (R)0;
        return;
    }
    RPSpeeds<R> sp;
    if (sizeof(R) == 4 || !wave_any(rp_needs_scaling(e)))
        rp_speeds<R, true>(e, sp);
    else
        rp_speeds<R, false>(e, sp);
    rp_solve<R, SED, FB>(P, bed_mobile, e, sp, Dm, Dp);
}

// Two edges' Riemann problems (values identical to two riemann calls):
// side[i] = (hl, zl, qnl, qtl, hcl, hr, zr, qnr, qtr, hcr) of edge i.  The
// wave-wide tests are taken over both edges, so that the two solves run
// side by side in one basic block and their dependency chains overlap; a
// lane whose two depths vanish on an edge gets that edge's zeros by a
// select instead of an early exit.
template <typename R, bool SED, bool FB>
__device__ __forceinline__ void riemann_pair(const DevParams& P, const bool (&mobile)[2], const R (&side)[2][10],
                                             R (&Dm)[2][5], R (&Dp)[2][5])
{
    constexpr int NF = SED ? 5 : 3;
    EdgeRP<R> e[2];
    bool dry[2], quiet = true;
    for (int i = 0; i < 2; i++) {
        rp_depths(side[i][0], side[i][1], side[i][5], side[i][6], e[i]);
        dry[i] = e[i].h0 == (R)0 && e[i].h1 == (R)0;
    }
    for (int i = 0; i < 2; i++) {
        const bool q = rp_velocities<R, SED>(side[i][0], side[i][2], side[i][3], side[i][4], side[i][5], side[i][7],
                                             side[i][8], side[i][9], e[i]);
        quiet = quiet && q;
    }
    if (quiet && __builtin_amdgcn_ballot_w64(rp_has_jump(e[0]) || rp_has_jump(e[1])) == 0) {
        for (int i = 0; i < 2; i++)
            for (int f = 0; f < NF; f++) Dm[i][f] = Dp[i][f] = (R)0;
        return;
    }
    RPSpeeds<R> sp[2];
    if (sizeof(R) == 4 || !wave_any(rp_needs_scaling(e[0]) || rp_needs_scaling(e[1]))) {
        rp_speeds<R, true>(e[0], sp[0]);
        rp_speeds<R, true>(e[1], sp[1]);
    } else {
        rp_speeds<R, false>(e[0], sp[0]);
        rp_speeds<R, false>(e[1], sp[1]);
    }
    for (int i = 0; i < 2; i++) rp_solve<R, SED, FB>(P, mobile[i], e[i], sp[i], Dm[i], Dp[i]);
    // dry lanes are on the coast: away from it no lane of the wave has one
    if (wave_any(dry[0] || dry[1]))
        for (int i = 0; i < 2; i++)
            for (int f = 0; f < NF; f++) {
                Dm[i][f] = dry[i] ? (R)0 : Dm[i][f];
                Dp[i][f] = dry[i] ? (R)0 : Dp[i][f];
            }
}

// Reconstructor.cu:57-88 with SedFlux.cuh:51-99: exchange rate phi between
// the bed (H) and the suspended load (hc), evaluated on U^n.
// still (uniform over the wave) says |u| = 0 on every lane: then Z = 0,
// Es = 0 / 1 = 0 and the erosion flux Fe is +0 without evaluating them.
template <typename R>
__device__ __forceinline__ R erosion_rate(const DevParams& P, R h, R modu, R hcv, bool still = false)
{
    R c = (R)0;
    if (hcv != (R)0) c = rdiv(hcv, h + (R)P.epsilon_h);
    R Fe = (R)0;
    if (!still) {
        R Z = qdiv((R)P.ka * (modu * (R)P.sqrtFR), Den<R>{(R)P.vs, (R)P.inv_vs});
        R Z5 = Z * Z * Z * Z * Z;
        R Es = rdiv((R)1.3e-7 * Z5, (R)1 + (R)4.3e-7 * Z5);
        Fe = (R)P.vs1p * Es;
        if (Fe > (R)0.1) Fe = (R)0.1;
        Fe = (Fe > (R)0) ? Fe : (R)0;
    }
    R Cb = (R)P.cbk * c;
    R Fd = (R)P.vs * Cb;
    if (Fd > (R)0.1) Fd = (R)0.1;
    Fd = (Fd > (R)0) ? Fd : (R)0;
    return Fe - Fd;
}

// sqrt(x) for x >= 0: the unscaled square root unless a lane of the wave
// has an argument below its range (FP64); the solver's square root (FP32).
__device__ __forceinline__ double sqrt_voted(double x)
{
    return wave_any(sqrt_needs_scaling(x)) ? sqrt(x) : sqrt_unscaled0(x);
}
__device__ __forceinline__ float sqrt_voted(float x) { return rsqrt_(x); }

// Volume_kernel.cu:19-43 in two parts: the wave speeds of one cell and the
// dt of given speeds, cell_dt = dt_of_speeds(cell_speeds).  Division and
// multiplication by a positive constant round monotonically, so the minimum
// of cell_dt over a set of cells is dt_of_speeds of the set's largest speeds
// (a dry cell has speeds 0 and dt = inf either way; speeds are >= 0): the
// stage kernel reduces speeds and divides once per wave.
template <typename R>
__device__ __forceinline__ void cell_speeds(R h, R qx, R qy, R& lx, R& ly)
{
    Vel<R> u = velocity(h, qx, qy);
    R cg = sqrt_voted((R)kG * h);
    lx = rabs(u.x) + cg;
    ly = rabs(u.y) + cg;
}

template <typename R>
__device__ __forceinline__ R dt_of_speeds(const DevParams& P, R lx, R ly)
{
    // square cells: min(dx/lx, dx/ly) is dx/max(lx, ly) by the same monotonicity
    if (P.dx == P.dy) return (R)P.cfl_half * rdiv((R)P.dx, rmax(lx, ly));
    return (R)P.cfl_half * rmin(rdiv((R)P.dx, lx), rdiv((R)P.dy, ly));
}

template <typename R>
__device__ __forceinline__ R cell_dt(const DevParams& P, R h, R qx, R qy)
{
    R lx, ly;
    cell_speeds(h, qx, qy, lx, ly);
    return dt_of_speeds(P, lx, ly);
}

// Reduction_kernel_min.cu:155-196 applied to the exact minimum: for
// nvol > 4096 the reference's partial minima pass through fminf (rounding is
// monotone, so this is the float-rounded minimum) and the host minimum starts
// from 1e30.
__device__ __forceinline__ double finalize_dt(const DevParams& P, unsigned long long bits)
{
    double m = __longlong_as_double((long long)bits);
    if (P.dt_narrow && P.nvol > 4096) m = (double)(float)m;
    return (m < 1e30) ? m : 1e30;
}

// VolumeSolver.cu:172-202: linear-in-time ghost state from the CC series at
// boundary slot idx (eta, u, v, H per slot), h = eta + H, q = u*h, hc = 0.
__device__ __forceinline__ void inflow_state(const DevParams& P, double ct, long long idx, double* g)
{
    const double* CC = P.CC;
    double dtc = P.dtcc;
    int t = int((ct - P.tcc0) / dtc);
    t *= (t > 0);
    // the reference assumes uniform spacing; keep t + 1 inside the series
    // for any spacing (no effect on uniform series while t < tCC[T-2])
    if (t > (int)P.timeCC - 2) t = (int)P.timeCC - 2;
    long long a = (long long)t * P.ring * 4 + idx * 4;
    long long b2 = (long long)(t + 1) * P.ring * 4 + idx * 4;
    double tt = P.tCC[t];
    double m, b;
    m = (CC[b2] - CC[a]) / dtc;
    b = CC[a] - m * tt;
    double h1 = m * ct + b;
    m = (CC[b2 + 3] - CC[a + 3]) / dtc;
    b = CC[a + 3] - m * tt;
    double zb1 = m * ct + b;
    h1 += zb1;
    m = (CC[b2 + 1] - CC[a + 1]) / dtc;
    b = CC[a + 1] - m * tt;
    double q1x = m * ct + b;
    q1x *= h1;
    m = (CC[b2 + 2] - CC[a + 2]) / dtc;
    b = CC[a + 2] - m * tt;
    double q1y = m * ct + b;
    q1y *= h1;
    g[0] = h1;
    g[1] = q1x;
    g[2] = q1y;
    g[3] = zb1;
    g[4] = 0.0;
}

// Ghost state of a boundary edge (VolumeSolver.cu:233-255 and its three
// analogues): copy of the inner face, normal momentum component ax (1 = qx on
// x sides, 2 = qy on y sides) negated for walls, inflow series while
// t < tCC[timeCC-2] and a wall afterwards.
template <typename R>
__device__ __forceinline__ void ghost_state(const DevParams& P, int side, const R* in, R* g, long long idx, int ax,
                                            double ct)
{
    int type = P.bc[side];
    for (int f = 0; f < 5; f++) g[f] = in[f];
    if (type == BC_WALL) g[ax] = -in[ax];
    if (type == BC_INFLOW) {
        if (ct < P.tcc_end) {
            double gd[5];
            inflow_state(P, ct, idx, gd);
            for (int f = 0; f < 5; f++) g[f] = (R)gd[f];
            if constexpr (sizeof(R) == 4) {   // FP32 faces carry velocities
                const Vel<float> u = velocity(g[0], g[1], g[2]);
                g[1] = u.x;
                g[2] = u.y;
            }
        } else {
            g[ax] = -in[ax];
        }
    }
}

}  // namespace tsu
