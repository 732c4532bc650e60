// tsu_device.h -- device-side arithmetic of the TSUNER per-step solver for
// gfx950 (CDNA4), shared by the fused stage kernels and the known-answer
// kernels.
//
// Every function computes the same IEEE value as the reference expression it
// cites (no FMA contraction: the library is built with -ffp-contract=off, as
// the reference is built with -fmad=false, GPU/makefile:10).  Where this code
// is shaped differently from the reference, the change is value-preserving:
//   * products with a 0/1 gate, e.g. x*(h>=EPS), become a select: the reference
//     yields +-0 for a closed gate, this code yields +0 (equal as values);
//   * rotations by the axis-aligned edge normals are signed permutations
//     (q*1 + q'*0 equals q as a value);
//   * a quantity the reference recomputes from identical operands (the cell
//     velocity, |u|) is computed once.
// Results therefore equal the reference's as values (+0 == -0), which is how
// the parity tests compare them.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tsu {

// Sim_parameters.hxx:28-36
constexpr double kEpsilon = 2.2204460492503131e-16;  // DBL_EPSILON
constexpr double kG = 9.81;
constexpr double kEps = 1e-6;
constexpr double kVmax = 10.0;

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
    double ka;              // alpha1 * pow(Rp, alpha2)   (SedFlux.cuh:56-65)
    double sqrtFR;
    double cbk;             // 1.64 + 0.4*pow(DS/Dsg, 1.64) (SedFlux.cuh:83)
    double vs1p;            // vs * (1 - porosity)  (SedFlux.cuh:70)
    double grass_k;         // A_g / (1 - porosity)
    double manning_n;
    int bedload, suspension, limiter, bedload_law, friction_law, order, dt_narrow, extrema;
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
    long long pad[3];
};

__device__ __forceinline__ int sgn(double x) { return (fabs(x) < kEpsilon) ? 0 : ((x > 0) ? 1 : -1); }

// Reconstructor.cu:19-31.  When a and b do not share a strict sign the
// numerator a|b| + b|a| is exactly zero (for finite products), so the
// division is only evaluated for same-sign slopes.
__device__ __forceinline__ double lim_vanleer(double a, double b)
{
    double aa = fabs(a), bb = fabs(b);
    double s = aa + bb;
    double d = 0.0;
    if ((a > 0.0 && b > 0.0) || (a < 0.0 && b < 0.0)) {
        if (s > kEpsilon) d = (a * bb + b * aa) / s;
    }
    return d;
}

__device__ __forceinline__ double lim_minmod(double a, double b)
{
    if (a > 0.0 && b > 0.0) return fmin(a, b);
    if (a < 0.0 && b < 0.0) return fmax(a, b);
    return 0.0;
}

__device__ __forceinline__ double limiter(int law, double a, double b)
{
    return law == LIM_MINMOD ? lim_minmod(a, b) : lim_vanleer(a, b);
}

// sqrt is monotone and correctly rounded, so sqrt(s) > 10 exactly when
// s > nextafter(100, +inf).
constexpr double kVmax2Cut = 100.00000000000001421085471520200371742248535156250;

struct Vel {
    double x, y;
    double s;   // u.x*u.x + u.y*u.y of the returned (possibly capped) velocity
};

// |u| as the reference evaluates it after u_calculator (Reconstructor.cu:43,74;
// SedFlux.cuh:157): sqrt(u.x*u.x + u.y*u.y).
__device__ __forceinline__ double speed(const Vel& u) { return sqrt(u.s); }

// SedFlux.cuh:26-49.  Divisions with a zero numerator are skipped (0/d is
// 0); the cap test uses s instead of sqrt(s) (see kVmax2Cut).
__device__ __forceinline__ Vel velocity(double h, double qx, double qy)
{
    Vel u;
    u.x = 0.0;
    u.y = 0.0;
    if (h >= kEps) {
        double d = h + kEpsilon;
        if (qx != 0.0) u.x = qx / d;
        if (qy != 0.0) u.y = qy / d;
    }
    double s = u.x * u.x + u.y * u.y;
    if (s > kVmax2Cut) {
        double m = sqrt(s);
        u.x = u.x / m * kVmax;
        u.y = u.y / m * kVmax;
        s = u.x * u.x + u.y * u.y;
    }
    u.s = s;
    return u;
}

// SedFlux.cuh:118-132
template <bool SED>
__device__ __forceinline__ void state_filter(double& h, double& qx, double& qy, double& hc)
{
    h = (h > 0.0) ? h : 0.0;
    double c = 0.0;
    if (SED) {
        if (hc != 0.0) c = hc / (h + kEpsilon);
        c = (c > 0.0) ? c : 0.0;
    }
    Vel u = velocity(h, qx, qy);
    qx = h * u.x;
    qy = h * u.y;
    if (SED) hc = h * c;
}

// Bed-plane flux before the mask: SedFlux.cuh:150-168 (Ashida-Michiue), or
// Grass q_b = A_g/(1-p) u_n |u|^2.  Sign as in the reference (returns -q_b).
__device__ __forceinline__ double bedload(const DevParams& P, double h, const Vel& u)
{
    if (!(h > P.HQB && h > P.fiveDS)) return 0.0;
    double m = speed(u);
    if (P.bedload_law == BEDLOAD_GRASS) return -(P.grass_k * u.x * m * m);
    double O = P.Opar * m * m;
    if (!(O > P.OC)) return 0.0;
    double K1 = 17.0 * u.x / (m + kEpsilon);
    double qb = K1 * P.Q * (O - P.OC) * fmax(sqrt(O) - P.sqrtOC, 0.0);
    return -qb;
}

// Unrotated (qx, qy) -> normal frame (qn, qt) for the four axis normals
// (VolumeSolver.cu:28-31 with normal_x, normal_y in {-1, 0, 1}).
__device__ __forceinline__ void to_normal(int kind, double qx, double qy, double& qn, double& qt)
{
    switch (kind) {
    case N_PX: qn = qx; qt = qy; break;
    case N_MX: qn = -qx; qt = -qy; break;
    case N_PY: qn = qy; qt = -qx; break;
    default:   qn = -qy; qt = qx; break;
    }
}

// Normal frame -> unrotated (VolumeSolver.cu:142-152).
__device__ __forceinline__ void from_normal(int kind, double a, double b, double& fx, double& fy)
{
    switch (kind) {
    case N_PX: fx = a; fy = b; break;
    case N_MX: fx = -a; fy = -b; break;
    case N_PY: fx = -b; fy = a; break;
    default:   fx = b; fy = -a; break;
    }
}

// VolumeSolver.cu:19-126 in the normal frame: hydrostatic reconstruction and
// the HLL/PVM path-conservative fluctuations Dm (to the left cell) and Dp (to
// the right cell).  Inputs: left/right face (h, H, qn, qt, hc).  With SED
// false the bed and concentration rows are identically zero (no bedload, no
// suspension, hc == 0) and are not computed.
template <bool SED>
__device__ __forceinline__ void riemann(const DevParams& P, bool bed_mobile,
                                        double hl, double zl, double qnl, double qtl, double hcl,
                                        double hr, double zr, double qnr, double qtr, double hcr,
                                        double* Dm, double* Dp)
{
    double zmH = fmin(zl, zr);
    double h0 = fmax(hl - zl + zmH, 0.0);
    double h1 = fmax(hr - zr + zmH, 0.0);
    if (h0 == 0.0 && h1 == 0.0) {
        // Both reconstructed depths vanish: every flux, jump and bed term
        // below is a product with h0, h1, their difference or a zero jump,
        // so Dm = Dp = 0 (the speeds stay finite: |u| <= VMAX).
        for (int f = 0; f < (SED ? 5 : 3); f++) Dm[f] = Dp[f] = 0.0;
        return;
    }
    Vel u0 = velocity(hl, qnl, qtl);
    Vel u1 = velocity(hr, qnr, qtr);
    double c0 = 0.0, c1 = 0.0;
    if (SED) {
        if (hcl != 0.0) c0 = hcl / (hl + kEpsilon);
        if (hcr != 0.0) c1 = hcr / (hr + kEpsilon);
    }
    double q0n = u0.x * h0, q0t = u0.y * h0;
    double q1n = u1.x * h1, q1t = u1.y * h1;
    double deta = h1 - h0;

    double F00 = h0 * u0.x, F01 = F00 * u0.x, F02 = F00 * u0.y;
    double F10 = h1 * u1.x, F11 = F10 * u1.x, F12 = F10 * u1.y;

    // Roe average (SedFlux.cuh:181-189) and HLL speeds (VolumeSolver.cu:63-73)
    double hm = 0.5 * (h0 + h1);
    double SL = u0.x - sqrt(kG * h0);
    double SR = u1.x + sqrt(kG * h1);
    double cm = sqrt(kG * hm);
    double sh0 = sqrt(h0), sh1 = sqrt(h1);
    double ssum = sh0 + sh1;
    if (ssum > 0.0) {  // both sides dry: the reference's 0/0 is dropped by fmin/fmax
        double umx = (sh0 * u0.x + sh1 * u1.x) / ssum;
        SL = fmin(SL, umx - cm);
        SR = fmax(SR, umx + cm);
    }

    double a0, a1;
    double dS = SR - SL;
    if (dS > kEpsilon) {
        a0 = (SR * fabs(SL) - SL * fabs(SR)) / dS;
        a1 = (fabs(SR) - fabs(SL)) / dS;
    } else {
        a0 = fmax(fabs(SL), fabs(SR));
        a1 = 0.0;
    }

    double df0 = F10 - F00, df1 = F11 - F01, df2 = F12 - F02;
    df1 += kG * hm * deta;
    double DU0 = h1 - h0, DU1 = q1n - q0n, DU2 = q1t - q0t;
    double v0 = a0 * DU0 + a1 * df0;
    double v1 = a0 * DU1 + a1 * df1;
    double v2 = a0 * DU2 + a1 * df2;
    Dm[0] = 0.5 * (df0 - v0) + F00;
    Dm[1] = 0.5 * (df1 - v1) + F01;
    Dm[2] = 0.5 * (df2 - v2) + F02;
    Dp[0] = 0.5 * (df0 + v0) - F10;
    Dp[1] = 0.5 * (df1 + v1) - F11;
    Dp[2] = 0.5 * (df2 + v2) - F12;

    if (SED) {
        double hc0 = h0 * c0, hc1 = h1 * c1;
        double F03 = 0.0, F13 = 0.0;
        if (bed_mobile) {
            F03 = bedload(P, h0, u0);
            F13 = bedload(P, h1, u1);
        }
        double F04 = hc0 * u0.x, F14 = hc1 * u1.x;
        double df3 = F13 - F03, df4 = F14 - F04;
        double DU4 = hc1 - hc0;
        int s0 = sgn(zr - zl), s1 = sgn(df3), s2 = sgn(a1);
        s1 = (s0 != 0) ? s1 * s0 : 0;
        double v3 = (s2 == s1) ? a1 * df3 : -a1 * df3;
        double v4 = a0 * DU4 + a1 * df4;
        Dm[3] = 0.5 * (df3 - v3) + F03;
        Dm[4] = 0.5 * (df4 - v4) + F04;
        Dp[3] = 0.5 * (df3 + v3) - F13;
        Dp[4] = 0.5 * (df4 + v4) - F14;
    }
}

// Reconstructor.cu:57-88 with SedFlux.cuh:51-99: exchange rate phi between
// the bed (H) and the suspended load (hc), evaluated on U^n.
__device__ __forceinline__ double erosion_rate(const DevParams& P, double h, double modu, double hcv)
{
    double c = 0.0;
    if (hcv != 0.0) c = hcv / (h + P.epsilon_h);
    double Z = P.ka * (modu * P.sqrtFR) / P.vs;
    double Z5 = Z * Z * Z * Z * Z;
    double Es = (1.3e-7 * Z5) / (1 + 4.3e-7 * Z5);
    double Fe = P.vs1p * Es;
    if (Fe > 0.1) Fe = 0.1;
    Fe = (Fe > 0.0) ? Fe : 0.0;
    double Cb = P.cbk * c;
    double Fd = P.vs * Cb;
    if (Fd > 0.1) Fd = 0.1;
    Fd = (Fd > 0.0) ? Fd : 0.0;
    return Fe - Fd;
}

// Volume_kernel.cu:19-43
__device__ __forceinline__ double cell_dt(const DevParams& P, double h, double qx, double qy)
{
    Vel u = velocity(h, qx, qy);
    double cg = sqrt(kG * h);
    double lx = fabs(u.x) + cg;
    double ly = fabs(u.y) + cg;
    return P.cfl_half * fmin(P.dx / lx, P.dy / ly);
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
__device__ __forceinline__ void ghost_state(const DevParams& P, int side, const double* in, double* g,
                                            long long idx, int ax, double ct)
{
    int type = P.bc[side];
    for (int f = 0; f < 5; f++) g[f] = in[f];
    if (type == BC_WALL) g[ax] = -in[ax];
    if (type == BC_INFLOW) {
        if (ct < P.tcc_end)
            inflow_state(P, ct, idx, g);
        else
            g[ax] = -in[ax];
    }
}

}  // namespace tsu
