/*
 * tsu_oracle.c -- CPU oracle for the TSUNER per-step solver.
 *
 * TEST INFRASTRUCTURE ONLY (see tsu_oracle.h): the checker for the HIP
 * product path and the timed serial CPU baseline.  Nothing in the product
 * library links, loads or calls this file.
 *
 * Every function restates one piece of the reference, cited file:line, and
 * keeps the reference's operation order so that results agree bit for bit
 * with an FMA-free FP64 evaluation of the reference kernels.  The arrays
 * mirror the reference pipeline (face arrays reconsL/R/D/U and per-cell edge
 * slots FL/FR/FD/FU, GPU/main_TSUNER.cu:57-91) on purpose: the oracle is
 * structured like the reference, the product is not.
 *
 * Run constants that the reference recomputes per cell (pow(), sqrt() of
 * parameters) are evaluated once here; they are pure functions of the
 * parameters, so the values are the same.
 */
#include "tsu_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Sim_parameters.hxx:28-36 */
#define O_EPSILON DBL_EPSILON
#define O_G       9.81
#define O_EPS     1e-6
#define O_VMAX    10.0

/* Sim_parameters.hxx:29 -- SGN with an absolute DBL_EPSILON dead band. */
static int o_sgn(double x) { return (fabs(x) < O_EPSILON) ? 0 : ((x) > 0) ? 1 : -1; }

typedef struct {
    /* erosion / deposition run constants (Reconstructor.cu:64-77, SedFlux.cuh:51-99) */
    double vs, ka, sqrtFR, cbk, one_minus_p;
    double sqrtOC, fiveDS;
    double grass_k;  /* A_g / (1 - porosity) */
} o_consts;

struct orc_sim {
    orc_sim_params p;
    orc_scheme s;
    o_consts k;
    long nvol;
    double *V, *V0, *Vk;
    double *rL, *rR, *rD, *rU;     /* reconsL/R/D/U */
    double *FL, *FR, *FD, *FU;
    double *dtv;
    double *hmax, *bmax, *etamax;
    int *fixH;
    double *CC, *tCC;
    long timeCC;
    double t, dt;
    long steps;
};

/* ------------------------------------------------------------------ */
/* Point functions                                                     */
/* ------------------------------------------------------------------ */

/* Reconstructor.cu:19-31 -- harmonic (van Leer) slope average. */
static double o_avg(double a, double b)
{
    double d, aa, bb;
    aa = fabs(a);
    bb = fabs(b);
    if ((aa + bb) > O_EPSILON)
        d = (a * bb + b * aa) / (aa + bb);
    else
        d = 0.0;
    return d;
}

/* minmod, the limiter named by the plain-SWE benchmark configuration. */
static double o_minmod(double a, double b)
{
    if (a > 0.0 && b > 0.0) return fmin(a, b);
    if (a < 0.0 && b < 0.0) return fmax(a, b);
    return 0.0;
}

double orc_limiter(int law, double a, double b)
{
    return law == ORC_LIM_MINMOD ? o_minmod(a, b) : o_avg(a, b);
}

/* SedFlux.cuh:26-49 -- velocity with near-dry desingularisation and cap. */
void orc_velocity(double h, double qx, double qy, double *ux, double *uy)
{
    double modu;
    double x = qx / (h + O_EPSILON) * (h >= O_EPS);
    double y = qy / (h + O_EPSILON) * (h >= O_EPS);
    modu = sqrt(x * x + y * y);
    if (modu > O_VMAX) {
        x = x / modu * O_VMAX;
        y = y / modu * O_VMAX;
    }
    *ux = x;
    *uy = y;
}

/* SedFlux.cuh:103-132 */
void orc_state_filter(double *h, double *qx, double *qy, double *hc)
{
    double ux, uy, c;
    *h = *h * (*h > 0.0);
    c = *hc / (*h + O_EPSILON);
    c = c * (c > 0.0);
    orc_velocity(*h, *qx, *qy, &ux, &uy);
    *qx = *h * ux;
    *qy = *h * uy;
    *hc = *h * c;
}

static void o_consts_init(o_consts *k, const orc_sim_params *p, const orc_scheme *s)
{
    /* Reconstructor.cu:64-77 */
    double nu = 1.002e-6;
    double sd = p->R - 1.0;
    double Rp;
    k->vs = sqrt((13.95 * nu / p->DS) * (13.95 * nu / p->DS) + 1.09 * sd * O_G * p->DS) - 13.95 * nu / p->DS;
    Rp = (p->DS * sqrt(sd * O_G * p->DS)) / nu;
    /* SedFlux.cuh:56-65: Z = alpha1 * pow(Rp, alpha2) * (modu*sqrt(FR)) / vs */
    if (Rp > 2.36)
        k->ka = 1.0 * pow(Rp, 0.6);
    else
        k->ka = 0.586 * pow(Rp, 1.23);
    k->sqrtFR = sqrt(p->FR);
    /* SedFlux.cuh:83 */
    k->cbk = 1.64 + 0.4 * pow(p->DS / p->Dsg, 1.64);
    k->one_minus_p = 1 - p->porosity;
    k->sqrtOC = sqrt(p->OC);
    k->fiveDS = 5 * p->DS;
    k->grass_k = s ? s->grass_ag / (1.0 - p->porosity) : 0.0;
}

/* SedFlux.cuh:150-168 (Ashida-Michiue), or Grass q_b = A_g/(1-p) u_n |u|^2.
 * Returns the bed-plane flux F3 before the mask (already negated: the
 * evolved plane is H = -z). */
static double o_bedload(const orc_sim_params *p, const orc_scheme *s, const o_consts *k,
                        double h, double unx, double uny)
{
    double qb, K1, O, u, modu;
    u = unx;
    modu = sqrt(unx * unx + uny * uny);
    if (s->bedload_law == ORC_BEDLOAD_GRASS) {
        qb = k->grass_k * u * modu * modu * (h > p->HQB) * (h > k->fiveDS);
        return -qb;
    }
    O = p->Opar * modu * modu;
    K1 = 17.0 * u / (modu + O_EPSILON);
    qb = K1 * p->Q * fmax(O - p->OC, 0.0) * fmax(sqrt(O) - k->sqrtOC, 0.0) * (h > p->HQB) * (h > k->fiveDS);
    return -qb;
}

double orc_bedload_flux(const orc_sim_params *p, const orc_scheme *s, double h, double un, double ut)
{
    o_consts k;
    o_consts_init(&k, p, s);
    return o_bedload(p, s, &k, h, un, ut);
}

/* SedFlux.cuh:171-179 */
static void o_cflux(const orc_sim_params *p, const orc_scheme *s, const o_consts *k,
                    double h, double ux, double uy, double hc, double *F, int fixed)
{
    F[0] = h * ux;
    F[1] = F[0] * ux;
    F[2] = F[0] * uy;
    F[3] = o_bedload(p, s, k, h, ux, uy) * fixed * s->bedload;
    F[4] = hc * ux;
}

/* VolumeSolver.cu:19-126 -- hydrostatic reconstruction + HLL/PVM
 * path-conservative flux for one edge.  left/right are (h, qx, qy, H, hc)
 * in the unrotated frame; Dm/Dp come back rotated to the unrotated frame
 * (VolumeSolver.cu:128-169). */
static void o_riemann(const orc_sim_params *p, const orc_scheme *s, const o_consts *k,
                      const double *L, const double *R, int normal_x, int normal_y, int fixed,
                      double *Dmo, double *Dpo)
{
    double hhl = L[0], zb0 = L[3], hcl = L[4];
    double hhr = R[0], zb1 = R[3], hcr = R[4];
    double qlx, qly, qrx, qry;
    double u0x, u0y, u1x, u1y;
    double h0, h1, c0, c1, hc0, hc1;
    double q0x, q0y, q1x, q1y;
    double zmH, deta;
    double F0[5], F1[5];
    double hm, umx, ev_c;
    double SL, SR, a0, a1;
    double df[5], vis[5], DU[5], Dm[5], Dp[5];
    double s0, s1, s2;
    int i;

    qlx = L[1] * normal_x + L[2] * normal_y;
    qly = L[2] * normal_x - L[1] * normal_y;
    qrx = R[1] * normal_x + R[2] * normal_y;
    qry = R[2] * normal_x - R[1] * normal_y;
    orc_velocity(hhl, qlx, qly, &u0x, &u0y);
    orc_velocity(hhr, qrx, qry, &u1x, &u1y);
    c0 = hcl / (hhl + O_EPSILON);
    c1 = hcr / (hhr + O_EPSILON);

    zmH = fmin(zb0, zb1);
    h0 = fmax(hhl - zb0 + zmH, 0.0);
    h1 = fmax(hhr - zb1 + zmH, 0.0);
    q0x = u0x * h0;
    q0y = u0y * h0;
    q1x = u1x * h1;
    q1y = u1y * h1;
    hc0 = h0 * c0;
    hc1 = h1 * c1;
    deta = h1 - h0;

    o_cflux(p, s, k, h0, u0x, u0y, hc0, F0, fixed);
    o_cflux(p, s, k, h1, u1x, u1y, hc1, F1, fixed);

    /* SedFlux.cuh:181-189 (only hm and um.x are used downstream) */
    hm = 0.5 * (h0 + h1);
    umx = (sqrt(h0) * u0x + sqrt(h1) * u1x) / (sqrt(h0) + sqrt(h1));

    /* SedFlux.cuh:192-199 and VolumeSolver.cu:63-73 */
    ev_c = sqrt(O_G * h0);
    SL = u0x - ev_c;
    ev_c = sqrt(O_G * h1);
    SR = u1x + ev_c;
    ev_c = sqrt(O_G * hm);
    SL = fmin(SL, umx - ev_c);
    SR = fmax(SR, umx + ev_c);

    if (SR - SL > O_EPSILON) {
        a0 = (SR * fabs(SL) - SL * fabs(SR)) / (SR - SL);
        a1 = (fabs(SR) - fabs(SL)) / (SR - SL);
    } else {
        a0 = fmax(fabs(SL), fabs(SR));
        a1 = 0.0;
    }

    for (i = 0; i < 5; i++) df[i] = F1[i] - F0[i];
    df[1] += O_G * hm * deta;

    DU[0] = h1 - h0;
    DU[1] = q1x - q0x;
    DU[2] = q1y - q0y;
    DU[3] = zb1 - zb0;
    DU[4] = hc1 - hc0;

    s0 = o_sgn(zb1 - zb0);
    s1 = o_sgn(df[3]);
    s2 = o_sgn(a1);
    if (fabs(s0) > O_EPSILON)
        s1 = s1 / s0;
    else
        s1 = 0.0;

    for (i = 0; i < 5; i++) vis[i] = a0 * DU[i] + a1 * df[i];
    if (s2 == s1)
        vis[3] = a1 * df[3];
    else
        vis[3] = -a1 * df[3];

    for (i = 0; i < 5; i++) {
        Dm[i] = 0.5 * (df[i] - vis[i]) + F0[i];
        Dp[i] = 0.5 * (df[i] + vis[i]) - F1[i];
    }
    Dmo[0] = Dm[0];
    Dmo[1] = (normal_x * Dm[1] - normal_y * Dm[2]);
    Dmo[2] = (normal_y * Dm[1] + normal_x * Dm[2]);
    Dmo[3] = Dm[3];
    Dmo[4] = Dm[4];
    Dpo[0] = Dp[0];
    Dpo[1] = (normal_x * Dp[1] - normal_y * Dp[2]);
    Dpo[2] = (normal_y * Dp[1] + normal_x * Dp[2]);
    Dpo[3] = Dp[3];
    Dpo[4] = Dp[4];
}

void orc_edge_flux(const orc_sim_params *p, const orc_scheme *s,
                   const double left[5], const double right[5],
                   int normal_x, int normal_y, int fixed, double Dm[5], double Dp[5])
{
    o_consts k;
    o_consts_init(&k, p, s);
    o_riemann(p, s, &k, left, right, normal_x, normal_y, fixed, Dm, Dp);
}

/* Volume_kernel.cu:19-43 */
double orc_cell_dt(const orc_sim_params *p, double h, double qx, double qy)
{
    double ux, uy, cg, lx, ly;
    orc_velocity(h, qx, qy, &ux, &uy);
    cg = sqrt(O_G * h);
    lx = fabs(ux) + cg;
    ly = fabs(uy) + cg;
    return p->CFL * 0.5 * fmin(p->dx / lx, p->dy / ly);
}

/* Reduction_kernel_min.cu:155-196.  For n > 4096 every partial minimum
 * passes through fminf, i.e. is rounded to float (rounding is monotone, so
 * this equals the float-rounded global minimum); the host then takes the
 * minimum of the partials starting from 1e30.  For n <= 4096 the minimum is
 * taken on the host in double, again starting from 1e30.  NaN never wins. */
double orc_min_reduction(const double *dtv, long n, int narrow)
{
    double m = INFINITY;
    long i;
    for (i = 0; i < n; i++)
        if (dtv[i] < m) m = dtv[i];
    if (narrow && n > 4096) m = (double)(float)m;
    return (m < 1e30) ? m : 1e30;
}

/* Reconstructor.cu:57-88 -- erosion/deposition exchange rate phi (before dt). */
static double o_erosion_rate(const orc_sim_params *p, const orc_scheme *s, const o_consts *k,
                             double h, double qx, double qy, double hcv)
{
    double Fe, Fd, c, ux, uy, modu, Z, Z5, Es, Cb;
    c = hcv / (h + p->epsilon_h);
    orc_velocity(h, qx, qy, &ux, &uy);
    modu = sqrt(ux * ux + uy * uy);
    /* SedFlux.cuh:51-77 */
    Z = k->ka * (modu * k->sqrtFR) / k->vs;
    Z5 = Z * Z * Z * Z * Z;
    Es = (1.3e-7 * Z5) / (1 + 4.3e-7 * Z5);
    Fe = k->vs * k->one_minus_p * Es;
    if (Fe > 0.1) Fe = 0.1;
    Fe = Fe * (Fe > 0.0);
    /* SedFlux.cuh:79-99 */
    Cb = k->cbk * c;
    Fd = k->vs * Cb;
    if (Fd > 0.1) Fd = 0.1;
    Fd = Fd * (Fd > 0.0);
    return (Fe - Fd) * s->suspension;
}

void orc_erosion_deposition(const orc_sim_params *p, const orc_scheme *s,
                            double h, double qx, double qy, double hc, double *phi)
{
    o_consts k;
    o_consts_init(&k, p, s);
    *phi = o_erosion_rate(p, s, &k, h, qx, qy, hc);
}

void orc_derive_params(orc_sim_params *p)
{
    /* Problem_loader.cxx:337-342 (Manning == 0 branch) */
    p->FR /= 8.0;
    p->HQB = 5.0 * p->DS;
    p->Q = p->DS * sqrt(O_G * (p->R - 1.0) * p->DS) / (1.0 - p->porosity);
    p->Opar = p->FR / ((p->R - 1.0) * O_G * p->DS);
}

/* ------------------------------------------------------------------ */
/* Grid kernels                                                        */
/* ------------------------------------------------------------------ */

/* Reconstructor.cu:33-55 (quadratic law), or semi-implicit Manning
 * c = dt g n^2 |u| / (h+eps)^(4/3) with the same 1/(1+c) update. */
static void o_friction(orc_sim *S, long pos, double dt)
{
    long nv = S->nvol;
    double h = S->V0[pos], ux, uy, modu, c;
    orc_velocity(h, S->V0[1 * nv + pos], S->V0[2 * nv + pos], &ux, &uy);
    modu = sqrt(ux * ux + uy * uy);
    if (S->s.friction_law == ORC_FRIC_MANNING) {
        double hp = h + O_EPSILON;
        c = dt * O_G * S->s.manning_n * S->s.manning_n * modu / (hp * cbrt(hp));
    } else {
        c = dt * S->p.FR * modu / (h + O_EPSILON);
    }
    S->Vk[1 * nv + pos] /= (1.0 + c);
    S->Vk[2 * nv + pos] /= (1.0 + c);
}

/* Reconstructor.cu:57-88 */
static void o_erosion(orc_sim *S, long pos, double dt)
{
    long nv = S->nvol;
    double phi = o_erosion_rate(&S->p, &S->s, &S->k, S->V0[pos], S->V0[1 * nv + pos],
                                S->V0[2 * nv + pos], S->V0[4 * nv + pos]);
    S->Vk[3 * nv + pos] -= dt * phi;
    S->Vk[4 * nv + pos] += dt * phi;
}

/* Neighbour index along one axis: clamped (Reconstructor.cu:128-129,
 * 207-208) or wrapped for periodic sides (:130-133, :209-212). */
static int o_nb(int i, int n, int dir, int periodic)
{
    if (dir < 0) {
        if (i > 0) return i - 1;
        return periodic ? n - 1 : 0;
    }
    if (i < n - 1) return i + 1;
    return periodic ? 0 : n - 1;
}

/* One direction of Reconstructor.cu:120-197 (x) / :203-274 (y).
 * Writes the two faces (minus side F1, plus side F2) and returns deta. */
static double o_recons_dir(orc_sim *S, long pos, long pl, long pr, double uix, double uiy,
                           double *Fm, double *Fp)
{
    long nv = S->nvol;
    const double *V = S->V;
    int lim = S->s.limiter;
    double hi = V[pos], Hi = V[3 * nv + pos], ci = V[4 * nv + pos] / (hi + O_EPSILON);
    double hl = V[pl], Hl = V[3 * nv + pl], cl = V[4 * nv + pl] / (hl + O_EPSILON);
    double hr = V[pr], Hr = V[3 * nv + pr], cr = V[4 * nv + pr] / (hr + O_EPSILON);
    double dl, dr, dh, zm, deta, dz, du, dq;
    double ulx, uly, urx, ury;

    dl = hi - hl;
    dr = hr - hi;
    dh = orc_limiter(lim, dl, dr);
    if (0.5 * fabs(dh) > hi) dh = 2.0 * hi * o_sgn(dh);
    Fm[0] = hi - 0.5 * dh;
    Fp[0] = hi + 0.5 * dh;

    zm = fmin(Hl, Hi);
    dl = fmax(hi - Hi + zm, 0.0) - fmax(hl - Hl + zm, 0.0);
    zm = fmin(Hr, Hi);
    dr = fmax(hr - Hr + zm, 0.0) - fmax(hi - Hi + zm, 0.0);
    deta = orc_limiter(lim, dl, dr);
    dz = dh - deta;
    Fm[3] = Hi - 0.5 * dz;
    Fp[3] = Hi + 0.5 * dz;

    orc_velocity(hl, V[1 * nv + pl], V[2 * nv + pl], &ulx, &uly);
    orc_velocity(hr, V[1 * nv + pr], V[2 * nv + pr], &urx, &ury);
    dl = uix - ulx;
    dr = urx - uix;
    du = orc_limiter(lim, dl, dr);
    dq = hi * du + uix * dh;
    Fm[1] = V[1 * nv + pos] - 0.5 * dq;
    Fp[1] = V[1 * nv + pos] + 0.5 * dq;

    dl = uiy - uly;
    dr = ury - uiy;
    du = orc_limiter(lim, dl, dr);
    dq = hi * du + uiy * dh;
    Fm[2] = V[2 * nv + pos] - 0.5 * dq;
    Fp[2] = V[2 * nv + pos] + 0.5 * dq;

    dl = ci - cl;
    dr = cr - ci;
    du = orc_limiter(lim, dl, dr);
    Fm[4] = (ci - 0.5 * du) * Fm[0];
    Fp[4] = (ci + 0.5 * du) * Fp[0];
    return deta;
}

/* One cell's two faces along one axis, as the reconstructor forms them
 * (Reconstructor.cu:120-197 for x, :203-274 for y, then the face filter of
 * :345-348).  c3 holds the minus-side neighbour, the cell and the plus-side
 * neighbour (h, q_axis, q_other, H, hc each, in the axis frame); the cell's
 * velocity is formed as Reconstructor.cu:112-118 does.  Returns deta. */
double orc_cell_faces(int limiter, const double c3[15], double Fm[5], double Fp[5])
{
    orc_sim S;
    double V[15], ux, uy, deta;
    int f, k;
    memset(&S, 0, sizeof(S));
    for (f = 0; f < 5; f++)
        for (k = 0; k < 3; k++) V[f * 3 + k] = c3[k * 5 + f];
    S.nvol = 3;
    S.V = V;
    S.s.limiter = limiter;
    orc_velocity(V[1], V[3 + 1], V[6 + 1], &ux, &uy);
    deta = o_recons_dir(&S, 1, 0, 2, ux, uy, Fm, Fp);
    orc_state_filter(&Fm[0], &Fm[1], &Fm[2], &Fm[4]);
    orc_state_filter(&Fp[0], &Fp[1], &Fp[2], &Fp[4]);
    return deta;
}

static void o_put5(double *A, long nv, long pos, const double *v)
{
    int f;
    for (f = 0; f < 5; f++) A[f * nv + pos] = v[f];
}
static void o_get5(const double *A, long nv, long pos, double *v)
{
    int f;
    for (f = 0; f < 5; f++) v[f] = A[f * nv + pos];
}

/* Reconstructor.cu:90-356 (RK == 1 path) */
static void o_reconstructor(orc_sim *S, double dt)
{
    const orc_sim_params *p = &S->p;
    long nv = S->nvol;
    int npx = p->npx, npy = p->npy;
    int perx = S->s.bc[0] == ORC_BC_PERIODIC, pery = S->s.bc[2] == ORC_BC_PERIODIC;
    double inv_dx = 1.0 / p->dx, inv_dy = 1.0 / p->dy;
    int i, j;
    for (j = 0; j < npy; j++) {
        for (i = 0; i < npx; i++) {
            long pos = (long)j * npx + i;
            if (S->s.order == 1) {
                /* Reconstructor.cu:99-107 */
                int f;
                for (f = 0; f < 5; f++)
                    S->rL[f * nv + pos] = S->rR[f * nv + pos] = S->rU[f * nv + pos] =
                        S->rD[f * nv + pos] = S->V[f * nv + pos];
            } else {
                double L[5], R[5], D[5], U[5], deta, uix, uiy, dta;
                long pl, pr;
                orc_velocity(S->V[pos], S->V[1 * nv + pos], S->V[2 * nv + pos], &uix, &uiy);
                pl = (long)j * npx + o_nb(i, npx, -1, perx);
                pr = (long)j * npx + o_nb(i, npx, +1, perx);
                dta = dt * inv_dx;
                deta = o_recons_dir(S, pos, pl, pr, uix, uiy, L, R);
                S->Vk[1 * nv + pos] -= dta * O_G * S->V[pos] * deta;

                pl = (long)o_nb(j, npy, -1, pery) * npx + i;
                pr = (long)o_nb(j, npy, +1, pery) * npx + i;
                dta = dt * inv_dy;
                deta = o_recons_dir(S, pos, pl, pr, uix, uiy, D, U);
                S->Vk[2 * nv + pos] -= dta * O_G * S->V[pos] * deta;

                /* Reconstructor.cu:345-348 (filter order L, R, U, D) */
                orc_state_filter(&L[0], &L[1], &L[2], &L[4]);
                orc_state_filter(&R[0], &R[1], &R[2], &R[4]);
                orc_state_filter(&U[0], &U[1], &U[2], &U[4]);
                orc_state_filter(&D[0], &D[1], &D[2], &D[4]);
                o_put5(S->rL, nv, pos, L);
                o_put5(S->rR, nv, pos, R);
                o_put5(S->rD, nv, pos, D);
                o_put5(S->rU, nv, pos, U);
            }
            /* Reconstructor.cu:353-354 */
            o_friction(S, pos, dt);
            o_erosion(S, pos, dt);
        }
    }
}

/* VolumeSolver.cu:172-202 -- linear-in-time boundary state from the CC
 * series.  idx is the boundary slot: bottom i, right npx+j, top npx+npy+i,
 * left 2npx+npy+j; four values per slot (eta, u, v, H). */
static void o_interpolate_time(const orc_sim *S, double ct, long idx, double *g)
{
    const double *CC = S->CC, *tCC = S->tCC;
    long ring = 2L * S->p.npx + 2L * S->p.npy;
    double dtc = tCC[1] - tCC[0];
    long t = (long)(int)((ct - tCC[0]) / dtc);
    long tp, a, b2;
    double m, b, h1, zb1, q1x, q1y;
    t *= (t > 0);
    /* bounds guard for non-uniform series (no effect on uniform ones) */
    if (t > S->timeCC - 2) t = S->timeCC - 2;
    tp = t + 1;
    a = t * ring * 4 + idx * 4;
    b2 = tp * ring * 4 + idx * 4;

    m = (CC[b2] - CC[a]) / dtc;
    b = CC[a] - m * tCC[t];
    h1 = m * ct + b;

    m = (CC[b2 + 3] - CC[a + 3]) / dtc;
    b = CC[a + 3] - m * tCC[t];
    zb1 = m * ct + b;

    h1 += zb1;

    m = (CC[b2 + 1] - CC[a + 1]) / dtc;
    b = CC[a + 1] - m * tCC[t];
    q1x = m * ct + b;
    q1x *= h1;

    m = (CC[b2 + 2] - CC[a + 2]) / dtc;
    b = CC[a + 2] - m * tCC[t];
    q1y = m * ct + b;
    q1y *= h1;

    g[0] = h1;
    g[1] = q1x;
    g[2] = q1y;
    g[3] = zb1;
    g[4] = 0.0;
}

/* Ghost state for a boundary edge (VolumeSolver.cu:233-255 and analogues).
 * ax = 1 negates qx (x sides), ax = 2 negates qy (y sides). */
static void o_ghost(const orc_sim *S, int side, const double *in, double *g, long idx, int ax)
{
    int f, type = S->s.bc[side];
    for (f = 0; f < 5; f++) g[f] = in[f];
    if (type == ORC_BC_WALL) g[ax] = -in[ax];
    if (type == ORC_BC_INFLOW) {
        if (S->t < S->tCC[S->timeCC - 2])
            o_interpolate_time(S, S->t, idx, g);
        else
            g[ax] = -in[ax];
    }
}

/* VolumeSolver.cu:217-408 */
static void o_volume_solver(orc_sim *S)
{
    const orc_sim_params *p = &S->p;
    const orc_scheme *s = &S->s;
    long nv = S->nvol;
    int npx = p->npx, npy = p->npy;
    double Lf[5], Rf[5], G5[5], Dm[5], Dp[5];
    int i, j;

    /* vertical edges i-1/2 */
    for (j = 0; j < npy; j++) {
        for (i = 0; i <= npx; i++) {
            long p0 = (long)j * npx + i - 1, p1 = (long)j * npx + i;
            if (i == 0) {
                if (s->bc[0] == ORC_BC_PERIODIC) continue;
                o_get5(S->rL, nv, p1, Lf);
                o_ghost(S, 0, Lf, G5, 2L * npx + npy + j, 1);
                o_riemann(p, s, &S->k, Lf, G5, -1, 0, S->fixH[p1], Dm, Dp);
                o_put5(S->FL, nv, p1, Dm);
            } else if (i == npx) {
                o_get5(S->rR, nv, p0, Lf);
                if (s->bc[1] == ORC_BC_PERIODIC) {
                    long q1 = (long)j * npx;
                    o_get5(S->rL, nv, q1, Rf);
                    o_riemann(p, s, &S->k, Lf, Rf, 1, 0, S->fixH[p0], Dm, Dp);
                    o_put5(S->FR, nv, p0, Dm);
                    o_put5(S->FL, nv, q1, Dp);
                } else {
                    o_ghost(S, 1, Lf, G5, (long)npx + j, 1);
                    o_riemann(p, s, &S->k, Lf, G5, 1, 0, S->fixH[p0], Dm, Dp);
                    o_put5(S->FR, nv, p0, Dm);
                }
            } else {
                o_get5(S->rR, nv, p0, Lf);
                o_get5(S->rL, nv, p1, Rf);
                o_riemann(p, s, &S->k, Lf, Rf, 1, 0, S->fixH[p0] * S->fixH[p1], Dm, Dp);
                o_put5(S->FR, nv, p0, Dm);
                o_put5(S->FL, nv, p1, Dp);
            }
        }
    }
    /* horizontal edges j-1/2 */
    for (j = 0; j <= npy; j++) {
        for (i = 0; i < npx; i++) {
            long p0 = (long)(j - 1) * npx + i, p1 = (long)j * npx + i;
            if (j == 0) {
                if (s->bc[2] == ORC_BC_PERIODIC) continue;
                o_get5(S->rD, nv, p1, Lf);
                o_ghost(S, 2, Lf, G5, (long)i, 2);
                o_riemann(p, s, &S->k, Lf, G5, 0, -1, S->fixH[p1], Dm, Dp);
                o_put5(S->FD, nv, p1, Dm);
            } else if (j == npy) {
                o_get5(S->rU, nv, p0, Lf);
                if (s->bc[3] == ORC_BC_PERIODIC) {
                    long q1 = (long)i;
                    o_get5(S->rD, nv, q1, Rf);
                    o_riemann(p, s, &S->k, Lf, Rf, 0, 1, S->fixH[p0], Dm, Dp);
                    o_put5(S->FU, nv, p0, Dm);
                    o_put5(S->FD, nv, q1, Dp);
                } else {
                    o_ghost(S, 3, Lf, G5, (long)npx + npy + i, 2);
                    o_riemann(p, s, &S->k, Lf, G5, 0, 1, S->fixH[p0], Dm, Dp);
                    o_put5(S->FU, nv, p0, Dm);
                }
            } else {
                o_get5(S->rU, nv, p0, Lf);
                o_get5(S->rD, nv, p1, Rf);
                o_riemann(p, s, &S->k, Lf, Rf, 0, 1, S->fixH[p0] * S->fixH[p1], Dm, Dp);
                o_put5(S->FU, nv, p0, Dm);
                o_put5(S->FD, nv, p1, Dp);
            }
        }
    }
}

/* Volume_kernel.cu:180-282 (RK == 1 path, no sponge layer) */
static void o_states_updater(orc_sim *S, int k, double dt)
{
    long nv = S->nvol, pos;
    double invx = dt * (1.0 / S->p.dx), invy = dt * (1.0 / S->p.dy);
    double coef0 = 0.0 * (k == 0) + 0.5 * (k == 1);
    double coef1 = 1.0 - coef0;
    for (pos = 0; pos < nv; pos++) {
        int f;
        double W[5];
        for (f = 0; f < 5; f++)
            S->Vk[f * nv + pos] -= invx * (S->FL[f * nv + pos] + S->FR[f * nv + pos]) +
                                   invy * (S->FD[f * nv + pos] + S->FU[f * nv + pos]);
        for (f = 0; f < 5; f++) W[f] = coef0 * S->V0[f * nv + pos] + coef1 * S->Vk[f * nv + pos];
        orc_state_filter(&W[0], &W[1], &W[2], &W[4]);
        o_put5(S->Vk, nv, pos, W);
        o_put5(S->V, nv, pos, W);
        if (k == S->s.order - 1) o_put5(S->V0, nv, pos, W);
    }
}

/* Volume_kernel.cu:79-110 */
static void o_extrema(orc_sim *S)
{
    long nv = S->nvol, pos;
    double eh = S->p.epsilon_h;
    for (pos = 0; pos < nv; pos++) {
        double h = S->V[pos], H = S->V[3 * nv + pos];
        if (h > eh) S->hmax[pos] = fmax(S->hmax[pos], h);
        S->bmax[pos] = fmax(S->bmax[pos], H);
        if (h > eh) S->etamax[pos] = fmax(S->etamax[pos], h - H);
    }
}

static double o_get_delta_t(orc_sim *S)
{
    long nv = S->nvol, pos;
    for (pos = 0; pos < nv; pos++)
        S->dtv[pos] = orc_cell_dt(&S->p, S->V[pos], S->V[1 * nv + pos], S->V[2 * nv + pos]);
    return orc_min_reduction(S->dtv, nv, S->s.dt_narrow);
}

/* ------------------------------------------------------------------ */
/* Driver                                                              */
/* ------------------------------------------------------------------ */

orc_sim *orc_create(const orc_sim_params *p, const orc_scheme *s,
                    const double *V, const int *fixH,
                    const double *CC, const double *tCC, long timeCC)
{
    orc_sim *S = (orc_sim *)calloc(1, sizeof(orc_sim));
    long nv = (long)p->npx * p->npy, i;
    size_t b5 = 5 * (size_t)nv * sizeof(double);
    int side;
    if (!S) return NULL;
    S->p = *p;
    S->p.nvol = (int)nv;
    S->s = *s;
    S->nvol = nv;
    o_consts_init(&S->k, &S->p, &S->s);
    S->V = (double *)malloc(b5);
    S->V0 = (double *)malloc(b5);
    S->Vk = (double *)malloc(b5);
    S->rL = (double *)malloc(b5);
    S->rR = (double *)malloc(b5);
    S->rD = (double *)malloc(b5);
    S->rU = (double *)malloc(b5);
    S->FL = (double *)malloc(b5);
    S->FR = (double *)malloc(b5);
    S->FD = (double *)malloc(b5);
    S->FU = (double *)malloc(b5);
    S->dtv = (double *)malloc(nv * sizeof(double));
    S->hmax = (double *)malloc(nv * sizeof(double));
    S->bmax = (double *)malloc(nv * sizeof(double));
    S->etamax = (double *)malloc(nv * sizeof(double));
    S->fixH = (int *)malloc(nv * sizeof(int));
    if (!S->V || !S->V0 || !S->Vk || !S->rL || !S->rR || !S->rD || !S->rU || !S->FL ||
        !S->FR || !S->FD || !S->FU || !S->dtv || !S->hmax || !S->bmax || !S->etamax || !S->fixH) {
        orc_destroy(S);
        return NULL;
    }
    memcpy(S->V, V, b5);
    memcpy(S->V0, V, b5);
    memcpy(S->Vk, V, b5);
    memcpy(S->fixH, fixH, nv * sizeof(int));
    for (i = 0; i < nv; i++) S->hmax[i] = S->bmax[i] = S->etamax[i] = -1e30;
    S->timeCC = 0;
    for (side = 0; side < 4; side++) {
        if (s->bc[side] == ORC_BC_INFLOW) {
            size_t ncc = (size_t)(2L * p->npx + 2L * p->npy) * 4 * timeCC;
            S->CC = (double *)malloc(ncc * sizeof(double));
            S->tCC = (double *)malloc(timeCC * sizeof(double));
            memcpy(S->CC, CC, ncc * sizeof(double));
            memcpy(S->tCC, tCC, timeCC * sizeof(double));
            S->timeCC = timeCC;
            break;
        }
    }
    S->t = 0.0;
    S->steps = 0;
    /* GPU/main_TSUNER.cu:147-148 */
    S->dt = o_get_delta_t(S);
    return S;
}

void orc_destroy(orc_sim *S)
{
    if (!S) return;
    free(S->V); free(S->V0); free(S->Vk);
    free(S->rL); free(S->rR); free(S->rD); free(S->rU);
    free(S->FL); free(S->FR); free(S->FD); free(S->FU);
    free(S->dtv); free(S->hmax); free(S->bmax); free(S->etamax);
    free(S->fixH); free(S->CC); free(S->tCC);
    free(S);
}

int orc_stage(orc_sim *S, int k)
{
    o_reconstructor(S, S->dt);
    o_volume_solver(S);
    o_states_updater(S, k, S->dt);
    return 0;
}

/* GPU/main_TSUNER.cu:169-192 */
int orc_step(orc_sim *S, int nsteps)
{
    int n, k;
    for (n = 0; n < nsteps; n++) {
        for (k = 0; k < S->s.order; k++) orc_stage(S, k);
        S->t += S->dt;
        if (S->s.extrema) o_extrema(S);
        S->dt = o_get_delta_t(S);
        S->steps++;
    }
    return 0;
}

void orc_get_state(const orc_sim *S, double *V) { memcpy(V, S->V, 5 * (size_t)S->nvol * sizeof(double)); }
void orc_get_state0(const orc_sim *S, double *V0) { memcpy(V0, S->V0, 5 * (size_t)S->nvol * sizeof(double)); }
void orc_get_extrema(const orc_sim *S, double *hmax, double *bmax, double *eta_max)
{
    size_t b = (size_t)S->nvol * sizeof(double);
    if (hmax) memcpy(hmax, S->hmax, b);
    if (bmax) memcpy(bmax, S->bmax, b);
    if (eta_max) memcpy(eta_max, S->etamax, b);
}
double orc_time(const orc_sim *S) { return S->t; }
double orc_dt(const orc_sim *S) { return S->dt; }
long orc_steps(const orc_sim *S) { return S->steps; }
void orc_set_time(orc_sim *S, double t, double dt)
{
    S->t = t;
    S->dt = dt;
}
