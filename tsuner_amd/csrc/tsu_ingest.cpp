// tsu_ingest.cpp -- the reference's real-data input path (the Read_files = 1
// build, Problem_loader.cxx:355-624 with interpolations.cxx:46-95,368-1114):
// a gridded bathymetry and an NCTR free-surface/velocity series (three files
// <stem>h.nc, <stem>u.nc, <stem>v.nc) in longitude/latitude become the
// Cartesian initial state, fixed-bed mask and boundary series CC/tCC of the
// solver.
//
// Steps, each in the reference's order of floating-point operations:
//   1. bathymetry: cell-centred grid read whole, domain = its cell edges,
//      resampled to npx x npy with the 3x3 quadratic stencil whose Lagrange
//      weights are evaluated at the GLOBAL normalised coordinate of the
//      target cell (xt, yt in [-1, 1] over the whole domain), not at the
//      offset inside the stencil (interpolations.cxx:779-865);
//   2. optional fixed-bed map, same resampling (Problem_loader.cxx:436-471);
//   3. initial free surface: time level 0 of the h file, resampled likewise
//      with the file's first coordinate taken as the grid origin and spacing
//      span/n (Problem_loader.cxx:473-505), then sea-level shift and fill
//      values replaced by the depth (interpolations.cxx:867-877);
//   4. boundary series: at each time level and each boundary cell centre, a
//      local 3x3 quadratic Lagrange interpolation of h, u, v (cm, cm/s),
//      scaled by 0.01, + sea-level shift for h, depth from step 1 in slot 3,
//      fill values (|v| > 1e20) replaced (interpolations.cxx:368-774);
//   5. longitude/latitude to metres: per-row cell widths at the row's
//      latitude on a sphere of radius 6371 km, each row centred in a domain as
//      wide as the widest row, and every field resampled onto the Cartesian
//      grid with the global-coordinate stencil of step 1
//      (interpolations.cxx:79-95,879-1114);
//   6. h = eta*0.01 + H + shift, clipped at 0; q = 0; fixed flag = map >= 0.5
//      or 1 without a map (Problem_loader.cxx:586-610).
//
// Two index rules are kept as the reference's code has them: in step 4 the
// source index is compared with the (unsigned) grid size, so a negative index
// becomes the last one; in the right-hand boundary of step 5 the upper
// neighbour test uses npx where the others use npy.
#define TSU_BUILDING_LIBRARY 1
#include "../../include/tsuner.h"

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "tsu_gridfile.h"
#include "tsu_host.h"

namespace {

const double kPi = 3.14159265358979323846;   // Sim_parameters.hxx:30
const double kEarthRadius = 6371e3;           // Sim_parameters.hxx:37
const double kUnits = 0.01;                   // Sim_parameters.hxx:42: cm -> m for eta, u, v

// Quadratic Lagrange basis on the nodes -1, 0, 1 (interpolations.cxx:46-54).
inline double q0(double s) { return s / 2.0 * (s - 1); }
inline double q1(double s) { return 1 - s * s; }
inline double q2(double s) { return s / 2.0 * (s + 1); }

// Quadratic Lagrange basis on arbitrary nodes (interpolations.cxx:67-75).
inline double n0(double x, double x0, double x1, double x2) { return ((x - x1) * (x - x2)) / ((x0 - x1) * (x0 - x2)); }
inline double n1(double x, double x0, double x1, double x2) { return ((x - x0) * (x - x2)) / ((x1 - x0) * (x1 - x2)); }
inline double n2(double x, double x0, double x1, double x2) { return ((x - x0) * (x - x1)) / ((x2 - x0) * (x2 - x1)); }

inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// 3x3 stencil around (ic, jc) of a row-major nx-wide field, in the order
// lower-left, lower, lower-right, left, centre, right, upper-left, upper,
// upper-right.  Two edge rules occur in the reference:
//  * collapse (interpolations.cxx:818-850, 921-952): on a clamped side the
//    three neighbours of that side, corners included, become the centre;
//  * per axis (interpolations.cxx:476-484): each index is clamped on its own
//    axis, so a corner becomes the edge-adjacent neighbour.
inline void stencil_collapse(int nx, int ny, int ic, int jc, size_t k[9])
{
    const size_t p = (size_t)jc * nx + ic;
    size_t L = p - 1, R = p + 1, U = p + nx, D = p - nx;
    size_t LU = U - 1, LD = D - 1, RU = U + 1, RD = D + 1;
    if (ic == 0) L = LU = LD = p;
    if (ic == nx - 1) R = RD = RU = p;
    if (jc == 0) D = LD = RD = p;
    if (jc == ny - 1) U = LU = RU = p;
    const size_t o[9] = {LD, D, RD, L, p, R, LU, U, RU};
    std::memcpy(k, o, sizeof(o));
}

inline void stencil_per_axis(int nx, int ny, int ic, int jc, size_t k[9])
{
    const int ii[3] = {ic - (ic > 0), ic, ic + (ic < nx - 1)};
    const int jj[3] = {jc - (jc > 0), jc, jc + (jc < ny - 1)};
    for (int b = 0; b < 3; b++)
        for (int a = 0; a < 3; a++) k[3 * b + a] = (size_t)jj[b] * nx + ii[a];
}

// Weighted sum over a stencil with weights wx[a]*wy[b], terms added in
// stencil order as the reference writes them.
inline double stencil_sum(const double* f, const size_t k[9], const double wx[3], const double wy[3])
{
    double s = f[k[0]] * (wx[0] * wy[0]);
    for (int m = 1; m < 9; m++) s = s + f[k[m]] * (wx[m % 3] * wy[m / 3]);
    return s;
}

// interpolations.cxx:779-865: fun (nxf x nyf, cell centres from
// (xf0 + dxf/2, yf0 + dyf/2)) onto npx x npy cells covering the box.
void resample_box(double xmin, double xmax, double ymin, double ymax, double xf0, double yf0, int nxf, int nyf,
                  double dxf, double dyf, int npx, int npy, const double* fun, double* sol)
{
    const double dx = (xmax - xmin) / (npx), dy = (ymax - ymin) / (npy);
    for (int j = 0; j < npy; j++) {
        const double y = ymin + 0.5 * dy + j * dy;
        const double yt = 1.0 / (ymax - ymin) * (2 * y - (ymax + ymin));
        const double wy[3] = {q0(yt), q1(yt), q2(yt)};
        const int jg = clampi(int((y - yf0 - 0.5 * dyf) / dyf), 0, nyf - 1);
        for (int i = 0; i < npx; i++) {
            const double x = xmin + 0.5 * dx + i * dx;
            const double xt = 1.0 / (xmax - xmin) * (2 * x - (xmax + xmin));
            const double wx[3] = {q0(xt), q1(xt), q2(xt)};
            const int ig = clampi(int((x - xf0 - 0.5 * dxf) / dxf), 0, nxf - 1);
            size_t k[9];
            stencil_collapse(nxf, nyf, ig, jg, k);
            sol[(size_t)j * npx + i] = stencil_sum(fun, k, wx, wy);
        }
    }
}

// interpolations.cxx:879-968: a field on the npx x npy longitude/latitude
// grid onto the npx x npy Cartesian grid [0, xmax] x [0, ymax]; row j of the
// source is dxE[j] metres per cell wide and centred.
void to_cartesian(double xmax, double ymax, const double* dxE, double dyE, int npx, int npy, const double* fun,
                  double* sol)
{
    const double xmin = 0, ymin = 0;
    const double dxG = (xmax - xmin) / npx, dyG = (ymax - ymin) / npy;
    for (int i = 0; i < npx; i++) {
        const double x = xmin + 0.5 * dxG + i * dxG;
        const double xt = 1.0 / (xmax - xmin) * (2 * x - (xmax + xmin));
        const double wx[3] = {q0(xt), q1(xt), q2(xt)};
        for (int j = 0; j < npy; j++) {
            const double y = ymin + 0.5 * dyG + j * dyG;
            const double yt = 1.0 / (ymax - ymin) * (2 * y - (ymax + ymin));
            const double wy[3] = {q0(yt), q1(yt), q2(yt)};
            const int jg = clampi(int((y - ymin - 0.5 * dyE) / dyE), 0, npy - 1);
            const double x0 = 0 + (xmax - dxE[jg] * npx) / 2;
            const int ig = clampi(int((x - x0 - 0.5 * dxE[jg]) / dxE[jg]), 0, npx - 1);
            size_t k[9];
            stencil_collapse(npx, npy, ig, jg, k);
            sol[(size_t)j * npx + i] = stencil_sum(fun, k, wx, wy);
        }
    }
}

// Boundary slot layout of one time level (VolumeSolver.cu:172-202): bottom
// i, right npx + j, top npx + npy + i, left 2*npx + npy + j; 4 values each.
struct Slots {
    int npx, npy;
    size_t level() const { return (size_t)(2 * npx + 2 * npy) * 4; }
    size_t bottom(int i) const { return (size_t)i * 4; }
    size_t right(int j) const { return (size_t)(npx + j) * 4; }
    size_t top(int i) const { return (size_t)(npx + npy + i) * 4; }
    size_t left(int j) const { return (size_t)(2 * npx + npy + j) * 4; }
};

// interpolations.cxx:970-1114: the boundary series onto the Cartesian
// boundary, 1-D global-coordinate quadratic along each side, all four values.
void series_to_cartesian(double xmax, double ymax, const double* dxE, double dyE, int npx, int npy, size_t T,
                         double* CC)
{
    const Slots S{npx, npy};
    std::vector<double> aux(S.level());
    const double xmin = 0, ymin = 0;
    const double dx = (xmax - xmin) / npx, dy = (ymax - ymin) / npy;
    auto mix = [](const double* c, size_t p0, size_t p, size_t p1, const double w[3], double* out) {
        for (int k = 0; k < 4; k++) out[k] = w[0] * c[p0 + k] + w[1] * c[p + k] + w[2] * c[p1 + k];
    };
    for (size_t t = 0; t < T; t++) {
        const double* lev = CC + t * S.level();
        for (int i = 0; i < npx; i++) {
            const double x = xmin + 0.5 * dx + i * dx;
            const double xt = 1.0 / (xmax - xmin) * (2 * x - (xmax + xmin));
            const double w[3] = {q0(xt), q1(xt), q2(xt)};
            for (int side = 0; side < 2; side++) {   // bottom (row 0 width), top (row npy-1 width)
                const double dxr = dxE[side == 0 ? 0 : npy - 1];
                const double x0 = 0 + (xmax - dxr * npx) / 2;
                const int ig = clampi(int((x - x0 - 0.5 * dxr) / dxr), 0, npx - 1);
                const size_t p = side == 0 ? S.bottom(ig) : S.top(ig);
                mix(lev, p - 4 * (ig > 0), p, p + 4 * (ig < npx - 1), w,
                    aux.data() + (side == 0 ? S.bottom(i) : S.top(i)));
            }
        }
        for (int j = 0; j < npy; j++) {
            const double y = ymin + 0.5 * dy + j * dy;
            const double yt = 1.0 / (ymax - ymin) * (2 * y - (ymax + ymin));
            const double w[3] = {q0(yt), q1(yt), q2(yt)};
            const double y0 = 0 + (ymax - dyE * npy) / 2;
            const int jg = clampi(int((y - y0 - 0.5 * dyE) / dyE), 0, npy - 1);
            size_t p = S.left(jg);
            mix(lev, p - 4 * (jg > 0), p, p + 4 * (jg < npy - 1), w, aux.data() + S.left(j));
            p = S.right(jg);
            mix(lev, p - 4 * (jg > 0), p, p + 4 * (jg < npx - 1), w, aux.data() + S.right(j));
        }
        std::memcpy(CC + t * S.level(), aux.data(), aux.size() * sizeof(double));
    }
}

// Source index of a boundary point in the forcing grid: truncation toward
// zero, then the reference's unsigned upper clamp (a negative index compares
// as a huge unsigned value and becomes n - 1).
inline int series_index(double v, size_t n)
{
    const int k = int(v);
    return (k < 0 || (size_t)k > n - 1) ? (int)(n - 1) : k;
}

struct Forcing {
    size_t nx = 0, ny = 0, T = 0;
    double x0 = 0, y0 = 0, dx = 0, dy = 0;   // origin (first coordinate) and span/n spacing
    std::vector<double> h, u, v, time;
};

// interpolations.cxx:368-774 for one boundary point (x, y): slots 0-2 of
// `out` from the forcing at level t, slot 3 the depth Hb, fill rules last.
void series_point(const Forcing& F, size_t t, double x, double y, double Hb, double slr, double* out)
{
    const int ig = series_index((x - F.x0 - 0.5 * F.dx) / F.dx, F.nx);
    const int jg = series_index((y - F.y0 - 0.5 * F.dy) / F.dy, F.ny);
    const double x1 = F.x0 + 0.5 * F.dx + ig * F.dx, xa = x1 - F.dx, xb = x1 + F.dx;
    const double y1 = F.y0 + 0.5 * F.dy + jg * F.dy, ya = y1 - F.dy, yb = y1 + F.dy;
    const double wx[3] = {n0(x, xa, x1, xb), n1(x, xa, x1, xb), n2(x, xa, x1, xb)};
    const double wy[3] = {n0(y, ya, y1, yb), n1(y, ya, y1, yb), n2(y, ya, y1, yb)};
    const size_t off = t * F.nx * F.ny;
    size_t k[9];
    stencil_per_axis((int)F.nx, (int)F.ny, ig, jg, k);
    out[0] = stencil_sum(F.h.data() + off, k, wx, wy);
    out[1] = stencil_sum(F.u.data() + off, k, wx, wy);
    out[2] = stencil_sum(F.v.data() + off, k, wx, wy);
    out[3] = Hb;
    out[0] *= kUnits;
    out[1] *= kUnits;
    out[2] *= kUnits;
    out[0] += slr;
    if (out[0] < -1e20) out[0] = Hb;
    if (out[1] < -1e20) out[1] = 0.0;
    if (out[2] < -1e20) out[2] = 0.0;
    if (out[0] > 1e20) out[0] = Hb;
    if (out[1] > 1e20) out[1] = 0.0;
    if (out[2] > 1e20) out[2] = 0.0;
}

std::string sibling(const std::string& hfile, char c)
{
    std::string s = hfile;   // interpolations.cxx:385-386: the 4th character from the end
    if (s.size() >= 4) s[s.size() - 4] = c;
    return s;
}

int fail(const std::string& msg) { return tsu::set_error(TSU_ERR_HOST, msg); }

// Grid spacings divide source offsets into indices: positive and finite.
bool spacing_ok(double d) { return std::isfinite(d) && d > 0.0; }

bool coords(const tsu::GridFile& g, const std::vector<std::string>& names, std::vector<double>* v)
{
    const std::string n = g.find_var(names);
    return !n.empty() && g.read_double(n, v) && !v->empty();
}

// A gridded field (lat, lon) with its coordinates (Problem_loader.cxx:57-151).
int read_grid(const std::string& path, const std::vector<std::string>& var, std::vector<double>* lon,
              std::vector<double>* lat, std::vector<double>* data)
{
    tsu::GridFile g;
    if (!g.open(path)) return fail(g.error());
    if (!coords(g, {"lon", "x", "LON"}, lon) || !coords(g, {"lat", "y", "LAT"}, lat))
        return fail(path + ": no lon/x/LON or lat/y/LAT coordinate");
    const std::string n = g.find_var(var);
    if (n.empty() || !g.read_double(n, data)) return fail(path + ": no " + var[0] + " variable");
    if (data->size() % (lon->size() * lat->size()) != 0) return fail(path + ": " + n + " is not (lat, lon) shaped");
    return TSU_OK;
}

}  // namespace

extern "C" {

TSU_API int tsu_input_time_levels(const char* eta_file, unsigned long* time_levels)
{
    if (!eta_file || !time_levels) return tsu::set_error(TSU_ERR_ARG, "null argument");
    tsu::GridFile g;
    if (!g.open(eta_file)) return fail(g.error());
    const std::string d = g.find_dim({"time", "TIME"});
    size_t T = 0;
    if (d.empty() || !g.dim_len(d, &T)) return fail(std::string(eta_file) + ": no time/TIME dimension");
    *time_levels = (unsigned long)T;
    return TSU_OK;
}

TSU_API int tsu_load_input_files(const char* bathy_file, const char* eta_file, const char* fixed_bed_file,
                                 double sea_level_correction, tsu_sim_params* params, double* dVh,
                                 tsu_double2* dVhu, double* dVH, int* fixH, double* CC, double* tCC,
                                 unsigned long time_levels)
{
    if (!bathy_file || !eta_file || !params || !dVh || !dVhu || !dVH || !fixH || !CC || !tCC)
        return tsu::set_error(TSU_ERR_ARG, "null argument");
    const int npx = params->npx, npy = params->npy;
    if (npx < 1 || npy < 1) return tsu::set_error(TSU_ERR_ARG, "npx and npy must be positive");
    const size_t nvol = (size_t)npx * npy;
    const double slr = sea_level_correction;
    const bool read_fix = fixed_bed_file && std::strcmp(fixed_bed_file, "0") != 0 && fixed_bed_file[0];

    // 1. bathymetry (Problem_loader.cxx:397-434)
    std::vector<double> lon, lat, bat;
    int e = read_grid(bathy_file, {"bathy"}, &lon, &lat, &bat);
    if (e) return e;
    const int nxb = (int)lon.size(), nyb = (int)lat.size();
    if (nxb < 2 || nyb < 2 || bat.size() != (size_t)nxb * nyb)
        return tsu::set_error(TSU_ERR_ARG, std::string(bathy_file) + ": bathymetry must be a (lat, lon) grid of at least 2 x 2");
    const double dxb = (lon[nxb - 1] - lon[0]) / (nxb - 1), dyb = (lat[nyb - 1] - lat[0]) / (nyb - 1);
    if (!spacing_ok(dxb) || !spacing_ok(dyb))
        return tsu::set_error(TSU_ERR_ARG, std::string(bathy_file) + ": coordinates must increase");
    double xmin = lon[0] - 0.5 * dxb, xmax = lon[nxb - 1] + 0.5 * dxb;
    double ymin = lat[0] - 0.5 * dyb, ymax = lat[nyb - 1] + 0.5 * dyb;
    std::vector<double> H_s(nvol), fix_s, eta_s(nvol);
    resample_box(xmin, xmax, ymin, ymax, xmin, ymin, nxb, nyb, dxb, dyb, npx, npy, bat.data(), H_s.data());
    std::vector<double>().swap(bat);

    // 2. fixed-bed map (Problem_loader.cxx:436-471)
    if (read_fix) {
        std::vector<double> flon, flat, fx;
        e = read_grid(fixed_bed_file, {"bathy"}, &flon, &flat, &fx);
        if (e) return e;
        const int nxf = (int)flon.size(), nyf = (int)flat.size();
        if (nxf < 2 || nyf < 2 || fx.size() != (size_t)nxf * nyf)
            return tsu::set_error(TSU_ERR_ARG, std::string(fixed_bed_file) + ": map must be a (lat, lon) grid of at least 2 x 2");
        if (!spacing_ok(flon[nxf - 1] - flon[0]) || !spacing_ok(flat[nyf - 1] - flat[0]))
            return tsu::set_error(TSU_ERR_ARG, std::string(fixed_bed_file) + ": coordinates must increase");
        fix_s.resize(nvol);
        resample_box(xmin, xmax, ymin, ymax, flon[0], flat[0], nxf, nyf, (flon[nxf - 1] - flon[0]) / (nxf - 1),
                     (flat[nyf - 1] - flat[0]) / (nyf - 1), npx, npy, fx.data(), fix_s.data());
    }

    // 3./4. free surface and boundary series (Problem_loader.cxx:473-517)
    Forcing F;
    {
        tsu::GridFile g;
        if (!g.open(eta_file)) return fail(g.error());
        std::vector<double> flon, flat;
        if (!coords(g, {"lon", "x", "LON"}, &flon) || !coords(g, {"lat", "y", "LAT"}, &flat))
            return fail(std::string(eta_file) + ": no lon/LON or lat/LAT coordinate");
        const std::string td = g.find_dim({"time", "TIME"});
        if (td.empty() || !g.dim_len(td, &F.T)) return fail(std::string(eta_file) + ": no time/TIME dimension");
        if (F.T != time_levels)
            return tsu::set_error(TSU_ERR_ARG, std::string(eta_file) + " holds " + std::to_string(F.T) +
                                                   " time levels, buffers were sized for " + std::to_string(time_levels));
        const std::string hv = g.find_var({"ha_lc", "HA", "ha"});
        if (hv.empty() || !g.read_double(hv, &F.h)) return fail(std::string(eta_file) + ": no ha_lc/HA/ha variable");
        F.nx = flon.size();
        F.ny = flat.size();
        F.x0 = flon[0];
        F.y0 = flat[0];
        F.dx = (flon[F.nx - 1] - flon[0]) / F.nx;   // Problem_loader.cxx:501-502: span / n
        F.dy = (flat[F.ny - 1] - flat[0]) / F.ny;
    }
    const size_t lev = F.nx * F.ny;
    if (F.T < 1 || F.nx < 2 || F.ny < 2 || F.h.size() != F.T * lev)
        return tsu::set_error(TSU_ERR_ARG, std::string(eta_file) + ": expected (time, lat, lon) free-surface levels");
    if (!spacing_ok(F.dx) || !spacing_ok(F.dy))
        return tsu::set_error(TSU_ERR_ARG, std::string(eta_file) + ": coordinates must increase");
    for (int c = 0; c < 2; c++) {
        const std::string path = sibling(eta_file, c == 0 ? 'u' : 'v');
        tsu::GridFile g;
        if (!g.open(path)) return fail(g.error());
        const std::string vn = g.find_var(c == 0 ? std::vector<std::string>{"ua_lc", "UA", "ua"}
                                                 : std::vector<std::string>{"va_lc", "VA", "va"});
        // the first T levels, T from the h file (interpolations.cxx:405-406,447)
        std::vector<double>& dst = c == 0 ? F.u : F.v;
        if (vn.empty() || !g.read_double(vn, &dst) || dst.size() % lev != 0 || dst.size() < F.T * lev)
            return fail(path + ": no velocity series on the free surface's grid with " + std::to_string(F.T) +
                        " levels");
        dst.resize(F.T * lev);
        if (c == 0) {   // the times come from the u file (interpolations.cxx:418-425)
            const std::string tv = g.find_var({"time", "TIME"});
            if (tv.empty() || !g.read_double(tv, &F.time) || F.time.size() < F.T)
                return fail(path + ": no time/TIME variable of " + std::to_string(F.T) + " levels");
            F.time.resize(F.T);
        }
    }
    resample_box(xmin, xmax, ymin, ymax, F.x0, F.y0, (int)F.nx, (int)F.ny, F.dx, F.dy, npx, npy, F.h.data(),
                 eta_s.data());
    for (size_t k = 0; k < nvol; k++) {   // interpolations.cxx:867-877
        eta_s[k] -= slr;
        if (eta_s[k] < -1e20) eta_s[k] = H_s[k];
        if (eta_s[k] > 1e20) eta_s[k] = H_s[k];
    }
    for (size_t t = 0; t < F.T; t++) tCC[t] = F.time[t];
    {
        const Slots S{npx, npy};
        const double dx = (xmax - xmin) / npx, dy = (ymax - ymin) / npy;
        for (size_t t = 0; t < F.T; t++) {
            double* L = CC + t * S.level();
            for (int i = 0; i < npx; i++) {
                const double x = xmin + 0.5 * dx + i * dx;
                series_point(F, t, x, ymin, H_s[(size_t)i], slr, L + S.bottom(i));
                series_point(F, t, x, ymax, H_s[(size_t)(npy - 1) * npx + i], slr, L + S.top(i));
            }
            for (int j = 0; j < npy; j++) {
                const double y = ymin + 0.5 * dy + j * dy;
                series_point(F, t, xmin, y, H_s[(size_t)j * npx], slr, L + S.left(j));
                series_point(F, t, xmax, y, H_s[(size_t)j * npx + npx - 1], slr, L + S.right(j));
            }
        }
    }

    // 5. longitude/latitude -> metres (Problem_loader.cxx:520-564)
    double dx = (xmax - xmin) / (npx), dy = (ymax - ymin) / npy;
    const double dy_unit = (2 * kPi * kEarthRadius) / 360;   // interpolations.cxx:79-81
    std::vector<double> dxE((size_t)npy);
    double dmax = -1.0;
    for (int j = 0; j < npy; j++) {   // interpolations.cxx:83-93
        const double la = ymin + 0.5 * dy + j * dy;
        dxE[(size_t)j] = (2 * kPi * kEarthRadius * std::cos(la * kPi / 180)) / 360;
        dmax = std::max(dmax, dxE[(size_t)j] * (xmax - xmin));
        dxE[(size_t)j] *= dx;
    }
    xmin = 0;
    xmax = dmax;
    ymax = dy_unit * (ymax - ymin);
    ymin = 0;
    const double dyE = dy * dy_unit;
    std::vector<double> H_c(nvol), eta_c(nvol), fix_c;
    to_cartesian(xmax, ymax, dxE.data(), dyE, npx, npy, H_s.data(), H_c.data());
    if (read_fix) {
        fix_c.resize(nvol);
        to_cartesian(xmax, ymax, dxE.data(), dyE, npx, npy, fix_s.data(), fix_c.data());
    }
    to_cartesian(xmax, ymax, dxE.data(), dyE, npx, npy, eta_s.data(), eta_c.data());
    series_to_cartesian(xmax, ymax, dxE.data(), dyE, npx, npy, F.T, CC);

    // 6. state (Problem_loader.cxx:575-624)
    dx = (xmax - xmin) / (npx);
    dy = (ymax - ymin) / npy;
    for (size_t k = 0; k < nvol; k++) {
        dVH[k] = H_c[k];
        double h = eta_c[k] * kUnits + dVH[k] + slr;
        h *= (h >= 0.0);
        dVh[k] = h;
        dVhu[k].x = 0.0;
        dVhu[k].y = 0.0;
        fixH[k] = read_fix ? (fix_c[k] >= 0.5 ? 1 : 0) : 1;
    }
    params->xmin = xmin;
    params->xmax = xmax;
    params->ymin = ymin;
    params->ymax = ymax;
    params->dx = dx;
    params->dy = dy;
    params->nvol = (int)nvol;
    return TSU_OK;
}

}  // extern "C"
