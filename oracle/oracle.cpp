// oracle.cpp — TEST INFRASTRUCTURE ONLY. FP64 CPU restatement of the reference solver's algorithm
// (backend.cu:33-1195), used by tests/ as the parity checker and by bench.py's cpu_baseline leg as the
// host-core baseline. The product library (libcavity_solver.so) never links, loads or calls this code.
//
// Every kernel of the reference is restated as a loop over its index space (OpenMP across the outer
// index; each iteration writes only its own output, so results do not depend on the thread count), in the
// reference's order of floating-point operations. The field type R is a template parameter:
//   R = double: float replaced by double everywhere (the "plain FP64" semantics: the grid is not rounded
//               through float, backend.cu:695-697, and the CFL rate is not rounded through float, :886);
//   R = float:  the reference's own precision: float velocity, impulse and stencil arithmetic, the grid
//               rounded through float, FP64 potential, residuals and refinement, float transforms.
// Compile with -ffp-contract=off (the reference builds with --fmad=false, build_windows.bat:4).
//
// The FDM transforms (cublasSgemmStridedBatched in the reference, backend.cu:792-798) are restated as
// contractions in R that sum over the contracted index in ascending order. cuBLAS's own summation order
// is unpublished: at that boundary parity rests on the GEMM definition only.
//
// How each function is pinned (closed forms, invariants, brute force) is listed in tests/test_oracle_*.py.
#include "../include/cavity_solver.h"

#include <algorithm>
#include <chrono>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

namespace {

using vec = std::vector<double>;
using i64 = long long;
enum { X = 0, Y = 1, Z = 2 };

[[noreturn]] void fail(const std::string& msg) { throw std::runtime_error(msg); }

// ------------------------------------------------------------------------------------------------ grid
// backend.cu:643-699
struct Axis1 {
    int n = 0;
    double length = 0;
    vec face, centre, width, spacing;
};

vec linspace(double a, double b, int count) {                          // backend.cu:643-649
    vec v(count);
    double step = (b - a) / (count - 1);
    for (int i = 0; i < count; i++) v[i] = i * step + a;
    v[count - 1] = b;
    return v;
}

vec tanh_faces(int n, double length, double dx_wall) {                 // backend.cu:651-666
    if (dx_wall >= length / n * (1 - 1e-12)) return linspace(0.0, length, n + 1);
    vec xi = linspace(0.0, 1.0, n + 1), f(n + 1);
    auto place = [&](double beta) {
        double tb = std::tanh(beta);
        for (int i = 0; i <= n; i++) f[i] = length * 0.5 * (1.0 + std::tanh(beta * (2 * xi[i] - 1)) / tb);
    };
    double lo = 1e-8, hi = 30.0;
    for (int it = 0; it < 200; it++) {
        double mid = 0.5 * (lo + hi);
        place(mid);
        if (f[1] - f[0] < dx_wall) hi = mid; else lo = mid;
    }
    place(0.5 * (lo + hi));
    f[0] = 0.0;
    f[n] = length;
    f[n / 2] = 0.5 * length;
    return f;
}

double largest_over_smallest(const vec& f) {                          // backend.cu:667-671
    double big = 0, small = 1e300;
    for (size_t i = 1; i < f.size(); i++) {
        big = std::max(big, f[i] - f[i - 1]);
        small = std::min(small, f[i] - f[i - 1]);
    }
    return big / small;
}

Axis1 build_axis(const CavConfig& c, int n, double length, bool round_float = false) {   // backend.cu:672-699
    vec f;
    if (!c.grid_stretch) {
        f = tanh_faces(n, length, 1e300);
    } else {
        double d = std::min(c.grid_dx_wall, length / n);
        auto ratio = [&](double dw) { return largest_over_smallest(tanh_faces(n, length, dw)); };
        if (c.grid_max_aspect > 0 && ratio(d) > c.grid_max_aspect) {
            double lo = d, hi = length / n;
            for (int it = 0; it < 60; it++) {
                double mid = 0.5 * (lo + hi);
                if (ratio(mid) > c.grid_max_aspect) lo = mid; else hi = mid;
            }
            d = hi;
        }
        f = tanh_faces(n, length, d);
    }
    Axis1 a;
    a.n = n;
    a.face = f;
    a.width.resize(n);
    a.centre.resize(n);
    a.spacing.resize(n + 1);
    for (int i = 0; i < n; i++) {
        a.width[i] = f[i + 1] - f[i];
        a.centre[i] = 0.5 * (f[i + 1] + f[i]);
    }
    for (int i = 1; i < n; i++) a.spacing[i] = a.centre[i] - a.centre[i - 1];
    a.spacing[0] = 0.5 * a.width[0];
    a.spacing[n] = 0.5 * a.width[n - 1];
    a.length = f[n] - f[0];
    if (round_float)                        // backend.cu:695-697 (the FP32 fields); FP64: no rounding
        for (vec* v : {&a.face, &a.centre, &a.width, &a.spacing})
            for (double& q : *v) q = (double)(float)q;
    return a;
}

struct Mesh {
    int nx = 0, ny = 0, nz = 0, lid_shaped = 0;
    Axis1 ax[3];
    i64 n_u() const { return (i64)(nx + 1) * ny * nz; }
    i64 n_v() const { return (i64)nx * (ny + 1) * nz; }
    i64 n_w() const { return (i64)nx * ny * (nz + 1); }
    i64 n_c() const { return (i64)nx * ny * nz; }
    i64 size(int k) const { return k == 0 ? n_u() : k == 1 ? n_v() : n_w(); }
    // backend.cu:50-53
    i64 iu(i64 x, i64 y, i64 z) const { return x + (nx + 1) * (y + (i64)ny * z); }
    i64 iv(i64 x, i64 y, i64 z) const { return x + nx * (y + (i64)(ny + 1) * z); }
    i64 iw(i64 x, i64 y, i64 z) const { return x + nx * (y + (i64)ny * z); }
    i64 ic(i64 x, i64 y, i64 z) const { return x + nx * (y + (i64)ny * z); }
    double wd(int a, int i) const { return ax[a].width[i]; }
    double sp(int a, int i) const { return ax[a].spacing[i]; }
    double ce(int a, int i) const { return ax[a].centre[i]; }
    double fc(int a, int i) const { return ax[a].face[i]; }
    // backend.cu:58-65
    static double lid_edge(double s) {
        double s2 = s * s, s4 = s2 * s2, s8 = s4 * s4, s16 = s8 * s8, a = 1.0 - s16 * s2;
        return a * a;
    }
    double lid_shape(double fx, double fz) const {
        if (!lid_shaped) return 1.0;
        return lid_edge(2.0 * fx - 1.0) * lid_edge(2.0 * fz - 1.0);
    }
    double lid_at_u_face(int x, int z) const { return lid_shape(fc(X, x) / ax[X].length, ce(Z, z) / ax[Z].length); }
    double lid_at_cell(int x, int z) const { return lid_shape(ce(X, x) / ax[X].length, ce(Z, z) / ax[Z].length); }
    // the same in the field type R (the grid values and the box length as R, backend.cu:41-45, 58-65)
    template <class R> static R lid_edge_t(R s) {
        R s2 = s * s, s4 = s2 * s2, s8 = s4 * s4, s16 = s8 * s8, a = R(1) - s16 * s2;
        return a * a;
    }
    template <class R> R lid_shape_t(R fx, R fz) const {
        if (!lid_shaped) return R(1);
        return lid_edge_t<R>(R(2) * fx - R(1)) * lid_edge_t<R>(R(2) * fz - R(1));
    }
    template <class R> R lid_at_u_face_t(int x, int z) const {
        return lid_shape_t<R>(R(fc(X, x)) / R(ax[X].length), R(ce(Z, z)) / R(ax[Z].length));
    }
    template <class R> R lid_at_cell_t(int x, int z) const {
        return lid_shape_t<R>(R(ce(X, x)) / R(ax[X].length), R(ce(Z, z)) / R(ax[Z].length));
    }
};

template <class R> struct F3T { const R* c[3]; };
using F3 = F3T<double>;

// backend.cu:67-72
template <class R>
R centred(const Mesh& g, const R* u, const R* v, const R* w, int c, int x, int y, int z) {
    if (c == 0) return R(0.5) * (u[g.iu(x, y, z)] + u[g.iu(x + 1, y, z)]);
    if (c == 1) return R(0.5) * (v[g.iv(x, y, z)] + v[g.iv(x, y + 1, z)]);
    return R(0.5) * (w[g.iw(x, y, z)] + w[g.iw(x, y, z + 1)]);
}

// backend.cu:76-96  g[i][c] = d u_c / d x_i at the cell centre, mirrored ghosts at the walls
template <class R>
void cell_gradient(const Mesh& g, const R* u, const R* v, const R* w, R lid, int x, int y, int z, R gr[3][3]) {
    int NX = g.nx, NY = g.ny, NZ = g.nz;
    auto sp = [&](int a, int i) { return R(g.sp(a, i)); };
    auto wd = [&](int a, int i) { return R(g.wd(a, i)); };
    R dxm = x > 0 ? sp(X, x) : wd(X, 0), dxp = x + 1 < NX ? sp(X, x + 1) : wd(X, NX - 1);
    R dym = y > 0 ? sp(Y, y) : wd(Y, 0), dyp = y + 1 < NY ? sp(Y, y + 1) : wd(Y, NY - 1);
    R dzm = z > 0 ? sp(Z, z) : wd(Z, 0), dzp = z + 1 < NZ ? sp(Z, z + 1) : wd(Z, NZ - 1);
    for (int c = 0; c < 3; c++) {
        R q = centred(g, u, v, w, c, x, y, z);
        R xm = x > 0 ? centred(g, u, v, w, c, x - 1, y, z) : -q;
        R xp = x + 1 < NX ? centred(g, u, v, w, c, x + 1, y, z) : -q;
        R ym = y > 0 ? centred(g, u, v, w, c, x, y - 1, z) : -q;
        R yp = y + 1 < NY ? centred(g, u, v, w, c, x, y + 1, z)
                          : R(2) * (c == 0 ? lid * g.lid_at_cell_t<R>(x, z) : R(0)) - q;
        R zp = z + 1 < NZ ? centred(g, u, v, w, c, x, y, z + 1) : -q;
        R zm = z > 0 ? centred(g, u, v, w, c, x, y, z - 1) : -q;
        gr[0][c] = (xp - xm) / (dxp + dxm);
        gr[1][c] = (yp - ym) / (dyp + dym);
        gr[2][c] = (zp - zm) / (dzp + dzm);
    }
    gr[0][0] = (u[g.iu(x + 1, y, z)] - u[g.iu(x, y, z)]) / wd(X, x);
    gr[1][1] = (v[g.iv(x, y + 1, z)] - v[g.iv(x, y, z)]) / wd(Y, y);
    gr[2][2] = (w[g.iw(x, y, z + 1)] - w[g.iw(x, y, z)]) / wd(Z, z);
}

// ------------------------------------------------------------------------------------------------ I(u, m)
// backend.cu:102-274: I = u x omega - grad(u.m), Lamb vector from edge vorticity, 4th order as narrow
// (q = 1) plus wide (q = 3) stencils with a wall ramp.
template <class R> struct Impulse {
    const Mesh& g;
    const R *u, *v, *w, *mx, *my, *mz;
    int o4;
    R K;
    R sp(int a, int i) const { return R(g.sp(a, i)); }
    R wd(int a, int i) const { return R(g.wd(a, i)); }
    R ce(int a, int i) const { return R(g.ce(a, i)); }

    static bool nodes(int e, int q, int n, int& lo, int& hi) {         // backend.cu:108-112
        lo = e - (q + 1) / 2;
        hi = e + (q - 1) / 2;
        return lo >= 0 && hi <= n - 1;
    }
    R node_gap(int a, int e, int q, int lo, int hi) const {       // backend.cu:114-117
        if (q == 1) return sp(a, e);
        return ce(a, hi) - ce(a, lo);
    }
    R at_edge(int a, int e, int q, int lo, int hi, R flo, R fhi) const {   // backend.cu:119-122
        if (q == 1) return (wd(a, lo) * flo + wd(a, hi) * fhi) / (R(2) * sp(a, e));
        return (wd(a, lo) * flo + wd(a, hi) * fhi) / (wd(a, lo) + wd(a, hi));
    }
    R edge_width(int a, int e, int q, int n) const {              // backend.cu:124-128
        int lo, hi;
        if (q == 1 || !nodes(e, q, n, lo, hi)) return sp(a, e);
        return R(0.5) * (wd(a, lo) + wd(a, hi));
    }
    R ramp(int e, int n) const {                                  // backend.cu:130-133
        R d = (R)(e < n - e ? e : n - e) - R(1);
        return std::fmin(std::fmax(d / K, R(0)), R(1));
    }
    R weight(int q, R s) const {                             // backend.cu:135-138
        if (q == 1) return o4 ? R(1) + s * R(0.125) : R(1);
        return -s * R(0.125);
    }
    // backend.cu:141-158
    R om_z(int x, int y, int z, int q) const {
        int xl, xh, yl, yh;
        if (!nodes(x, q, g.nx, xl, xh) || !nodes(y, q, g.ny, yl, yh)) return R(0);
        return (v[g.iv(xh, y, z)] - v[g.iv(xl, y, z)]) / node_gap(X, x, q, xl, xh)
             - (u[g.iu(x, yh, z)] - u[g.iu(x, yl, z)]) / node_gap(Y, y, q, yl, yh);
    }
    R om_x(int x, int y, int z, int q) const {
        int yl, yh, zl, zh;
        if (!nodes(y, q, g.ny, yl, yh) || !nodes(z, q, g.nz, zl, zh)) return R(0);
        return (w[g.iw(x, yh, z)] - w[g.iw(x, yl, z)]) / node_gap(Y, y, q, yl, yh)
             - (v[g.iv(x, y, zh)] - v[g.iv(x, y, zl)]) / node_gap(Z, z, q, zl, zh);
    }
    R om_y(int x, int y, int z, int q) const {
        int xl, xh, zl, zh;
        if (!nodes(x, q, g.nx, xl, xh) || !nodes(z, q, g.nz, zl, zh)) return R(0);
        return (u[g.iu(x, y, zh)] - u[g.iu(x, y, zl)]) / node_gap(Z, z, q, zl, zh)
             - (w[g.iw(xh, y, z)] - w[g.iw(xl, y, z)]) / node_gap(X, x, q, xl, xh);
    }
    // backend.cu:162-215, the six (vorticity, interpolated velocity) pairings
    R zx(int x, int y, int z, int q) const {     // omega_z * v along x
        int xl, xh;
        if (y < 0 || y > g.ny || !nodes(x, q, g.nx, xl, xh)) return R(0);
        R wq = weight(q, ramp(y, g.ny) * ramp(x, g.nx));
        if (wq == R(0)) return R(0);
        R o = om_z(x, y, z, q);
        if (o == R(0)) return R(0);
        return wq * o * at_edge(X, x, q, xl, xh, v[g.iv(xl, y, z)], v[g.iv(xh, y, z)]);
    }
    R zy(int x, int y, int z, int q) const {     // omega_z * u along y
        int yl, yh;
        if (x < 0 || x > g.nx || !nodes(y, q, g.ny, yl, yh)) return R(0);
        R wq = weight(q, ramp(y, g.ny) * ramp(x, g.nx));
        if (wq == R(0)) return R(0);
        R o = om_z(x, y, z, q);
        if (o == R(0)) return R(0);
        return wq * o * at_edge(Y, y, q, yl, yh, u[g.iu(x, yl, z)], u[g.iu(x, yh, z)]);
    }
    R xy(int x, int y, int z, int q) const {     // omega_x * w along y
        int yl, yh;
        if (z < 0 || z > g.nz || !nodes(y, q, g.ny, yl, yh)) return R(0);
        R wq = weight(q, ramp(z, g.nz) * ramp(y, g.ny));
        if (wq == R(0)) return R(0);
        R o = om_x(x, y, z, q);
        if (o == R(0)) return R(0);
        return wq * o * at_edge(Y, y, q, yl, yh, w[g.iw(x, yl, z)], w[g.iw(x, yh, z)]);
    }
    R xz(int x, int y, int z, int q) const {     // omega_x * v along z
        int zl, zh;
        if (y < 0 || y > g.ny || !nodes(z, q, g.nz, zl, zh)) return R(0);
        R wq = weight(q, ramp(z, g.nz) * ramp(y, g.ny));
        if (wq == R(0)) return R(0);
        R o = om_x(x, y, z, q);
        if (o == R(0)) return R(0);
        return wq * o * at_edge(Z, z, q, zl, zh, v[g.iv(x, y, zl)], v[g.iv(x, y, zh)]);
    }
    R yx(int x, int y, int z, int q) const {     // omega_y * w along x
        int xl, xh;
        if (z < 0 || z > g.nz || !nodes(x, q, g.nx, xl, xh)) return R(0);
        R wq = weight(q, ramp(z, g.nz) * ramp(x, g.nx));
        if (wq == R(0)) return R(0);
        R o = om_y(x, y, z, q);
        if (o == R(0)) return R(0);
        return wq * o * at_edge(X, x, q, xl, xh, w[g.iw(xl, y, z)], w[g.iw(xh, y, z)]);
    }
    R yz(int x, int y, int z, int q) const {     // omega_y * u along z
        int zl, zh;
        if (x < 0 || x > g.nx || !nodes(z, q, g.nz, zl, zh)) return R(0);
        R wq = weight(q, ramp(z, g.nz) * ramp(x, g.nx));
        if (wq == R(0)) return R(0);
        R o = om_y(x, y, z, q);
        if (o == R(0)) return R(0);
        return wq * o * at_edge(Z, z, q, zl, zh, u[g.iu(x, y, zl)], u[g.iu(x, y, zh)]);
    }
    R Kc(int x, int y, int z) const {                              // backend.cu:218-222
        i64 a = g.iu(x, y, z), b = g.iv(x, y, z), c = g.iw(x, y, z), c1 = g.iw(x, y, z + 1);
        i64 NX = g.nx;
        return R(0.5) * ((u[a] * mx[a] + u[a + 1] * mx[a + 1]) + (v[b] * my[b] + v[b + NX] * my[b + NX])
                      + (w[c] * mz[c] + w[c1] * mz[c1]));
    }
    R Ix(int x, int y, int z) const {                              // backend.cu:225-240 (interior face)
        R L = R(0);
        for (int q = 1; q <= (o4 ? 3 : 1); q += 2) {
            int up = (q + 1) / 2, down = (q - 1) / 2;
            R c = edge_width(X, x, q, g.nx) / (R(2) * sp(X, x));
            L += c * (zx(x, y + up, z, q) + zx(x, y - down, z, q));
            L -= c * (yx(x, y, z + up, q) + yx(x, y, z - down, q));
        }
        L -= (Kc(x, y, z) - Kc(x - 1, y, z)) / sp(X, x);
        return L;
    }
    R Iy(int x, int y, int z) const {                              // backend.cu:242-257
        R L = R(0);
        for (int q = 1; q <= (o4 ? 3 : 1); q += 2) {
            int up = (q + 1) / 2, down = (q - 1) / 2;
            R c = edge_width(Y, y, q, g.ny) / (R(2) * sp(Y, y));
            L -= c * (zy(x + up, y, z, q) + zy(x - down, y, z, q));
            L += c * (xy(x, y, z + up, q) + xy(x, y, z - down, q));
        }
        L -= (Kc(x, y, z) - Kc(x, y - 1, z)) / sp(Y, y);
        return L;
    }
    R Iz(int x, int y, int z) const {                              // backend.cu:259-274
        R L = R(0);
        for (int q = 1; q <= (o4 ? 3 : 1); q += 2) {
            int up = (q + 1) / 2, down = (q - 1) / 2;
            R c = edge_width(Z, z, q, g.nz) / (R(2) * sp(Z, z));
            L -= c * (xz(x, y + up, z, q) + xz(x, y - down, z, q));
            L += c * (yz(x + up, y, z, q) + yz(x - down, y, z, q));
        }
        L -= (Kc(x, y, z) - Kc(x, y, z - 1)) / sp(Z, z);
        return L;
    }
};

template <class R>
void inviscid(const Mesh& g, const F3T<R>& u, const F3T<R>& m, int o4, R K, R* const out[3]) {
    Impulse<R> im{g, u.c[0], u.c[1], u.c[2], m.c[0], m.c[1], m.c[2], o4, K};
    int NX = g.nx, NY = g.ny, NZ = g.nz;
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < NZ; z++)
        for (int y = 0; y < NY; y++)
            for (int x = 0; x <= NX; x++)
                out[0][g.iu(x, y, z)] = (x == 0 || x == NX) ? R(0) : im.Ix(x, y, z);
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < NZ; z++)
        for (int y = 0; y <= NY; y++)
            for (int x = 0; x < NX; x++)
                out[1][g.iv(x, y, z)] = (y == 0 || y == NY) ? R(0) : im.Iy(x, y, z);
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z <= NZ; z++)
        for (int y = 0; y < NY; y++)
            for (int x = 0; x < NX; x++)
                out[2][g.iw(x, y, z)] = (z == 0 || z == NZ) ? R(0) : im.Iz(x, y, z);
}

// ------------------------------------------------------------------------------------------------ RK3 stage
// backend.cu:279-330: out = [a0 m0 +] b (m + dt (nu lap(m) + I)), flux-form Laplacian, wall fluxes from vel.
template <class R>
void rk3_stage(const Mesh& g, int comp, const R* m, const R* m0, const R* vel, const R* I, R nu, R lid, R dt, R a0, R b,
               int has_base, R* out) {
    int NX = g.nx, NY = g.ny, NZ = g.nz;
    auto sp = [&](int a, int i) { return R(g.sp(a, i)); };
    auto wd = [&](int a, int i) { return R(g.wd(a, i)); };
    int sx = NX + (comp == 0), sy = NY + (comp == 1), sz = NZ + (comp == 2);
    i64 dy = sx, dz = (i64)sx * sy;
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < sz; z++)
        for (int y = 0; y < sy; y++)
            for (int x = 0; x < sx; x++) {
                i64 i = x + dy * y + dz * z;
                if ((comp == 0 && (x == 0 || x == NX)) || (comp == 1 && (y == 0 || y == NY))
                    || (comp == 2 && (z == 0 || z == NZ))) {
                    out[i] = R(0);
                    continue;
                }
                R c = m[i];
                R wall = vel[i];
                R lap = R(0);
                if (comp == 0) {
                    lap += ((m[i + 1] - c) / wd(X, x) - (c - m[i - 1]) / wd(X, x - 1)) / sp(X, x);
                } else {
                    R east = x < sx - 1 ? (m[i + 1] - c) / sp(X, x + 1) : R(0);
                    R west = x > 0 ? (c - m[i - 1]) / sp(X, x) : R(0);
                    if (x == 0) west = wall / sp(X, 0);
                    if (x == sx - 1) east = -wall / sp(X, NX);
                    lap += (east - west) / wd(X, x);
                }
                if (comp == 1) {
                    lap += ((m[i + dy] - c) / wd(Y, y) - (c - m[i - dy]) / wd(Y, y - 1)) / sp(Y, y);
                } else {
                    R north = y < sy - 1 ? (m[i + dy] - c) / sp(Y, y + 1) : R(0);
                    R south = y > 0 ? (c - m[i - dy]) / sp(Y, y) : R(0);
                    if (y == 0) south = wall / sp(Y, 0);
                    if (y == sy - 1) {
                        R U = comp == 0 ? lid * g.lid_at_u_face_t<R>(x, z) : R(0);
                        north = (U - wall) / sp(Y, NY);
                    }
                    lap += (north - south) / wd(Y, y);
                }
                if (comp == 2) {
                    lap += ((m[i + dz] - c) / wd(Z, z) - (c - m[i - dz]) / wd(Z, z - 1)) / sp(Z, z);
                } else {
                    R top = z < sz - 1 ? (m[i + dz] - c) / sp(Z, z + 1) : R(0);
                    R bottom = z > 0 ? (c - m[i - dz]) / sp(Z, z) : R(0);
                    if (z == 0) bottom = wall / sp(Z, 0);
                    if (z == sz - 1) top = -wall / sp(Z, NZ);
                    lap += (top - bottom) / wd(Z, z);
                }
                R rhs = nu * lap + R(0);
                R q = c + dt * (rhs + I[i]);
                if (has_base) q = a0 * m0[i] + b * q;
                out[i] = q;
            }
}

// ------------------------------------------------------------------------------------------------ projection
template <class R>
void divergence(const Mesh& g, const F3T<R>& f, double* rhs) {        // backend.cu:334-342
    const R *u = f.c[0], *v = f.c[1], *w = f.c[2];
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < g.nz; z++)
        for (int y = 0; y < g.ny; y++)
            for (int x = 0; x < g.nx; x++) {
                double dx = g.wd(X, x), dy = g.wd(Y, y), dz = g.wd(Z, z);
                rhs[g.ic(x, y, z)] = ((double)u[g.iu(x + 1, y, z)] - (double)u[g.iu(x, y, z)]) * (dy * dz)
                                   + ((double)v[g.iv(x, y + 1, z)] - (double)v[g.iv(x, y, z)]) * (dx * dz)
                                   + ((double)w[g.iw(x, y, z + 1)] - (double)w[g.iw(x, y, z)]) * (dx * dy);
            }
}

void pressure_residual(const Mesh& g, const double* phi, const double* rhs, double* r) {   // backend.cu:345-359
    int NX = g.nx, NY = g.ny, NZ = g.nz;
    i64 sy = NX, sz = (i64)NX * NY;
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < NZ; z++)
        for (int y = 0; y < NY; y++)
            for (int x = 0; x < NX; x++) {
                i64 i = g.ic(x, y, z);
                double dx = g.wd(X, x), dy = g.wd(Y, y), dz = g.wd(Z, z), c = phi[i];
                double s = 0.0;
                if (x + 1 < NX) s += (dy * dz) * (phi[i + 1] - c) / g.sp(X, x + 1);
                if (x > 0) s -= (dy * dz) * (c - phi[i - 1]) / g.sp(X, x);
                if (y + 1 < NY) s += (dx * dz) * (phi[i + sy] - c) / g.sp(Y, y + 1);
                if (y > 0) s -= (dx * dz) * (c - phi[i - sy]) / g.sp(Y, y);
                if (z + 1 < NZ) s += (dx * dy) * (phi[i + sz] - c) / g.sp(Z, z + 1);
                if (z > 0) s -= (dx * dy) * (c - phi[i - sz]) / g.sp(Z, z);
                r[i] = rhs[i] - s;
            }
}

template <class R>
void subtract_gradient(const Mesh& g, int comp, const R* m, const double* phi, R* out) {   // backend.cu:361-371
    int NX = g.nx, NY = g.ny, NZ = g.nz;
    int sx = NX + (comp == 0), sy = NY + (comp == 1), sz = NZ + (comp == 2);
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < sz; z++)
        for (int y = 0; y < sy; y++)
            for (int x = 0; x < sx; x++) {
                i64 i = x + (i64)sx * (y + (i64)sy * z);
                double d;
                if (comp == 0) {
                    if (x == 0 || x == NX) { out[i] = R(0); continue; }
                    d = (phi[g.ic(x, y, z)] - phi[g.ic(x - 1, y, z)]) / g.sp(X, x);
                } else if (comp == 1) {
                    if (y == 0 || y == NY) { out[i] = R(0); continue; }
                    d = (phi[g.ic(x, y, z)] - phi[g.ic(x, y - 1, z)]) / g.sp(Y, y);
                } else {
                    if (z == 0 || z == NZ) { out[i] = R(0); continue; }
                    d = (phi[g.ic(x, y, z)] - phi[g.ic(x, y, z - 1)]) / g.sp(Z, z);
                }
                out[i] = m[i] - (R)d;
            }
}

template <class R>
void zero_walls(const Mesh& g, int comp, R* f) {                       // backend.cu:373-380
    int NX = g.nx, NY = g.ny, NZ = g.nz;
    int sx = NX + (comp == 0), sy = NY + (comp == 1), sz = NZ + (comp == 2);
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < sz; z++)
        for (int y = 0; y < sy; y++)
            for (int x = 0; x < sx; x++)
                if ((comp == 0 && (x == 0 || x == NX)) || (comp == 1 && (y == 0 || y == NY))
                    || (comp == 2 && (z == 0 || z == NZ)))
                    f[x + (i64)sx * (y + (i64)sy * z)] = R(0);
}

// ------------------------------------------------------------------------------------------------ reductions
// backend.cu:424-443, 629-640: 65536 strided partials, then 256, then the host; NaN wins a max.
// op: 0 max, 1 sum, 2 max |x|, 3 sum x^2.
double combine(int op, double a, double b) { return (op == 0 || op == 2) ? ((a != a || a > b) ? a : b) : a + b; }

double reduce(const double* x, i64 n, int op) {
    const int T1 = 65536, T2 = 256;
    std::vector<double> p1(T1), p2(T2);
#pragma omp parallel for schedule(static)
    for (int t = 0; t < T1; t++) {
        double r = (op == 0) ? -INFINITY : 0.0;
        for (i64 i = t; i < n; i += T1) {
            double v = x[i];
            if (op == 2) v = std::fabs(v);
            if (op == 3) v = v * v;
            r = combine(op, r, v);
        }
        p1[t] = r;
    }
    int op2 = (op == 0 || op == 2) ? 0 : 1;
    for (int t = 0; t < T2; t++) {
        double r = (op2 == 0) ? -INFINITY : 0.0;
        for (int i = t; i < T1; i += T2) r = combine(op2, r, p1[i]);
        p2[t] = r;
    }
    double r = (op2 == 0) ? -INFINITY : 0.0;
    for (int k = 0; k < T2; k++) r = (op2 == 0) ? ((r != r || r > p2[k]) ? r : p2[k]) : r + p2[k];
    return r;
}

// ------------------------------------------------------------------------------------------------ FDM
// backend.cu:723-746: cyclic Jacobi on a symmetric row-major n x n matrix; A's diagonal -> eigenvalues,
// columns of V -> eigenvectors.
void jacobi_eigen(int n, vec& A, vec& V) {
    auto a = [&](int i, int j) -> double& { return A[(size_t)i * n + j]; };
    auto v = [&](int i, int j) -> double& { return V[(size_t)i * n + j]; };
    V.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) v(i, i) = 1.0;
    double total = 0;
    for (double q : A) total += q * q;
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++)
                if (i != j) off += a(i, j) * a(i, j);
        if (off <= 1e-28 * total) return;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) {
                if (a(p, q) == 0.0) continue;
                double theta = (a(q, q) - a(p, p)) / (2 * a(p, q));
                double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1));
                double cs = 1 / std::sqrt(t * t + 1), sn = t * cs;
                for (int k = 0; k < n; k++) { double x = a(k, p), y = a(k, q); a(k, p) = cs * x - sn * y; a(k, q) = sn * x + cs * y; }
                for (int k = 0; k < n; k++) { double x = a(p, k), y = a(q, k); a(p, k) = cs * x - sn * y; a(q, k) = sn * x + cs * y; }
                for (int k = 0; k < n; k++) { double x = v(k, p), y = v(k, q); v(k, p) = cs * x - sn * y; v(k, q) = sn * x + cs * y; }
            }
    }
    fail("pressure solver set-up: eigenvalues did not converge");
}

struct Eigen1 {
    int n = 0, zero_mode = 0;
    vec S, lambda;   // S[i*n + j] = eigenvector j at cell i (W-orthonormal)
};

Eigen1 axis_eigen(const Axis1& g) {                                    // backend.cu:753-776
    int n = g.n;
    vec A((size_t)n * n, 0.0), V;
    for (int i = 0; i < n - 1; i++) {
        double k = 1.0 / g.spacing[i + 1];
        A[i * n + i] -= k;
        A[(i + 1) * n + i + 1] -= k;
        A[i * n + i + 1] += k;
        A[(i + 1) * n + i] += k;
    }
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) A[i * n + j] /= std::sqrt(g.width[i] * g.width[j]);
    jacobi_eigen(n, A, V);
    Eigen1 e;
    e.n = n;
    e.lambda.resize(n);
    for (int i = 0; i < n; i++) e.lambda[i] = A[i * n + i];
    e.zero_mode = (int)(std::max_element(e.lambda.begin(), e.lambda.end()) - e.lambda.begin());
    e.lambda[e.zero_mode] = 0.0;
    e.S.resize((size_t)n * n);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) e.S[i * n + j] = V[i * n + j] / std::sqrt(g.width[i]);
    return e;
}

bool same_axis(const Axis1& a, const Axis1& b) {
    return a.n == b.n && a.width == b.width && a.spacing == b.spacing;
}

// Transforms of a cell array [nz][ny][nx] along one axis with an n x n matrix M given row-major:
// out[.., j, ..] = sum_i M(j, i) in[.., i, ..], i ascending.  transpose_m: use M(i, j) instead.
template <class T>
void apply_x(const Mesh& g, const T* M, bool transpose_m, const T* in, T* out) {
    int n = g.nx;
    i64 rows = (i64)g.ny * g.nz;
#pragma omp parallel for schedule(static)
    for (i64 r = 0; r < rows; r++) {
        const T* a = in + r * n;
        T* o = out + r * n;
        for (int j = 0; j < n; j++) o[j] = T(0);
        for (int i = 0; i < n; i++) {
            T ai = a[i];
            if (transpose_m) {                                  // o[j] += a[i] * M[i][j]
                const T* mr = M + (size_t)i * n;
#pragma omp simd
                for (int j = 0; j < n; j++) o[j] += ai * mr[j];
            } else {                                            // o[j] += a[i] * M[j][i]
                for (int j = 0; j < n; j++) o[j] += ai * M[(size_t)j * n + i];
            }
        }
    }
}

// out[a][j][b] = sum_i M(j, i) in[a][i][b] for a cell array viewed as [A][n][B]
template <class T>
void apply_mid(i64 A, int n, i64 B, const T* M, const T* in, T* out) {
    const i64 CB = 512;
    i64 nb = (B + CB - 1) / CB;
#pragma omp parallel for collapse(2) schedule(static)
    for (i64 a = 0; a < A; a++)
        for (i64 blk = 0; blk < nb; blk++) {
            i64 b0 = blk * CB, b1 = std::min(B, b0 + CB);
            const T* src = in + a * n * B;
            T* dst = out + a * n * B;
            for (int j = 0; j < n; j++) {
                T* o = dst + (i64)j * B;
                for (i64 b = b0; b < b1; b++) o[b] = T(0);
                for (int i = 0; i < n; i++) {
                    T mji = M[(size_t)j * n + i];
                    const T* s = src + (i64)i * B;
#pragma omp simd
                    for (i64 b = b0; b < b1; b++) o[b] += mji * s[b];
                }
            }
        }
}

template <class T>
void transpose(int n, const std::vector<T>& S, std::vector<T>& ST) {
    ST.resize((size_t)n * n);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) ST[(size_t)j * n + i] = S[(size_t)i * n + j];
}

// ------------------------------------------------------------------------------------------------ solver
struct History { double t, scale; int buf; };

struct HandleBase {
    bool fp32 = false;
    std::string err;
    virtual ~HandleBase() = default;
};

// R = double: the plain FP64 solver; R = float: the reference's precision (float fields, grid and stencils)
template <class R> struct OracleT : HandleBase {
    using fvec = std::vector<R>;
    static constexpr bool F32 = std::is_same<R, float>::value;
    CavConfig c{};
    Mesh g;
    i64 n_cell = 0, size[3] = {0, 0, 0};
    double volume = 0, viscous_rate = 0;
    fvec u[3], m[3];
    vec phi;
    fvec m0[3], I[3], new_m[2][3], new_u[2][3], tmp_f[3];
    vec phi0, phi_stage[2], base, guess, rhs, res, hist_buf[4], stats;
    fvec t1, t2;
    fvec S[3], ST[3], inv_lambda;
    std::vector<int> probe_cells;
    double t = 0, last_dt = 0, last_cfl = 0, phi_rel = 0, prev_rates[2] = {0, 0};
    i64 steps = 0, projections = 0, pass_total = 0;
    int dt_reason = 0, phi_passes = 0;
    double xform_seconds = 0;        // wall time inside the six FDM transforms of every pass (CPU-baseline split)
    bool have_prev_rates = false;
    std::vector<History> hist;

    double lid(double time) const {                                   // backend.cu:623-626
        if (c.lid_ramp_t > 0) return c.u_lid * std::min(1.0, std::max(time, 0.0) / c.lid_ramp_t);
        return c.u_lid;
    }
    F3T<R> view(const fvec* a) const { return F3T<R>{{a[0].data(), a[1].data(), a[2].data()}}; }
    // a face array as FP64 (itself when R is double)
    static const double* wide(const fvec& a, vec& tmp) {
        if constexpr (F32) {
            tmp.assign(a.begin(), a.end());
            return tmp.data();
        } else {
            (void)tmp;
            return a.data();
        }
    }

    void create(const CavConfig& cfg) {                                // backend.cu:1009-1073
        c = cfg;
        int nx = c.nx, ny = c.ny, nz = c.nz;
        if (nx % 2 || ny % 2 || nx < 6 || ny < 6) fail("nx and ny must be even and >= 6");
        if (nz < 5) fail("need nz >= 5");
        if (std::max(nx, std::max(ny, nz)) > 1024) fail("at most 1024 cells per axis");
        if (c.impulse_order != 2 && c.impulse_order != 4) fail("IMPULSE_ORDER must be 2 or 4");
        if ((double)(nx + 1) * ny * (nz + 1) > 2e9) fail("grid too large");
        if (c.n_probes < 0 || c.n_probes > 8) fail("n_probes must be 0..8");
        g.nx = nx; g.ny = ny; g.nz = nz; g.lid_shaped = c.lid_shaped;
        n_cell = g.n_c();
        for (int k = 0; k < 3; k++) size[k] = g.size(k);
        fp32 = F32;
        g.ax[X] = build_axis(c, nx, c.length, F32);
        g.ax[Y] = build_axis(c, ny, c.length, F32);
        g.ax[Z] = build_axis(c, nz, nz * (c.lz / nz), F32);
        volume = g.ax[X].length * g.ax[Y].length * g.ax[Z].length;
        // backend.cu:1030-1034 (loop over every cell, as the reference does)
        auto geo = [&](int a, int i) {
            return R(4) / (R(g.wd(a, i)) * std::min(R(g.sp(a, i)), R(g.sp(a, i + 1))));
        };
        R vmax = R(0);
        for (int z = 0; z < nz; z++)
            for (int y = 0; y < ny; y++)
                for (int x = 0; x < nx; x++) vmax = std::max(vmax, R(c.nu) * (geo(X, x) + geo(Y, y) + geo(Z, z)));
        viscous_rate = (double)vmax;
        for (int k = 0; k < 3; k++) {
            u[k].assign(size[k], R(0)); m[k].assign(size[k], R(0)); m0[k].assign(size[k], R(0));
            I[k].assign(size[k], R(0)); tmp_f[k].assign(size[k], R(0));
            for (int j = 0; j < 2; j++) { new_m[j][k].assign(size[k], R(0)); new_u[j][k].assign(size[k], R(0)); }
        }
        for (vec* p : {&phi, &phi0, &phi_stage[0], &phi_stage[1], &base, &guess, &rhs, &res}) p->assign(n_cell, 0.0);
        t1.assign(n_cell, R(0));
        t2.assign(n_cell, R(0));
        for (int k = 0; k < 4; k++) hist_buf[k].assign(n_cell, 0.0);
        stats.assign(6 * (i64)nx * ny, 0.0);
        setup_fdm();
        for (int p = 0; p < c.n_probes; p++) {                          // backend.cu:1058-1067
            auto cell_of = [](const Axis1& a, double frac) {
                int k = (int)(std::upper_bound(a.face.begin(), a.face.end(), frac * a.length) - a.face.begin()) - 1;
                return std::min(a.n - 1, std::max(0, k));
            };
            probe_cells.push_back(cell_of(g.ax[X], c.probe_x[p]));
            probe_cells.push_back(cell_of(g.ax[Y], c.probe_y[p]));
            probe_cells.push_back(nz / 2);
        }
    }

    void setup_fdm() {                                                 // backend.cu:750-787
        Eigen1 e[3];
        for (int a = 0; a < 3; a++) {
            int same = -1;
            for (int b = 0; b < a; b++) if (same_axis(g.ax[a], g.ax[b])) same = b;
            e[a] = same >= 0 ? e[same] : axis_eigen(g.ax[a]);           // identical input, identical result
            S[a].assign(e[a].S.begin(), e[a].S.end());                  // (float) in the FP32 fields, backend.cu:768
            transpose(e[a].n, S[a], ST[a]);
        }
        inv_lambda.assign(n_cell, R(0));
        for (int z = 0; z < g.nz; z++)
            for (int y = 0; y < g.ny; y++)
                for (int x = 0; x < g.nx; x++) {
                    bool zero = x == e[X].zero_mode && y == e[Y].zero_mode && z == e[Z].zero_mode;
                    inv_lambda[g.ic(x, y, z)] = zero ? R(0) : R(1.0 / (e[Z].lambda[z] + e[Y].lambda[y] + e[X].lambda[x]));
                }
    }

    // backend.cu:801-814
    void pressure_pass(const double* r, double* ph) {
        i64 n = n_cell;
        double rms = std::sqrt(reduce(r, n, 3) / n);
        if (!(rms > 0)) return;
        double s = 1.0 / rms;
        R *a = t1.data(), *b = t2.data();
#pragma omp parallel for schedule(static)
        for (i64 i = 0; i < n; i++) a[i] = (R)(r[i] * s);
        i64 plane = (i64)g.nx * g.ny;
        const auto t0 = std::chrono::steady_clock::now();
        apply_x(g, S[X].data(), true, a, b);                          // t2 = t1 S_x
        apply_mid(g.nz, g.ny, g.nx, ST[Y].data(), b, a);               // t1 = S_y^T t2 (per z)
        apply_mid(1, g.nz, plane, ST[Z].data(), a, b);                 // t2 = S_z^T t1
        const R* il = inv_lambda.data();
#pragma omp parallel for schedule(static)
        for (i64 i = 0; i < n; i++) b[i] = b[i] * il[i];
        apply_mid(1, g.nz, plane, S[Z].data(), b, a);                  // t1 = S_z t2
        apply_mid(g.nz, g.ny, g.nx, S[Y].data(), a, b);                // t2 = S_y t1
        apply_x(g, ST[X].data(), true, b, a);                          // t1 = t2 S_x^T
        xform_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
#pragma omp parallel for schedule(static)
        for (i64 i = 0; i < n; i++) ph[i] = ph[i] + (double)a[i] * rms;
    }

    // backend.cu:816-847
    void solve_pressure(const double* rh, const double* gs, double* ph) {
        i64 n = n_cell;
        double rhs_norm = std::sqrt(reduce(rh, n, 3));
        phi_passes = 0;
        if (!std::isfinite(rhs_norm)) fail("pressure solve: right-hand side is NaN/inf (the run has blown up)");
        if (rhs_norm == 0.0) { std::fill(ph, ph + n, 0.0); phi_rel = 0; return; }
        const double* r = rh;
        double rel = 1.0;
        std::fill(ph, ph + n, 0.0);
        if (gs) {
            std::copy(gs, gs + n, ph);
            pressure_residual(g, ph, rh, res.data());
            r = res.data();
            rel = std::sqrt(reduce(r, n, 3)) / rhs_norm;
            if (rel >= 1.0) { std::fill(ph, ph + n, 0.0); r = rh; rel = 1.0; }
        }
        while (rel > c.fdm_rtol && phi_passes < c.fdm_max_passes) {
            pressure_pass(r, ph);
            phi_passes++;
            pressure_residual(g, ph, rh, res.data());
            r = res.data();
            rel = std::sqrt(reduce(r, n, 3)) / rhs_norm;
            if (!std::isfinite(rel)) fail("pressure solve: residual is NaN/inf");
        }
        projections++;
        pass_total += phi_passes;
        if (rel > c.fdm_rtol) {
            char msg[200];
            std::snprintf(msg, sizeof msg, "pressure solve: residual %.2e > FDM_RTOL after %d passes", rel, phi_passes);
            fail(msg);
        }
        phi_rel = rel;
        double mean = reduce(ph, n, 1) / n;
#pragma omp parallel for schedule(static)
        for (i64 i = 0; i < n; i++) ph[i] = ph[i] - mean;
    }

    // backend.cu:849-854
    void project(fvec* mm, const double* gs, double* ph, fvec* out) {
        for (int k = 0; k < 3; k++) zero_walls(g, k, mm[k].data());
        divergence(g, view(mm), rhs.data());
        solve_pressure(rhs.data(), gs, ph);
        for (int k = 0; k < 3; k++) subtract_gradient(g, k, mm[k].data(), ph, out[k].data());
    }

    // backend.cu:857-873
    bool guess_phi(double tau, double scale, double* out) {
        if (!c.phi_extrapolate || hist.empty()) return false;
        std::vector<int> order(hist.size());
        for (size_t k = 0; k < order.size(); k++) order[k] = (int)k;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
            double da = std::fabs(hist[a].t - tau), db = std::fabs(hist[b].t - tau);
            return da != db ? da < db : a > b;
        });
        const History& h1 = hist[order[0]];
        const History* h2 = nullptr;
        for (size_t k = 1; k < order.size() && !h2; k++)
            if (std::fabs(hist[order[k]].t - h1.t) > 1e-12 * std::max(1.0, std::fabs(h1.t))) h2 = &hist[order[k]];
        double w2 = h2 ? (tau - h1.t) / (h2->t - h1.t) : 0.0, w1 = 1.0 - w2;
        i64 n = n_cell;
        const double* d1 = hist_buf[h1.buf].data();
        double c1 = scale * w1 / h1.scale;
        if (h2) {
            const double* d2 = hist_buf[h2->buf].data();
            double c2 = scale * w2 / h2->scale;
#pragma omp parallel for schedule(static)
            for (i64 i = 0; i < n; i++) out[i] = c1 * d1[i] + c2 * d2[i];
        } else {
#pragma omp parallel for schedule(static)
            for (i64 i = 0; i < n; i++) out[i] = c1 * d1[i];
        }
        return true;
    }

    double cfl_max(const fvec* uu) {                                   // backend.cu:465-472 + reduce op 0
        vec rate(n_cell);
        const R *u0 = uu[0].data(), *v0 = uu[1].data(), *w0 = uu[2].data();
#pragma omp parallel for collapse(2) schedule(static)
        for (int z = 0; z < g.nz; z++)
            for (int y = 0; y < g.ny; y++)
                for (int x = 0; x < g.nx; x++)
                    rate[g.ic(x, y, z)] = (double)(
                        std::fmax(std::fabs(u0[g.iu(x, y, z)]), std::fabs(u0[g.iu(x + 1, y, z)])) / R(g.wd(X, x))
                        + std::fmax(std::fabs(v0[g.iv(x, y, z)]), std::fabs(v0[g.iv(x, y + 1, z)])) / R(g.wd(Y, y))
                        + std::fmax(std::fabs(w0[g.iw(x, y, z)]), std::fabs(w0[g.iw(x, y, z + 1)])) / R(g.wd(Z, z)));
        return reduce(rate.data(), n_cell, 0);
    }

    // backend.cu:877-955
    void step() {
        i64 n = n_cell;
        double r = c.reinit_r;
        const R rf = R(c.reinit_r), omrf = R(1.0 - c.reinit_r);       // backend.cu:881: (float) in the FP32 fields
        for (int k = 0; k < 3; k++) {
            const R *mk = m[k].data(), *uk = u[k].data();
            R* o = m0[k].data();
#pragma omp parallel for schedule(static)
            for (i64 i = 0; i < size[k]; i++) o[i] = rf * mk[i] + omrf * uk[i];
        }
#pragma omp parallel for schedule(static)
        for (i64 i = 0; i < n; i++) phi0[i] = r * phi[i];

        double rates[2] = {cfl_max(u), viscous_rate};
        double use[2] = {rates[0], rates[1]};
        if (c.dt_lag && have_prev_rates) { use[0] = prev_rates[0]; use[1] = prev_rates[1]; }
        prev_rates[0] = rates[0]; prev_rates[1] = rates[1]; have_prev_rates = true;
        if (!std::isfinite(use[0])) fail("velocity is NaN/inf: the solution has blown up");
        double dt;
        if (!c.adaptive_dt) { dt = c.dt_fixed; dt_reason = 0; }
        else {
            double caps[3] = {c.dt_max, c.cfl_target / std::max(use[0], 1e-300),
                              c.diffusion_limit / std::max(use[1], 1e-300)};
            int best = 0;
            for (int k = 1; k < 3; k++) if (caps[k] < caps[best]) best = k;
            dt = caps[best];
            dt_reason = best + 1;
        }
        if (c.dt_lag && last_dt > 0) dt = std::min(dt, 1.1 * last_dt);

        const double a0s[3] = {0.0, 0.75, 1.0 / 3.0}, bs[3] = {1.0, 0.25, 2.0 / 3.0};
        const double times[3] = {t, t + dt, t + 0.5 * dt};
        fvec* cur_u = u;
        fvec* cur_m = m0;
        vec* cur_phi = &phi0;
        int o4 = c.impulse_order == 4;
        const R K = R(std::max(c.impulse_order4_ramp, 1.0));
        for (int stage = 0; stage < 3; stage++) {
            double a0 = a0s[stage], b = bs[stage], ts = times[stage];
            fvec* nm = new_m[stage % 2];
            fvec* nu = new_u[stage % 2];
            R* Iout[3] = {I[0].data(), I[1].data(), I[2].data()};
            inviscid(g, view(cur_u), view(cur_m), o4, K, Iout);
            for (int k = 0; k < 3; k++)
                rk3_stage(g, k, cur_m[k].data(), m0[k].data(), cur_u[k].data(), I[k].data(), R(c.nu), R(lid(ts)), R(dt),
                          R(a0), R(b), stage > 0 ? 1 : 0, nm[k].data());
            const double *p0 = phi0.data(), *pc = cur_phi->data();
            double* bs_ = base.data();
            if (a0 != 0.0) {
#pragma omp parallel for schedule(static)
                for (i64 i = 0; i < n; i++) bs_[i] = a0 * p0[i] + b * pc[i];
            } else {
#pragma omp parallel for schedule(static)
                for (i64 i = 0; i < n; i++) bs_[i] = b * pc[i];
            }
            double* gs = guess.data();
            if (guess_phi(ts, b * dt, gs)) {
#pragma omp parallel for schedule(static)
                for (i64 i = 0; i < n; i++) gs[i] = gs[i] + bs_[i];
            } else {
                std::copy(bs_, bs_ + n, gs);
            }
            vec* new_phi = &phi_stage[stage % 2];
            project(nm, gs, new_phi->data(), nu);
            if (c.phi_extrapolate) {
                int buf = -1;
                for (int k = 0; k < 4 && buf < 0; k++) {
                    bool used = false;
                    for (const History& h : hist) used |= (h.buf == k);
                    if (!used) buf = k;
                }
                double* d = hist_buf[buf].data();
                const double* np = new_phi->data();
#pragma omp parallel for schedule(static)
                for (i64 i = 0; i < n; i++) d[i] = 1.0 * np[i] + -1.0 * bs_[i];
                hist.push_back({ts, b * dt, buf});
                if (hist.size() > 3) hist.erase(hist.begin());
            }
            cur_u = nu;
            cur_m = nm;
            cur_phi = new_phi;
        }
        for (int k = 0; k < 3; k++) {
            std::swap(u[k], new_u[0][k]);
            std::swap(m[k], new_m[0][k]);
        }
        std::swap(phi, phi_stage[0]);
        t += dt;
        steps++;
        last_dt = dt;
        last_cfl = dt * use[0];
    }

    // backend.cu:958-978
    void kick(const double* const noise[3], double amp, int passes) {
        for (int k = 0; k < 3; k++) if (!noise[k]) fail("kick: noise arrays must not be NULL");
        vec na, nb;
        for (int k = 0; k < 3; k++) {
            i64 n = size[k];
            int sx = g.nx + (k == 0), sy = g.ny + (k == 1), sz = g.nz + (k == 2);
            na.assign(noise[k], noise[k] + n);
            nb.assign(n, 0.0);
            for (int p = 0; p < std::max(1, passes); p++)
                for (int axis = 2; axis >= 0; axis--) {
                    smooth(na.data(), nb.data(), axis, sx, sy, sz);
                    std::swap(na, nb);
                }
            double rms = std::sqrt(reduce(na.data(), n, 3) / n);
            double s = amp * c.u_lid / std::max(rms, 1e-30);
            const R* uk = u[k].data();
            R* o = tmp_f[k].data();
#pragma omp parallel for schedule(static)
            for (i64 i = 0; i < n; i++) o[i] = uk[i] + (R)(na[i] * s);
        }
        project(tmp_f, nullptr, phi_stage[0].data(), new_u[0]);
        for (int k = 0; k < 3; k++) {
            R* mk = m[k].data();
            const R *a = new_u[0][k].data(), *b = u[k].data();
#pragma omp parallel for schedule(static)
            for (i64 i = 0; i < size[k]; i++) mk[i] = mk[i] + (a[i] - b[i]);
            std::swap(u[k], new_u[0][k]);
        }
    }

    static void smooth(const double* in, double* out, int axis, int sx, int sy, int sz) {   // backend.cu:544-553
#pragma omp parallel for collapse(2) schedule(static)
        for (int z = 0; z < sz; z++)
            for (int y = 0; y < sy; y++)
                for (int x = 0; x < sx; x++) {
                    i64 i = x + (i64)sx * (y + (i64)sy * z), lo, hi;
                    if (axis == 0) { lo = (x + sx - 1) % sx + (i64)sx * (y + (i64)sy * z); hi = (x + 1) % sx + (i64)sx * (y + (i64)sy * z); }
                    else if (axis == 1) { lo = x + (i64)sx * ((y + sy - 1) % sy + (i64)sy * z); hi = x + (i64)sx * ((y + 1) % sy + (i64)sy * z); }
                    else { lo = x + (i64)sx * (y + (i64)sy * ((z + sz - 1) % sz)); hi = x + (i64)sx * (y + (i64)sy * ((z + 1) % sz)); }
                    out[i] = 0.25 * in[lo] + 0.5 * in[i] + 0.25 * in[hi];
                }
    }

    void initialize(const double* vn, const double* wn) {             // backend.cu:1075-1088
        const double* src[3] = {nullptr, vn, wn};
        for (int k = 0; k < 3; k++) {
            m[k].assign(size[k], R(0));
            if (src[k])
                for (i64 i = 0; i < size[k]; i++) m[k][i] = (R)src[k][i];
        }
        project(m, nullptr, phi.data(), u);
        t = 0; steps = 0; last_dt = 0; hist.clear(); have_prev_rates = false;
    }

    void fill(CavReport* r) const {                                   // backend.cu:995-999
        r->steps = steps; r->time = t; r->dt = last_dt; r->cfl = last_cfl; r->phi_rel = phi_rel;
        r->dt_reason = dt_reason; r->phi_passes = phi_passes;
    }

    void diagnostics(CavDiagnostics* d) {                             // backend.cu:1096-1129, 447-463
        i64 n = n_cell;
        const R lidv = R(lid(t));
        vec ke(n), ens(n), div2(n);
        const R *uu = u[0].data(), *vv = u[1].data(), *ww = u[2].data();
#pragma omp parallel for collapse(2) schedule(static)
        for (int z = 0; z < g.nz; z++)
            for (int y = 0; y < g.ny; y++)
                for (int x = 0; x < g.nx; x++) {
                    i64 i = g.ic(x, y, z);
                    double vol = g.wd(X, x) * g.wd(Y, y) * g.wd(Z, z);
                    double a = centred(g, uu, vv, ww, 0, x, y, z), b = centred(g, uu, vv, ww, 1, x, y, z),
                           cc = centred(g, uu, vv, ww, 2, x, y, z);
                    ke[i] = 0.5 * (a * a + b * b + cc * cc) * vol;
                    R gr[3][3];
                    cell_gradient(g, uu, vv, ww, lidv, x, y, z, gr);
                    R ox = gr[1][2] - gr[2][1], oy = gr[2][0] - gr[0][2], oz = gr[0][1] - gr[1][0];
                    ens[i] = (double)(ox * ox + oy * oy + oz * oz) * vol;
                    double dd = (double)(uu[g.iu(x + 1, y, z)] - uu[g.iu(x, y, z)]) / g.wd(X, x)
                              + (double)(vv[g.iv(x, y + 1, z)] - vv[g.iv(x, y, z)]) / g.wd(Y, y)
                              + (double)(ww[g.iw(x, y, z + 1)] - ww[g.iw(x, y, z)]) / g.wd(Z, z);
                    div2[i] = dd * dd * vol;
                }
        d->kinetic_energy = reduce(ke.data(), n, 1) / volume;
        d->enstrophy = 0.5 * reduce(ens.data(), n, 1) / volume;
        d->div_rms = std::sqrt(reduce(div2.data(), n, 1) / volume);
        double umax = 0;
        vec tmp;
        for (int k = 0; k < 3; k++) umax = std::max(umax, reduce(wide(u[k], tmp), size[k], 2));
        d->umax = umax;
        d->cfl_rate = cfl_max(u);
        double best = -1;
        int best_comp = 0;
        i64 best_i = 0;
        for (int k = 0; k < 3; k++) {
            vec diff(size[k]);
            for (i64 i = 0; i < size[k]; i++) diff[i] = (double)std::fabs(m[k][i] - u[k][i]);
            double val = reduce(diff.data(), size[k], 0);
            if (val > best) {
                i64 first = 2147483647;
                for (i64 i = 0; i < size[k]; i++) if (diff[i] == val) { first = std::min(first, i); break; }
                best_i = first;
                best = val;
                best_comp = k;
            }
        }
        int sx = g.nx + (best_comp == 0), sy = g.ny + (best_comp == 1);
        i64 ix = best_i % sx, iy = (best_i / sx) % sy;
        d->gauge_ratio = best / std::max(umax, 1e-30);
        d->gauge_x = (best_comp == 0 ? g.fc(X, (int)ix) : g.ce(X, (int)std::min<i64>(ix, g.nx - 1))) / c.length;
        d->gauge_y = (best_comp == 1 ? g.fc(Y, (int)iy) : g.ce(Y, (int)std::min<i64>(iy, g.ny - 1))) / c.length;
    }

    void probes(double* out) const {                                  // backend.cu:534-540
        const R *uu = u[0].data(), *vv = u[1].data();
        for (int p = 0; p < c.n_probes; p++) {
            int x = probe_cells[3 * p], y = probe_cells[3 * p + 1], z = probe_cells[3 * p + 2];
            out[2 * p] = (double)(R(0.5) * (uu[g.iu(x, y, z)] + uu[g.iu(x + 1, y, z)]));
            out[2 * p + 1] = (double)(R(0.5) * (vv[g.iv(x, y, z)] + vv[g.iv(x, y + 1, z)]));
        }
    }

    void midspan(int& k1, int& k2) const { k2 = g.nz / 2; k1 = (g.nz % 2) ? k2 : k2 - 1; }   // backend.cu:488

    void add_statistics(double weight, double* line) {                // backend.cu:507-532
        const R *uu = u[0].data(), *vv = u[1].data(), *ww = u[2].data();
        int k1, k2;
        midspan(k1, k2);
        i64 n = (i64)g.nx * g.ny;
        for (int y = 0; y < g.ny; y++)
            for (int x = 0; x < g.nx; x++) {
                i64 i = x + (i64)g.nx * y;
                double a = 0.5 * ((double)centred(g, uu, vv, ww, 0, x, y, k1) + (double)centred(g, uu, vv, ww, 0, x, y, k2));
                double b = 0.5 * ((double)centred(g, uu, vv, ww, 1, x, y, k1) + (double)centred(g, uu, vv, ww, 1, x, y, k2));
                double cc = 0.5 * ((double)centred(g, uu, vv, ww, 2, x, y, k1) + (double)centred(g, uu, vv, ww, 2, x, y, k2));
                stats[i] += weight * a; stats[n + i] += weight * b; stats[2 * n + i] += weight * (a * a);
                stats[3 * n + i] += weight * (b * b); stats[4 * n + i] += weight * (cc * cc);
                stats[5 * n + i] += weight * (a * b);
            }
        if (!line) return;
        if (g.nx != g.ny) fail("centreline sample needs nx == ny");
        int NX = g.nx, mid = NX / 2;
        auto mean4 = [&](int cc, int xa, int ya, int xb, int yb) {
            return 0.5 * (0.5 * ((double)centred(g, uu, vv, ww, cc, xa, ya, k1) + (double)centred(g, uu, vv, ww, cc, xa, ya, k2))
                          + 0.5 * ((double)centred(g, uu, vv, ww, cc, xb, yb, k1) + (double)centred(g, uu, vv, ww, cc, xb, yb, k2)));
        };
        for (int j = 0; j < NX; j++) {
            double a = mean4(0, mid - 1, j, mid, j), b = mean4(1, j, mid - 1, j, mid), cc = mean4(2, mid - 1, j, mid, j);
            line[j] = a; line[NX + j] = b; line[2 * NX + j] = a * a; line[3 * NX + j] = b * b; line[4 * NX + j] = cc * cc;
        }
    }

    void midplane(double* out) const {                                // backend.cu:490-505
        const R *uu = u[0].data(), *vv = u[1].data(), *ww = u[2].data();
        const R lidv = R(lid(t));
        int k1, k2;
        midspan(k1, k2);
        i64 P = (i64)g.nx * g.ny;
        for (int y = 0; y < g.ny; y++)
            for (int x = 0; x < g.nx; x++) {
                R val[2][4];
                for (int s = 0; s < 2; s++) {
                    int z = s ? k2 : k1;
                    for (int cc = 0; cc < 3; cc++) val[s][cc] = centred(g, uu, vv, ww, cc, x, y, z);
                    R gr[3][3];
                    cell_gradient(g, uu, vv, ww, lidv, x, y, z, gr);
                    R a = gr[1][2] - gr[2][1], b = gr[2][0] - gr[0][2], cz = gr[0][1] - gr[1][0];
                    val[s][3] = std::sqrt(a * a + b * b + cz * cz);
                }
                i64 i = x + (i64)g.nx * y;
                for (int cc = 0; cc < 4; cc++) out[cc * P + i] = (double)((k1 == k2) ? val[1][cc] : R(0.5) * (val[0][cc] + val[1][cc]));
            }
    }

    void q_volume(int colour_w, double* q, double* colour) const {    // backend.cu:474-486, 1185-1194
        const R *uu = u[0].data(), *vv = u[1].data(), *ww = u[2].data();
        const R lidv = R(lid(t));
        double LU = c.length / c.u_lid;
        const R qscale = R(LU * LU), inv_U = R(1.0 / c.u_lid);        // backend.cu:1189-1190
#pragma omp parallel for collapse(2) schedule(static)
        for (int z = 0; z < g.nz; z++)
            for (int y = 0; y < g.ny; y++)
                for (int x = 0; x < g.nx; x++) {
                    i64 i = g.ic(x, y, z);
                    R gr[3][3];
                    cell_gradient(g, uu, vv, ww, lidv, x, y, z, gr);
                    R s = R(0);
                    for (int a = 0; a < 3; a++) for (int b = 0; b < 3; b++) s += gr[a][b] * gr[b][a];
                    q[i] = (double)(R(-0.5) * s * qscale);
                    R a = centred(g, uu, vv, ww, 0, x, y, z), b = centred(g, uu, vv, ww, 1, x, y, z),
                      cc = centred(g, uu, vv, ww, 2, x, y, z);
                    colour[i] = (double)(colour_w ? cc * inv_U : std::sqrt(a * a + b * b + cc * cc) * inv_U);
                }
    }

    // field I/O in FP64 (the FP32 fields convert at the boundary)
    void get_field(int f, double* out) {
        if (f < 0 || f > 6) fail("field must be 0..6");
        if (f == 6) { std::copy(phi.begin(), phi.end(), out); return; }
        const fvec& a = f < 3 ? u[f] : m[f - 3];
        for (size_t i = 0; i < a.size(); i++) out[i] = (double)a[i];
    }
    void set_field(int f, const double* in) {
        if (f < 0 || f > 6) fail("field must be 0..6");
        if (f == 6) { std::copy(in, in + n_cell, phi.begin()); return; }
        fvec& a = f < 3 ? u[f] : m[f - 3];
        for (size_t i = 0; i < a.size(); i++) a[i] = (R)in[i];
    }
    i64 field_size(int f) const { return f < 3 ? size[f] : f < 6 ? size[f - 3] : n_cell; }
};

thread_local std::string g_err;

using Oracle = OracleT<double>;
using Oracle32 = OracleT<float>;

// Run fn on the handle's oracle (FP64 or FP32 fields), errors into g_err
template <class Fn> int with(void* h, Fn fn) {
    HandleBase* b = (HandleBase*)h;
    try {
        if (!b) fail("no oracle");
        if (b->fp32) fn(*static_cast<Oracle32*>(b));
        else fn(*static_cast<Oracle*>(b));
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        if (b) b->err = e.what();
        return 1;
    }
}
// The kernel-level test hooks restate the FP64 kernels only
template <class Fn> int with64(void* h, Fn fn) {
    return with(h, [&](auto& o) {
        if constexpr (std::is_same<std::decay_t<decltype(o)>, Oracle>::value) fn(o);
        else fail("the kernel-level entry points run on an FP64 oracle");
    });
}

}  // namespace

// ------------------------------------------------------------------------------------------------ C API
// Mirrors the solver_* boundary with a handle (orc_*), all field I/O in float64, plus entry points to
// the individual kernels for the invariant tests.
extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

// fp32 = 0: the plain FP64 oracle; 1: the reference's precision (float fields and stencils)
void* orc_create_mode(const CavConfig* c, int fp32) {
    HandleBase* o = fp32 ? (HandleBase*)new Oracle32() : (HandleBase*)new Oracle();
    try {
        if (fp32) static_cast<Oracle32*>(o)->create(*c);
        else static_cast<Oracle*>(o)->create(*c);
    } catch (const std::exception& e) {
        g_err = e.what();
        delete o;
        return nullptr;
    }
    return o;
}
void* orc_create(const CavConfig* c) { return orc_create_mode(c, 0); }
void orc_destroy(void* h) { delete (HandleBase*)h; }
int orc_initialize(void* h, const double* v, const double* w) {
    return with(h, [&](auto& o) { o.initialize(v, w); });
}
int orc_step(void* h, CavReport* r) { return with(h, [&](auto& o) { o.step(); o.fill(r); }); }
int orc_steps(void* h, int n, CavReport* r) {
    return with(h, [&](auto& o) { for (int i = 0; i < n; i++) o.step(); o.fill(r); });
}
int orc_kick(void* h, const double* nu, const double* nv, const double* nw, double amp, int passes) {
    return with(h, [&](auto& o) { const double* n[3] = {nu, nv, nw}; o.kick(n, amp, passes); });
}
int orc_report(void* h, CavReport* r) { return with(h, [&](auto& o) { o.fill(r); }); }
int orc_diagnostics(void* h, CavDiagnostics* d) { return with(h, [&](auto& o) { o.diagnostics(d); }); }
int orc_grid(void* h, int axis, double* faces, double* centres, double* widths, double* spacings) {
    return with(h, [&](auto& o) {
        if (axis < 0 || axis > 2) fail("axis must be 0..2");
        const Axis1& a = o.g.ax[axis];
        if (faces) std::copy(a.face.begin(), a.face.end(), faces);
        if (centres) std::copy(a.centre.begin(), a.centre.end(), centres);
        if (widths) std::copy(a.width.begin(), a.width.end(), widths);
        if (spacings) std::copy(a.spacing.begin(), a.spacing.end(), spacings);
    });
}
int orc_get_field(void* h, int f, double* out) { return with(h, [&](auto& o) { o.get_field(f, out); }); }
int orc_set_field(void* h, int f, const double* in) {
    return with(h, [&](auto& o) {
        o.set_field(f, in);
        o.hist.clear();
        o.have_prev_rates = false;
    });
}
int orc_set_clock(void* h, double t, long long steps, double last_dt) {
    return with(h, [&](auto& o) { o.t = t; o.steps = steps; o.last_dt = last_dt; });
}
int orc_probes(void* h, double* out) { return with(h, [&](auto& o) { o.probes(out); }); }
int orc_add_statistics(void* h, double weight, double* line) {
    return with(h, [&](auto& o) { o.add_statistics(weight, line); });
}
int orc_get_statistics(void* h, double* out) {
    return with(h, [&](auto& o) { std::copy(o.stats.begin(), o.stats.end(), out); });
}
int orc_set_statistics(void* h, const double* in) {
    return with(h, [&](auto& o) { std::copy(in, in + o.stats.size(), o.stats.begin()); });
}
int orc_midplane(void* h, double* out) { return with(h, [&](auto& o) { o.midplane(out); }); }
int orc_q_volume(void* h, int colour_w, double* q, double* colour) {
    return with(h, [&](auto& o) { o.q_volume(colour_w, q, colour); });
}
int orc_pass_counters(void* h, long long* projections, long long* passes) {
    return with(h, [&](auto& o) { *projections = o.projections; *passes = o.pass_total; });
}
// Seconds spent in the FDM transforms since create, and the OpenMP threads the oracle runs on
int orc_timing(void* h, double* xform_seconds, int* threads) {
    return with(h, [&](auto& o) {
        *xform_seconds = o.xform_seconds;
        int nt = 1;
#ifdef _OPENMP
        nt = omp_get_max_threads();
#endif
        *threads = nt;
    });
}
int orc_viscous_rate(void* h, double* out) { return with(h, [&](auto& o) { *out = o.viscous_rate; }); }

// ---- kernel-level entry points (same mesh and settings as the handle's config; FP64 oracle)
int orc_inviscid(void* h, const double* u, const double* v, const double* w, const double* mx, const double* my,
                 const double* mz, double* ox, double* oy, double* oz) {
    return with64(h, [&](Oracle& o) {
        double* out[3] = {ox, oy, oz};
        inviscid(o.g, F3{{u, v, w}}, F3{{mx, my, mz}}, o.c.impulse_order == 4, std::max(o.c.impulse_order4_ramp, 1.0),
                 out);
    });
}
int orc_rk3_stage(void* h, int comp, const double* m, const double* m0, const double* vel, const double* I,
                  double nu, double lid, double dt, double a0, double b, int has_base, double* out) {
    return with64(h, [&](Oracle& o) { rk3_stage(o.g, comp, m, m0, vel, I, nu, lid, dt, a0, b, has_base, out); });
}
int orc_divergence(void* h, const double* u, const double* v, const double* w, double* rhs) {
    return with64(h, [&](Oracle& o) { divergence(o.g, F3{{u, v, w}}, rhs); });
}
int orc_residual(void* h, const double* phi, const double* rhs, double* r) {
    return with(h, [&](auto& o) { pressure_residual(o.g, phi, rhs, r); });
}
int orc_pressure_pass(void* h, const double* r, double* phi) {
    return with(h, [&](auto& o) { o.pressure_pass(r, phi); });
}
int orc_solve_pressure(void* h, const double* rhs, const double* guess, double* phi, double* rel, int* passes) {
    return with(h, [&](auto& o) { o.solve_pressure(rhs, guess, phi); *rel = o.phi_rel; *passes = o.phi_passes; });
}
// Test hook: replace the warm-start history by n entries (times t, scales sc, data n x n_cell, oldest first)
// and return guess_phi's output for (tau, scale); *used = 0 when it makes no guess.
int orc_guess_phi(void* h, int n, const double* t, const double* sc, const double* data, double tau, double scale,
                  double* out, int* used) {
    return with(h, [&](auto& o) {
        if (n < 0 || n > 3) fail("the history holds at most 3 entries");
        o.hist.clear();
        for (int k = 0; k < n; k++) {
            std::copy(data + (i64)k * o.n_cell, data + (i64)(k + 1) * o.n_cell, o.hist_buf[k].begin());
            o.hist.push_back({t[k], sc[k], k});
        }
        *used = o.guess_phi(tau, scale, out) ? 1 : 0;
    });
}
int orc_subtract_gradient(void* h, int comp, const double* m, const double* phi, double* out) {
    return with64(h, [&](Oracle& o) { subtract_gradient(o.g, comp, m, phi, out); });
}
int orc_eigen(void* h, int axis, double* S, double* lambda, int* zero_mode) {
    return with(h, [&](auto& o) {
        if (axis < 0 || axis > 2) fail("axis must be 0..2");
        Eigen1 e = axis_eigen(o.g.ax[axis]);
        std::copy(e.S.begin(), e.S.end(), S);
        std::copy(e.lambda.begin(), e.lambda.end(), lambda);
        *zero_mode = e.zero_mode;
    });
}
int orc_inv_lambda(void* h, double* out) {
    return with(h, [&](auto& o) { for (size_t i = 0; i < o.inv_lambda.size(); i++) out[i] = (double)o.inv_lambda[i]; });
}
int orc_cfl_max(void* h, double* out) { return with(h, [&](auto& o) { *out = o.cfl_max(o.u); }); }
double orc_reduce(const double* x, long long n, int op) { return reduce(x, n, op); }
int orc_smooth(const double* in, double* out, int axis, int sx, int sy, int sz) {
    Oracle::smooth(in, out, axis, sx, sy, sz);
    return 0;
}
// stretched/uniform face positions on their own (the grid restatement, backend.cu:651-699)
int orc_axis_faces(const CavConfig* c, int n, double length, double* faces) {
    try {
        Axis1 a = build_axis(*c, n, length);
        std::copy(a.face.begin(), a.face.end(), faces);
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

}  // extern "C"
