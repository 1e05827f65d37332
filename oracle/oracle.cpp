// oracle.cpp — TEST INFRASTRUCTURE ONLY. FP64 CPU restatement of the reference solver's algorithm
// (backend.cu:33-1195), used by tests/ as the parity checker and by bench.py's cpu_baseline leg as the
// host-core baseline. The product library (libcavity_solver.so) never links, loads or calls this code.
//
// Every kernel of the reference is restated as a loop over its index space (OpenMP across the outer
// index; each iteration writes only its own output, so results do not depend on the thread count), in the
// reference's order of floating-point operations, with float replaced by double everywhere (the "plain
// FP64" semantics: the grid is not rounded through float, backend.cu:695-697, and the CFL rate is not
// rounded through float, backend.cu:886). Compile with -ffp-contract=off (the reference builds with
// --fmad=false, build_windows.bat:4).
//
// The FDM transforms (cublasSgemmStridedBatched in the reference, backend.cu:792-798) are restated as
// FP64 contractions that sum over the contracted index in ascending order. cuBLAS's own summation order
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

Axis1 build_axis(const CavConfig& c, int n, double length) {           // backend.cu:672-699
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
    return a;                                                           // FP64: no rounding through float
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
};

struct F3 { const double* c[3]; };

// backend.cu:67-72
double centred(const Mesh& g, const double* u, const double* v, const double* w, int c, int x, int y, int z) {
    if (c == 0) return 0.5 * (u[g.iu(x, y, z)] + u[g.iu(x + 1, y, z)]);
    if (c == 1) return 0.5 * (v[g.iv(x, y, z)] + v[g.iv(x, y + 1, z)]);
    return 0.5 * (w[g.iw(x, y, z)] + w[g.iw(x, y, z + 1)]);
}

// backend.cu:76-96  g[i][c] = d u_c / d x_i at the cell centre, mirrored ghosts at the walls
void cell_gradient(const Mesh& g, const double* u, const double* v, const double* w, double lid, int x, int y,
                   int z, double gr[3][3]) {
    int NX = g.nx, NY = g.ny, NZ = g.nz;
    double dxm = x > 0 ? g.sp(X, x) : g.wd(X, 0), dxp = x + 1 < NX ? g.sp(X, x + 1) : g.wd(X, NX - 1);
    double dym = y > 0 ? g.sp(Y, y) : g.wd(Y, 0), dyp = y + 1 < NY ? g.sp(Y, y + 1) : g.wd(Y, NY - 1);
    double dzm = z > 0 ? g.sp(Z, z) : g.wd(Z, 0), dzp = z + 1 < NZ ? g.sp(Z, z + 1) : g.wd(Z, NZ - 1);
    for (int c = 0; c < 3; c++) {
        double q = centred(g, u, v, w, c, x, y, z);
        double xm = x > 0 ? centred(g, u, v, w, c, x - 1, y, z) : -q;
        double xp = x + 1 < NX ? centred(g, u, v, w, c, x + 1, y, z) : -q;
        double ym = y > 0 ? centred(g, u, v, w, c, x, y - 1, z) : -q;
        double yp = y + 1 < NY ? centred(g, u, v, w, c, x, y + 1, z)
                               : 2 * (c == 0 ? lid * g.lid_at_cell(x, z) : 0) - q;
        double zp = z + 1 < NZ ? centred(g, u, v, w, c, x, y, z + 1) : -q;
        double zm = z > 0 ? centred(g, u, v, w, c, x, y, z - 1) : -q;
        gr[0][c] = (xp - xm) / (dxp + dxm);
        gr[1][c] = (yp - ym) / (dyp + dym);
        gr[2][c] = (zp - zm) / (dzp + dzm);
    }
    gr[0][0] = (u[g.iu(x + 1, y, z)] - u[g.iu(x, y, z)]) / g.wd(X, x);
    gr[1][1] = (v[g.iv(x, y + 1, z)] - v[g.iv(x, y, z)]) / g.wd(Y, y);
    gr[2][2] = (w[g.iw(x, y, z + 1)] - w[g.iw(x, y, z)]) / g.wd(Z, z);
}

// ------------------------------------------------------------------------------------------------ I(u, m)
// backend.cu:102-274: I = u x omega - grad(u.m), Lamb vector from edge vorticity, 4th order as narrow
// (q = 1) plus wide (q = 3) stencils with a wall ramp.
struct Impulse {
    const Mesh& g;
    const double *u, *v, *w, *mx, *my, *mz;
    int o4;
    double K;

    static bool nodes(int e, int q, int n, int& lo, int& hi) {         // backend.cu:108-112
        lo = e - (q + 1) / 2;
        hi = e + (q - 1) / 2;
        return lo >= 0 && hi <= n - 1;
    }
    double node_gap(int a, int e, int q, int lo, int hi) const {       // backend.cu:114-117
        if (q == 1) return g.sp(a, e);
        return g.ce(a, hi) - g.ce(a, lo);
    }
    double at_edge(int a, int e, int q, int lo, int hi, double flo, double fhi) const {   // backend.cu:119-122
        if (q == 1) return (g.wd(a, lo) * flo + g.wd(a, hi) * fhi) / (2.0 * g.sp(a, e));
        return (g.wd(a, lo) * flo + g.wd(a, hi) * fhi) / (g.wd(a, lo) + g.wd(a, hi));
    }
    double edge_width(int a, int e, int q, int n) const {              // backend.cu:124-128
        int lo, hi;
        if (q == 1 || !nodes(e, q, n, lo, hi)) return g.sp(a, e);
        return 0.5 * (g.wd(a, lo) + g.wd(a, hi));
    }
    double ramp(int e, int n) const {                                  // backend.cu:130-133
        double d = (double)(e < n - e ? e : n - e) - 1.0;
        return std::fmin(std::fmax(d / K, 0.0), 1.0);
    }
    double weight(int q, double s) const {                             // backend.cu:135-138
        if (q == 1) return o4 ? 1.0 + s * 0.125 : 1.0;
        return -s * 0.125;
    }
    // backend.cu:141-158
    double om_z(int x, int y, int z, int q) const {
        int xl, xh, yl, yh;
        if (!nodes(x, q, g.nx, xl, xh) || !nodes(y, q, g.ny, yl, yh)) return 0.0;
        return (v[g.iv(xh, y, z)] - v[g.iv(xl, y, z)]) / node_gap(X, x, q, xl, xh)
             - (u[g.iu(x, yh, z)] - u[g.iu(x, yl, z)]) / node_gap(Y, y, q, yl, yh);
    }
    double om_x(int x, int y, int z, int q) const {
        int yl, yh, zl, zh;
        if (!nodes(y, q, g.ny, yl, yh) || !nodes(z, q, g.nz, zl, zh)) return 0.0;
        return (w[g.iw(x, yh, z)] - w[g.iw(x, yl, z)]) / node_gap(Y, y, q, yl, yh)
             - (v[g.iv(x, y, zh)] - v[g.iv(x, y, zl)]) / node_gap(Z, z, q, zl, zh);
    }
    double om_y(int x, int y, int z, int q) const {
        int xl, xh, zl, zh;
        if (!nodes(x, q, g.nx, xl, xh) || !nodes(z, q, g.nz, zl, zh)) return 0.0;
        return (u[g.iu(x, y, zh)] - u[g.iu(x, y, zl)]) / node_gap(Z, z, q, zl, zh)
             - (w[g.iw(xh, y, z)] - w[g.iw(xl, y, z)]) / node_gap(X, x, q, xl, xh);
    }
    // backend.cu:162-215, the six (vorticity, interpolated velocity) pairings
    double zx(int x, int y, int z, int q) const {     // omega_z * v along x
        int xl, xh;
        if (y < 0 || y > g.ny || !nodes(x, q, g.nx, xl, xh)) return 0.0;
        double wq = weight(q, ramp(y, g.ny) * ramp(x, g.nx));
        if (wq == 0.0) return 0.0;
        double o = om_z(x, y, z, q);
        if (o == 0.0) return 0.0;
        return wq * o * at_edge(X, x, q, xl, xh, v[g.iv(xl, y, z)], v[g.iv(xh, y, z)]);
    }
    double zy(int x, int y, int z, int q) const {     // omega_z * u along y
        int yl, yh;
        if (x < 0 || x > g.nx || !nodes(y, q, g.ny, yl, yh)) return 0.0;
        double wq = weight(q, ramp(y, g.ny) * ramp(x, g.nx));
        if (wq == 0.0) return 0.0;
        double o = om_z(x, y, z, q);
        if (o == 0.0) return 0.0;
        return wq * o * at_edge(Y, y, q, yl, yh, u[g.iu(x, yl, z)], u[g.iu(x, yh, z)]);
    }
    double xy(int x, int y, int z, int q) const {     // omega_x * w along y
        int yl, yh;
        if (z < 0 || z > g.nz || !nodes(y, q, g.ny, yl, yh)) return 0.0;
        double wq = weight(q, ramp(z, g.nz) * ramp(y, g.ny));
        if (wq == 0.0) return 0.0;
        double o = om_x(x, y, z, q);
        if (o == 0.0) return 0.0;
        return wq * o * at_edge(Y, y, q, yl, yh, w[g.iw(x, yl, z)], w[g.iw(x, yh, z)]);
    }
    double xz(int x, int y, int z, int q) const {     // omega_x * v along z
        int zl, zh;
        if (y < 0 || y > g.ny || !nodes(z, q, g.nz, zl, zh)) return 0.0;
        double wq = weight(q, ramp(z, g.nz) * ramp(y, g.ny));
        if (wq == 0.0) return 0.0;
        double o = om_x(x, y, z, q);
        if (o == 0.0) return 0.0;
        return wq * o * at_edge(Z, z, q, zl, zh, v[g.iv(x, y, zl)], v[g.iv(x, y, zh)]);
    }
    double yx(int x, int y, int z, int q) const {     // omega_y * w along x
        int xl, xh;
        if (z < 0 || z > g.nz || !nodes(x, q, g.nx, xl, xh)) return 0.0;
        double wq = weight(q, ramp(z, g.nz) * ramp(x, g.nx));
        if (wq == 0.0) return 0.0;
        double o = om_y(x, y, z, q);
        if (o == 0.0) return 0.0;
        return wq * o * at_edge(X, x, q, xl, xh, w[g.iw(xl, y, z)], w[g.iw(xh, y, z)]);
    }
    double yz(int x, int y, int z, int q) const {     // omega_y * u along z
        int zl, zh;
        if (x < 0 || x > g.nx || !nodes(z, q, g.nz, zl, zh)) return 0.0;
        double wq = weight(q, ramp(z, g.nz) * ramp(x, g.nx));
        if (wq == 0.0) return 0.0;
        double o = om_y(x, y, z, q);
        if (o == 0.0) return 0.0;
        return wq * o * at_edge(Z, z, q, zl, zh, u[g.iu(x, y, zl)], u[g.iu(x, y, zh)]);
    }
    double Kc(int x, int y, int z) const {                              // backend.cu:218-222
        i64 a = g.iu(x, y, z), b = g.iv(x, y, z), c = g.iw(x, y, z), c1 = g.iw(x, y, z + 1);
        i64 NX = g.nx;
        return 0.5 * ((u[a] * mx[a] + u[a + 1] * mx[a + 1]) + (v[b] * my[b] + v[b + NX] * my[b + NX])
                      + (w[c] * mz[c] + w[c1] * mz[c1]));
    }
    double Ix(int x, int y, int z) const {                              // backend.cu:225-240 (interior face)
        double L = 0.0;
        for (int q = 1; q <= (o4 ? 3 : 1); q += 2) {
            int up = (q + 1) / 2, down = (q - 1) / 2;
            double c = edge_width(X, x, q, g.nx) / (2.0 * g.sp(X, x));
            L += c * (zx(x, y + up, z, q) + zx(x, y - down, z, q));
            L -= c * (yx(x, y, z + up, q) + yx(x, y, z - down, q));
        }
        L -= (Kc(x, y, z) - Kc(x - 1, y, z)) / g.sp(X, x);
        return L;
    }
    double Iy(int x, int y, int z) const {                              // backend.cu:242-257
        double L = 0.0;
        for (int q = 1; q <= (o4 ? 3 : 1); q += 2) {
            int up = (q + 1) / 2, down = (q - 1) / 2;
            double c = edge_width(Y, y, q, g.ny) / (2.0 * g.sp(Y, y));
            L -= c * (zy(x + up, y, z, q) + zy(x - down, y, z, q));
            L += c * (xy(x, y, z + up, q) + xy(x, y, z - down, q));
        }
        L -= (Kc(x, y, z) - Kc(x, y - 1, z)) / g.sp(Y, y);
        return L;
    }
    double Iz(int x, int y, int z) const {                              // backend.cu:259-274
        double L = 0.0;
        for (int q = 1; q <= (o4 ? 3 : 1); q += 2) {
            int up = (q + 1) / 2, down = (q - 1) / 2;
            double c = edge_width(Z, z, q, g.nz) / (2.0 * g.sp(Z, z));
            L -= c * (xz(x, y + up, z, q) + xz(x, y - down, z, q));
            L += c * (yz(x + up, y, z, q) + yz(x - down, y, z, q));
        }
        L -= (Kc(x, y, z) - Kc(x, y, z - 1)) / g.sp(Z, z);
        return L;
    }
};

void inviscid(const Mesh& g, const F3& u, const F3& m, int o4, double K, double* const out[3]) {
    Impulse im{g, u.c[0], u.c[1], u.c[2], m.c[0], m.c[1], m.c[2], o4, K};
    int NX = g.nx, NY = g.ny, NZ = g.nz;
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < NZ; z++)
        for (int y = 0; y < NY; y++)
            for (int x = 0; x <= NX; x++)
                out[0][g.iu(x, y, z)] = (x == 0 || x == NX) ? 0.0 : im.Ix(x, y, z);
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < NZ; z++)
        for (int y = 0; y <= NY; y++)
            for (int x = 0; x < NX; x++)
                out[1][g.iv(x, y, z)] = (y == 0 || y == NY) ? 0.0 : im.Iy(x, y, z);
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z <= NZ; z++)
        for (int y = 0; y < NY; y++)
            for (int x = 0; x < NX; x++)
                out[2][g.iw(x, y, z)] = (z == 0 || z == NZ) ? 0.0 : im.Iz(x, y, z);
}

// ------------------------------------------------------------------------------------------------ RK3 stage
// backend.cu:279-330: out = [a0 m0 +] b (m + dt (nu lap(m) + I)), flux-form Laplacian, wall fluxes from vel.
void rk3_stage(const Mesh& g, int comp, const double* m, const double* m0, const double* vel, const double* I,
               double nu, double lid, double dt, double a0, double b, int has_base, double* out) {
    int NX = g.nx, NY = g.ny, NZ = g.nz;
    int sx = NX + (comp == 0), sy = NY + (comp == 1), sz = NZ + (comp == 2);
    i64 dy = sx, dz = (i64)sx * sy;
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < sz; z++)
        for (int y = 0; y < sy; y++)
            for (int x = 0; x < sx; x++) {
                i64 i = x + dy * y + dz * z;
                if ((comp == 0 && (x == 0 || x == NX)) || (comp == 1 && (y == 0 || y == NY))
                    || (comp == 2 && (z == 0 || z == NZ))) {
                    out[i] = 0.0;
                    continue;
                }
                double c = m[i];
                double wall = vel[i];
                double lap = 0.0;
                if (comp == 0) {
                    lap += ((m[i + 1] - c) / g.wd(X, x) - (c - m[i - 1]) / g.wd(X, x - 1)) / g.sp(X, x);
                } else {
                    double east = x < sx - 1 ? (m[i + 1] - c) / g.sp(X, x + 1) : 0.0;
                    double west = x > 0 ? (c - m[i - 1]) / g.sp(X, x) : 0.0;
                    if (x == 0) west = wall / g.sp(X, 0);
                    if (x == sx - 1) east = -wall / g.sp(X, NX);
                    lap += (east - west) / g.wd(X, x);
                }
                if (comp == 1) {
                    lap += ((m[i + dy] - c) / g.wd(Y, y) - (c - m[i - dy]) / g.wd(Y, y - 1)) / g.sp(Y, y);
                } else {
                    double north = y < sy - 1 ? (m[i + dy] - c) / g.sp(Y, y + 1) : 0.0;
                    double south = y > 0 ? (c - m[i - dy]) / g.sp(Y, y) : 0.0;
                    if (y == 0) south = wall / g.sp(Y, 0);
                    if (y == sy - 1) {
                        double U = comp == 0 ? lid * g.lid_at_u_face(x, z) : 0.0;
                        north = (U - wall) / g.sp(Y, NY);
                    }
                    lap += (north - south) / g.wd(Y, y);
                }
                if (comp == 2) {
                    lap += ((m[i + dz] - c) / g.wd(Z, z) - (c - m[i - dz]) / g.wd(Z, z - 1)) / g.sp(Z, z);
                } else {
                    double top = z < sz - 1 ? (m[i + dz] - c) / g.sp(Z, z + 1) : 0.0;
                    double bottom = z > 0 ? (c - m[i - dz]) / g.sp(Z, z) : 0.0;
                    if (z == 0) bottom = wall / g.sp(Z, 0);
                    if (z == sz - 1) top = -wall / g.sp(Z, NZ);
                    lap += (top - bottom) / g.wd(Z, z);
                }
                double rhs = nu * lap + 0.0;
                double q = c + dt * (rhs + I[i]);
                if (has_base) q = a0 * m0[i] + b * q;
                out[i] = q;
            }
}

// ------------------------------------------------------------------------------------------------ projection
void divergence(const Mesh& g, const F3& f, double* rhs) {             // backend.cu:334-342
    const double *u = f.c[0], *v = f.c[1], *w = f.c[2];
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < g.nz; z++)
        for (int y = 0; y < g.ny; y++)
            for (int x = 0; x < g.nx; x++) {
                double dx = g.wd(X, x), dy = g.wd(Y, y), dz = g.wd(Z, z);
                rhs[g.ic(x, y, z)] = (u[g.iu(x + 1, y, z)] - u[g.iu(x, y, z)]) * (dy * dz)
                                   + (v[g.iv(x, y + 1, z)] - v[g.iv(x, y, z)]) * (dx * dz)
                                   + (w[g.iw(x, y, z + 1)] - w[g.iw(x, y, z)]) * (dx * dy);
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

void subtract_gradient(const Mesh& g, int comp, const double* m, const double* phi, double* out) {   // 361-371
    int NX = g.nx, NY = g.ny, NZ = g.nz;
    int sx = NX + (comp == 0), sy = NY + (comp == 1), sz = NZ + (comp == 2);
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < sz; z++)
        for (int y = 0; y < sy; y++)
            for (int x = 0; x < sx; x++) {
                i64 i = x + (i64)sx * (y + (i64)sy * z);
                double d;
                if (comp == 0) {
                    if (x == 0 || x == NX) { out[i] = 0.0; continue; }
                    d = (phi[g.ic(x, y, z)] - phi[g.ic(x - 1, y, z)]) / g.sp(X, x);
                } else if (comp == 1) {
                    if (y == 0 || y == NY) { out[i] = 0.0; continue; }
                    d = (phi[g.ic(x, y, z)] - phi[g.ic(x, y - 1, z)]) / g.sp(Y, y);
                } else {
                    if (z == 0 || z == NZ) { out[i] = 0.0; continue; }
                    d = (phi[g.ic(x, y, z)] - phi[g.ic(x, y, z - 1)]) / g.sp(Z, z);
                }
                out[i] = m[i] - d;
            }
}

void zero_walls(const Mesh& g, int comp, double* f) {                  // backend.cu:373-380
    int NX = g.nx, NY = g.ny, NZ = g.nz;
    int sx = NX + (comp == 0), sy = NY + (comp == 1), sz = NZ + (comp == 2);
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < sz; z++)
        for (int y = 0; y < sy; y++)
            for (int x = 0; x < sx; x++)
                if ((comp == 0 && (x == 0 || x == NX)) || (comp == 1 && (y == 0 || y == NY))
                    || (comp == 2 && (z == 0 || z == NZ)))
                    f[x + (i64)sx * (y + (i64)sy * z)] = 0.0;
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
void apply_x(const Mesh& g, const double* M, bool transpose_m, const double* in, double* out) {
    int n = g.nx;
    i64 rows = (i64)g.ny * g.nz;
#pragma omp parallel for schedule(static)
    for (i64 r = 0; r < rows; r++) {
        const double* a = in + r * n;
        double* o = out + r * n;
        for (int j = 0; j < n; j++) o[j] = 0.0;
        for (int i = 0; i < n; i++) {
            double ai = a[i];
            if (transpose_m) {                                  // o[j] += a[i] * M[i][j]
                const double* mr = M + (size_t)i * n;
#pragma omp simd
                for (int j = 0; j < n; j++) o[j] += ai * mr[j];
            } else {                                            // o[j] += a[i] * M[j][i]
                for (int j = 0; j < n; j++) o[j] += ai * M[(size_t)j * n + i];
            }
        }
    }
}

// out[a][j][b] = sum_i M(j, i) in[a][i][b] for a cell array viewed as [A][n][B]
void apply_mid(i64 A, int n, i64 B, const double* M, const double* in, double* out) {
    const i64 CB = 512;
    i64 nb = (B + CB - 1) / CB;
#pragma omp parallel for collapse(2) schedule(static)
    for (i64 a = 0; a < A; a++)
        for (i64 blk = 0; blk < nb; blk++) {
            i64 b0 = blk * CB, b1 = std::min(B, b0 + CB);
            const double* src = in + a * n * B;
            double* dst = out + a * n * B;
            for (int j = 0; j < n; j++) {
                double* o = dst + (i64)j * B;
                for (i64 b = b0; b < b1; b++) o[b] = 0.0;
                for (int i = 0; i < n; i++) {
                    double mji = M[(size_t)j * n + i];
                    const double* s = src + (i64)i * B;
#pragma omp simd
                    for (i64 b = b0; b < b1; b++) o[b] += mji * s[b];
                }
            }
        }
}

void transpose(int n, const vec& S, vec& ST) {
    ST.resize((size_t)n * n);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) ST[(size_t)j * n + i] = S[(size_t)i * n + j];
}

// ------------------------------------------------------------------------------------------------ solver
struct History { double t, scale; int buf; };

struct Oracle {
    CavConfig c{};
    Mesh g;
    i64 n_cell = 0, size[3] = {0, 0, 0};
    double volume = 0, viscous_rate = 0;
    vec u[3], m[3], phi;
    vec m0[3], I[3], new_m[2][3], new_u[2][3], tmp_f[3];
    vec phi0, phi_stage[2], base, guess, rhs, res, hist_buf[4], t1, t2, stats;
    vec S[3], ST[3], inv_lambda;
    std::vector<int> probe_cells;
    double t = 0, last_dt = 0, last_cfl = 0, phi_rel = 0, prev_rates[2] = {0, 0};
    i64 steps = 0, projections = 0, pass_total = 0;
    int dt_reason = 0, phi_passes = 0;
    double xform_seconds = 0;        // wall time inside the six FDM transforms of every pass (CPU-baseline split)
    bool have_prev_rates = false;
    std::vector<History> hist;
    std::string err;

    double lid(double time) const {                                   // backend.cu:623-626
        if (c.lid_ramp_t > 0) return c.u_lid * std::min(1.0, std::max(time, 0.0) / c.lid_ramp_t);
        return c.u_lid;
    }
    F3 view(const vec* a) const { return F3{{a[0].data(), a[1].data(), a[2].data()}}; }

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
        g.ax[X] = build_axis(c, nx, c.length);
        g.ax[Y] = build_axis(c, ny, c.length);
        g.ax[Z] = build_axis(c, nz, nz * (c.lz / nz));
        volume = g.ax[X].length * g.ax[Y].length * g.ax[Z].length;
        // backend.cu:1030-1034 (loop over every cell, as the reference does)
        auto geo = [&](int a, int i) {
            return 4.0 / (g.wd(a, i) * std::min(g.sp(a, i), g.sp(a, i + 1)));
        };
        double vmax = 0.0;
        for (int z = 0; z < nz; z++)
            for (int y = 0; y < ny; y++)
                for (int x = 0; x < nx; x++) vmax = std::max(vmax, c.nu * (geo(X, x) + geo(Y, y) + geo(Z, z)));
        viscous_rate = vmax;
        for (int k = 0; k < 3; k++) {
            u[k].assign(size[k], 0.0); m[k].assign(size[k], 0.0); m0[k].assign(size[k], 0.0);
            I[k].assign(size[k], 0.0); tmp_f[k].assign(size[k], 0.0);
            for (int j = 0; j < 2; j++) { new_m[j][k].assign(size[k], 0.0); new_u[j][k].assign(size[k], 0.0); }
        }
        for (vec* p : {&phi, &phi0, &phi_stage[0], &phi_stage[1], &base, &guess, &rhs, &res, &t1, &t2})
            p->assign(n_cell, 0.0);
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
            S[a] = e[a].S;
            transpose(e[a].n, S[a], ST[a]);
        }
        inv_lambda.assign(n_cell, 0.0);
        for (int z = 0; z < g.nz; z++)
            for (int y = 0; y < g.ny; y++)
                for (int x = 0; x < g.nx; x++) {
                    bool zero = x == e[X].zero_mode && y == e[Y].zero_mode && z == e[Z].zero_mode;
                    inv_lambda[g.ic(x, y, z)] = zero ? 0.0 : 1.0 / (e[Z].lambda[z] + e[Y].lambda[y] + e[X].lambda[x]);
                }
    }

    // backend.cu:801-814
    void pressure_pass(const double* r, double* ph) {
        i64 n = n_cell;
        double rms = std::sqrt(reduce(r, n, 3) / n);
        if (!(rms > 0)) return;
        double s = 1.0 / rms;
        double *a = t1.data(), *b = t2.data();
#pragma omp parallel for schedule(static)
        for (i64 i = 0; i < n; i++) a[i] = r[i] * s;
        i64 plane = (i64)g.nx * g.ny;
        const auto t0 = std::chrono::steady_clock::now();
        apply_x(g, S[X].data(), true, a, b);                          // t2 = t1 S_x
        apply_mid(g.nz, g.ny, g.nx, ST[Y].data(), b, a);               // t1 = S_y^T t2 (per z)
        apply_mid(1, g.nz, plane, ST[Z].data(), a, b);                 // t2 = S_z^T t1
        const double* il = inv_lambda.data();
#pragma omp parallel for schedule(static)
        for (i64 i = 0; i < n; i++) b[i] = b[i] * il[i];
        apply_mid(1, g.nz, plane, S[Z].data(), b, a);                  // t1 = S_z t2
        apply_mid(g.nz, g.ny, g.nx, S[Y].data(), a, b);                // t2 = S_y t1
        apply_x(g, ST[X].data(), true, b, a);                          // t1 = t2 S_x^T
        xform_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
#pragma omp parallel for schedule(static)
        for (i64 i = 0; i < n; i++) ph[i] = ph[i] + a[i] * rms;
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
    void project(vec* mm, const double* gs, double* ph, vec* out) {
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

    double cfl_max(const vec* uu) {                                    // backend.cu:465-472 + reduce op 0
        vec rate(n_cell);
        const double *u0 = uu[0].data(), *v0 = uu[1].data(), *w0 = uu[2].data();
#pragma omp parallel for collapse(2) schedule(static)
        for (int z = 0; z < g.nz; z++)
            for (int y = 0; y < g.ny; y++)
                for (int x = 0; x < g.nx; x++)
                    rate[g.ic(x, y, z)] =
                        std::fmax(std::fabs(u0[g.iu(x, y, z)]), std::fabs(u0[g.iu(x + 1, y, z)])) / g.wd(X, x)
                        + std::fmax(std::fabs(v0[g.iv(x, y, z)]), std::fabs(v0[g.iv(x, y + 1, z)])) / g.wd(Y, y)
                        + std::fmax(std::fabs(w0[g.iw(x, y, z)]), std::fabs(w0[g.iw(x, y, z + 1)])) / g.wd(Z, z);
        return reduce(rate.data(), n_cell, 0);
    }

    // backend.cu:877-955
    void step() {
        i64 n = n_cell;
        double r = c.reinit_r, omr = 1.0 - c.reinit_r;
        for (int k = 0; k < 3; k++) {
            const double *mk = m[k].data(), *uk = u[k].data();
            double* o = m0[k].data();
#pragma omp parallel for schedule(static)
            for (i64 i = 0; i < size[k]; i++) o[i] = r * mk[i] + omr * uk[i];
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
        vec* cur_u = u;
        vec* cur_m = m0;
        vec* cur_phi = &phi0;
        int o4 = c.impulse_order == 4;
        double K = std::max(c.impulse_order4_ramp, 1.0);
        for (int stage = 0; stage < 3; stage++) {
            double a0 = a0s[stage], b = bs[stage], ts = times[stage];
            vec* nm = new_m[stage % 2];
            vec* nu = new_u[stage % 2];
            double* Iout[3] = {I[0].data(), I[1].data(), I[2].data()};
            inviscid(g, view(cur_u), view(cur_m), o4, K, Iout);
            for (int k = 0; k < 3; k++)
                rk3_stage(g, k, cur_m[k].data(), m0[k].data(), cur_u[k].data(), I[k].data(), c.nu, lid(ts), dt, a0,
                          b, stage > 0 ? 1 : 0, nm[k].data());
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
            const double* uk = u[k].data();
            double* o = tmp_f[k].data();
#pragma omp parallel for schedule(static)
            for (i64 i = 0; i < n; i++) o[i] = uk[i] + na[i] * s;
        }
        project(tmp_f, nullptr, phi_stage[0].data(), new_u[0]);
        for (int k = 0; k < 3; k++) {
            double* mk = m[k].data();
            const double *a = new_u[0][k].data(), *b = u[k].data();
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
            m[k].assign(size[k], 0.0);
            if (src[k]) std::copy(src[k], src[k] + size[k], m[k].begin());
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
        double lidv = lid(t);
        vec ke(n), ens(n), div2(n);
        const double *uu = u[0].data(), *vv = u[1].data(), *ww = u[2].data();
#pragma omp parallel for collapse(2) schedule(static)
        for (int z = 0; z < g.nz; z++)
            for (int y = 0; y < g.ny; y++)
                for (int x = 0; x < g.nx; x++) {
                    i64 i = g.ic(x, y, z);
                    double vol = g.wd(X, x) * g.wd(Y, y) * g.wd(Z, z);
                    double a = centred(g, uu, vv, ww, 0, x, y, z), b = centred(g, uu, vv, ww, 1, x, y, z),
                           cc = centred(g, uu, vv, ww, 2, x, y, z);
                    ke[i] = 0.5 * (a * a + b * b + cc * cc) * vol;
                    double gr[3][3];
                    cell_gradient(g, uu, vv, ww, lidv, x, y, z, gr);
                    double ox = gr[1][2] - gr[2][1], oy = gr[2][0] - gr[0][2], oz = gr[0][1] - gr[1][0];
                    ens[i] = (ox * ox + oy * oy + oz * oz) * vol;
                    double dd = (uu[g.iu(x + 1, y, z)] - uu[g.iu(x, y, z)]) / g.wd(X, x)
                              + (vv[g.iv(x, y + 1, z)] - vv[g.iv(x, y, z)]) / g.wd(Y, y)
                              + (ww[g.iw(x, y, z + 1)] - ww[g.iw(x, y, z)]) / g.wd(Z, z);
                    div2[i] = dd * dd * vol;
                }
        d->kinetic_energy = reduce(ke.data(), n, 1) / volume;
        d->enstrophy = 0.5 * reduce(ens.data(), n, 1) / volume;
        d->div_rms = std::sqrt(reduce(div2.data(), n, 1) / volume);
        double umax = 0;
        for (int k = 0; k < 3; k++) umax = std::max(umax, reduce(u[k].data(), size[k], 2));
        d->umax = umax;
        d->cfl_rate = cfl_max(u);
        double best = -1;
        int best_comp = 0;
        i64 best_i = 0;
        for (int k = 0; k < 3; k++) {
            vec diff(size[k]);
            for (i64 i = 0; i < size[k]; i++) diff[i] = std::fabs(m[k][i] - u[k][i]);
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
        const double *uu = u[0].data(), *vv = u[1].data();
        for (int p = 0; p < c.n_probes; p++) {
            int x = probe_cells[3 * p], y = probe_cells[3 * p + 1], z = probe_cells[3 * p + 2];
            out[2 * p] = 0.5 * (uu[g.iu(x, y, z)] + uu[g.iu(x + 1, y, z)]);
            out[2 * p + 1] = 0.5 * (vv[g.iv(x, y, z)] + vv[g.iv(x, y + 1, z)]);
        }
    }

    void midspan(int& k1, int& k2) const { k2 = g.nz / 2; k1 = (g.nz % 2) ? k2 : k2 - 1; }   // backend.cu:488

    void add_statistics(double weight, double* line) {                // backend.cu:507-532
        const double *uu = u[0].data(), *vv = u[1].data(), *ww = u[2].data();
        int k1, k2;
        midspan(k1, k2);
        i64 n = (i64)g.nx * g.ny;
        for (int y = 0; y < g.ny; y++)
            for (int x = 0; x < g.nx; x++) {
                i64 i = x + (i64)g.nx * y;
                double a = 0.5 * (centred(g, uu, vv, ww, 0, x, y, k1) + centred(g, uu, vv, ww, 0, x, y, k2));
                double b = 0.5 * (centred(g, uu, vv, ww, 1, x, y, k1) + centred(g, uu, vv, ww, 1, x, y, k2));
                double cc = 0.5 * (centred(g, uu, vv, ww, 2, x, y, k1) + centred(g, uu, vv, ww, 2, x, y, k2));
                stats[i] += weight * a; stats[n + i] += weight * b; stats[2 * n + i] += weight * (a * a);
                stats[3 * n + i] += weight * (b * b); stats[4 * n + i] += weight * (cc * cc);
                stats[5 * n + i] += weight * (a * b);
            }
        if (!line) return;
        if (g.nx != g.ny) fail("centreline sample needs nx == ny");
        int NX = g.nx, mid = NX / 2;
        auto mean4 = [&](int cc, int xa, int ya, int xb, int yb) {
            return 0.5 * (0.5 * (centred(g, uu, vv, ww, cc, xa, ya, k1) + centred(g, uu, vv, ww, cc, xa, ya, k2))
                          + 0.5 * (centred(g, uu, vv, ww, cc, xb, yb, k1) + centred(g, uu, vv, ww, cc, xb, yb, k2)));
        };
        for (int j = 0; j < NX; j++) {
            double a = mean4(0, mid - 1, j, mid, j), b = mean4(1, j, mid - 1, j, mid), cc = mean4(2, mid - 1, j, mid, j);
            line[j] = a; line[NX + j] = b; line[2 * NX + j] = a * a; line[3 * NX + j] = b * b; line[4 * NX + j] = cc * cc;
        }
    }

    void midplane(double* out) const {                                // backend.cu:490-505
        const double *uu = u[0].data(), *vv = u[1].data(), *ww = u[2].data();
        double lidv = lid(t);
        int k1, k2;
        midspan(k1, k2);
        i64 P = (i64)g.nx * g.ny;
        for (int y = 0; y < g.ny; y++)
            for (int x = 0; x < g.nx; x++) {
                double val[2][4];
                for (int s = 0; s < 2; s++) {
                    int z = s ? k2 : k1;
                    for (int cc = 0; cc < 3; cc++) val[s][cc] = centred(g, uu, vv, ww, cc, x, y, z);
                    double gr[3][3];
                    cell_gradient(g, uu, vv, ww, lidv, x, y, z, gr);
                    double a = gr[1][2] - gr[2][1], b = gr[2][0] - gr[0][2], cz = gr[0][1] - gr[1][0];
                    val[s][3] = std::sqrt(a * a + b * b + cz * cz);
                }
                i64 i = x + (i64)g.nx * y;
                for (int cc = 0; cc < 4; cc++) out[cc * P + i] = (k1 == k2) ? val[1][cc] : 0.5 * (val[0][cc] + val[1][cc]);
            }
    }

    void q_volume(int colour_w, double* q, double* colour) const {    // backend.cu:474-486, 1185-1194
        const double *uu = u[0].data(), *vv = u[1].data(), *ww = u[2].data();
        double lidv = lid(t);
        double LU = c.length / c.u_lid, qscale = LU * LU, inv_U = 1.0 / c.u_lid;
#pragma omp parallel for collapse(2) schedule(static)
        for (int z = 0; z < g.nz; z++)
            for (int y = 0; y < g.ny; y++)
                for (int x = 0; x < g.nx; x++) {
                    i64 i = g.ic(x, y, z);
                    double gr[3][3];
                    cell_gradient(g, uu, vv, ww, lidv, x, y, z, gr);
                    double s = 0.0;
                    for (int a = 0; a < 3; a++) for (int b = 0; b < 3; b++) s += gr[a][b] * gr[b][a];
                    q[i] = -0.5 * s * qscale;
                    double a = centred(g, uu, vv, ww, 0, x, y, z), b = centred(g, uu, vv, ww, 1, x, y, z),
                           cc = centred(g, uu, vv, ww, 2, x, y, z);
                    colour[i] = colour_w ? cc * inv_U : std::sqrt(a * a + b * b + cc * cc) * inv_U;
                }
    }

    double* field(int f) {
        if (f < 0 || f > 6) fail("field must be 0..6");
        if (f < 3) return u[f].data();
        if (f < 6) return m[f - 3].data();
        return phi.data();
    }
    i64 field_size(int f) const { return f < 3 ? size[f] : f < 6 ? size[f - 3] : n_cell; }
};

thread_local std::string g_err;

template <class F> int guarded(Oracle* o, F f) {
    try {
        if (!o) fail("no oracle");
        f();
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        if (o) o->err = e.what();
        return 1;
    }
}

}  // namespace

// ------------------------------------------------------------------------------------------------ C API
// Mirrors the solver_* boundary with a handle (orc_*), all field I/O in float64, plus entry points to
// the individual kernels for the invariant tests.
extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

void* orc_create(const CavConfig* c) {
    Oracle* o = new Oracle();
    try {
        o->create(*c);
    } catch (const std::exception& e) {
        g_err = e.what();
        delete o;
        return nullptr;
    }
    return o;
}
void orc_destroy(void* h) { delete (Oracle*)h; }
int orc_initialize(void* h, const double* v, const double* w) {
    Oracle* o = (Oracle*)h; return guarded(o, [&] { o->initialize(v, w); });
}
int orc_step(void* h, CavReport* r) {
    Oracle* o = (Oracle*)h; return guarded(o, [&] { o->step(); o->fill(r); });
}
int orc_steps(void* h, int n, CavReport* r) {
    Oracle* o = (Oracle*)h; return guarded(o, [&] { for (int i = 0; i < n; i++) o->step(); o->fill(r); });
}
int orc_kick(void* h, const double* nu, const double* nv, const double* nw, double amp, int passes) {
    Oracle* o = (Oracle*)h; return guarded(o, [&] { const double* n[3] = {nu, nv, nw}; o->kick(n, amp, passes); });
}
int orc_report(void* h, CavReport* r) { Oracle* o = (Oracle*)h; return guarded(o, [&] { o->fill(r); }); }
int orc_diagnostics(void* h, CavDiagnostics* d) { Oracle* o = (Oracle*)h; return guarded(o, [&] { o->diagnostics(d); }); }
int orc_grid(void* h, int axis, double* faces, double* centres, double* widths, double* spacings) {
    Oracle* o = (Oracle*)h;
    return guarded(o, [&] {
        if (axis < 0 || axis > 2) fail("axis must be 0..2");
        const Axis1& a = o->g.ax[axis];
        if (faces) std::copy(a.face.begin(), a.face.end(), faces);
        if (centres) std::copy(a.centre.begin(), a.centre.end(), centres);
        if (widths) std::copy(a.width.begin(), a.width.end(), widths);
        if (spacings) std::copy(a.spacing.begin(), a.spacing.end(), spacings);
    });
}
int orc_get_field(void* h, int f, double* out) {
    Oracle* o = (Oracle*)h;
    return guarded(o, [&] { double* p = o->field(f); std::copy(p, p + o->field_size(f), out); });
}
int orc_set_field(void* h, int f, const double* in) {
    Oracle* o = (Oracle*)h;
    return guarded(o, [&] {
        double* p = o->field(f);
        std::copy(in, in + o->field_size(f), p);
        o->hist.clear();
        o->have_prev_rates = false;
    });
}
int orc_set_clock(void* h, double t, long long steps, double last_dt) {
    Oracle* o = (Oracle*)h; return guarded(o, [&] { o->t = t; o->steps = steps; o->last_dt = last_dt; });
}
int orc_probes(void* h, double* out) { Oracle* o = (Oracle*)h; return guarded(o, [&] { o->probes(out); }); }
int orc_add_statistics(void* h, double weight, double* line) {
    Oracle* o = (Oracle*)h; return guarded(o, [&] { o->add_statistics(weight, line); });
}
int orc_get_statistics(void* h, double* out) {
    Oracle* o = (Oracle*)h; return guarded(o, [&] { std::copy(o->stats.begin(), o->stats.end(), out); });
}
int orc_set_statistics(void* h, const double* in) {
    Oracle* o = (Oracle*)h; return guarded(o, [&] { std::copy(in, in + o->stats.size(), o->stats.begin()); });
}
int orc_midplane(void* h, double* out) { Oracle* o = (Oracle*)h; return guarded(o, [&] { o->midplane(out); }); }
int orc_q_volume(void* h, int colour_w, double* q, double* colour) {
    Oracle* o = (Oracle*)h; return guarded(o, [&] { o->q_volume(colour_w, q, colour); });
}
int orc_pass_counters(void* h, long long* projections, long long* passes) {
    Oracle* o = (Oracle*)h; return guarded(o, [&] { *projections = o->projections; *passes = o->pass_total; });
}
// Seconds spent in the FDM transforms since create, and the OpenMP threads the oracle runs on
int orc_timing(void* h, double* xform_seconds, int* threads) {
    Oracle* o = (Oracle*)h;
    return guarded(o, [&] {
        *xform_seconds = o->xform_seconds;
        int nt = 1;
#ifdef _OPENMP
        nt = omp_get_max_threads();
#endif
        *threads = nt;
    });
}
int orc_viscous_rate(void* h, double* out) { Oracle* o = (Oracle*)h; return guarded(o, [&] { *out = o->viscous_rate; }); }

// ---- kernel-level entry points (same mesh and settings as the handle's config)
int orc_inviscid(void* h, const double* u, const double* v, const double* w, const double* mx, const double* my,
                 const double* mz, double* ox, double* oy, double* oz) {
    Oracle* o = (Oracle*)h;
    return guarded(o, [&] {
        double* out[3] = {ox, oy, oz};
        inviscid(o->g, F3{{u, v, w}}, F3{{mx, my, mz}}, o->c.impulse_order == 4,
                 std::max(o->c.impulse_order4_ramp, 1.0), out);
    });
}
int orc_rk3_stage(void* h, int comp, const double* m, const double* m0, const double* vel, const double* I,
                  double nu, double lid, double dt, double a0, double b, int has_base, double* out) {
    Oracle* o = (Oracle*)h;
    return guarded(o, [&] { rk3_stage(o->g, comp, m, m0, vel, I, nu, lid, dt, a0, b, has_base, out); });
}
int orc_divergence(void* h, const double* u, const double* v, const double* w, double* rhs) {
    Oracle* o = (Oracle*)h; return guarded(o, [&] { divergence(o->g, F3{{u, v, w}}, rhs); });
}
int orc_residual(void* h, const double* phi, const double* rhs, double* r) {
    Oracle* o = (Oracle*)h; return guarded(o, [&] { pressure_residual(o->g, phi, rhs, r); });
}
int orc_pressure_pass(void* h, const double* r, double* phi) {
    Oracle* o = (Oracle*)h; return guarded(o, [&] { o->pressure_pass(r, phi); });
}
int orc_solve_pressure(void* h, const double* rhs, const double* guess, double* phi, double* rel, int* passes) {
    Oracle* o = (Oracle*)h;
    return guarded(o, [&] { o->solve_pressure(rhs, guess, phi); *rel = o->phi_rel; *passes = o->phi_passes; });
}
// Test hook: replace the warm-start history by n entries (times t, scales sc, data n x n_cell, oldest first)
// and return guess_phi's output for (tau, scale); *used = 0 when it makes no guess.
int orc_guess_phi(void* h, int n, const double* t, const double* sc, const double* data, double tau, double scale,
                  double* out, int* used) {
    Oracle* o = (Oracle*)h;
    return guarded(o, [&] {
        if (n < 0 || n > 3) fail("the history holds at most 3 entries");
        o->hist.clear();
        for (int k = 0; k < n; k++) {
            std::copy(data + (i64)k * o->n_cell, data + (i64)(k + 1) * o->n_cell, o->hist_buf[k].begin());
            o->hist.push_back({t[k], sc[k], k});
        }
        *used = o->guess_phi(tau, scale, out) ? 1 : 0;
    });
}
int orc_subtract_gradient(void* h, int comp, const double* m, const double* phi, double* out) {
    Oracle* o = (Oracle*)h; return guarded(o, [&] { subtract_gradient(o->g, comp, m, phi, out); });
}
int orc_eigen(void* h, int axis, double* S, double* lambda, int* zero_mode) {
    Oracle* o = (Oracle*)h;
    return guarded(o, [&] {
        if (axis < 0 || axis > 2) fail("axis must be 0..2");
        Eigen1 e = axis_eigen(o->g.ax[axis]);
        std::copy(e.S.begin(), e.S.end(), S);
        std::copy(e.lambda.begin(), e.lambda.end(), lambda);
        *zero_mode = e.zero_mode;
    });
}
int orc_inv_lambda(void* h, double* out) {
    Oracle* o = (Oracle*)h; return guarded(o, [&] { std::copy(o->inv_lambda.begin(), o->inv_lambda.end(), out); });
}
int orc_cfl_max(void* h, double* out) { Oracle* o = (Oracle*)h; return guarded(o, [&] { *out = o->cfl_max(o->u); }); }
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
