// kernels.hip — gfx950 kernels of the cavity time step and its outputs.
//
// Every stencil follows the reference's order of floating-point operations (backend.cu:98-561) in FP64,
// compiled with -ffp-contract=off, so a stencil output equals the host oracle's bit for bit; only the
// reductions (block/wavefront trees instead of the reference's 65536-strided partials) and the FDM
// transforms (MFMA accumulation order) round differently.
#include <climits>

#include "kernels.h"

namespace cav {

const char* family_name(int f) {
    static const char* names[F_COUNT] = {"gauge_reset", "cfl_max", "stage", "divergence", "base_guess",
                                         "residual", "reduce", "xform_x", "xform_y", "xform_z", "finalize_phi",
                                         "subtract_gradient", "fill", "output", "kick"};
    return (f >= 0 && f < F_COUNT) ? names[f] : "?";
}

namespace {

constexpr int TPB = 256;

inline int blocks_for(i64 n, int tpb = TPB) { return (int)((n + tpb - 1) / tpb); }
inline int red_blocks(i64 n) { return (int)std::max<i64>(1, std::min<i64>(RED_BLOCKS, (n + TPB - 1) / TPB)); }

template <class K, class... Args>
void run(const Exec& ex, int family, K kernel, dim3 grid, dim3 block, Args... args) {
    if (grid.x == 0 || grid.y == 0 || grid.z == 0) return;
    if (ex.timer) ex.timer->before(family, ex.stream);
    hipLaunchKernelGGL(kernel, grid, block, 0, ex.stream, args...);
    HIP_CHECK(hipGetLastError());
    if (ex.timer) ex.timer->after(family, ex.stream);
}

// ------------------------------------------------------------------------------------------------ indexing
// backend.cu:50-53, x fastest
struct Ix {
    int nx, ny, nz;
    __device__ int u(int x, int y, int z) const { return x + (nx + 1) * (y + ny * z); }
    __device__ int v(int x, int y, int z) const { return x + nx * (y + (ny + 1) * z); }
    __device__ int w(int x, int y, int z) const { return x + nx * (y + ny * z); }
    __device__ int c(int x, int y, int z) const { return x + nx * (y + ny * z); }
};

__device__ __forceinline__ double wd(const Geo& g, int a, int i) { return g.width[a][i]; }
__device__ __forceinline__ double sp(const Geo& g, int a, int i) { return g.spacing[a][i]; }
__device__ __forceinline__ double ce(const Geo& g, int a, int i) { return g.centre[a][i]; }

// backend.cu:58-65
__device__ __forceinline__ double lid_edge(double s) {
    double s2 = s * s, s4 = s2 * s2, s8 = s4 * s4, s16 = s8 * s8, a = 1.0 - s16 * s2;
    return a * a;
}
__device__ __forceinline__ double lid_shape(const Geo& g, double fx, double fz) {
    if (!g.lid_shaped) return 1.0;
    return lid_edge(2.0 * fx - 1.0) * lid_edge(2.0 * fz - 1.0);
}
__device__ __forceinline__ double lid_at_u_face(const Geo& g, int x, int z) {
    return lid_shape(g, g.face[X][x] / g.length[X], g.centre[Z][z] / g.length[Z]);
}
__device__ __forceinline__ double lid_at_cell(const Geo& g, int x, int z) {
    return lid_shape(g, g.centre[X][x] / g.length[X], g.centre[Z][z] / g.length[Z]);
}

// ------------------------------------------------------------------------------------------------ reductions
__device__ __forceinline__ double comb(int op, double a, double b) {
    return (op == 0 || op == 2) ? ((a != a || a > b) ? a : b) : a + b;
}
__device__ __forceinline__ double init_of(int op) { return op == 0 ? -INFINITY : 0.0; }

// Block of TPB threads -> one value (valid in thread 0). Fixed combination order.
__device__ double block_reduce(int op, double v) {
    __shared__ double part[TPB / 64];
    for (int off = 32; off > 0; off >>= 1) v = comb(op, v, __shfl_down(v, off, 64));
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) part[wid] = v;
    __syncthreads();
    double r = part[0];
    for (int k = 1; k < TPB / 64; k++) r = comb(op, r, part[k]);
    return r;
}

__global__ __launch_bounds__(TPB) void k_reduce_array(const double* __restrict__ x, i64 n, int op,
                                                      double* __restrict__ partial) {
    double r = init_of(op);
    const int sop = (op == 2 || op == 0) ? 0 : 1;
    for (i64 i = (i64)blockIdx.x * TPB + threadIdx.x; i < n; i += (i64)gridDim.x * TPB) {
        double v = x[i];
        if (op == 2) v = fabs(v);
        if (op == 3) v = v * v;
        r = comb(sop, r, v);
    }
    r = block_reduce(sop, r);
    if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

__global__ __launch_bounds__(TPB) void k_reduce_final(const double* __restrict__ partial, int count, int op,
                                                      double* __restrict__ out) {
    const int sop = (op == 2 || op == 0) ? 0 : 1;
    double r = init_of(sop);
    for (int i = threadIdx.x; i < count; i += TPB) r = comb(sop, r, partial[i]);
    r = block_reduce(sop, r);
    if (threadIdx.x == 0) out[0] = r;
}

__global__ __launch_bounds__(TPB) void k_fill(double* __restrict__ p, i64 n, double value) {
    for (i64 i = (i64)blockIdx.x * TPB + threadIdx.x; i < n; i += (i64)gridDim.x * TPB) p[i] = value;
}

// ------------------------------------------------------------------------------------------------ gauge reset
// backend.cu:880-882: m0 = r m + (1-r) u per face, phi0 = r phi per cell. Blocks are split by segment.
struct GaugeArgs {
    const double* m[3];
    const double* u[3];
    const double* phi;
    double* m0[3];
    double* phi0;
    i64 n[4];
    int blocks[4];
    double r, omr;
};
__global__ __launch_bounds__(TPB) void k_gauge(GaugeArgs a) {
    int blk = blockIdx.x, seg = 0;
    while (seg < 3 && blk >= a.blocks[seg]) { blk -= a.blocks[seg]; seg++; }
    const i64 i = (i64)blk * TPB + threadIdx.x;
    if (i >= a.n[seg]) return;
    if (seg < 3) a.m0[seg][i] = a.r * a.m[seg][i] + a.omr * a.u[seg][i];
    else a.phi0[i] = a.r * a.phi[i];
}

// ------------------------------------------------------------------------------------------------ CFL
// backend.cu:465-472 per cell, reduced with max (NaN wins)
__global__ __launch_bounds__(TPB) void k_cfl(Geo g, const double* __restrict__ u, const double* __restrict__ v,
                                             const double* __restrict__ w, double* __restrict__ partial) {
    const Ix ix{g.nx, g.ny, g.nz};
    const i64 n = (i64)g.nx * g.ny * g.nz;
    double r = -INFINITY;
    for (i64 i = (i64)blockIdx.x * TPB + threadIdx.x; i < n; i += (i64)gridDim.x * TPB) {
        const int x = (int)(i % g.nx), y = (int)((i / g.nx) % g.ny), z = (int)(i / ((i64)g.nx * g.ny));
        const double rate = fmax(fabs(u[ix.u(x, y, z)]), fabs(u[ix.u(x + 1, y, z)])) / wd(g, X, x)
                          + fmax(fabs(v[ix.v(x, y, z)]), fabs(v[ix.v(x, y + 1, z)])) / wd(g, Y, y)
                          + fmax(fabs(w[ix.w(x, y, z)]), fabs(w[ix.w(x, y, z + 1)])) / wd(g, Z, z);
        r = comb(0, r, rate);
    }
    r = block_reduce(0, r);
    if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

// ------------------------------------------------------------------------------------------------ I(u, m)
// backend.cu:102-274: Lamb vector from edge vorticity (narrow q = 1 plus wide q = 3 with a wall ramp in
// 4th order) minus the gradient of K = u.m.
struct Impulse {
    const Geo& g;
    Ix ix;
    const double *u, *v, *w, *mx, *my, *mz;
    int o4;
    double K;

    __device__ static bool nodes(int e, int q, int n, int& lo, int& hi) {
        lo = e - (q + 1) / 2;
        hi = e + (q - 1) / 2;
        return lo >= 0 && hi <= n - 1;
    }
    __device__ double node_gap(int a, int e, int q, int lo, int hi) const {
        if (q == 1) return sp(g, a, e);
        return ce(g, a, hi) - ce(g, a, lo);
    }
    __device__ double at_edge(int a, int e, int q, int lo, int hi, double flo, double fhi) const {
        if (q == 1) return (wd(g, a, lo) * flo + wd(g, a, hi) * fhi) / (2.0 * sp(g, a, e));
        return (wd(g, a, lo) * flo + wd(g, a, hi) * fhi) / (wd(g, a, lo) + wd(g, a, hi));
    }
    __device__ double edge_width(int a, int e, int q, int n) const {
        int lo, hi;
        if (q == 1 || !nodes(e, q, n, lo, hi)) return sp(g, a, e);
        return 0.5 * (wd(g, a, lo) + wd(g, a, hi));
    }
    __device__ double ramp(int e, int n) const {
        const double d = (double)(e < n - e ? e : n - e) - 1.0;
        return fmin(fmax(d / K, 0.0), 1.0);
    }
    __device__ double weight(int q, double s) const {
        if (q == 1) return o4 ? 1.0 + s * 0.125 : 1.0;
        return -s * 0.125;
    }
    __device__ double om_z(int x, int y, int z, int q) const {
        int xl, xh, yl, yh;
        if (!nodes(x, q, g.nx, xl, xh) || !nodes(y, q, g.ny, yl, yh)) return 0.0;
        return (v[ix.v(xh, y, z)] - v[ix.v(xl, y, z)]) / node_gap(X, x, q, xl, xh)
             - (u[ix.u(x, yh, z)] - u[ix.u(x, yl, z)]) / node_gap(Y, y, q, yl, yh);
    }
    __device__ double om_x(int x, int y, int z, int q) const {
        int yl, yh, zl, zh;
        if (!nodes(y, q, g.ny, yl, yh) || !nodes(z, q, g.nz, zl, zh)) return 0.0;
        return (w[ix.w(x, yh, z)] - w[ix.w(x, yl, z)]) / node_gap(Y, y, q, yl, yh)
             - (v[ix.v(x, y, zh)] - v[ix.v(x, y, zl)]) / node_gap(Z, z, q, zl, zh);
    }
    __device__ double om_y(int x, int y, int z, int q) const {
        int xl, xh, zl, zh;
        if (!nodes(x, q, g.nx, xl, xh) || !nodes(z, q, g.nz, zl, zh)) return 0.0;
        return (u[ix.u(x, y, zh)] - u[ix.u(x, y, zl)]) / node_gap(Z, z, q, zl, zh)
             - (w[ix.w(xh, y, z)] - w[ix.w(xl, y, z)]) / node_gap(X, x, q, xl, xh);
    }
    __device__ double zx(int x, int y, int z, int q) const {
        int xl, xh;
        if (y < 0 || y > g.ny || !nodes(x, q, g.nx, xl, xh)) return 0.0;
        const double wq = weight(q, ramp(y, g.ny) * ramp(x, g.nx));
        if (wq == 0.0) return 0.0;
        const double o = om_z(x, y, z, q);
        if (o == 0.0) return 0.0;
        return wq * o * at_edge(X, x, q, xl, xh, v[ix.v(xl, y, z)], v[ix.v(xh, y, z)]);
    }
    __device__ double zy(int x, int y, int z, int q) const {
        int yl, yh;
        if (x < 0 || x > g.nx || !nodes(y, q, g.ny, yl, yh)) return 0.0;
        const double wq = weight(q, ramp(y, g.ny) * ramp(x, g.nx));
        if (wq == 0.0) return 0.0;
        const double o = om_z(x, y, z, q);
        if (o == 0.0) return 0.0;
        return wq * o * at_edge(Y, y, q, yl, yh, u[ix.u(x, yl, z)], u[ix.u(x, yh, z)]);
    }
    __device__ double xy(int x, int y, int z, int q) const {
        int yl, yh;
        if (z < 0 || z > g.nz || !nodes(y, q, g.ny, yl, yh)) return 0.0;
        const double wq = weight(q, ramp(z, g.nz) * ramp(y, g.ny));
        if (wq == 0.0) return 0.0;
        const double o = om_x(x, y, z, q);
        if (o == 0.0) return 0.0;
        return wq * o * at_edge(Y, y, q, yl, yh, w[ix.w(x, yl, z)], w[ix.w(x, yh, z)]);
    }
    __device__ double xz(int x, int y, int z, int q) const {
        int zl, zh;
        if (y < 0 || y > g.ny || !nodes(z, q, g.nz, zl, zh)) return 0.0;
        const double wq = weight(q, ramp(z, g.nz) * ramp(y, g.ny));
        if (wq == 0.0) return 0.0;
        const double o = om_x(x, y, z, q);
        if (o == 0.0) return 0.0;
        return wq * o * at_edge(Z, z, q, zl, zh, v[ix.v(x, y, zl)], v[ix.v(x, y, zh)]);
    }
    __device__ double yx(int x, int y, int z, int q) const {
        int xl, xh;
        if (z < 0 || z > g.nz || !nodes(x, q, g.nx, xl, xh)) return 0.0;
        const double wq = weight(q, ramp(z, g.nz) * ramp(x, g.nx));
        if (wq == 0.0) return 0.0;
        const double o = om_y(x, y, z, q);
        if (o == 0.0) return 0.0;
        return wq * o * at_edge(X, x, q, xl, xh, w[ix.w(xl, y, z)], w[ix.w(xh, y, z)]);
    }
    __device__ double yz(int x, int y, int z, int q) const {
        int zl, zh;
        if (x < 0 || x > g.nx || !nodes(z, q, g.nz, zl, zh)) return 0.0;
        const double wq = weight(q, ramp(z, g.nz) * ramp(x, g.nx));
        if (wq == 0.0) return 0.0;
        const double o = om_y(x, y, z, q);
        if (o == 0.0) return 0.0;
        return wq * o * at_edge(Z, z, q, zl, zh, u[ix.u(x, y, zl)], u[ix.u(x, y, zh)]);
    }
    __device__ double Kc(int x, int y, int z) const {
        const int a = ix.u(x, y, z), b = ix.v(x, y, z), c = ix.w(x, y, z), c1 = ix.w(x, y, z + 1), NX = g.nx;
        return 0.5 * ((u[a] * mx[a] + u[a + 1] * mx[a + 1]) + (v[b] * my[b] + v[b + NX] * my[b + NX])
                      + (w[c] * mz[c] + w[c1] * mz[c1]));
    }
    __device__ double Ix_(int x, int y, int z) const {
        double L = 0.0;
        for (int q = 1; q <= (o4 ? 3 : 1); q += 2) {
            const int up = (q + 1) / 2, down = (q - 1) / 2;
            const double c = edge_width(X, x, q, g.nx) / (2.0 * sp(g, X, x));
            L += c * (zx(x, y + up, z, q) + zx(x, y - down, z, q));
            L -= c * (yx(x, y, z + up, q) + yx(x, y, z - down, q));
        }
        L -= (Kc(x, y, z) - Kc(x - 1, y, z)) / sp(g, X, x);
        return L;
    }
    __device__ double Iy_(int x, int y, int z) const {
        double L = 0.0;
        for (int q = 1; q <= (o4 ? 3 : 1); q += 2) {
            const int up = (q + 1) / 2, down = (q - 1) / 2;
            const double c = edge_width(Y, y, q, g.ny) / (2.0 * sp(g, Y, y));
            L -= c * (zy(x + up, y, z, q) + zy(x - down, y, z, q));
            L += c * (xy(x, y, z + up, q) + xy(x, y, z - down, q));
        }
        L -= (Kc(x, y, z) - Kc(x, y - 1, z)) / sp(g, Y, y);
        return L;
    }
    __device__ double Iz_(int x, int y, int z) const {
        double L = 0.0;
        for (int q = 1; q <= (o4 ? 3 : 1); q += 2) {
            const int up = (q + 1) / 2, down = (q - 1) / 2;
            const double c = edge_width(Z, z, q, g.nz) / (2.0 * sp(g, Z, z));
            L -= c * (xz(x, y + up, z, q) + xz(x, y - down, z, q));
            L += c * (yz(x + up, y, z, q) + yz(x - down, y, z, q));
        }
        L -= (Kc(x, y, z) - Kc(x, y, z - 1)) / sp(g, Z, z);
        return L;
    }
};

// ------------------------------------------------------------------------------------------------ stage
// Fused inviscid term (backend.cu:225-274) and RK3 stage (backend.cu:279-330) per face: the I value never
// leaves registers. Blocks [0, b0) u-faces, [b0, b0+b1) v-faces, rest w-faces.
template <int COMP>
__device__ void stage_face(const Geo& g, const StageArgs& a, int i) {
    const int NX = g.nx, NY = g.ny, NZ = g.nz;
    const int sx = NX + (COMP == 0), sy = NY + (COMP == 1), sz = NZ + (COMP == 2);
    if (i >= sx * sy * sz) return;
    const int x = i % sx, y = (i / sx) % sy, z = i / (sx * sy);
    double* out = a.out[COMP];
    if ((COMP == 0 && (x == 0 || x == NX)) || (COMP == 1 && (y == 0 || y == NY)) || (COMP == 2 && (z == 0 || z == NZ))) {
        out[i] = 0.0;
        return;
    }
    const Impulse im{g, Ix{NX, NY, NZ}, a.u[0], a.u[1], a.u[2], a.m[0], a.m[1], a.m[2], a.o4, a.K};
    const double Iv = COMP == 0 ? im.Ix_(x, y, z) : COMP == 1 ? im.Iy_(x, y, z) : im.Iz_(x, y, z);
    const double* m = a.m[COMP];
    const double c = m[i];
    const double wall = a.u[COMP][i];
    const int dy = sx, dz = sx * sy;
    double lap = 0.0;
    if (COMP == 0) {
        lap += ((m[i + 1] - c) / wd(g, X, x) - (c - m[i - 1]) / wd(g, X, x - 1)) / sp(g, X, x);
    } else {
        double east = x < sx - 1 ? (m[i + 1] - c) / sp(g, X, x + 1) : 0.0;
        double west = x > 0 ? (c - m[i - 1]) / sp(g, X, x) : 0.0;
        if (x == 0) west = wall / sp(g, X, 0);
        if (x == sx - 1) east = -wall / sp(g, X, NX);
        lap += (east - west) / wd(g, X, x);
    }
    if (COMP == 1) {
        lap += ((m[i + dy] - c) / wd(g, Y, y) - (c - m[i - dy]) / wd(g, Y, y - 1)) / sp(g, Y, y);
    } else {
        double north = y < sy - 1 ? (m[i + dy] - c) / sp(g, Y, y + 1) : 0.0;
        double south = y > 0 ? (c - m[i - dy]) / sp(g, Y, y) : 0.0;
        if (y == 0) south = wall / sp(g, Y, 0);
        if (y == sy - 1) {
            const double U = COMP == 0 ? a.lid * lid_at_u_face(g, x, z) : 0.0;
            north = (U - wall) / sp(g, Y, NY);
        }
        lap += (north - south) / wd(g, Y, y);
    }
    if (COMP == 2) {
        lap += ((m[i + dz] - c) / wd(g, Z, z) - (c - m[i - dz]) / wd(g, Z, z - 1)) / sp(g, Z, z);
    } else {
        double top = z < sz - 1 ? (m[i + dz] - c) / sp(g, Z, z + 1) : 0.0;
        double bottom = z > 0 ? (c - m[i - dz]) / sp(g, Z, z) : 0.0;
        if (z == 0) bottom = wall / sp(g, Z, 0);
        if (z == sz - 1) top = -wall / sp(g, Z, NZ);
        lap += (top - bottom) / wd(g, Z, z);
    }
    const double rhs = a.nu * lap + 0.0;
    double q = c + a.dt * (rhs + Iv);
    if (a.has_base) q = a.a0 * a.m0[COMP][i] + a.b * q;
    out[i] = q;
}

__global__ __launch_bounds__(TPB) void k_stage(Geo g, StageArgs a, int b0, int b1) {
    const int blk = blockIdx.x;
    if (blk < b0) stage_face<0>(g, a, blk * TPB + threadIdx.x);
    else if (blk < b0 + b1) stage_face<1>(g, a, (blk - b0) * TPB + threadIdx.x);
    else stage_face<2>(g, a, (blk - b0 - b1) * TPB + threadIdx.x);
}

// ------------------------------------------------------------------------------------------------ projection
// backend.cu:334-342 with the rhs^2 partial sum fused
__global__ __launch_bounds__(TPB) void k_divergence(Geo g, const double* __restrict__ u, const double* __restrict__ v,
                                                    const double* __restrict__ w, double* __restrict__ rhs,
                                                    double* __restrict__ partial) {
    const Ix ix{g.nx, g.ny, g.nz};
    const i64 n = (i64)g.nx * g.ny * g.nz;
    double acc = 0.0;
    for (i64 i = (i64)blockIdx.x * TPB + threadIdx.x; i < n; i += (i64)gridDim.x * TPB) {
        const int x = (int)(i % g.nx), y = (int)((i / g.nx) % g.ny), z = (int)(i / ((i64)g.nx * g.ny));
        const double dx = wd(g, X, x), dy = wd(g, Y, y), dz = wd(g, Z, z);
        const double r = (u[ix.u(x + 1, y, z)] - u[ix.u(x, y, z)]) * (dy * dz)
                       + (v[ix.v(x, y + 1, z)] - v[ix.v(x, y, z)]) * (dx * dz)
                       + (w[ix.w(x, y, z + 1)] - w[ix.w(x, y, z)]) * (dx * dy);
        rhs[i] = r;
        acc += r * r;
    }
    acc = block_reduce(1, acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

// backend.cu:373-380
__global__ __launch_bounds__(TPB) void k_zero_walls(Geo g, double* f0, double* f1, double* f2) {
    // thread per wall face: u walls x = 0, NX; v walls y = 0, NY; w walls z = 0, NZ
    const int NX = g.nx, NY = g.ny, NZ = g.nz;
    const int nu = 2 * NY * NZ, nv = 2 * NX * NZ, nw = 2 * NX * NY;
    int t = blockIdx.x * TPB + threadIdx.x;
    if (t < nu) {
        const int side = t % 2, r = t / 2, y = r % NY, z = r / NY;
        f0[(side ? NX : 0) + (NX + 1) * (y + NY * z)] = 0.0;
        return;
    }
    t -= nu;
    if (t < nv) {
        const int side = t % 2, r = t / 2, x = r % NX, z = r / NX;
        f1[x + NX * ((side ? NY : 0) + (NY + 1) * z)] = 0.0;
        return;
    }
    t -= nv;
    if (t < nw) {
        const int side = t % 2, r = t / 2, x = r % NX, y = r / NX;
        f2[x + NX * (y + NY * (side ? NZ : 0))] = 0.0;
    }
}

// backend.cu:922-925 (+ guess_phi's lincomb/scaled, backend.cu:870-871, and add_to)
__global__ __launch_bounds__(TPB) void k_base_guess(i64 n, const double* __restrict__ phi0,
                                                    const double* __restrict__ cur, double a0, double b, int has_a0,
                                                    const double* __restrict__ h1, double c1,
                                                    const double* __restrict__ h2, double c2, int nh,
                                                    double* __restrict__ base, double* __restrict__ phi) {
    const i64 i = (i64)blockIdx.x * TPB + threadIdx.x;
    if (i >= n) return;
    const double bs = has_a0 ? a0 * phi0[i] + b * cur[i] : b * cur[i];
    base[i] = bs;
    double gs = bs;
    if (nh == 1) gs = c1 * h1[i] + bs;
    else if (nh == 2) gs = (c1 * h1[i] + c2 * h2[i]) + bs;
    phi[i] = gs;
}

// backend.cu:345-359 with the r^2 partial sum fused
__global__ __launch_bounds__(TPB) void k_residual(Geo g, const double* __restrict__ phi,
                                                  const double* __restrict__ rhs, double* __restrict__ r,
                                                  double* __restrict__ partial) {
    const int NX = g.nx, NY = g.ny, NZ = g.nz;
    const i64 n = (i64)NX * NY * NZ, sy = NX, sz = (i64)NX * NY;
    double acc = 0.0;
    for (i64 i = (i64)blockIdx.x * TPB + threadIdx.x; i < n; i += (i64)gridDim.x * TPB) {
        const int x = (int)(i % NX), y = (int)((i / NX) % NY), z = (int)(i / sz);
        const double dx = wd(g, X, x), dy = wd(g, Y, y), dz = wd(g, Z, z), c = phi[i];
        double s = 0.0;
        if (x + 1 < NX) s += (dy * dz) * (phi[i + 1] - c) / sp(g, X, x + 1);
        if (x > 0) s -= (dy * dz) * (c - phi[i - 1]) / sp(g, X, x);
        if (y + 1 < NY) s += (dx * dz) * (phi[i + sy] - c) / sp(g, Y, y + 1);
        if (y > 0) s -= (dx * dz) * (c - phi[i - sy]) / sp(g, Y, y);
        if (z + 1 < NZ) s += (dx * dy) * (phi[i + sz] - c) / sp(g, Z, z + 1);
        if (z > 0) s -= (dx * dy) * (c - phi[i - sz]) / sp(g, Z, z);
        const double ri = rhs[i] - s;
        r[i] = ri;
        acc += ri * ri;
    }
    acc = block_reduce(1, acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

// ------------------------------------------------------------------------------------------------ FDM transform
// One axis of phi += S (Lambda^-1 o (S^T r)) as a dense contraction on MFMA:
//   MODE 0: FP64 operands, v_mfma_f64_16x16x4f64 (plain FP64 path)
//   MODE 1: FP32 operands, v_mfma_f32_16x16x4f32 (the reference's precision: FP32 transforms, backend.cu:795)
//   MODE 2: FP32 values split into bf16 hi + lo, three v_mfma_f32_16x16x32_bf16 products (hi*hi, hi*lo,
//           lo*hi) accumulated in FP32
// Block tile: 64 data columns x 64 matrix rows j, 4 waves of 32 x 32 (2 x 2 fragments). The data tile
// [BK i][64 cols] is staged in LDS; matrix fragments come pre-packed in operand order (L2-resident).
// Operand layout of the 16x16xK instructions: lane l holds row l%16 and k-group l/16 (k = l/16 for K = 4,
// k = 8 (l/16) + 0..7 for K = 32); result lane l, register r holds column l%16 and row 4 (l/16) + r for the
// FP32-accumulating instructions, row 4 r + l/16 for v_mfma_f64_16x16x4f64.
// With XL (x axis, rows of length n contiguous) the data is the A operand and the matrix the B operand;
// otherwise ([A][n][B] layout, contraction over the middle index) the roles swap. Both operand fragments
// use the same packed matrix because the A and B lane layouts mirror each other.
typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

template <int MODE> struct XT;
template <> struct XT<0> { using T = double; using Acc = dbl4; static constexpr int BK = 16; static constexpr int KT = 4; };
template <> struct XT<1> { using T = float; using Acc = flt4; static constexpr int BK = 16; static constexpr int KT = 4; };
template <> struct XT<2> { using T = float; using Acc = flt4; static constexpr int BK = 32; static constexpr int KT = 32; };

__device__ __forceinline__ dbl4 mfma(double a, double b, dbl4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
__device__ __forceinline__ flt4 mfma(float a, float b, flt4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
__device__ __forceinline__ flt4 mfma(bf8 a, bf8 b, flt4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }

template <bool XL, int MODE, class TIN, class TOUT>
__global__ __launch_bounds__(256) void k_axis(AxisArgs a, i64 ncols, int npad) {
    using T = typename XT<MODE>::T;
    using Acc = typename XT<MODE>::Acc;
    constexpr int BK = XT<MODE>::BK, KT = XT<MODE>::KT;
    __shared__ T tile[BK][64 + 2];
    const TIN* __restrict__ in = (const TIN*)a.in;
    TOUT* __restrict__ out = (TOUT*)a.out;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const i64 c0 = (i64)blockIdx.x * 64;
    const int wr = wave >> 1, wc = wave & 1;
    const int n = a.n, kts = npad / KT;
    const int jt0 = (blockIdx.y * 64 + wc * 32) / 16;
    const bool jv[2] = {16 * jt0 < npad, 16 * (jt0 + 1) < npad};
    Acc acc[2][2];
    for (int f = 0; f < 2; f++)
        for (int jf = 0; jf < 2; jf++)
            for (int r = 0; r < 4; r++) acc[f][jf][r] = 0;
    // [A][n][B] layout: this thread stages the same 4 columns at every K step; decode them once.
    const int cg = (tid & 15) * 4;
    i64 cbase[4];
    bool cok[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const i64 col = c0 + cg + q;
        cok[q] = !XL && col < ncols;
        const i64 aa = XL ? 0 : col / a.B, bb = col - aa * a.B;
        cbase[q] = aa * n * a.B + bb;
    }
    for (int i0 = 0; i0 < npad; i0 += BK) {
        __syncthreads();
        if (XL) {
            constexpr int PER = BK / 4;                  // consecutive i per thread, 4 threads per row
            const int col = tid >> 2, ig = (tid & 3) * PER;
            const i64 row = c0 + col;
#pragma unroll
            for (int q = 0; q < PER; q++) {
                const int i = i0 + ig + q;
                T val = 0;
                if (row < ncols && i < n) val = (T)(in[row * n + i] * a.in_scale);
                tile[ig + q][col] = val;
            }
        } else {
#pragma unroll
            for (int e = tid; e < BK * 16; e += 256) {
                const int ii = e >> 4, i = i0 + ii;
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    T val = 0;
                    if (cok[q] && i < n) val = (T)(in[cbase[q] + (i64)i * a.B] * a.in_scale);
                    tile[ii][cg + q] = val;
                }
            }
        }
        __syncthreads();
        if constexpr (MODE < 2) {
            const T* mp = (const T*)a.mpk;
#pragma unroll
            for (int ks = 0; ks < BK / 4; ks++) {
                const int kt = i0 / 4 + ks;
                const T d[2] = {tile[ks * 4 + lane / 16][wr * 32 + lane % 16],
                                tile[ks * 4 + lane / 16][wr * 32 + 16 + lane % 16]};
#pragma unroll
                for (int jf = 0; jf < 2; jf++) {
                    if (!jv[jf]) continue;
                    const T mf = mp[((i64)(jt0 + jf) * kts + kt) * 64 + lane];
#pragma unroll
                    for (int f = 0; f < 2; f++)
                        acc[f][jf] = XL ? mfma(d[f], mf, acc[f][jf]) : mfma(mf, d[f], acc[f][jf]);
                }
            }
        } else {
            const bf8* mhi = (const bf8*)a.mpk;
            const bf8* mlo = (const bf8*)a.mpk_lo;
            const int kt = i0 / 32;
            bf8 dh[2], dl[2];
#pragma unroll
            for (int f = 0; f < 2; f++)
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    const float x = tile[8 * (lane / 16) + q][wr * 32 + f * 16 + lane % 16];
                    const __bf16 h = (__bf16)x;
                    dh[f][q] = h;
                    dl[f][q] = (__bf16)(x - (float)h);
                }
#pragma unroll
            for (int jf = 0; jf < 2; jf++) {
                if (!jv[jf]) continue;
                const i64 at = ((i64)(jt0 + jf) * kts + kt) * 64 + lane;
                const bf8 mh = mhi[at], ml = mlo[at];
#pragma unroll
                for (int f = 0; f < 2; f++) {
                    if (XL) {
                        acc[f][jf] = mfma(dl[f], mh, acc[f][jf]);
                        acc[f][jf] = mfma(dh[f], ml, acc[f][jf]);
                        acc[f][jf] = mfma(dh[f], mh, acc[f][jf]);
                    } else {
                        acc[f][jf] = mfma(ml, dh[f], acc[f][jf]);
                        acc[f][jf] = mfma(mh, dl[f], acc[f][jf]);
                        acc[f][jf] = mfma(mh, dh[f], acc[f][jf]);
                    }
                }
            }
        }
    }
    const T* il = (const T*)a.il;
#pragma unroll
    for (int f = 0; f < 2; f++)
#pragma unroll
        for (int jf = 0; jf < 2; jf++) {
            if (!jv[jf]) continue;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                i64 idx;
                // result row within the 16 x 16 tile: FP64 interleaves the lane groups (4 r + l/16),
                // the FP32-accumulating instructions stack them (4 (l/16) + r)
                const int rr = MODE == 0 ? 4 * r + lane / 16 : 4 * (lane / 16) + r;
                if (XL) {
                    const i64 row = c0 + wr * 32 + f * 16 + rr;
                    const int j = 16 * (jt0 + jf) + lane % 16;
                    if (row >= ncols || j >= n) continue;
                    idx = row * n + j;
                } else {
                    const int j = 16 * (jt0 + jf) + rr;
                    const i64 col = c0 + wr * 32 + f * 16 + lane % 16;
                    if (col >= ncols || j >= n) continue;
                    const i64 aa = col / a.B, bb = col - aa * a.B;
                    idx = (aa * n + j) * a.B + bb;
                }
                const T val = acc[f][jf][r];
                if (a.epi == 0) out[idx] = (TOUT)val;
                else if (a.epi == 1) out[idx] = (TOUT)(val * il[idx]);
                else out[idx] = out[idx] + (TOUT)((double)val * a.out_scale);
            }
        }
}

// ------------------------------------------------------------------------------------------------ finalise
// backend.cu:846 (subtract the mean) and backend.cu:936 (dphi = phi - base)
__global__ __launch_bounds__(TPB) void k_finalize(i64 n, double* __restrict__ phi, const double* __restrict__ dsum,
                                                  const double* __restrict__ base, double* __restrict__ dphi) {
    const i64 i = (i64)blockIdx.x * TPB + threadIdx.x;
    if (i >= n) return;
    const double mean = dsum[0] / n;
    const double p = phi[i] - mean;
    phi[i] = p;
    if (dphi) dphi[i] = 1.0 * p + -1.0 * base[i];
}

// backend.cu:361-371
template <int COMP>
__device__ void gradient_face(const Geo& g, const double* __restrict__ m, const double* __restrict__ phi,
                              double* __restrict__ out, int i) {
    const int NX = g.nx, NY = g.ny, NZ = g.nz;
    const int sx = NX + (COMP == 0), sy = NY + (COMP == 1), sz = NZ + (COMP == 2);
    if (i >= sx * sy * sz) return;
    const int x = i % sx, y = (i / sx) % sy, z = i / (sx * sy);
    const Ix ix{NX, NY, NZ};
    double d;
    if (COMP == 0) {
        if (x == 0 || x == NX) { out[i] = 0.0; return; }
        d = (phi[ix.c(x, y, z)] - phi[ix.c(x - 1, y, z)]) / sp(g, X, x);
    } else if (COMP == 1) {
        if (y == 0 || y == NY) { out[i] = 0.0; return; }
        d = (phi[ix.c(x, y, z)] - phi[ix.c(x, y - 1, z)]) / sp(g, Y, y);
    } else {
        if (z == 0 || z == NZ) { out[i] = 0.0; return; }
        d = (phi[ix.c(x, y, z)] - phi[ix.c(x, y, z - 1)]) / sp(g, Z, z);
    }
    out[i] = m[i] - d;
}

struct Face3 {
    const double* m[3];
    double* out[3];
};
__global__ __launch_bounds__(TPB) void k_gradient(Geo g, Face3 f, const double* __restrict__ phi, int b0, int b1) {
    const int blk = blockIdx.x;
    if (blk < b0) gradient_face<0>(g, f.m[0], phi, f.out[0], blk * TPB + threadIdx.x);
    else if (blk < b0 + b1) gradient_face<1>(g, f.m[1], phi, f.out[1], (blk - b0) * TPB + threadIdx.x);
    else gradient_face<2>(g, f.m[2], phi, f.out[2], (blk - b0 - b1) * TPB + threadIdx.x);
}

// ------------------------------------------------------------------------------------------------ outputs
// backend.cu:67-72
__device__ __forceinline__ double centred(const Ix& ix, const double* u, const double* v, const double* w, int c,
                                          int x, int y, int z) {
    if (c == 0) return 0.5 * (u[ix.u(x, y, z)] + u[ix.u(x + 1, y, z)]);
    if (c == 1) return 0.5 * (v[ix.v(x, y, z)] + v[ix.v(x, y + 1, z)]);
    return 0.5 * (w[ix.w(x, y, z)] + w[ix.w(x, y, z + 1)]);
}

// backend.cu:76-96
__device__ void cell_gradient(const Geo& g, const double* u, const double* v, const double* w, double lid, int x,
                              int y, int z, double gr[3][3]) {
    const Ix ix{g.nx, g.ny, g.nz};
    const int NX = g.nx, NY = g.ny, NZ = g.nz;
    const double dxm = x > 0 ? sp(g, X, x) : wd(g, X, 0), dxp = x + 1 < NX ? sp(g, X, x + 1) : wd(g, X, NX - 1);
    const double dym = y > 0 ? sp(g, Y, y) : wd(g, Y, 0), dyp = y + 1 < NY ? sp(g, Y, y + 1) : wd(g, Y, NY - 1);
    const double dzm = z > 0 ? sp(g, Z, z) : wd(g, Z, 0), dzp = z + 1 < NZ ? sp(g, Z, z + 1) : wd(g, Z, NZ - 1);
    for (int c = 0; c < 3; c++) {
        const double q = centred(ix, u, v, w, c, x, y, z);
        const double xm = x > 0 ? centred(ix, u, v, w, c, x - 1, y, z) : -q;
        const double xp = x + 1 < NX ? centred(ix, u, v, w, c, x + 1, y, z) : -q;
        const double ym = y > 0 ? centred(ix, u, v, w, c, x, y - 1, z) : -q;
        const double yp = y + 1 < NY ? centred(ix, u, v, w, c, x, y + 1, z)
                                     : 2 * (c == 0 ? lid * lid_at_cell(g, x, z) : 0) - q;
        const double zp = z + 1 < NZ ? centred(ix, u, v, w, c, x, y, z + 1) : -q;
        const double zm = z > 0 ? centred(ix, u, v, w, c, x, y, z - 1) : -q;
        gr[0][c] = (xp - xm) / (dxp + dxm);
        gr[1][c] = (yp - ym) / (dyp + dym);
        gr[2][c] = (zp - zm) / (dzp + dzm);
    }
    gr[0][0] = (u[ix.u(x + 1, y, z)] - u[ix.u(x, y, z)]) / wd(g, X, x);
    gr[1][1] = (v[ix.v(x, y + 1, z)] - v[ix.v(x, y, z)]) / wd(g, Y, y);
    gr[2][2] = (w[ix.w(x, y, z + 1)] - w[ix.w(x, y, z)]) / wd(g, Z, z);
}

// backend.cu:447-463, the three volume sums fused: partial[k*RED_BLOCKS + block]
__global__ __launch_bounds__(TPB) void k_cell_diag(Geo g, const double* __restrict__ u, const double* __restrict__ v,
                                                   const double* __restrict__ w, double lid, double* __restrict__ partial) {
    const Ix ix{g.nx, g.ny, g.nz};
    const i64 n = (i64)g.nx * g.ny * g.nz;
    double ke = 0, ens = 0, dv = 0;
    for (i64 i = (i64)blockIdx.x * TPB + threadIdx.x; i < n; i += (i64)gridDim.x * TPB) {
        const int x = (int)(i % g.nx), y = (int)((i / g.nx) % g.ny), z = (int)(i / ((i64)g.nx * g.ny));
        const double vol = wd(g, X, x) * wd(g, Y, y) * wd(g, Z, z);
        const double a = centred(ix, u, v, w, 0, x, y, z), b = centred(ix, u, v, w, 1, x, y, z),
                     c = centred(ix, u, v, w, 2, x, y, z);
        ke += 0.5 * (a * a + b * b + c * c) * vol;
        double gr[3][3];
        cell_gradient(g, u, v, w, lid, x, y, z, gr);
        const double ox = gr[1][2] - gr[2][1], oy = gr[2][0] - gr[0][2], oz = gr[0][1] - gr[1][0];
        ens += (ox * ox + oy * oy + oz * oz) * vol;
        const double d = (u[ix.u(x + 1, y, z)] - u[ix.u(x, y, z)]) / wd(g, X, x)
                       + (v[ix.v(x, y + 1, z)] - v[ix.v(x, y, z)]) / wd(g, Y, y)
                       + (w[ix.w(x, y, z + 1)] - w[ix.w(x, y, z)]) / wd(g, Z, z);
        dv += d * d * vol;
    }
    ke = block_reduce(1, ke);
    ens = block_reduce(1, ens);
    dv = block_reduce(1, dv);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = ke;
        partial[RED_BLOCKS + blockIdx.x] = ens;
        partial[2 * RED_BLOCKS + blockIdx.x] = dv;
    }
}

// backend.cu:415-422 (|m - u| max, then its first index)
__global__ __launch_bounds__(TPB) void k_max_abs_diff(const double* __restrict__ a, const double* __restrict__ b, i64 n,
                                                      double* __restrict__ partial) {
    double r = -INFINITY;
    for (i64 i = (i64)blockIdx.x * TPB + threadIdx.x; i < n; i += (i64)gridDim.x * TPB) r = comb(0, r, fabs(a[i] - b[i]));
    r = block_reduce(0, r);
    if (threadIdx.x == 0) partial[blockIdx.x] = r;
}
__global__ __launch_bounds__(TPB) void k_first_index(const double* __restrict__ a, const double* __restrict__ b, i64 n,
                                                     const double* __restrict__ value, int* __restrict__ first) {
    const i64 i = (i64)blockIdx.x * TPB + threadIdx.x;
    if (i >= n) return;
    if (fabs(a[i] - b[i]) == value[0]) atomicMin(first, (int)i);
}
__global__ void k_set_int(int* p, int v) { *p = v; }

// backend.cu:534-540
__global__ void k_probes(Geo g, const double* __restrict__ u, const double* __restrict__ v, const int* __restrict__ cells,
                         int n, float* __restrict__ out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const Ix ix{g.nx, g.ny, g.nz};
    const int x = cells[3 * p], y = cells[3 * p + 1], z = cells[3 * p + 2];
    out[2 * p] = (float)(0.5 * (u[ix.u(x, y, z)] + u[ix.u(x + 1, y, z)]));
    out[2 * p + 1] = (float)(0.5 * (v[ix.v(x, y, z)] + v[ix.v(x, y + 1, z)]));
}

__device__ __forceinline__ void midspan_layers(int nz, int& k1, int& k2) { k2 = nz / 2; k1 = (nz % 2) ? k2 : k2 - 1; }

// backend.cu:507-517
__global__ __launch_bounds__(TPB) void k_add_statistics(Geo g, const double* __restrict__ u, const double* __restrict__ v,
                                                        const double* __restrict__ w, double weight,
                                                        double* __restrict__ sums) {
    const int i = blockIdx.x * TPB + threadIdx.x, n = g.nx * g.ny;
    if (i >= n) return;
    const Ix ix{g.nx, g.ny, g.nz};
    const int x = i % g.nx, y = i / g.nx;
    int k1, k2;
    midspan_layers(g.nz, k1, k2);
    const double a = 0.5 * (centred(ix, u, v, w, 0, x, y, k1) + centred(ix, u, v, w, 0, x, y, k2));
    const double b = 0.5 * (centred(ix, u, v, w, 1, x, y, k1) + centred(ix, u, v, w, 1, x, y, k2));
    const double c = 0.5 * (centred(ix, u, v, w, 2, x, y, k1) + centred(ix, u, v, w, 2, x, y, k2));
    sums[i] += weight * a;
    sums[n + i] += weight * b;
    sums[2 * n + i] += weight * (a * a);
    sums[3 * n + i] += weight * (b * b);
    sums[4 * n + i] += weight * (c * c);
    sums[5 * n + i] += weight * (a * b);
}

// backend.cu:519-532 (nx == ny)
__global__ void k_centreline(Geo g, const double* __restrict__ u, const double* __restrict__ v,
                             const double* __restrict__ w, double* __restrict__ out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x, NX = g.nx;
    if (j >= NX) return;
    const Ix ix{g.nx, g.ny, g.nz};
    const int mid = NX / 2;
    int k1, k2;
    midspan_layers(g.nz, k1, k2);
    auto mean4 = [&](int c, int xa, int ya, int xb, int yb) {
        return 0.5 * (0.5 * (centred(ix, u, v, w, c, xa, ya, k1) + centred(ix, u, v, w, c, xa, ya, k2))
                      + 0.5 * (centred(ix, u, v, w, c, xb, yb, k1) + centred(ix, u, v, w, c, xb, yb, k2)));
    };
    const double a = mean4(0, mid - 1, j, mid, j), b = mean4(1, j, mid - 1, j, mid), c = mean4(2, mid - 1, j, mid, j);
    out[j] = a;
    out[NX + j] = b;
    out[2 * NX + j] = a * a;
    out[3 * NX + j] = b * b;
    out[4 * NX + j] = c * c;
}

// backend.cu:490-505
__global__ __launch_bounds__(TPB) void k_midplane(Geo g, const double* __restrict__ u, const double* __restrict__ v,
                                                  const double* __restrict__ w, double lid, float* __restrict__ out) {
    const int i = blockIdx.x * TPB + threadIdx.x, P = g.nx * g.ny;
    if (i >= P) return;
    const Ix ix{g.nx, g.ny, g.nz};
    const int x = i % g.nx, y = i / g.nx;
    int k1, k2;
    midspan_layers(g.nz, k1, k2);
    double val[2][4];
    for (int s = 0; s < 2; s++) {
        const int z = s ? k2 : k1;
        for (int c = 0; c < 3; c++) val[s][c] = centred(ix, u, v, w, c, x, y, z);
        double gr[3][3];
        cell_gradient(g, u, v, w, lid, x, y, z, gr);
        const double a = gr[1][2] - gr[2][1], b = gr[2][0] - gr[0][2], c = gr[0][1] - gr[1][0];
        val[s][3] = sqrt(a * a + b * b + c * c);
    }
    for (int c = 0; c < 4; c++) out[c * P + i] = (float)((k1 == k2) ? val[1][c] : 0.5 * (val[0][c] + val[1][c]));
}

// backend.cu:474-486
__global__ __launch_bounds__(TPB) void k_q_criterion(Geo g, const double* __restrict__ u, const double* __restrict__ v,
                                                     const double* __restrict__ w, double lid, double qscale,
                                                     double inv_U, int colour_w, float* __restrict__ q,
                                                     float* __restrict__ colour) {
    const i64 n = (i64)g.nx * g.ny * g.nz;
    const i64 i = (i64)blockIdx.x * TPB + threadIdx.x;
    if (i >= n) return;
    const Ix ix{g.nx, g.ny, g.nz};
    const int x = (int)(i % g.nx), y = (int)((i / g.nx) % g.ny), z = (int)(i / ((i64)g.nx * g.ny));
    double gr[3][3];
    cell_gradient(g, u, v, w, lid, x, y, z, gr);
    double s = 0.0;
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) s += gr[a][b] * gr[b][a];
    q[i] = (float)(-0.5 * s * qscale);
    const double a = centred(ix, u, v, w, 0, x, y, z), b = centred(ix, u, v, w, 1, x, y, z),
                 c = centred(ix, u, v, w, 2, x, y, z);
    colour[i] = (float)(colour_w ? c * inv_U : sqrt(a * a + b * b + c * c) * inv_U);
}

// ------------------------------------------------------------------------------------------------ kick
// backend.cu:544-553
__global__ __launch_bounds__(TPB) void k_smooth(const double* __restrict__ in, double* __restrict__ out, int axis,
                                                int sx, int sy, int sz) {
    const i64 i = (i64)blockIdx.x * TPB + threadIdx.x;
    if (i >= (i64)sx * sy * sz) return;
    const int x = (int)(i % sx), y = (int)((i / sx) % sy), z = (int)(i / ((i64)sx * sy));
    i64 lo, hi;
    if (axis == 0) {
        lo = (x + sx - 1) % sx + (i64)sx * (y + (i64)sy * z);
        hi = (x + 1) % sx + (i64)sx * (y + (i64)sy * z);
    } else if (axis == 1) {
        lo = x + (i64)sx * ((y + sy - 1) % sy + (i64)sy * z);
        hi = x + (i64)sx * ((y + 1) % sy + (i64)sy * z);
    } else {
        lo = x + (i64)sx * (y + (i64)sy * ((z + sz - 1) % sz));
        hi = x + (i64)sx * (y + (i64)sy * ((z + 1) % sz));
    }
    out[i] = 0.25 * in[lo] + 0.5 * in[i] + 0.25 * in[hi];
}
__global__ __launch_bounds__(TPB) void k_add_noise(const double* __restrict__ f, const double* __restrict__ noise,
                                                   double s, double* __restrict__ out, i64 n) {
    const i64 i = (i64)blockIdx.x * TPB + threadIdx.x;
    if (i < n) out[i] = f[i] + noise[i] * s;
}
__global__ __launch_bounds__(TPB) void k_add_difference(double* __restrict__ m, const double* __restrict__ a,
                                                        const double* __restrict__ b, i64 n) {
    const i64 i = (i64)blockIdx.x * TPB + threadIdx.x;
    if (i < n) m[i] = m[i] + (a[i] - b[i]);
}

}  // namespace

// ================================================================================================ wrappers
void reduce_array(const Exec& ex, const double* x, i64 n, int op, double* partial, double* out) {
    const int nb = red_blocks(n);
    run(ex, F_REDUCE, k_reduce_array, dim3(nb), dim3(TPB), x, n, op, partial);
    reduce_final(ex, partial, nb, op, out);
}
void reduce_final(const Exec& ex, const double* partial, int count, int op, double* out) {
    run(ex, F_REDUCE, k_reduce_final, dim3(1), dim3(TPB), partial, count, op, out);
}
void fill(const Exec& ex, double* p, i64 n, double value) {
    run(ex, F_FILL, k_fill, dim3(red_blocks(n)), dim3(TPB), p, n, value);
}

void gauge_reset(const Exec& ex, const Geo& g, const double* const m[3], const double* const u[3], const double* phi,
                 double r, double omr, double* const m0[3], double* phi0) {
    GaugeArgs a;
    const i64 n[4] = {(i64)(g.nx + 1) * g.ny * g.nz, (i64)g.nx * (g.ny + 1) * g.nz, (i64)g.nx * g.ny * (g.nz + 1),
                      (i64)g.nx * g.ny * g.nz};
    int total = 0;
    for (int k = 0; k < 4; k++) {
        a.n[k] = n[k];
        a.blocks[k] = blocks_for(n[k]);
        total += a.blocks[k];
    }
    for (int k = 0; k < 3; k++) { a.m[k] = m[k]; a.u[k] = u[k]; a.m0[k] = m0[k]; }
    a.phi = phi;
    a.phi0 = phi0;
    a.r = r;
    a.omr = omr;
    run(ex, F_GAUGE, k_gauge, dim3(total), dim3(TPB), a);
}

void cfl_max(const Exec& ex, const Geo& g, const double* const u[3], double* partial, double* out) {
    const int nb = red_blocks((i64)g.nx * g.ny * g.nz);
    run(ex, F_CFL, k_cfl, dim3(nb), dim3(TPB), g, u[0], u[1], u[2], partial);
    reduce_final(ex, partial, nb, 0, out);
}

void stage(const Exec& ex, const Geo& g, const StageArgs& a) {
    const int b0 = blocks_for((i64)(g.nx + 1) * g.ny * g.nz), b1 = blocks_for((i64)g.nx * (g.ny + 1) * g.nz),
              b2 = blocks_for((i64)g.nx * g.ny * (g.nz + 1));
    run(ex, F_STAGE, k_stage, dim3(b0 + b1 + b2), dim3(TPB), g, a, b0, b1);
}

int divergence(const Exec& ex, const Geo& g, const double* const f[3], double* rhs, double* partial) {
    const int nb = red_blocks((i64)g.nx * g.ny * g.nz);
    run(ex, F_DIVERGENCE, k_divergence, dim3(nb), dim3(TPB), g, f[0], f[1], f[2], rhs, partial);
    return nb;
}

void zero_walls(const Exec& ex, const Geo& g, double* const f[3]) {
    const int total = 2 * g.ny * g.nz + 2 * g.nx * g.nz + 2 * g.nx * g.ny;
    run(ex, F_FILL, k_zero_walls, dim3(blocks_for(total)), dim3(TPB), g, f[0], f[1], f[2]);
}

void base_guess(const Exec& ex, i64 n, const double* phi0, const double* cur, double a0, double b, int has_a0,
                const double* h1, double c1, const double* h2, double c2, int nh, double* base, double* phi) {
    run(ex, F_GUESS, k_base_guess, dim3(blocks_for(n)), dim3(TPB), n, phi0, cur, a0, b, has_a0, h1, c1, h2, c2, nh,
        base, phi);
}

int residual(const Exec& ex, const Geo& g, const double* phi, const double* rhs, double* r, double* partial) {
    const int nb = red_blocks((i64)g.nx * g.ny * g.nz);
    run(ex, F_RESIDUAL, k_residual, dim3(nb), dim3(TPB), g, phi, rhs, r, partial);
    return nb;
}

int pad_for_mode(int n, int mode) { return mode == 2 ? (n + 31) / 32 * 32 : pad16(n); }

void axis_transform(const Exec& ex, bool layout_x, int mode, bool in_f64, bool out_f64, const AxisArgs& a,
                    int family) {
    const int npad = pad_for_mode(a.n, mode);
    const i64 ncols = layout_x ? a.A : a.A * a.B;
    const dim3 grid((unsigned)((ncols + 63) / 64), (unsigned)((npad + 63) / 64)), block(256);
    if (mode == 0) {
        if (layout_x) run(ex, family, k_axis<true, 0, double, double>, grid, block, a, ncols, npad);
        else run(ex, family, k_axis<false, 0, double, double>, grid, block, a, ncols, npad);
        return;
    }
#define CAV_AXIS_MODE(M)                                                                                        \
    if (layout_x && in_f64) run(ex, family, k_axis<true, M, double, float>, grid, block, a, ncols, npad);      \
    else if (layout_x && out_f64) run(ex, family, k_axis<true, M, float, double>, grid, block, a, ncols, npad); \
    else if (layout_x) run(ex, family, k_axis<true, M, float, float>, grid, block, a, ncols, npad);            \
    else run(ex, family, k_axis<false, M, float, float>, grid, block, a, ncols, npad);
    if (mode == 1) { CAV_AXIS_MODE(1) }
    else { CAV_AXIS_MODE(2) }
#undef CAV_AXIS_MODE
}

void finalize_phi(const Exec& ex, i64 n, double* phi, const double* dsum, const double* base, double* dphi) {
    run(ex, F_FINALIZE, k_finalize, dim3(blocks_for(n)), dim3(TPB), n, phi, dsum, base, dphi);
}

void subtract_gradient(const Exec& ex, const Geo& g, const double* const m[3], const double* phi,
                       double* const out[3]) {
    const int b0 = blocks_for((i64)(g.nx + 1) * g.ny * g.nz), b1 = blocks_for((i64)g.nx * (g.ny + 1) * g.nz),
              b2 = blocks_for((i64)g.nx * g.ny * (g.nz + 1));
    Face3 f;
    for (int k = 0; k < 3; k++) { f.m[k] = m[k]; f.out[k] = out[k]; }
    run(ex, F_GRADIENT, k_gradient, dim3(b0 + b1 + b2), dim3(TPB), g, f, phi, b0, b1);
}

void cell_diagnostics(const Exec& ex, const Geo& g, const double* const u[3], double lid, double* partial3,
                      double* out3) {
    const int nb = red_blocks((i64)g.nx * g.ny * g.nz);
    run(ex, F_OUTPUT, k_cell_diag, dim3(nb), dim3(TPB), g, u[0], u[1], u[2], lid, partial3);
    for (int k = 0; k < 3; k++) reduce_final(ex, partial3 + k * RED_BLOCKS, nb, 1, out3 + k);
}

void max_abs_diff(const Exec& ex, const double* a, const double* b, i64 n, double* partial, double* out) {
    const int nb = red_blocks(n);
    run(ex, F_OUTPUT, k_max_abs_diff, dim3(nb), dim3(TPB), a, b, n, partial);
    reduce_final(ex, partial, nb, 0, out);
}

void first_index_of_abs_diff(const Exec& ex, const double* a, const double* b, i64 n, const double* value, int* first) {
    run(ex, F_OUTPUT, k_set_int, dim3(1), dim3(1), first, INT_MAX);
    run(ex, F_OUTPUT, k_first_index, dim3(blocks_for(n)), dim3(TPB), a, b, n, value, first);
}

void probe_values(const Exec& ex, const Geo& g, const double* u, const double* v, const int* cells, int n, float* out) {
    run(ex, F_OUTPUT, k_probes, dim3(1), dim3(64), g, u, v, cells, n, out);
}

void add_statistics(const Exec& ex, const Geo& g, const double* const u[3], double weight, double* sums) {
    run(ex, F_OUTPUT, k_add_statistics, dim3(blocks_for((i64)g.nx * g.ny)), dim3(TPB), g, u[0], u[1], u[2], weight,
        sums);
}

void centreline_sample(const Exec& ex, const Geo& g, const double* const u[3], double* out) {
    run(ex, F_OUTPUT, k_centreline, dim3(blocks_for(g.nx)), dim3(TPB), g, u[0], u[1], u[2], out);
}

void midplane(const Exec& ex, const Geo& g, const double* const u[3], double lid, float* out) {
    run(ex, F_OUTPUT, k_midplane, dim3(blocks_for((i64)g.nx * g.ny)), dim3(TPB), g, u[0], u[1], u[2], lid, out);
}

void q_criterion(const Exec& ex, const Geo& g, const double* const u[3], double lid, double qscale, double inv_U,
                 int colour_w, float* q, float* colour) {
    run(ex, F_OUTPUT, k_q_criterion, dim3(blocks_for((i64)g.nx * g.ny * g.nz)), dim3(TPB), g, u[0], u[1], u[2], lid,
        qscale, inv_U, colour_w, q, colour);
}

void smooth(const Exec& ex, const double* in, double* out, int axis, int sx, int sy, int sz) {
    run(ex, F_KICK, k_smooth, dim3(blocks_for((i64)sx * sy * sz)), dim3(TPB), in, out, axis, sx, sy, sz);
}
void add_noise(const Exec& ex, const double* f, const double* noise, double s, double* out, i64 n) {
    run(ex, F_KICK, k_add_noise, dim3(blocks_for(n)), dim3(TPB), f, noise, s, out, n);
}
void add_difference(const Exec& ex, double* m, const double* a, const double* b, i64 n) {
    run(ex, F_KICK, k_add_difference, dim3(blocks_for(n)), dim3(TPB), m, a, b, n);
}

}  // namespace cav
