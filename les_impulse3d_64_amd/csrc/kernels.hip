// kernels.hip — gfx950 kernels of the cavity time step and its outputs.
//
// The stencils follow the reference's formulas and order of operations (backend.cu:98-561) in FP64,
// compiled with -ffp-contract=off, with one change: every division by a grid quantity (width, spacing,
// node distance, edge weights, the wall ramp) is a multiplication by a reciprocal or coefficient table
// built once on the host (axis_tables). That moves each result by at most about one ulp per operation
// against the host oracle; the reductions (block/wavefront trees) and the FDM transforms (MFMA
// accumulation order) round differently as well.
#include <climits>

#include "kernels.h"

namespace cav {

const char* family_name(int f) {
    static const char* names[F_COUNT] = {"gauge_reset", "cfl_max", "edges", "stage", "divergence", "base_guess",
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

// one block of 1024 threads folds the partials of one reduction (strided, fixed order)
__global__ __launch_bounds__(1024) void k_reduce_final(const double* __restrict__ partial, int count, int op,
                                                       double* __restrict__ out) {
    __shared__ double part[16];
    const int sop = (op == 2 || op == 0) ? 0 : 1;
    double r = init_of(sop);
    for (int i = threadIdx.x; i < count; i += 1024) r = comb(sop, r, partial[i]);
    for (int off = 32; off > 0; off >>= 1) r = comb(sop, r, __shfl_down(r, off, 64));
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = r;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = part[0];
        for (int k = 1; k < 16; k++) t = comb(sop, t, part[k]);
        out[0] = t;
    }
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

// Table entry kind at index e of axis a (Geo::tab rows of n_a + 1 entries)
__device__ __forceinline__ int axis_n(const Geo& g, int a) { return a == 0 ? g.nx : a == 1 ? g.ny : g.nz; }
__device__ __forceinline__ double tb(const Geo& g, int a, int kind, int e) {
    return g.tab[a][kind * (axis_n(g, a) + 1) + e];
}

// Workgroups are dealt round-robin to the 8 XCDs (block b runs on XCD b % 8), each with its own L2.
// xcd_block maps a hardware block id to a logical one so that XCD k gets the contiguous logical range
// [k G/8, (k+1) G/8): neighbouring rows and planes of a stencil then share an L2. G must be a multiple of 8.
__device__ __forceinline__ int xcd_block(int b, int G) { return (b % 8) * (G / 8) + b / 8; }
inline int round8(i64 n) { return (int)((n + 7) / 8 * 8); }

// z-march: every thread owns one (x, y) column of a px x py plane and walks a chunk of z, keeping the
// z-neighbours it needs in registers, so each plane of an array is read from memory once (x/y neighbours
// come from the same plane, just read by the neighbouring threads). Blocks cover 32 x 8 columns; the
// logical block order is (z chunk, tile y, tile x) and is dealt to the XCDs in contiguous ranges.
constexpr int MTX = 32, MTY = 8;
struct March {
    int tx, ty, chunks, zc;    // tiles along x and y, number of z chunks, chunk length
};
inline March march_of(int px, int py, int pz) {
    March m;
    m.tx = (px + MTX - 1) / MTX;
    m.ty = (py + MTY - 1) / MTY;
    const int tiles = m.tx * m.ty;
    int chunks = std::max(1, std::min(2048 / std::max(1, tiles), std::max(1, pz / 8)));
    m.zc = (pz + chunks - 1) / chunks;
    m.chunks = (pz + m.zc - 1) / m.zc;
    return m;
}
inline int march_blocks(const March& m) { return round8((i64)m.tx * m.ty * m.chunks); }
// this thread's column (x, y) and z range [z0, z1) of a plane extent pz; empty range for padding blocks
__device__ __forceinline__ void march_point(const March& m, int pz, int& x, int& y, int& z0, int& z1) {
    const int L = xcd_block(blockIdx.x, gridDim.x), tiles = m.tx * m.ty;
    const int chunk = L / tiles, t = L - chunk * tiles;
    x = (t % m.tx) * MTX + (threadIdx.x & (MTX - 1));
    y = (t / m.tx) * MTY + threadIdx.x / MTX;
    z0 = chunk * m.zc;
    z1 = chunk < m.chunks ? min(pz, z0 + m.zc) : z0;
}

// ------------------------------------------------------------------------------------------------ CFL
// backend.cu:465-472 per cell, reduced with max (NaN wins)
__global__ __launch_bounds__(TPB) void k_cfl(Geo g, const double* __restrict__ u, const double* __restrict__ v,
                                             const double* __restrict__ w, double* __restrict__ partial, March M) {
    const Ix ix{g.nx, g.ny, g.nz};
    double r = -INFINITY;
    int x, y, z0, z1;
    march_point(M, g.nz, x, y, z0, z1);
    if (x < g.nx && y < g.ny && z0 < z1) {
        const double rwx = tb(g, X, T_RW, x), rwy = tb(g, Y, T_RW, y);
        double w0 = fabs(w[ix.w(x, y, z0)]);
        for (int z = z0; z < z1; z++) {
            const double w1 = fabs(w[ix.w(x, y, z + 1)]);
            const double rate = fmax(fabs(u[ix.u(x, y, z)]), fabs(u[ix.u(x + 1, y, z)])) * rwx
                              + fmax(fabs(v[ix.v(x, y, z)]), fabs(v[ix.v(x, y + 1, z)])) * rwy
                              + fmax(w0, w1) * tb(g, Z, T_RW, z);
            r = comb(0, r, rate);
            w0 = w1;
        }
    }
    r = block_reduce(0, r);
    if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

// ------------------------------------------------------------------------------------------------ I(u, m)
// backend.cu:102-274: I = u x omega - grad(K), K = u.m. The Lamb vector is assembled from products
// (vorticity on an edge) x (velocity interpolated to that edge), narrow q = 1 and wide q = 3 stencils, with a
// wall ramp on the 4th-order weights. Each product depends on its edge only, so k_edges evaluates every
// edge once (the two velocity differences of the vorticity are also the two nodes of both interpolations)
// and k_faces combines them per face. Grid-only factors come from the axis tables.
constexpr int STX = 32, STY = 8;        // 2-D tile of one plane per 256-thread block

__device__ __forceinline__ bool nodes(int e, int q, int n, int& lo, int& hi) {    // backend.cu:108-112
    lo = e - (q + 1) / 2;
    hi = e + (q - 1) / 2;
    return lo >= 0 && hi <= n - 1;
}
__device__ __forceinline__ double stencil_weight(int q, int o4, double s) {        // backend.cu:135-138
    if (q == 1) return o4 ? 1.0 + s * 0.125 : 1.0;
    return -s * 0.125;
}
// width-weighted value at edge e of axis a from nodes lo, hi (backend.cu:119-122)
__device__ __forceinline__ double at_edge(const Geo& g, int a, int e, int q, int lo, int hi, double flo, double fhi) {
    return (tb(g, a, T_W, lo) * flo + tb(g, a, T_W, hi) * fhi) * tb(g, a, q == 1 ? T_R2SP : T_RWS3, e);
}
__device__ __forceinline__ double rgap(const Geo& g, int a, int e, int q) { return tb(g, a, q == 1 ? T_RSP : T_RGAP3, e); }

i64 edge_size_of(int nx, int ny, int nz, int which) {
    switch (which % 6) {
        case 0: case 1: return (i64)(nx + 1) * (ny + 1) * nz;        // z-edges
        case 2: case 3: return (i64)nx * (ny + 1) * (nz + 1);        // x-edges
        default: return (i64)(nx + 1) * ny * (nz + 1);               // y-edges
    }
}

// Tile -> (x, y, z) of an sx x sy x sz array: 32 x 8 tiles, plane by plane. False outside the array.
__device__ __forceinline__ bool tile_point(int tile, int sx, int sy, int sz, int& x, int& y, int& z) {
    const int tx = (sx + STX - 1) / STX, ty = (sy + STY - 1) / STY;
    z = tile / (tx * ty);
    const int rem = tile - z * tx * ty;
    x = (rem % tx) * STX + (threadIdx.x & (STX - 1));
    y = (rem / tx) * STY + threadIdx.x / STX;
    return z < sz && x < sx && y < sy;
}
inline int tiles_of(int sx, int sy, int sz) {
    return round8((i64)((sx + STX - 1) / STX) * ((sy + STY - 1) / STY) * sz);
}

// One edge family: a along the edge, (p, r) the two axes the edge sits on a face position of.
//   z-edges (ex, ey, z): omega_z = dv/dx - du/dy; E_ZX = omega_z v|x (backend.cu:162-170), E_ZY = omega_z u|y (171-179)
//   x-edges (x, ey, ez): omega_x = dw/dy - dv/dz; E_XY = omega_x w|y (180-188), E_XZ = omega_x v|z (189-197)
//   y-edges (ex, y, ez): omega_y = du/dz - dw/dx; E_YX = omega_y w|x (198-206), E_YZ = omega_y u|z (207-215)
// The products are 0 where a stencil leaves the box (backend.cu:143, 149, 155).
template <int FAM>
__device__ void edge_point(const Geo& g, const StageArgs& a, int i0, int i1, int i2) {
    const Ix ix{g.nx, g.ny, g.nz};
    const double *u = a.u[0], *v = a.u[1], *w = a.u[2];
    for (int qi = 0; qi < (a.o4 ? 2 : 1); qi++) {
        const int q = qi ? 3 : 1;
        double p0 = 0.0, p1 = 0.0;
        if (FAM == 0) {                                   // z-edge at (ex, ey) = (i0, i1), cell layer z = i2
            int xl, xh, yl, yh;
            if (nodes(i0, q, g.nx, xl, xh) && nodes(i1, q, g.ny, yl, yh)) {
                const double vl = v[ix.v(xl, i1, i2)], vh = v[ix.v(xh, i1, i2)];
                const double ul = u[ix.u(i0, yl, i2)], uh = u[ix.u(i0, yh, i2)];
                const double o = (vh - vl) * rgap(g, X, i0, q) - (uh - ul) * rgap(g, Y, i1, q);
                const double wo = stencil_weight(q, a.o4, tb(g, Y, T_RAMP, i1) * tb(g, X, T_RAMP, i0)) * o;
                p0 = wo * at_edge(g, X, i0, q, xl, xh, vl, vh);
                p1 = wo * at_edge(g, Y, i1, q, yl, yh, ul, uh);
            }
            const int e = i0 + (g.nx + 1) * (i1 + (g.ny + 1) * i2);
            a.edge[qi ? E_ZX3 : E_ZX1][e] = p0;
            a.edge[qi ? E_ZY3 : E_ZY1][e] = p1;
        } else if (FAM == 1) {                            // x-edge at cell x = i0, (ey, ez) = (i1, i2)
            int yl, yh, zl, zh;
            if (nodes(i1, q, g.ny, yl, yh) && nodes(i2, q, g.nz, zl, zh)) {
                const double wl = w[ix.w(i0, yl, i2)], wh = w[ix.w(i0, yh, i2)];
                const double vl = v[ix.v(i0, i1, zl)], vh = v[ix.v(i0, i1, zh)];
                const double o = (wh - wl) * rgap(g, Y, i1, q) - (vh - vl) * rgap(g, Z, i2, q);
                const double wo = stencil_weight(q, a.o4, tb(g, Z, T_RAMP, i2) * tb(g, Y, T_RAMP, i1)) * o;
                p0 = wo * at_edge(g, Y, i1, q, yl, yh, wl, wh);
                p1 = wo * at_edge(g, Z, i2, q, zl, zh, vl, vh);
            }
            const int e = i0 + g.nx * (i1 + (g.ny + 1) * i2);
            a.edge[qi ? E_XY3 : E_XY1][e] = p0;
            a.edge[qi ? E_XZ3 : E_XZ1][e] = p1;
        } else {                                          // y-edge at (ex, ez) = (i0, i2), cell y = i1
            int xl, xh, zl, zh;
            if (nodes(i0, q, g.nx, xl, xh) && nodes(i2, q, g.nz, zl, zh)) {
                const double ul = u[ix.u(i0, i1, zl)], uh = u[ix.u(i0, i1, zh)];
                const double wl = w[ix.w(xl, i1, i2)], wh = w[ix.w(xh, i1, i2)];
                const double o = (uh - ul) * rgap(g, Z, i2, q) - (wh - wl) * rgap(g, X, i0, q);
                const double wo = stencil_weight(q, a.o4, tb(g, Z, T_RAMP, i2) * tb(g, X, T_RAMP, i0)) * o;
                p0 = wo * at_edge(g, X, i0, q, xl, xh, wl, wh);
                p1 = wo * at_edge(g, Z, i2, q, zl, zh, ul, uh);
            }
            const int e = i0 + (g.nx + 1) * (i1 + g.ny * i2);
            a.edge[qi ? E_YX3 : E_YX1][e] = p0;
            a.edge[qi ? E_YZ3 : E_YZ1][e] = p1;
        }
    }
}

// backend.cu:218-222: K = u.m at a cell centre
__device__ void kinetic_point(const Geo& g, const StageArgs& a, int x, int y, int z) {
    const Ix ix{g.nx, g.ny, g.nz};
    const int ia = ix.u(x, y, z), ib = ix.v(x, y, z), ic = ix.w(x, y, z), ic1 = ix.w(x, y, z + 1), NX = g.nx;
    const double *u = a.u[0], *v = a.u[1], *w = a.u[2], *mx = a.m[0], *my = a.m[1], *mz = a.m[2];
    a.edge[E_K][ix.c(x, y, z)] = 0.5 * ((u[ia] * mx[ia] + u[ia + 1] * mx[ia + 1]) + (v[ib] * my[ib] + v[ib + NX] * my[ib + NX])
                                        + (w[ic] * mz[ic] + w[ic1] * mz[ic1]));
}

// segments (multiples of 8 blocks): z-edges, x-edges, y-edges, cells
__global__ __launch_bounds__(TPB) void k_edges(Geo g, StageArgs a, int b0, int b1, int b2, int b3) {
    int blk = blockIdx.x, x, y, z;
    if (blk < b0) {
        if (tile_point(xcd_block(blk, b0), g.nx + 1, g.ny + 1, g.nz, x, y, z)) edge_point<0>(g, a, x, y, z);
        return;
    }
    blk -= b0;
    if (blk < b1) {
        if (tile_point(xcd_block(blk, b1), g.nx, g.ny + 1, g.nz + 1, x, y, z)) edge_point<1>(g, a, x, y, z);
        return;
    }
    blk -= b1;
    if (blk < b2) {
        if (tile_point(xcd_block(blk, b2), g.nx + 1, g.ny, g.nz + 1, x, y, z)) edge_point<2>(g, a, x, y, z);
        return;
    }
    blk -= b2;
    if (tile_point(xcd_block(blk, b3), g.nx, g.ny, g.nz, x, y, z)) kinetic_point(g, a, x, y, z);
}

// ------------------------------------------------------------------------------------------------ stage
// Per face: I from the edge products (backend.cu:225-274, same association as the reference's
// accumulation), then the RK3 stage (backend.cu:279-330). An edge index outside the box counts 0.
template <int COMP>
__device__ void stage_face(const Geo& g, const StageArgs& a, int x, int y, int z) {
    const int NX = g.nx, NY = g.ny, NZ = g.nz;
    const int sx = NX + (COMP == 0), sy = NY + (COMP == 1), sz = NZ + (COMP == 2);
    const int i = x + sx * (y + sy * z);
    double* out = a.out[COMP];
    if ((COMP == 0 && (x == 0 || x == NX)) || (COMP == 1 && (y == 0 || y == NY)) || (COMP == 2 && (z == 0 || z == NZ))) {
        out[i] = 0.0;
        return;
    }
    const Ix ix{NX, NY, NZ};
    const double* const* E = a.edge;
    const double* K = E[E_K];
    // z-edge (ex, ey, z), x-edge (x, ey, ez), y-edge (ex, y, ez); 0 outside
    auto ez_ = [&](int k, int ex, int ey, int zz) {
        return (ey >= 0 && ey <= NY && ex >= 0 && ex <= NX) ? E[k][ex + (NX + 1) * (ey + (NY + 1) * zz)] : 0.0;
    };
    auto ex_ = [&](int k, int xx, int ey, int ez) {
        return (ey >= 0 && ey <= NY && ez >= 0 && ez <= NZ) ? E[k][xx + NX * (ey + (NY + 1) * ez)] : 0.0;
    };
    auto ey_ = [&](int k, int ex, int yy, int ez) {
        return (ex >= 0 && ex <= NX && ez >= 0 && ez <= NZ) ? E[k][ex + (NX + 1) * (yy + NY * ez)] : 0.0;
    };
    double Iv;
    if (COMP == 0) {
        double L = 0.5 * (ez_(E_ZX1, x, y + 1, z) + ez_(E_ZX1, x, y, z)) - 0.5 * (ey_(E_YX1, x, y, z + 1) + ey_(E_YX1, x, y, z));
        if (a.o4) {
            const double c = tb(g, X, T_C3, x);
            L += c * (ez_(E_ZX3, x, y + 2, z) + ez_(E_ZX3, x, y - 1, z));
            L -= c * (ey_(E_YX3, x, y, z + 2) + ey_(E_YX3, x, y, z - 1));
        }
        Iv = L - (K[ix.c(x, y, z)] - K[ix.c(x - 1, y, z)]) * tb(g, X, T_RSP, x);
    } else if (COMP == 1) {
        double L = -0.5 * (ez_(E_ZY1, x + 1, y, z) + ez_(E_ZY1, x, y, z)) + 0.5 * (ex_(E_XY1, x, y, z + 1) + ex_(E_XY1, x, y, z));
        if (a.o4) {
            const double c = tb(g, Y, T_C3, y);
            L -= c * (ez_(E_ZY3, x + 2, y, z) + ez_(E_ZY3, x - 1, y, z));
            L += c * (ex_(E_XY3, x, y, z + 2) + ex_(E_XY3, x, y, z - 1));
        }
        Iv = L - (K[ix.c(x, y, z)] - K[ix.c(x, y - 1, z)]) * tb(g, Y, T_RSP, y);
    } else {
        double L = -0.5 * (ex_(E_XZ1, x, y + 1, z) + ex_(E_XZ1, x, y, z)) + 0.5 * (ey_(E_YZ1, x + 1, y, z) + ey_(E_YZ1, x, y, z));
        if (a.o4) {
            const double c = tb(g, Z, T_C3, z);
            L -= c * (ex_(E_XZ3, x, y + 2, z) + ex_(E_XZ3, x, y - 1, z));
            L += c * (ey_(E_YZ3, x + 2, y, z) + ey_(E_YZ3, x - 1, y, z));
        }
        Iv = L - (K[ix.c(x, y, z)] - K[ix.c(x, y, z - 1)]) * tb(g, Z, T_RSP, z);
    }
    const double* m = a.m[COMP];
    const double c = m[i];
    // the face velocity enters only the wall and lid fluxes of faces next to a tangential wall
    const bool by_wall = (COMP != 0 && (x == 0 || x == sx - 1)) || (COMP != 1 && (y == 0 || y == sy - 1))
                      || (COMP != 2 && (z == 0 || z == sz - 1));
    const double wall = by_wall ? a.u[COMP][i] : 0.0;
    const int dy = sx, dz = sx * sy;
    double lap = 0.0;
    if (COMP == 0) {
        lap += ((m[i + 1] - c) * tb(g, X, T_RW, x) - (c - m[i - 1]) * tb(g, X, T_RW, x - 1)) * tb(g, X, T_RSP, x);
    } else {
        double east = x < sx - 1 ? (m[i + 1] - c) * tb(g, X, T_RSP, x + 1) : 0.0;
        double west = x > 0 ? (c - m[i - 1]) * tb(g, X, T_RSP, x) : 0.0;
        if (x == 0) west = wall * tb(g, X, T_RSP, 0);
        if (x == sx - 1) east = -wall * tb(g, X, T_RSP, NX);
        lap += (east - west) * tb(g, X, T_RW, x);
    }
    if (COMP == 1) {
        lap += ((m[i + dy] - c) * tb(g, Y, T_RW, y) - (c - m[i - dy]) * tb(g, Y, T_RW, y - 1)) * tb(g, Y, T_RSP, y);
    } else {
        double north = y < sy - 1 ? (m[i + dy] - c) * tb(g, Y, T_RSP, y + 1) : 0.0;
        double south = y > 0 ? (c - m[i - dy]) * tb(g, Y, T_RSP, y) : 0.0;
        if (y == 0) south = wall * tb(g, Y, T_RSP, 0);
        if (y == sy - 1) {
            const double U = COMP == 0 ? a.lid * lid_at_u_face(g, x, z) : 0.0;
            north = (U - wall) * tb(g, Y, T_RSP, NY);
        }
        lap += (north - south) * tb(g, Y, T_RW, y);
    }
    if (COMP == 2) {
        lap += ((m[i + dz] - c) * tb(g, Z, T_RW, z) - (c - m[i - dz]) * tb(g, Z, T_RW, z - 1)) * tb(g, Z, T_RSP, z);
    } else {
        double top = z < sz - 1 ? (m[i + dz] - c) * tb(g, Z, T_RSP, z + 1) : 0.0;
        double bottom = z > 0 ? (c - m[i - dz]) * tb(g, Z, T_RSP, z) : 0.0;
        if (z == 0) bottom = wall * tb(g, Z, T_RSP, 0);
        if (z == sz - 1) top = -wall * tb(g, Z, T_RSP, NZ);
        lap += (top - bottom) * tb(g, Z, T_RW, z);
    }
    const double rhs = a.nu * lap + 0.0;
    double q = c + a.dt * (rhs + Iv);
    if (a.has_base) q = a.a0 * a.m0[COMP][i] + a.b * q;
    out[i] = q;
}

// b0, b1, b2: block counts of the three components, each a multiple of 8 so every segment starts on XCD 0
__global__ __launch_bounds__(TPB) void k_stage(Geo g, StageArgs a, int b0, int b1, int b2) {
    int blk = blockIdx.x, x, y, z;
    if (blk < b0) {
        if (tile_point(xcd_block(blk, b0), g.nx + 1, g.ny, g.nz, x, y, z)) stage_face<0>(g, a, x, y, z);
    } else if ((blk -= b0) < b1) {
        if (tile_point(xcd_block(blk, b1), g.nx, g.ny + 1, g.nz, x, y, z)) stage_face<1>(g, a, x, y, z);
    } else if (tile_point(xcd_block(blk - b1, b2), g.nx, g.ny, g.nz + 1, x, y, z)) {
        stage_face<2>(g, a, x, y, z);
    }
}

// ------------------------------------------------------------------------------------------------ projection
// backend.cu:334-342 with the rhs^2 partial sum fused (z-march, w(z+1) carried to the next layer)
__global__ __launch_bounds__(TPB) void k_divergence(Geo g, const double* __restrict__ u, const double* __restrict__ v,
                                                    const double* __restrict__ w, double* __restrict__ rhs,
                                                    double* __restrict__ partial, March M) {
    const Ix ix{g.nx, g.ny, g.nz};
    double acc = 0.0;
    int x, y, z0, z1;
    march_point(M, g.nz, x, y, z0, z1);
    if (x < g.nx && y < g.ny && z0 < z1) {
        const double dx = tb(g, X, T_W, x), dy = tb(g, Y, T_W, y);
        double w0 = w[ix.w(x, y, z0)];
        for (int z = z0; z < z1; z++) {
            const double dz = tb(g, Z, T_W, z), w1 = w[ix.w(x, y, z + 1)];
            const double r = (u[ix.u(x + 1, y, z)] - u[ix.u(x, y, z)]) * (dy * dz)
                           + (v[ix.v(x, y + 1, z)] - v[ix.v(x, y, z)]) * (dx * dz)
                           + (w1 - w0) * (dx * dy);
            rhs[ix.c(x, y, z)] = r;
            acc += r * r;
            w0 = w1;
        }
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

// backend.cu:345-359 with the r^2 partial sum fused (z-march, phi(z-1), phi(z), phi(z+1) in registers)
__global__ __launch_bounds__(TPB) void k_residual(Geo g, const double* __restrict__ phi,
                                                  const double* __restrict__ rhs, double* __restrict__ r,
                                                  double* __restrict__ partial, March M) {
    const int NX = g.nx, NY = g.ny, NZ = g.nz;
    const int sy = NX, sz = NX * NY;
    double acc = 0.0;
    int x, y, z0, z1;
    march_point(M, NZ, x, y, z0, z1);
    if (x < NX && y < NY && z0 < z1) {
        const double dx = tb(g, X, T_W, x), dy = tb(g, Y, T_W, y);
        const double rx1 = x + 1 < NX ? tb(g, X, T_RSP, x + 1) : 0.0, rx0 = tb(g, X, T_RSP, x);
        const double ry1 = y + 1 < NY ? tb(g, Y, T_RSP, y + 1) : 0.0, ry0 = tb(g, Y, T_RSP, y);
        int i = x + NX * (y + NY * z0);
        double prev = z0 > 0 ? phi[i - sz] : 0.0, c = phi[i];
        for (int z = z0; z < z1; z++, i += sz) {
            const double next = z + 1 < NZ ? phi[i + sz] : 0.0;
            const double dz = tb(g, Z, T_W, z);
            double s = 0.0;
            if (x + 1 < NX) s += (dy * dz) * (phi[i + 1] - c) * rx1;
            if (x > 0) s -= (dy * dz) * (c - phi[i - 1]) * rx0;
            if (y + 1 < NY) s += (dx * dz) * (phi[i + sy] - c) * ry1;
            if (y > 0) s -= (dx * dz) * (c - phi[i - sy]) * ry0;
            if (z + 1 < NZ) s += (dx * dy) * (next - c) * tb(g, Z, T_RSP, z + 1);
            if (z > 0) s -= (dx * dy) * (c - prev) * tb(g, Z, T_RSP, z);
            const double ri = rhs[i] - s;
            r[i] = ri;
            acc += ri * ri;
            prev = c;
            c = next;
        }
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
template <> struct XT<0> { using T = double; using Acc = dbl4; };
template <> struct XT<1> { using T = float; using Acc = flt4; };
template <> struct XT<2> { using T = float; using Acc = flt4; };
constexpr int XBK = 32;                 // contraction indices staged per LDS tile

__device__ __forceinline__ dbl4 mfma(double a, double b, dbl4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
__device__ __forceinline__ flt4 mfma(float a, float b, flt4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
__device__ __forceinline__ flt4 mfma(bf8 a, bf8 b, flt4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }

// Accumulate one staged chunk of XBK contraction indices into a wave's 2 x 2 fragment block: data from
// tile columns dc .. dc+31, matrix fragments of row tiles mt0, mt0 + 1 (packed, contraction padded to npad).
template <bool XL, int MODE, class T, class Acc>
__device__ __forceinline__ void mma_chunk(Acc (&acc)[2][2], T (*tile)[64 + 2], int dc, int lane, const void* mpk,
                                          const void* mpk_lo, int mt0, const bool (&jv)[2], int npad, int k0) {
    constexpr int BK = XBK;
    if constexpr (MODE < 2) {
        const T* mp = (const T*)mpk;
        const int kts = npad / 4;
        T mf[2][BK / 4];
#pragma unroll
        for (int jf = 0; jf < 2; jf++)
#pragma unroll
            for (int ks = 0; ks < BK / 4; ks++)
                mf[jf][ks] = jv[jf] ? mp[((i64)(mt0 + jf) * kts + k0 / 4 + ks) * 64 + lane] : (T)0;
#pragma unroll
        for (int ks = 0; ks < BK / 4; ks++) {
            const T d[2] = {tile[ks * 4 + lane / 16][dc + lane % 16], tile[ks * 4 + lane / 16][dc + 16 + lane % 16]};
#pragma unroll
            for (int jf = 0; jf < 2; jf++) {
                if (!jv[jf]) continue;
#pragma unroll
                for (int f = 0; f < 2; f++)
                    acc[f][jf] = XL ? mfma(d[f], mf[jf][ks], acc[f][jf]) : mfma(mf[jf][ks], d[f], acc[f][jf]);
            }
        }
    } else {
        const bf8* mhi = (const bf8*)mpk;
        const bf8* mlo = (const bf8*)mpk_lo;
        const int kts = npad / 32, kt = k0 / 32;
        bf8 mh[2], ml[2];
#pragma unroll
        for (int jf = 0; jf < 2; jf++) {
            const i64 at = ((i64)(mt0 + jf) * kts + kt) * 64 + lane;
            if (jv[jf]) { mh[jf] = mhi[at]; ml[jf] = mlo[at]; }
        }
        bf8 dh[2], dl[2];
#pragma unroll
        for (int f = 0; f < 2; f++)
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const float x = tile[8 * (lane / 16) + q][dc + f * 16 + lane % 16];
                const __bf16 h = (__bf16)x;
                dh[f][q] = h;
                dl[f][q] = (__bf16)(x - (float)h);
            }
#pragma unroll
        for (int jf = 0; jf < 2; jf++) {
            if (!jv[jf]) continue;
#pragma unroll
            for (int f = 0; f < 2; f++) {
                if (XL) {
                    acc[f][jf] = mfma(dl[f], mh[jf], acc[f][jf]);
                    acc[f][jf] = mfma(dh[f], ml[jf], acc[f][jf]);
                    acc[f][jf] = mfma(dh[f], mh[jf], acc[f][jf]);
                } else {
                    acc[f][jf] = mfma(ml[jf], dh[f], acc[f][jf]);
                    acc[f][jf] = mfma(mh[jf], dl[f], acc[f][jf]);
                    acc[f][jf] = mfma(mh[jf], dh[f], acc[f][jf]);
                }
            }
        }
    }
}
// result row within the 16 x 16 tile of register r: FP64 interleaves the lane groups (4 r + l/16), the
// FP32-accumulating instructions stack them (4 (l/16) + r)
template <int MODE> __device__ __forceinline__ int result_row(int r, int lane) {
    return MODE == 0 ? 4 * r + lane / 16 : 4 * (lane / 16) + r;
}

template <bool XL, int MODE, class TIN, class TOUT>
__global__ __launch_bounds__(256) void k_axis(AxisArgs a, i64 ncols, int npad) {
    using T = typename XT<MODE>::T;
    using Acc = typename XT<MODE>::Acc;
    constexpr int BK = XBK, PER = BK * 64 / 256;          // staged values per thread and tile
    __shared__ T tile[BK][64 + 2];
    const TIN* __restrict__ in = (const TIN*)a.in;
    TOUT* __restrict__ out = (TOUT*)a.out;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // 1-D grid: logical block = (column tile, j tile) with the j tiles of one column tile adjacent and on
    // one XCD, so the input tile they share is fetched once into that L2
    const int nby = (npad + 63) / 64;
    const int L = xcd_block(blockIdx.x, gridDim.x), by = L % nby;
    const i64 c0 = (i64)(L / nby) * 64;
    if (c0 >= ncols) return;
    const int wr = wave >> 1, wc = wave & 1;
    const int n = a.n;
    const int jt0 = (by * 64 + wc * 32) / 16;
    const bool jv[2] = {16 * jt0 < npad, 16 * (jt0 + 1) < npad};
    Acc acc[2][2];
    for (int f = 0; f < 2; f++)
        for (int jf = 0; jf < 2; jf++)
            for (int r = 0; r < 4; r++) acc[f][jf][r] = 0;

    // Staging map. XL: thread -> row c0 + tid/4, PER consecutive i from (tid%4)*PER.
    // [A][n][B]: thread -> 4 consecutive columns (tid%16)*4, rows tid/16 and tid/16 + 16.
    const int xcol = tid >> 2, xig = (tid & 3) * PER;
    const i64 xrow = c0 + xcol;
    const int cg = (tid & 15) * 4, ii0 = tid >> 4;
    i64 cbase[4];
    bool cok[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const i64 col = c0 + cg + q;
        cok[q] = !XL && col < ncols;
        const i64 aa = XL ? 0 : col / a.B, bb = col - aa * a.B;
        cbase[q] = aa * n * a.B + bb;
    }
    T pre[PER];
    auto load = [&](int i0) {
        if (XL) {
#pragma unroll
            for (int q = 0; q < PER; q++) {
                const int i = i0 + xig + q;
                pre[q] = (xrow < ncols && i < n) ? (T)(in[xrow * n + i] * a.in_scale) : (T)0;
            }
        } else {
#pragma unroll
            for (int h = 0; h < 2; h++)
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int i = i0 + ii0 + 16 * h;
                    pre[4 * h + q] = (cok[q] && i < n) ? (T)(in[cbase[q] + (i64)i * a.B] * a.in_scale) : (T)0;
                }
        }
    };
    auto store = [&]() {
        if (XL) {
#pragma unroll
            for (int q = 0; q < PER; q++) tile[xig + q][xcol] = pre[q];
        } else {
#pragma unroll
            for (int h = 0; h < 2; h++)
#pragma unroll
                for (int q = 0; q < 4; q++) tile[ii0 + 16 * h][cg + q] = pre[4 * h + q];
        }
    };

    load(0);
    for (int i0 = 0; i0 < npad; i0 += BK) {
        __syncthreads();
        store();
        __syncthreads();
        if (i0 + BK < npad) load(i0 + BK);                  // next tile in flight during the MFMAs
        mma_chunk<XL, MODE, T, Acc>(acc, tile, wr * 32, lane, a.mpk, a.mpk_lo, jt0, jv, npad, i0);
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
                const int rr = result_row<MODE>(r, lane);
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

// ------------------------------------------------------------------------------------------------ symmetric FDM transform
// On a grid symmetric about its centre (the tanh stretching is) every eigenvector is even or odd, so with the
// spectral modes ordered [even | odd] one axis transform is two contractions of half length h = n/2:
//   forward:  y_e = Me (r_lo + rev r_hi),  y_o = Mo (r_lo - rev r_hi)          (Me, Mo: h x h)
//   backward: e = Me^T y_e, o = Mo^T y_o,  phi_lo = e + o,  rev phi_hi = e - o
// which halves the MFMA work. The forward kernel folds while staging; the backward kernel computes the
// even and odd halves of the same rows in one block (waves 0-1 even, 2-3 odd) and unfolds in the epilogue.
template <bool XL, int MODE, class TIN, class TOUT>
__global__ __launch_bounds__(256) void k_axis_symf(AxisArgs a, i64 ncols, int hpad) {
    using T = typename XT<MODE>::T;
    using Acc = typename XT<MODE>::Acc;
    constexpr int BK = XBK, PER = BK * 64 / 256;
    __shared__ T tile[BK][64 + 2];
    const TIN* __restrict__ in = (const TIN*)a.in;
    TOUT* __restrict__ out = (TOUT*)a.out;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int n = a.n, h = n / 2, nbh = (hpad + 63) / 64;
    const int L = xcd_block(blockIdx.x, gridDim.x), by = L % (2 * nbh), odd = by / nbh, byh = by % nbh;
    const i64 c0 = (i64)(L / (2 * nbh)) * 64;
    if (c0 >= ncols) return;
    const int wr = wave >> 1, wc = wave & 1;
    const int jt0 = (byh * 64 + wc * 32) / 16;
    const bool jv[2] = {16 * jt0 < hpad, 16 * (jt0 + 1) < hpad};
    const double sgn = odd ? -1.0 : 1.0;
    Acc acc[2][2];
    for (int f = 0; f < 2; f++)
        for (int jf = 0; jf < 2; jf++)
            for (int r = 0; r < 4; r++) acc[f][jf][r] = 0;
    const int xcol = tid >> 2, xig = (tid & 3) * PER;
    const i64 xrow = c0 + xcol;
    const int cg = (tid & 15) * 4, ii0 = tid >> 4;
    i64 cbase[4];
    bool cok[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const i64 col = c0 + cg + q;
        cok[q] = !XL && col < ncols;
        const i64 aa = XL ? 0 : col / a.B, bb = col - aa * a.B;
        cbase[q] = aa * n * a.B + bb;
    }
    T pre[PER];
    auto load = [&](int i0) {
        if (XL) {
#pragma unroll
            for (int q = 0; q < PER; q++) {
                const int i = i0 + xig + q;
                T val = 0;
                if (xrow < ncols && i < h)
                    val = (T)(in[xrow * n + i] * a.in_scale + sgn * (in[xrow * n + (n - 1 - i)] * a.in_scale));
                pre[q] = val;
            }
        } else {
#pragma unroll
            for (int hh = 0; hh < 2; hh++)
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int i = i0 + ii0 + 16 * hh;
                    T val = 0;
                    if (cok[q] && i < h)
                        val = (T)(in[cbase[q] + (i64)i * a.B] * a.in_scale
                                  + sgn * (in[cbase[q] + (i64)(n - 1 - i) * a.B] * a.in_scale));
                    pre[4 * hh + q] = val;
                }
        }
    };
    auto store = [&]() {
        if (XL) {
#pragma unroll
            for (int q = 0; q < PER; q++) tile[xig + q][xcol] = pre[q];
        } else {
#pragma unroll
            for (int hh = 0; hh < 2; hh++)
#pragma unroll
                for (int q = 0; q < 4; q++) tile[ii0 + 16 * hh][cg + q] = pre[4 * hh + q];
        }
    };
    const void* mp = odd ? a.mpk2 : a.mpk;
    const void* mpl = odd ? a.mpk2_lo : a.mpk_lo;
    load(0);
    for (int i0 = 0; i0 < hpad; i0 += BK) {
        __syncthreads();
        store();
        __syncthreads();
        if (i0 + BK < hpad) load(i0 + BK);
        mma_chunk<XL, MODE, T, Acc>(acc, tile, wr * 32, lane, mp, mpl, jt0, jv, hpad, i0);
    }
    const T* il = (const T*)a.il;
#pragma unroll
    for (int f = 0; f < 2; f++)
#pragma unroll
        for (int jf = 0; jf < 2; jf++) {
            if (!jv[jf]) continue;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int rr = result_row<MODE>(r, lane);
                i64 idx;
                if (XL) {
                    const i64 row = c0 + wr * 32 + f * 16 + rr;
                    const int j = 16 * (jt0 + jf) + lane % 16;
                    if (row >= ncols || j >= h) continue;
                    idx = row * n + odd * h + j;
                } else {
                    const int j = 16 * (jt0 + jf) + rr;
                    const i64 col = c0 + wr * 32 + f * 16 + lane % 16;
                    if (col >= ncols || j >= h) continue;
                    const i64 aa = col / a.B, bb = col - aa * a.B;
                    idx = (aa * n + odd * h + j) * a.B + bb;
                }
                const T val = acc[f][jf][r];
                if (a.epi == 1) out[idx] = (TOUT)(val * il[idx]);
                else out[idx] = (TOUT)val;
            }
        }
}

template <bool XL, int MODE, class TIN, class TOUT>
__global__ __launch_bounds__(256) void k_axis_symb(AxisArgs a, i64 ncols, int hpad) {
    using T = typename XT<MODE>::T;
    using Acc = typename XT<MODE>::Acc;
    constexpr int BK = XBK, PER = BK * 64 / 256;
    __shared__ T tile[2][BK][64 + 2];               // [even | odd] spectral chunk; reused for the results
    const TIN* __restrict__ in = (const TIN*)a.in;
    TOUT* __restrict__ out = (TOUT*)a.out;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int n = a.n, h = n / 2, nbi = (hpad + 31) / 32;
    const int L = xcd_block(blockIdx.x, gridDim.x), by = L % nbi;
    const i64 c0 = (i64)(L / nbi) * 64;
    if (c0 >= ncols) return;
    const int part = wave >> 1, dsub = wave & 1;   // part 0 even, 1 odd; data columns dsub*32 .. +31
    const int mt0 = by * 2;                          // output rows i = 32 by .. 32 by + 31
    const bool jv[2] = {16 * mt0 < hpad, 16 * (mt0 + 1) < hpad};
    Acc acc[2][2];
    for (int f = 0; f < 2; f++)
        for (int jf = 0; jf < 2; jf++)
            for (int r = 0; r < 4; r++) acc[f][jf][r] = 0;
    const int xcol = tid >> 2, xig = (tid & 3) * PER;
    const i64 xrow = c0 + xcol;
    const int cg = (tid & 15) * 4, ii0 = tid >> 4;
    i64 cbase[4];
    bool cok[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const i64 col = c0 + cg + q;
        cok[q] = !XL && col < ncols;
        const i64 aa = XL ? 0 : col / a.B, bb = col - aa * a.B;
        cbase[q] = aa * n * a.B + bb;
    }
    T pre[2][PER];
    auto load = [&](int k0) {
#pragma unroll
        for (int p = 0; p < 2; p++) {
            if (XL) {
#pragma unroll
                for (int q = 0; q < PER; q++) {
                    const int k = k0 + xig + q;
                    pre[p][q] = (xrow < ncols && k < h) ? (T)in[xrow * n + p * h + k] : (T)0;
                }
            } else {
#pragma unroll
                for (int hh = 0; hh < 2; hh++)
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        const int k = k0 + ii0 + 16 * hh;
                        pre[p][4 * hh + q] = (cok[q] && k < h) ? (T)in[cbase[q] + (i64)(p * h + k) * a.B] : (T)0;
                    }
            }
        }
    };
    auto store = [&]() {
#pragma unroll
        for (int p = 0; p < 2; p++) {
            if (XL) {
#pragma unroll
                for (int q = 0; q < PER; q++) tile[p][xig + q][xcol] = pre[p][q];
            } else {
#pragma unroll
                for (int hh = 0; hh < 2; hh++)
#pragma unroll
                    for (int q = 0; q < 4; q++) tile[p][ii0 + 16 * hh][cg + q] = pre[p][4 * hh + q];
            }
        }
    };
    const void* mp = part ? a.mpk2 : a.mpk;
    const void* mpl = part ? a.mpk2_lo : a.mpk_lo;
    load(0);
    for (int k0 = 0; k0 < hpad; k0 += BK) {
        __syncthreads();
        store();
        __syncthreads();
        if (k0 + BK < hpad) load(k0 + BK);
        mma_chunk<XL, MODE, T, Acc>(acc, tile[part], dsub * 32, lane, mp, mpl, mt0, jv, hpad, k0);
    }
    // exchange: res[part][row i within 32][data column within 64], aliased onto the staging tiles
    __syncthreads();
    T (*res)[32][64 + 2] = (T (*)[32][64 + 2])&tile[0][0][0];
#pragma unroll
    for (int f = 0; f < 2; f++)
#pragma unroll
        for (int jf = 0; jf < 2; jf++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int rr = result_row<MODE>(r, lane);
                if (XL) res[part][jf * 16 + lane % 16][dsub * 32 + f * 16 + rr] = acc[f][jf][r];
                else res[part][jf * 16 + rr][dsub * 32 + f * 16 + lane % 16] = acc[f][jf][r];
            }
    __syncthreads();
    for (int e = tid; e < 32 * 64; e += 256) {
        const int il_ = XL ? e % 32 : e / 64, cl = XL ? e / 32 : e % 64;
        const int i = by * 32 + il_;
        const i64 col = c0 + cl;
        if (i >= h || col >= ncols) continue;
        const T ev = res[0][il_][cl], ov = res[1][il_][cl];
        const T vlo = ev + ov, vhi = ev - ov;
        i64 lo, hi;
        if (XL) {
            lo = col * n + i;
            hi = col * n + (n - 1 - i);
        } else {
            const i64 aa = col / a.B, bb = col - aa * a.B;
            lo = (aa * n + i) * a.B + bb;
            hi = (aa * n + (n - 1 - i)) * a.B + bb;
        }
        if (a.epi == 2) {
            out[lo] = out[lo] + (TOUT)((double)vlo * a.out_scale);
            out[hi] = out[hi] + (TOUT)((double)vhi * a.out_scale);
        } else {
            out[lo] = (TOUT)vlo;
            out[hi] = (TOUT)vhi;
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

// backend.cu:361-371: out = m - grad(phi), walls 0, all three components in one z-march over the
// (nx+1) x (ny+1) plane; phi(z-1) of the column is carried in a register.
struct Face3 {
    const double* m[3];
    double* out[3];
};
__global__ __launch_bounds__(TPB) void k_gradient(Geo g, Face3 f, const double* __restrict__ phi, March M) {
    const int NX = g.nx, NY = g.ny, NZ = g.nz;
    const Ix ix{NX, NY, NZ};
    int x, y, z0, z1;
    march_point(M, NZ + 1, x, y, z0, z1);
    if (x > NX || y > NY) return;
    const bool cell = x < NX && y < NY;
    double pprev = (cell && z0 > 0 && z0 - 1 < NZ) ? phi[ix.c(x, y, z0 - 1)] : 0.0;
    for (int z = z0; z < z1; z++) {
        const double pc = (cell && z < NZ) ? phi[ix.c(x, y, z)] : 0.0;
        if (y < NY && z < NZ) {                                       // u-face
            const int i = ix.u(x, y, z);
            if (x == 0 || x == NX) f.out[0][i] = 0.0;
            else f.out[0][i] = f.m[0][i] - (pc - phi[ix.c(x - 1, y, z)]) * tb(g, X, T_RSP, x);
        }
        if (x < NX && z < NZ) {                                       // v-face
            const int i = ix.v(x, y, z);
            if (y == 0 || y == NY) f.out[1][i] = 0.0;
            else f.out[1][i] = f.m[1][i] - (pc - phi[ix.c(x, y - 1, z)]) * tb(g, Y, T_RSP, y);
        }
        if (cell) {                                                   // w-face, z = 0 .. NZ
            const int i = ix.w(x, y, z);
            if (z == 0 || z == NZ) f.out[2][i] = 0.0;
            else f.out[2][i] = f.m[2][i] - (pc - pprev) * tb(g, Z, T_RSP, z);
        }
        pprev = pc;
    }
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

// Plain copy (16 B per lane per access): the calibration kernel for HBM byte counters.
__global__ __launch_bounds__(TPB) void k_copy(const double2* __restrict__ src, double2* __restrict__ dst, i64 n) {
    for (i64 i = (i64)blockIdx.x * TPB + threadIdx.x; i < n; i += (i64)gridDim.x * TPB) dst[i] = src[i];
}

// ================================================================================================ wrappers
void reduce_array(const Exec& ex, const double* x, i64 n, int op, double* partial, double* out) {
    const int nb = red_blocks(n);
    run(ex, F_REDUCE, k_reduce_array, dim3(nb), dim3(TPB), x, n, op, partial);
    reduce_final(ex, partial, nb, op, out);
}
void reduce_final(const Exec& ex, const double* partial, int count, int op, double* out) {
    run(ex, F_REDUCE, k_reduce_final, dim3(1), dim3(1024), partial, count, op, out);
}
void copy_probe(const Exec& ex, const double* src, double* dst, i64 n_doubles) {
    const i64 n2 = n_doubles / 2;
    run(ex, F_OUTPUT, k_copy, dim3((unsigned)std::min<i64>(65536, (n2 + TPB - 1) / TPB)), dim3(TPB), (const double2*)src,
        (double2*)dst, n2);
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
    const March M = march_of(g.nx, g.ny, g.nz);
    const int nb = march_blocks(M);
    run(ex, F_CFL, k_cfl, dim3(nb), dim3(TPB), g, u[0], u[1], u[2], partial, M);
    reduce_final(ex, partial, nb, 0, out);
}

i64 edge_array_size(int nx, int ny, int nz, int which) {
    return which == E_K ? (i64)nx * ny * nz : edge_size_of(nx, ny, nz, which);
}

void stage(const Exec& ex, const Geo& g, const StageArgs& a) {
    const int e0 = tiles_of(g.nx + 1, g.ny + 1, g.nz), e1 = tiles_of(g.nx, g.ny + 1, g.nz + 1),
              e2 = tiles_of(g.nx + 1, g.ny, g.nz + 1), e3 = tiles_of(g.nx, g.ny, g.nz);
    run(ex, F_EDGES, k_edges, dim3(e0 + e1 + e2 + e3), dim3(TPB), g, a, e0, e1, e2, e3);
    const int b0 = tiles_of(g.nx + 1, g.ny, g.nz), b1 = tiles_of(g.nx, g.ny + 1, g.nz), b2 = tiles_of(g.nx, g.ny, g.nz + 1);
    run(ex, F_STAGE, k_stage, dim3(b0 + b1 + b2), dim3(TPB), g, a, b0, b1, b2);
}

int divergence(const Exec& ex, const Geo& g, const double* const f[3], double* rhs, double* partial) {
    const March M = march_of(g.nx, g.ny, g.nz);
    const int nb = march_blocks(M);
    run(ex, F_DIVERGENCE, k_divergence, dim3(nb), dim3(TPB), g, f[0], f[1], f[2], rhs, partial, M);
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
    const March M = march_of(g.nx, g.ny, g.nz);
    const int nb = march_blocks(M);
    run(ex, F_RESIDUAL, k_residual, dim3(nb), dim3(TPB), g, phi, rhs, r, partial, M);
    return nb;
}

int pad_for_mode(int n, int mode) { return (n + XBK - 1) / XBK * XBK; }

void axis_transform(const Exec& ex, bool layout_x, int mode, bool in_f64, bool out_f64, const AxisArgs& a,
                    int family) {
    const i64 ncols = layout_x ? a.A : a.A * a.B;
    const dim3 block(256);
    if (a.sym) {
        const int hpad = pad_for_mode(a.n / 2, mode);
        const bool fwd = a.sym == 1;
        const unsigned nb = fwd ? (unsigned)round8(((ncols + 63) / 64) * 2 * ((hpad + 63) / 64))
                                : (unsigned)round8(((ncols + 63) / 64) * ((hpad + 31) / 32));
        const dim3 grid(nb);
#define CAV_SYM(M, TI, TO)                                                                                         \
    if (fwd) {                                                                                                     \
        if (layout_x) run(ex, family, k_axis_symf<true, M, TI, TO>, grid, block, a, ncols, hpad);                 \
        else run(ex, family, k_axis_symf<false, M, TI, TO>, grid, block, a, ncols, hpad);                         \
    } else {                                                                                                       \
        if (layout_x) run(ex, family, k_axis_symb<true, M, TI, TO>, grid, block, a, ncols, hpad);                 \
        else run(ex, family, k_axis_symb<false, M, TI, TO>, grid, block, a, ncols, hpad);                         \
    }
        if (mode == 0) { CAV_SYM(0, double, double) }
        else if (mode == 1) {
            if (in_f64) { CAV_SYM(1, double, float) } else if (out_f64) { CAV_SYM(1, float, double) } else { CAV_SYM(1, float, float) }
        } else {
            if (in_f64) { CAV_SYM(2, double, float) } else if (out_f64) { CAV_SYM(2, float, double) } else { CAV_SYM(2, float, float) }
        }
#undef CAV_SYM
        return;
    }
    const int npad = pad_for_mode(a.n, mode);
    const dim3 grid((unsigned)round8(((ncols + 63) / 64) * ((npad + 63) / 64)));
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
    Face3 f;
    for (int k = 0; k < 3; k++) { f.m[k] = m[k]; f.out[k] = out[k]; }
    const March M = march_of(g.nx + 1, g.ny + 1, g.nz + 1);
    run(ex, F_GRADIENT, k_gradient, dim3(march_blocks(M)), dim3(TPB), g, f, phi, M);
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
