// kernels.hip — gfx950 kernels of the cavity time step and its outputs.
//
// The stencils follow the reference's formulas and order of operations (backend.cu:98-561) in FP64,
// compiled with -ffp-contract=off, with one change: every division by a grid quantity (width, spacing,
// node distance, edge weights, the wall ramp) is a multiplication by a reciprocal or coefficient table
// built once on the host (axis_tables). That moves each result by at most about one ulp per operation
// against the host oracle; the reductions (block/wavefront trees) and the FDM transforms (MFMA
// accumulation order) round differently as well.
#include <climits>
#include <type_traits>

#include "kernels.h"

namespace cav {

const char* family_name(int f) {
    static const char* names[F_COUNT] = {"gauge_reset", "cfl_max", "stage", "divergence", "base_guess",
                                         "residual", "reduce", "xform_x", "xform_y", "xform_z", "finalize_phi",
                                         "subtract_gradient", "fill", "output", "kick", "control"};
    return (f >= 0 && f < F_COUNT) ? names[f] : "?";
}

namespace {

constexpr int TPB = 256;

// Global-memory (address space 1) views of pointers. Pointers read from a descriptor in device memory are
// generic to the compiler; casting them lets it issue global (and, for the descriptors, scalar) loads
// instead of flat ones.
#define GLB __attribute__((address_space(1)))
template <class T> __device__ __forceinline__ GLB T* gp(T* p) { return (GLB T*)p; }
template <class T> __device__ __forceinline__ const GLB T* gp(const T* p) { return (const GLB T*)p; }
// Constant-memory (address space 4) view of an array no kernel writes while it runs (the grid tables): its
// loads at uniform positions may then go through the scalar cache even where the kernel stores to other
// arrays (global loads of possibly clobbered memory must stay vector loads).
#define CNS __attribute__((address_space(4)))
template <class T> __device__ __forceinline__ const CNS T* cnp(const T* p) { return (const CNS T*)p; }
// buffers of a member by role, under the member's current parity (state and stage-1 buffers swap each step)
#define fbuf(M, par, role, k) gp((M).F[face_buffer((role), (par))][k])
#define cbuf(M, rot, role) gp((M).C[cell_buffer((role), (rot))])
// a descriptor's grid view copied field by field from global memory
__device__ __forceinline__ Geo geo_of(const GLB Geo& s) {
    Geo g;
    g.nx = s.nx; g.ny = s.ny; g.nz = s.nz; g.lid_shaped = s.lid_shaped;
    for (int a = 0; a < 3; a++) {
        g.face[a] = s.face[a]; g.centre[a] = s.centre[a]; g.width[a] = s.width[a]; g.spacing[a] = s.spacing[a];
        g.length[a] = s.length[a]; g.tab[a] = s.tab[a];
    }
    return g;
}

__host__ __device__ inline int blocks_for(i64 n, int tpb = TPB) { return (int)((n + tpb - 1) / tpb); }
__host__ __device__ inline int red_blocks(i64 n) {
    const i64 b = (n + TPB - 1) / TPB;
    return (int)(b < 1 ? 1 : b > RED_BLOCKS ? RED_BLOCKS : b);
}
__host__ __device__ inline int round8(i64 n) { return (int)((n + 7) / 8 * 8); }

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

// Arguments of the stencil bodies of one RK stage (built per member from its descriptor)
struct StageArgs {
    const GLB double* u[3];     // velocity the inviscid term and the wall fluxes use
    const GLB double* m[3];     // impulse being advanced
    const GLB double* m0[3];    // gauge-reset impulse (base of stages 2, 3)
    GLB double* out[3];
    double nu, lid, dt, a0, b;
    int has_base, o4;
};

// One axis of the FDM transform: out[.., j, ..] = sum_i M[j][i] in[.., i, ..] with M packed for the mode
// (pack_fragments / pack_bf16_fragments). Rows of length n contiguous on the x axis; else [A][n][B] with B
// the inner stride. epi: 0 store, 1 store * il[idx] (il in the mode's type), 2 out[idx] += acc * out_scale
// (out FP64). sym: 0 dense; 1 symmetric forward, 2 symmetric backward (mpk / mpk2 the even / odd half
// matrices, the spectral index ordered [even | odd]).
struct AxisArgs {
    const void* in;
    void* out;
    const void* mpk;       // packed matrix (mode 2: the bf16 hi parts); symmetric: the even half
    const void* mpk_lo;    // mode 2: the bf16 lo parts
    const void* mpk2;      // symmetric: the odd half
    const void* mpk2_lo;
    const void* il;
    int n;
    i64 A, B;
    double in_scale, out_scale;
    int epi;
    int sym;
};

__device__ __forceinline__ double wd(const Geo& g, int a, int i) { return gp(g.width[a])[i]; }
__device__ __forceinline__ double sp(const Geo& g, int a, int i) { return gp(g.spacing[a])[i]; }

// backend.cu:58-65
__device__ __forceinline__ double lid_edge(double s) {
    double s2 = s * s, s4 = s2 * s2, s8 = s4 * s4, s16 = s8 * s8, a = 1.0 - s16 * s2;
    return a * a;
}
__device__ __forceinline__ double lid_shape(const Geo& g, double fx, double fz) {
    if (!g.lid_shaped) return 1.0;
    return lid_edge(2.0 * fx - 1.0) * lid_edge(2.0 * fz - 1.0);
}
__device__ __forceinline__ double lid_at_cell(const Geo& g, int x, int z) {
    return lid_shape(g, gp(g.centre[X])[x] / g.length[X], gp(g.centre[Z])[z] / g.length[Z]);
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

// ------------------------------------------------------------------------------------------------ batches
// Block -> member of a batched launch. seg: this shape's n + 1 block offsets (device). Every segment is a
// multiple of 8 blocks, so local % 8 == blockIdx.x % 8 and the XCD remap works per member.
struct Where { int m, local, G; };
__device__ __forceinline__ Where locate(const int* __restrict__ seg, int n) {
    const int b = blockIdx.x;
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (seg[mid] <= b) lo = mid;
        else hi = mid;
    }
    return {lo, b - seg[lo], seg[lo + 1] - seg[lo]};
}

// Fold `count` partials in a fixed order with 1024 threads (strided, then the wavefront tree, then the
// 16 wavefront results in sequence); the result is valid in thread 0.
__device__ double fold1024(const GLB double* __restrict__ partial, int count, int sop) {
    __shared__ double part[16];
    double r = init_of(sop);
    for (int i = threadIdx.x; i < count; i += 1024) r = comb(sop, r, partial[i]);
    for (int off = 32; off > 0; off >>= 1) r = comb(sop, r, __shfl_down(r, off, 64));
    __syncthreads();
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = r;
    __syncthreads();
    double t = part[0];
    for (int k = 1; k < 16; k++) t = comb(sop, t, part[k]);
    return t;
}

// ------------------------------------------------------------------------------------------------ gauge reset
// backend.cu:880-882: m0 = r m + (1-r) u per face, phi0 = r phi per cell, grid-stride over the three face arrays
// and the cells one after another. Skipped (at the cost of a few thousand empty blocks) when the previous step's
// last gradient has formed them already (kb_gradient, gz).
constexpr int GAUGE_BLOCKS = 2048;
template <class F>
__global__ __launch_bounds__(TPB) void kb_gauge(const Mem* const* __restrict__ D, const int* __restrict__ seg, int nm) {
    const Where w = locate(seg, nm);
    const GLB Mem& M = *gp(D[w.m]);
    if (gp(M.ctl)->m0_ready) return;
    const int par = gp(M.ctl)->parity;
    const int rot = gp(M.ctl)->rot;
    const int nx = M.g.nx, ny = M.g.ny, nz = M.g.nz;
    const i64 n0 = (i64)(nx + 1) * ny * nz, n1 = (i64)nx * (ny + 1) * nz, n2 = (i64)nx * ny * (nz + 1);
    const i64 e1 = n0 + n1, e2 = e1 + n2, total = e2 + M.n_cell;
    const F r = (F)M.r, omr = (F)M.omr;   // (float) in the FP32 fields (backend.cu:881)
    for (i64 k = (i64)w.local * TPB + threadIdx.x; k < total; k += (i64)w.G * TPB) {
        if (k < e2) {
            const int s = k < n0 ? 0 : k < e1 ? 1 : 2;
            const i64 i = k - (s == 0 ? 0 : s == 1 ? n0 : e1);
            ((GLB F*)fbuf(M, par, FR_M0, s))[i] = r * ((const GLB F*)fbuf(M, par, FR_M, s))[i] + omr * ((const GLB F*)fbuf(M, par, FR_U, s))[i];
        } else {
            const i64 i = k - e2;
            cbuf(M, rot, CR_PHI0)[i] = M.r * cbuf(M, rot, CR_PHI)[i];
        }
    }
}

// Table entry kind at index e of axis a (Geo::tab rows of n_a + 1 entries)
__device__ __forceinline__ int axis_n(const Geo& g, int a) { return a == 0 ? g.nx : a == 1 ? g.ny : g.nz; }
__device__ __forceinline__ double tb(const Geo& g, int a, int kind, int e) {
    return cnp(g.tab[a])[kind * (axis_n(g, a) + 1) + e];
}
// Workgroups are dealt round-robin to the 8 XCDs (block b runs on XCD b % 8), each with its own L2.
// xcd_block maps a hardware block id to a logical one so that XCD k gets the contiguous logical range
// [k G/8, (k+1) G/8): neighbouring rows and planes of a stencil then share an L2. G must be a multiple of 8.
__device__ __forceinline__ int xcd_block(int b, int G) { return (b % 8) * (G / 8) + b / 8; }

// z-march: every thread owns one (x, y) column of a px x py plane and walks a chunk of z, keeping the
// z-neighbours it needs in registers, so each plane of an array is read from memory once (x/y neighbours
// come from the same plane, just read by the neighbouring threads). Blocks cover 32 x 8 columns; the
// logical block order is (z chunk, tile y, tile x) and is dealt to the XCDs in contiguous ranges.
#ifndef MARCH_TX
#define MARCH_TX 32
#define MARCH_TY 8
#endif
constexpr int MTX = MARCH_TX, MTY = MARCH_TY;
// the stage kernel's built-in column tile (its halo bookkeeping is written for 32 x 8 and 16 x 16)
constexpr int STX = 32, STY = 8;
#ifndef MARCH_TARGET
#define MARCH_TARGET 1792
#endif
struct March {
    int tx, ty, chunks, zc;    // tiles along x and y, number of z chunks, chunk length
    int tw, th;                // tile width and height (columns)
};
// target: blocks the march aims at, whole rounds of resident 4-wave blocks on 256 CUs (MARCH_TARGET: seven
// waves per SIMD, the residual's occupancy: one round of it); minz: shortest z chunk; tw x th: the tile
__host__ __device__ inline March march_of(int px, int py, int pz, int target = MARCH_TARGET, int minz = 8,
                                          int tw = MTX, int th = MTY) {
    March m;
    m.tw = tw;
    m.th = th;
    m.tx = (px + tw - 1) / tw;
    m.ty = (py + th - 1) / th;
    const int tiles = m.tx * m.ty;
    const int by_tiles = target / (tiles > 1 ? tiles : 1), by_z = pz / minz > 1 ? pz / minz : 1;
    int chunks = by_tiles < by_z ? by_tiles : by_z;
    if (chunks < 1) chunks = 1;
    m.zc = (pz + chunks - 1) / chunks;
    m.chunks = (pz + m.zc - 1) / m.zc;
    return m;
}
__host__ __device__ inline int march_blocks(const March& m) { return round8((i64)m.tx * m.ty * m.chunks); }
// The stage kernel runs three waves per SIMD (768 resident blocks): one whole round, so the z chunks are as
// long as the round allows and the levels every chunk marches early (its halo) are few (86-level chunks at
// 256^3)
constexpr int STAGE_MARCH_TARGET = 768;
// the marches of the stage and gradient kernels with a member's tuned choices (0: the defaults)
__host__ __device__ inline March stage_march_of(int nx, int ny, int nz, int tile, int target, int minz) {
    return march_of(nx, ny, nz, target > 0 ? target : STAGE_MARCH_TARGET, minz > 0 ? minz : 8, tile ? 16 : STX,
                    tile ? 16 : STY);
}
// (the gradient's column tile: 0 32 x 8, 1 64 x 4, 2 128 x 2, 3 256 x 1; wider tiles give each wavefront whole
// 512-byte row pieces of the face and cell arrays it reads and writes, and a block longer runs of a row)
__host__ __device__ inline March grad_march_of(int nx, int ny, int nz, int target, int minz, int tile) {
    const int tw = tile == 3 ? 256 : tile == 2 ? 128 : tile == 1 ? 64 : MTX;
    return march_of(nx, ny, nz + 1, target > 0 ? target : MARCH_TARGET, minz > 0 ? minz : 8, tw, TPB / tw);
}
// this thread's column (x, y) and z range [z0, z1) of a plane extent pz for logical block `local` of G;
// empty range for padding blocks
__device__ __forceinline__ void march_point(const March& m, int pz, int local, int G, int& x, int& y, int& z0,
                                            int& z1) {
    const int L = xcd_block(local, G), tiles = m.tx * m.ty;
    const int chunk = L / tiles, t = L - chunk * tiles;
    x = (t % m.tx) * MTX + (threadIdx.x & (MTX - 1));
    y = (t / m.tx) * MTY + threadIdx.x / MTX;
    z0 = chunk * m.zc;
    z1 = chunk < m.chunks ? min(pz, z0 + m.zc) : z0;
}

// the same for a march whose tile (m.tw x m.th, m.tw a power of two, TPB columns) is chosen at run time
__device__ __forceinline__ void march_point_tiled(const March& m, int pz, int local, int G, int& x, int& y, int& z0,
                                                  int& z1) {
    const int L = xcd_block(local, G), tiles = m.tx * m.ty;
    const int chunk = L / tiles, t = L - chunk * tiles, lg = __builtin_ctz((unsigned)m.tw);
    x = (t % m.tx) * m.tw + (threadIdx.x & (m.tw - 1));
    y = (t / m.tx) * m.th + (threadIdx.x >> lg);
    z0 = chunk * m.zc;
    z1 = chunk < m.chunks ? min(pz, z0 + m.zc) : z0;
}

// the tile origin (x0, y0) and z range of logical block `local` of G (uniform over the block)
__device__ __forceinline__ void march_tile(const March& m, int pz, int local, int G, int& x0, int& y0, int& z0,
                                           int& z1) {
    const int L = xcd_block(local, G), tiles = m.tx * m.ty;
    const int chunk = L / tiles, t = L - chunk * tiles;
    x0 = (t % m.tx) * m.tw;
    y0 = (t / m.tx) * m.th;
    z0 = chunk * m.zc;
    z1 = chunk < m.chunks ? min(pz, z0 + m.zc) : z0;
}

// a / (grid quantity): the FP64 fields multiply by the table's reciprocal t; the FP32 fields divide by the
// table's value, which there holds the grid quantity itself (the reference's float division, backend.cu:102-330)
template <class F> __device__ __forceinline__ F dv(F a, F t) {
    if constexpr (sizeof(F) == 8) return a * t;
    else return a / t;
}

// ------------------------------------------------------------------------------------------------ CFL and dt
// backend.cu:465-472 per cell, reduced with max (NaN wins); partials into region P_MAIN. Skipped when the
// previous step's last gradient has formed the rate of the state already (Ctl::cfl_ready).
template <class F>
__global__ __launch_bounds__(TPB) void kb_cfl(const Mem* const* __restrict__ D, const int* __restrict__ seg, int nm) {
    const Where w = locate(seg, nm);
    const GLB Mem& M = *gp(D[w.m]);
    if (gp(M.ctl)->cfl_ready) return;
    const int par = gp(M.ctl)->parity;
    const Geo g = geo_of(M.g);
    const GLB F* u = (const GLB F*)fbuf(M, par, FR_U, 0), *v = (const GLB F*)fbuf(M, par, FR_U, 1),
               *wv = (const GLB F*)fbuf(M, par, FR_U, 2);
    const Ix ix{g.nx, g.ny, g.nz};
    const March Mh = march_of(g.nx, g.ny, g.nz);
    double r = -INFINITY;
    int x, y, z0, z1;
    march_point(Mh, g.nz, w.local, w.G, x, y, z0, z1);
    if (x < g.nx && y < g.ny && z0 < z1) {
        const F rwx = (F)tb(g, X, T_RW, x), rwy = (F)tb(g, Y, T_RW, y);
        F w0 = fabs(wv[ix.w(x, y, z0)]);
        for (int z = z0; z < z1; z++) {
            const F w1 = fabs(wv[ix.w(x, y, z + 1)]);
            const F rate = dv<F>(fmax(fabs(u[ix.u(x, y, z)]), fabs(u[ix.u(x + 1, y, z)])), rwx)
                         + dv<F>(fmax(fabs(v[ix.v(x, y, z)]), fabs(v[ix.v(x, y + 1, z)])), rwy)
                         + dv<F>(fmax(w0, w1), (F)tb(g, Z, T_RW, z));
            r = comb(0, r, (double)rate);
            w0 = w1;
        }
    }
    r = block_reduce(0, r);
    if (threadIdx.x == 0) gp(M.partial)[P_MAIN * RED_BLOCKS + w.local] = r;
}

__device__ void dt_member(const GLB Mem& M, double rate, int control);
__device__ void stage_control_member(const GLB Mem& M, int st);
// One block of 1024 per member: the CFL rate into scal[S_CFL] (kb_cfl's partials, or the rate the previous
// step's last gradient formed); with control, the time step (backend.cu:884-899): lagged rates, adaptive caps,
// the 1.1 growth limit, the NaN check, and then the first stage's control (stage_control_member).
__global__ __launch_bounds__(1024) void kb_dt(const Mem* const* __restrict__ D, const int* __restrict__ seg_march,
                                              int control) {
    const GLB Mem& M = *gp(D[blockIdx.x]);
    const bool ready = gp(M.ctl)->cfl_ready;                // uniform over the block
    const double rate = ready ? __builtin_bit_cast(double, gp(M.ctl)->cfl_bits)
                              : fold1024(gp(M.partial) + P_MAIN * RED_BLOCKS, seg_march[blockIdx.x + 1] - seg_march[blockIdx.x], 0);
    if (threadIdx.x != 0) return;
    dt_member(M, rate, control);
    if (control) stage_control_member(M, 0);      // then the first stage's control (it reads the new dt)
}
__device__ void dt_member(const GLB Mem& M, double rate, int control) {
    gp(M.scal)[S_CFL] = rate;
    if (!control) return;
    GLB Ctl& c = *gp(M.ctl);
    if (c.err || c.stalled) return;
    {   // the control block as the step begins, restored if the step has to be redone
        const GLB long long* from = (const GLB long long*)&c;
        GLB long long* to = (GLB long long*)gp(M.ctl_bak);
        for (int k = 0; k < (int)(sizeof(Ctl) / sizeof(long long)); k++) to[k] = from[k];
        // a redo refines its first stage from a fresh guess (this attempt refines the formed one in place)
        gp(M.ctl_bak)->g0_ready = 0;
    }
    // this step's last gradient forms the next rate: from +0 up (every cell rate is >= +0 or NaN)
    c.cfl_ready = 0;
    c.cfl_bits = 0ull;
    double use0 = rate, use1 = c.viscous_rate;
    if (c.dt_lag && c.have_prev) { use0 = c.prev_rate[0]; use1 = c.prev_rate[1]; }
    c.prev_rate[0] = rate;
    c.prev_rate[1] = c.viscous_rate;
    c.have_prev = 1;
    if (!isfinite(use0)) { c.err = ERR_VELOCITY_NAN; return; }
    double dt;
    if (!c.adaptive) {
        dt = c.dt_fixed;
        c.dt_reason = 0;
    } else {
        const double caps[3] = {c.dt_max, c.cfl_target / fmax(use0, 1e-300), c.diffusion_limit / fmax(use1, 1e-300)};
        int best = 0;
        for (int k = 1; k < 3; k++) if (caps[k] < caps[best]) best = k;
        dt = caps[best];
        c.dt_reason = best + 1;
    }
    if (c.dt_lag && c.last_dt > 0) dt = fmin(dt, 1.1 * c.last_dt);
    c.dt = dt;
    c.use_rate = use0;
}

// backend.cu:623-626
__device__ __forceinline__ double lid_speed(const GLB Ctl& c, double time) {
    if (c.lid_ramp_t > 0) return c.u_lid * fmin(1.0, fmax(time, 0.0) / c.lid_ramp_t);
    return c.u_lid;
}

// Stage control (backend.cu:903-921): stage time, lid speed, SSP-RK3 coefficients, the warm-start plan of
// guess_phi (backend.cu:857-873: the history entry nearest in time, and the nearest one at a different time
// for linear extrapolation, ties to the newer entry) and a free history buffer for this stage's dphi.
// The warm-start plan of guess_phi (backend.cu:857-873) for a stage at time ts whose dphi scale is `scale`: the
// history entry nearest in time, and the nearest one at a different time for linear extrapolation, ties to the
// newer entry
__device__ void guess_plan(const GLB Ctl& c, double ts, double scale, int& gnh, int& gb1, int& gb2, double& gc1,
                           double& gc2) {
    gnh = 0;
    gb1 = gb2 = 0;
    gc1 = gc2 = 0.0;
    if (c.extrapolate && c.hn > 0) {
        int order[3] = {0, 1, 2};
        for (int i = 1; i < c.hn; i++)
            for (int j = i; j > 0; j--) {
                const int a = order[j], b = order[j - 1];
                const double da = fabs(c.ht[a] - ts), db = fabs(c.ht[b] - ts);
                if (da != db ? da < db : a > b) { order[j] = b; order[j - 1] = a; }
                else break;
            }
        const int h1 = order[0];
        int h2 = -1;
        for (int k = 1; k < c.hn && h2 < 0; k++)
            if (fabs(c.ht[order[k]] - c.ht[h1]) > 1e-12 * fmax(1.0, fabs(c.ht[h1]))) h2 = order[k];
        const double w2 = h2 >= 0 ? (ts - c.ht[h1]) / (c.ht[h2] - c.ht[h1]) : 0.0, w1 = 1.0 - w2;
        gnh = h2 >= 0 ? 2 : 1;
        gb1 = c.hb[h1];
        gc1 = scale * w1 / c.hs[h1];
        if (h2 >= 0) { gb2 = c.hb[h2]; gc2 = scale * w2 / c.hs[h2]; }
    }
}
__device__ void stage_control_member(const GLB Mem& M, int st) {
    GLB Ctl& c = *gp(M.ctl);
    if (c.err || c.stalled) return;
    const double a0s[3] = {0.0, 0.75, 1.0 / 3.0}, bs[3] = {1.0, 0.25, 2.0 / 3.0};
    const double ts = st == 0 ? c.t : st == 1 ? c.t + c.dt : c.t + 0.5 * c.dt;
    c.ts = ts;
    c.lid = lid_speed(c, ts);
    c.a0 = a0s[st];
    c.b = bs[st];
    int gnh, gb1, gb2;
    double gc1, gc2;
    guess_plan(c, ts, c.b * c.dt, gnh, gb1, gb2, gc1, gc2);
    // a first guess the previous step formed stands only if it was formed by this very plan (the host may have
    // changed the clock, the time step or the history since)
    if (st == 0 && c.g0_ready &&
        (gnh != c.ngnh || gc1 != c.ngc1 || gc2 != c.ngc2 || (gnh >= 1 && gb1 != c.ngb1) || (gnh == 2 && gb2 != c.ngb2)))
        c.g0_ready = 0;
    c.gnh = gnh;
    c.gb1 = gb1;
    c.gb2 = gb2;
    c.gc1 = gc1;
    c.gc2 = gc2;
    // a buffer neither in the history nor in the history the step began with (kept for a redo)
    c.out_buf = -1;
    const GLB Ctl& b0 = *gp(M.ctl_bak);
    if (c.extrapolate)
        for (int k = 0; k < HIST_BUFFERS && c.out_buf < 0; k++) {
            bool used = false;
            for (int h = 0; h < c.hn; h++) used |= c.hb[h] == k;
            for (int h = 0; h < b0.hn; h++) used |= b0.hb[h] == k;
            if (!used) c.out_buf = k;
        }
}
// At the end of a step's last projection: the next step's first-stage plan, when the next dt is known already
// (fixed dt, or lagged dt on the rates the next step's kb_dt will use: the same formula on the same values), so
// that the step's last kernel forms that stage's base and guess. The next step's stage control forms the same
// plan again (time t + dt, b = 1, the same history).
__device__ void next_step_plan_member(const GLB Mem& M) {
    GLB Ctl& c = *gp(M.ctl);
    c.nguess = 0;
    if (c.err || c.stalled) return;
    double dt;
    if (!c.adaptive) {
        dt = c.dt_fixed;
    } else {
        if (!(c.dt_lag && c.have_prev)) return;
        const double use0 = c.prev_rate[0], use1 = c.prev_rate[1];
        if (!isfinite(use0)) return;
        const double caps[3] = {c.dt_max, c.cfl_target / fmax(use0, 1e-300), c.diffusion_limit / fmax(use1, 1e-300)};
        int best = 0;
        for (int k = 1; k < 3; k++) if (caps[k] < caps[best]) best = k;
        dt = caps[best];
    }
    if (c.dt_lag && c.dt > 0) dt = fmin(dt, 1.1 * c.dt);    // (the next step's last_dt is this step's dt)
    int gnh, gb1, gb2;
    double gc1, gc2;
    guess_plan(c, c.t + c.dt, 1.0 * dt, gnh, gb1, gb2, gc1, gc2);
    c.ngnh = gnh;
    c.ngb1 = gb1;
    c.ngb2 = gb2;
    c.ngc1 = gc1;
    c.ngc2 = gc2;
    c.nguess = 1;
}

// ------------------------------------------------------------------------------------------------ I(u, m)
// backend.cu:102-274: I = u x omega - grad(K), K = u.m. The Lamb vector is assembled from products
// (vorticity on an edge) x (velocity interpolated to that edge), narrow q = 1 and wide q = 3 stencils, with a
// wall ramp on the 4th-order weights; every product depends on its edge only. kb_stage evaluates each product
// once per block, keeps it in LDS or registers, and uses I at the faces in the stage update; grid-only factors
// come from the axis tables.
//   z-edges (ex, ey, z): omega_z = dv/dx - du/dy; E_ZX = omega_z v|x (backend.cu:162-170), E_ZY = omega_z u|y (171-179)
//   x-edges (x, ey, ez): omega_x = dw/dy - dv/dz; E_XY = omega_x w|y (180-188), E_XZ = omega_x v|z (189-197)
//   y-edges (ex, y, ez): omega_y = du/dz - dw/dx; E_YX = omega_y w|x (198-206), E_YZ = omega_y u|z (207-215)
// Node k of a stencil of width q at edge e runs from e - (q+1)/2 to e + (q-1)/2 (backend.cu:108-112). A
// product whose stencil leaves the box is 0 (backend.cu:143, 149, 155), and so is a product at an edge outside
// the box. The product functions take any position and a select gives the 0, so the loads issue together; in
// the box the expressions are the reference's.
template <class F> __device__ __forceinline__ F stencil_weight(int q, int o4, F s) {   // backend.cu:135-138
    if (q == 1) return o4 ? F(1) + s * F(0.125) : F(1);
    return -s * F(0.125);
}
__device__ __forceinline__ int clampr(int k, int lo, int hi) { return k < lo ? lo : k > hi ? hi : k; }

// The axis-table entries a tile's edges use, for positions b .. b + N - 1 of the x or y axis (b = tile origin
// - 3), copied to LDS: the edge kinds at the position clamped to [0, n], T_W at the cell clamped to [0, n - 1].
// Indexed by unclamped positions, an entry is the table value the clamped index reads.
template <int N, class F = double> struct AxisSlice {
    F ramp[N], rsp[N], rgap3[N], r2sp[N], rws3[N], c3[N], lidf[N], w[N], rw[N];
};
template <int N, class F> __device__ __forceinline__ void fill_slice(AxisSlice<N, F>& A, const Geo& g, int a, int b) {
    const int n = axis_n(g, a);
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        const int e = clampr(b + i, 0, n), c = clampr(b + i, 0, n - 1);
        A.ramp[i] = (F)tb(g, a, T_RAMP, e);
        A.rsp[i] = (F)tb(g, a, T_RSP, e);
        A.rgap3[i] = (F)tb(g, a, T_RGAP3, e);
        A.r2sp[i] = (F)tb(g, a, T_R2SP, e);
        A.rws3[i] = (F)tb(g, a, T_RWS3, e);
        A.c3[i] = (F)tb(g, a, T_C3, e);
        A.lidf[i] = (F)tb(g, a, T_LIDF, e);
        A.w[i] = (F)tb(g, a, T_W, c);
        A.rw[i] = (F)tb(g, a, T_RW, c);
    }
}
// node distance factor and width-weighted value at edge e (slice index) of a stencil of width q from nodes at
// slice indices lo, hi (backend.cu:114-122)
template <class F, class A> __device__ __forceinline__ F rgap_s(const A& s, int e, int q) { return q == 1 ? s.rsp[e] : s.rgap3[e]; }
template <class F, class A>
__device__ __forceinline__ F at_edge_s(const A& s, int e, int q, int lo, int hi, F flo, F fhi) {
    return dv<F>(s.w[lo] * flo + s.w[hi] * fhi, q == 1 ? s.r2sp[e] : s.rws3[e]);
}

// Raw buffer view of n doubles at p (n * 8 < 2^32): 32-bit byte offsets, and a read whose offset is at or
// beyond the end returns 0 instead of faulting. kb_stage reads through these views without clamping positions:
// a value read for an edge or cell outside the box only enters products that a select replaces with 0, and
// every value a kept product uses lies in the box, where the view reads the reference's element. Offsets are
// whole element offsets (in-plane part and level part added in 32 bits) passed as the per-lane offset: the
// hardware range check covers only that offset, so no part of an offset goes into the uniform one.
typedef __amdgpu_buffer_rsrc_t Rsrc;
__device__ __forceinline__ Rsrc rsrc_of(const GLB double* p, i64 n) {
    return __builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, (int)(n * 8), 0x00020000);
}
// the double at byte offset vo
__device__ __forceinline__ double bload(Rsrc r, unsigned vo) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, vo, 0, 0));
}
constexpr unsigned OOB = 0xFFFFFFF8u;                 // a byte offset past the end of every view
__device__ __forceinline__ void bstore(Rsrc r, unsigned vo, double v) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned, v), r, vo, 0, 0);
}
// Face values of element type F (double: the FP64 fields; float: the FP32 fields) through buffer views
template <class F> __device__ __forceinline__ F fld(Rsrc r, unsigned vo);
template <> __device__ __forceinline__ double fld<double>(Rsrc r, unsigned vo) { return bload(r, vo); }
template <> __device__ __forceinline__ float fld<float>(Rsrc r, unsigned vo) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, vo, 0, 0));
}
__device__ __forceinline__ void fst(Rsrc r, unsigned vo, double v) { bstore(r, vo, v); }
__device__ __forceinline__ void fst(Rsrc r, unsigned vo, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, vo, 0, 0);
}
// The residual of the potential (CR_RES) at FP64 byte offset vo: FP64, or rounded to float (at half the offset)
// in the modes whose transforms take float operands, the only reader of the residual there
__device__ __forceinline__ void rstore(Rsrc r, unsigned vo, double v, bool f32) {
    if (f32) fst(r, vo >> 1, (float)v);
    else bstore(r, vo, v);
}
// A face array of a role as a view of the role's three arrays, which follow one another in the cavity's
// arena (u, v, w): one buffer view per role, and the array's byte offset in it. Four views instead of twelve
// keep kb_stage's scalar registers from spilling. A read at a position outside the box may now return a value
// of the neighbouring array instead of 0; such values only enter terms a select discards. The group stays
// below 4 GiB (checked at creation).
struct FaceView {
    Rsrc r;
    unsigned off;
};
template <class F> __device__ __forceinline__ F fld(const FaceView& f, unsigned vo) { return fld<F>(f.r, vo + f.off); }
template <class F> __device__ __forceinline__ void fst(const FaceView& f, unsigned vo, F v) { fst(f.r, vo + f.off, v); }
// The arrays of one RK stage that kb_stage reads and writes, with byte strides (elements of SZ bytes)
struct StageViews {
    FaceView u, v, w, mx, my, mz, Ou, Ov, Ow, M0u, M0v, M0w;
    unsigned su, sv, sw;        // z strides (bytes) of u, v, w
    unsigned ru, rw;            // y strides (bytes) of u and of v, w (nx + 1, nx values)
};
// (every role lays out its arrays alike, so the offsets of v and w are the same in every group)
template <class P>
__device__ __forceinline__ void face_group(const P (&p)[3], unsigned offv, unsigned offw, unsigned bytes,
                                           FaceView& fu, FaceView& fv, FaceView& fw) {
    const Rsrc r = __builtin_amdgcn_make_buffer_rsrc((void*)p[0], (short)0, (int)bytes, 0x00020000);
    fu = FaceView{r, 0u};
    fv = FaceView{r, offv};
    fw = FaceView{r, offw};
}
template <class F> __device__ __forceinline__ StageViews stage_views(const Geo& g, const StageArgs& a) {
    constexpr unsigned SZ = sizeof(F);
    const i64 NX = g.nx, NY = g.ny, NZ = g.nz;
    const i64 nw = NX * NY * (NZ + 1);
    const unsigned offv = (unsigned)((const GLB char*)a.u[1] - (const GLB char*)a.u[0]);
    const unsigned offw = (unsigned)((const GLB char*)a.u[2] - (const GLB char*)a.u[0]);
    const unsigned bytes = offw + (unsigned)(nw * SZ);
    StageViews b;
    face_group(a.u, offv, offw, bytes, b.u, b.v, b.w);
    face_group(a.m, offv, offw, bytes, b.mx, b.my, b.mz);
    face_group(a.out, offv, offw, bytes, b.Ou, b.Ov, b.Ow);
    face_group(a.m0, offv, offw, bytes, b.M0u, b.M0v, b.M0w);
    b.su = (unsigned)((NX + 1) * NY * SZ); b.sv = (unsigned)(NX * (NY + 1) * SZ); b.sw = (unsigned)(NX * NY * SZ);
    b.ru = (unsigned)((NX + 1) * SZ); b.rw = (unsigned)(NX * SZ);
    return b;
}

// z-edge (ex, ey) of cell layer z: p = {E_ZX q=1, E_ZY q=1, E_ZX q=3, E_ZY q=3}, the wide pair with o4 only.
// SX, SY: table slices starting at positions xb, yb. With zx >= 0 only E_ZX (zx = 1) or E_ZY (zx = 0) is
// formed, into p[0] (q = 1) and p[2] (q = 3).
template <int Q, bool OWN, class F, class A>
__device__ __forceinline__ void zedge_q(const Geo& g, const StageViews& B, int o4, const A& SX, int xb, const A& SY,
                                        int yb, int ex, int ey, int zx, F rampXY, unsigned vb, unsigned ub,
                                        F vc, F uc, F& p0, F& p1) {
    constexpr unsigned SZ = sizeof(F);
    const int NX = g.nx, NY = g.ny, sx = ex - xb, sy = ey - yb;
    constexpr int kl = Q == 3 ? 0 : 1, kh = Q == 3 ? 3 : 2;
    const int xl = ex - (Q + 1) / 2, xh = ex + (Q - 1) / 2, yl = ey - (Q + 1) / 2, yh = ey + (Q - 1) / 2;
    const bool ok = xl >= 0 && xh <= NX - 1 && yl >= 0 && yh <= NY - 1;
    // (OWN, narrow stencil: the nodes v(ex, ey), u(ex, ey) are the thread's own column values vc, uc)
    const F vl = fld<F>(B.v, vb + SZ * kl), vh = OWN && Q == 1 ? vc : fld<F>(B.v, vb + SZ * kh);
    const F ul = fld<F>(B.u, ub + kl * B.ru), uh = OWN && Q == 1 ? uc : fld<F>(B.u, ub + kh * B.ru);
    const F o = dv<F>(vh - vl, rgap_s<F>(SX, sx, Q)) - dv<F>(uh - ul, rgap_s<F>(SY, sy, Q));
    const F wo = stencil_weight<F>(Q, o4, rampXY) * o;
    if (zx < 0) {
        const F r0 = wo * at_edge_s<F>(SX, sx, Q, xl - xb, xh - xb, vl, vh);
        const F r1 = wo * at_edge_s<F>(SY, sy, Q, yl - yb, yh - yb, ul, uh);
        p0 = ok ? r0 : F(0);
        p1 = ok ? r1 : F(0);
    } else {
        const A& S = zx ? SX : SY;
        const F r = wo * (zx ? at_edge_s<F>(S, sx, Q, xl - xb, xh - xb, vl, vh)
                             : at_edge_s<F>(S, sy, Q, yl - yb, yh - yb, ul, uh));
        p0 = ok ? r : F(0);
    }
}
template <bool OWN = false, class F, class A>
__device__ __forceinline__ void zedge_products(const Geo& g, const StageViews& B, int o4, const A& SX, int xb,
                                               const A& SY, int yb, int ex, int ey, int z, int zx, F p[4],
                                               F vc = F(0), F uc = F(0)) {
    constexpr unsigned SZ = sizeof(F);
    const Ix ix{g.nx, g.ny, g.nz};
    const F rampXY = SY.ramp[ey - yb] * SX.ramp[ex - xb];
    const unsigned vb = (unsigned)ix.v(ex - 2, ey, z) * SZ;     // v(ex-2 .. ex+1, ey, z): vb + SZ k
    const unsigned ub = (unsigned)ix.u(ex, ey - 2, z) * SZ;     // u(ex, ey-2 .. ey+1, z): ub + k ru
    p[2] = p[3] = F(0);
    zedge_q<1, OWN, F>(g, B, o4, SX, xb, SY, yb, ex, ey, zx, rampXY, vb, ub, vc, uc, p[0], p[1]);
    if (o4) zedge_q<3, OWN, F>(g, B, o4, SX, xb, SY, yb, ex, ey, zx, rampXY, vb, ub, vc, uc, p[2], p[3]);
}

// z-axis factors of the x- and y-edges of level ez (the same for a whole block)
template <class F> struct EdgeLevel {
    F ramp, rsp, rgap3, r2sp, rws3, wm2, wm1, w0, wp1;        // T_W at the cells ez-2 .. ez+1 (clamped)
    bool ok1, ok3;                                            // the level keeps the narrow / wide stencil
};
// (from the level's record of the z factors, level_records: one contiguous scalar read)
template <class F> __device__ __forceinline__ EdgeLevel<F> edge_level(const CNS double* zl, int NZ, int ez) {
    EdgeLevel<F> l;
    l.ramp = (F)zl[ZR_RAMP];
    l.rsp = (F)zl[ZR_RSP];
    l.rgap3 = (F)zl[ZR_RGAP3];
    l.r2sp = (F)zl[ZR_R2SP];
    l.rws3 = (F)zl[ZR_RWS3];
    l.wm2 = (F)zl[ZR_WM2];
    l.wm1 = (F)zl[ZR_WM1];
    l.w0 = (F)zl[ZR_W0];
    l.wp1 = (F)zl[ZR_WP1];
    l.ok1 = ez >= 1 && ez <= NZ - 1;
    l.ok3 = ez >= 2 && ez + 1 <= NZ - 1;
    return l;
}
// the record of march level L (L = -1 .. nz + 2) after the z tables
__device__ __forceinline__ const CNS double* level_record(const Geo& g, int L) {
    return cnp(g.tab[Z]) + (TAB_KINDS * (g.nz + 1) + (L + 1) * ZREC);
}

// An x-edge (x, e) or a y-edge (e, y) of level ez, from the column values c = v(x, e, ez-2 .. ez+1), resp.
// u(e, y, ez-2 .. ez+1), and wa .. wd = w at the in-plane nodes e-2 .. e+1; S is the slice of the in-plane
// axis from position b, n its cell count.
//   x-edge (ydir): p = {E_XY q=1, E_XZ q=1, E_XY q=3, E_XZ q=3}, omega_x = dw/dy - dv/dz
//   y-edge:        p = {E_YX q=1, E_YZ q=1, E_YX q=3, E_YZ q=3}, omega_y = du/dz - dw/dx
template <class F, class A>
__device__ __forceinline__ void side_edge_products(int o4, const EdgeLevel<F>& l, const A& S, int b, int n, int e,
                                                   bool ydir, const F c[4], F wa, F wb, F wc, F wd, F p[4]) {
    const int s = e - b;
    {
        const bool ok = e >= 1 && e <= n - 1 && l.ok1;
        const F dw = dv<F>(wc - wb, S.rsp[s]), dc = dv<F>(c[2] - c[1], l.rsp);
        const F o = ydir ? dw - dc : dc - dw;
        const F wo = stencil_weight<F>(1, o4, l.ramp * S.ramp[s]) * o;
        const F p0 = wo * dv<F>(S.w[s - 1] * wb + S.w[s] * wc, S.r2sp[s]);
        const F p1 = wo * dv<F>(l.wm1 * c[1] + l.w0 * c[2], l.r2sp);
        p[0] = ok ? p0 : F(0);
        p[1] = ok ? p1 : F(0);
    }
    p[2] = p[3] = F(0);
    if (o4) {
        const bool ok = e >= 2 && e + 1 <= n - 1 && l.ok3;
        const F dw = dv<F>(wd - wa, S.rgap3[s]), dc = dv<F>(c[3] - c[0], l.rgap3);
        const F o = ydir ? dw - dc : dc - dw;
        const F wo = stencil_weight<F>(3, o4, l.ramp * S.ramp[s]) * o;
        const F p0 = wo * dv<F>(S.w[s - 2] * wa + S.w[s + 1] * wd, S.rws3[s]);
        const F p1 = wo * dv<F>(l.wm2 * c[0] + l.wp1 * c[3], l.rws3);
        p[2] = ok ? p0 : F(0);
        p[3] = ok ? p1 : F(0);
    }
}

// backend.cu:218-222: K = u.m at cell (x, y, z)
template <class F> __device__ __forceinline__ F kinetic(const Geo& g, const StageViews& B, int x, int y, int z) {
    constexpr unsigned SZ = sizeof(F);
    const Ix ix{g.nx, g.ny, g.nz};
    const unsigned ia = (unsigned)ix.u(x, y, z) * SZ, ib = (unsigned)ix.v(x, y, z) * SZ, ic = (unsigned)ix.w(x, y, z) * SZ;
    const F u0 = fld<F>(B.u, ia), u1 = fld<F>(B.u, ia + SZ), m0 = fld<F>(B.mx, ia), m1 = fld<F>(B.mx, ia + SZ);
    const F v0 = fld<F>(B.v, ib), v1 = fld<F>(B.v, ib + B.rw), n0 = fld<F>(B.my, ib), n1 = fld<F>(B.my, ib + B.rw);
    const F w0 = fld<F>(B.w, ic), w1 = fld<F>(B.w, ic + B.sw), o0 = fld<F>(B.mz, ic), o1 = fld<F>(B.mz, ic + B.sw);
    return F(0.5) * ((u0 * m0 + u1 * m1) + (v0 * n0 + v1 * n1) + (w0 * o0 + w1 * o1));
}

// Roles of RK stage st (backend.cu:903-921): velocity and impulse it reads, impulse it writes (stage 3 into
// the stage-1 buffer, the next state)
__device__ __forceinline__ void stage_roles(int st, int& ur, int& mr, int& outr) {
    ur = st == 0 ? FR_U : st == 1 ? FR_NU0 : FR_NU1;
    mr = st == 0 ? FR_M0 : st == 1 ? FR_NM0 : FR_NM1;
    outr = st == 1 ? FR_NM1 : FR_NM0;
}
__device__ __forceinline__ StageArgs stage_args(const GLB Mem& M, int st) {
    int ur, mr, outr;
    stage_roles(st, ur, mr, outr);
    const GLB Ctl& c = *gp(M.ctl);
    const int par = c.parity;
    StageArgs a;
    for (int k = 0; k < 3; k++) {
        a.u[k] = fbuf(M, par, ur, k);
        a.m[k] = fbuf(M, par, mr, k);
        a.m0[k] = fbuf(M, par, FR_M0, k);
        a.out[k] = fbuf(M, par, outr, k);
    }
    a.nu = M.nu;
    a.lid = c.lid;
    a.dt = c.dt;
    a.a0 = c.a0;
    a.b = c.b;
    a.has_base = st > 0;
    a.o4 = M.o4;
    return a;
}

// The stage kernel's column tile LTX x LTY (one thread per column, TPB threads) and the halo it needs. The
// tile is a template parameter: the autotuner picks 32 x 8 or 16 x 16 per cavity.
template <int LTX, int LTY> struct StageDims {
    static constexpr int LZX = LTX + 3, LZY = LTY + 3;
    static constexpr int LSN = (LTX > LTY ? LTX : LTY) + 6;       // table slice length
    static constexpr int STG_SIDE = 3 * LTX + 3 * LTY;            // halo x-edges (3 rows) and y-edges (3 columns)
    static constexpr int STG_ZHALO = 3 * LTX + 3 * LTY;           // halo z-edges: E_ZX rows, E_ZY columns
    static constexpr int STG_KHALO = LTX + LTY;                   // halo cells of K
    static constexpr int STG_PHALO = 2 * LTX + 2 * LTY;           // halo faces of one impulse plane
    static_assert(LTX * LTY == TPB, "one thread per column of the tile");
    static_assert((LTX & (LTX - 1)) == 0 && (LTY & (LTY - 1)) == 0, "power-of-two tile sides");
    static_assert(STG_SIDE <= TPB && STG_ZHALO <= TPB && STG_KHALO <= TPB && 3 * STG_PHALO <= TPB,
                  "every halo item has a thread");
};
template <int LTX, int LTY, class F>
struct StageTile {
    static constexpr int LZX = StageDims<LTX, LTY>::LZX, LZY = StageDims<LTX, LTY>::LZY;
    using Slice = AxisSlice<StageDims<LTX, LTY>::LSN, F>;
    F z[4][LZY][LZX];                 // z-edge products (zedge_products order) of layer j at (x0-1+i, y0-1+r)
    F xz[2][LZY][LTX];                // E_XZ q=1, 3 of level L at (x0+i, y0-1+r)
    F yz[2][LTY][LZX];                // E_YZ q=1, 3 of level L at (x0-1+i, y0+r)
    F k[3][LTY + 1][LTX + 1];         // K of three cell layers at (x0-1+i, y0-1+r)
    Slice sx, sy;                     // x- and y-axis tables from x0 - 3, y0 - 3
};
// The impulse planes the fused stage update reads its in-plane neighbours from: u and v faces of layer j,
// w faces of level L, at (x0-1+i, y0-1+r)
template <int LTX, int LTY, class F>
struct StagePlanes {
    F p[3][LTY + 2][LTX + 2];
};

// Halo work of thread t in kb_stage, the same at every step: a side edge (t < STG_SIDE: x-edges of rows
// y0 - 1, y0 + LTY, y0 + LTY + 1, then y-edges of columns x0 - 1, x0 + LTX, x0 + LTX + 1), a K cell (t <
// STG_KHALO: column x0 - 1, then row y0 - 1) and, on the last waves, a z-edge product (E_ZX on the rows,
// E_ZY on the columns). Positions relative to the LDS arrays.
struct StageHalo {
    bool side, ydir;       // has a side edge; it is an x-edge (w|y products)
    int se, sr, sc;        // its in-plane position along its axis, LDS row and column
    int kr, kc;            // K cell (row, column) in LDS (t < STG_KHALO)
    int zr, zc;            // z-edge product (row, column) in LDS
    bool zx, zon;          // E_ZX (rows) or E_ZY (columns); has one
};
template <int LTX, int LTY>
__device__ __forceinline__ StageHalo stage_halo(int t, int x0, int y0) {
    constexpr int STG_SIDE = StageDims<LTX, LTY>::STG_SIDE, STG_ZHALO = StageDims<LTX, LTY>::STG_ZHALO;
    StageHalo h;
    h.side = t < STG_SIDE;
    h.ydir = t < 3 * LTX;
    if (h.ydir) {
        const int r = t / LTX;
        h.sr = r == 0 ? 0 : LTY + r;
        h.sc = t & (LTX - 1);
        h.se = y0 - 1 + h.sr;
    } else {
        const int k = t - 3 * LTX, c = k / LTY;
        h.sc = c == 0 ? 0 : LTX + c;
        h.sr = k & (LTY - 1);
        h.se = x0 - 1 + h.sc;
    }
    const bool col = t < LTY;
    h.kr = col ? t + 1 : 0;
    h.kc = col ? 0 : t - LTY + 1;
    const int zt = t - (TPB - STG_ZHALO);
    h.zon = zt >= 0;
    h.zx = zt < 3 * LTX;
    if (h.zx) {
        const int r = zt / LTX;
        h.zr = r == 0 ? 0 : LTY + r;
        h.zc = 1 + (zt & (LTX - 1));
    } else {
        const int k = zt - 3 * LTX, c = k / LTY;
        h.zc = c == 0 ? 0 : LTX + c;
        h.zr = 1 + (k & (LTY - 1));
    }
    return h;
}

// One RK stage of one member (backend.cu:225-330): I = Lamb vector - grad K at every interior face, then the
// stage update q = m + dt (nu lap(m) + I) [a0 m0 + b q in stages 2, 3], marched along z by 32 x 8 column tiles;
// thread (lx, ly) owns the column (x0 + lx, y0 + ly): the u, v faces of its cell layers and the w faces of its
// levels. Step L of the march evaluates, for the tile and the halo its faces reach,
//   - the x- and y-edge products of level L: the E_XY / E_YX of a thread's own column stay in registers for
//     the four levels its u and v faces use, the E_XZ / E_YZ go to LDS for the w faces of level L;
//   - K of cell layer L (LDS, three layers kept; a thread's own K also in registers for its w faces);
//   - the z-edge products of cell layer j = L - 2 (LDS);
//   - the impulse planes of the u, v faces of layer j and the w faces of level L (LDS, for the Laplacian's
//     in-plane neighbours; the z neighbours are carried in registers);
// and then, after a barrier, the w faces of level L and the u and v faces of layer j. No edge product, K or I
// reaches HBM: the kernel reads u, m (and m0) and writes the stage's impulse. The sums follow the reference's
// accumulation order, and the update the reference's formula (the same as a separate face kernel would use).
// (168 VGPRs, three waves per SIMD: the register budget is set to keep the third wave)
template <int LTX, int LTY, class F>
__global__ __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(3))) void kb_stage(const Mem* const* __restrict__ D,
                                                                                        const int* __restrict__ seg,
                                                                                        int nm, int st) {
    using Dm = StageDims<LTX, LTY>;
    constexpr int STG_KHALO = Dm::STG_KHALO, STG_PHALO = Dm::STG_PHALO, LZX = Dm::LZX, LZY = Dm::LZY;
    constexpr unsigned SZ = sizeof(F);                       // bytes of a face value
    __shared__ StageTile<LTX, LTY, F> S;
    __shared__ StagePlanes<LTX, LTY, F> P;
    const Where wh = locate(seg, nm);
    const GLB Mem& M = *gp(D[wh.m]);
    const Geo g = geo_of(M.g);
    const StageArgs a = stage_args(M, st);
    const StageViews B = stage_views<F>(g, a);
    const Ctl* const ctlp = M.ctl;
    const Mem* const memp = D[wh.m];
    const int NX = g.nx, NY = g.ny, NZ = g.nz, o4 = a.o4;
    const Ix ix{NX, NY, NZ};
    const March Mh = stage_march_of(NX, NY, NZ, LTX == STX ? 0 : 1, M.tune[TU_STAGE_TARGET], M.tune[TU_STAGE_MINZ]);
    int x0, y0, z0, z1;
    march_tile(Mh, NZ, wh.local, wh.G, x0, y0, z0, z1);
    if (z0 >= z1) return;                                    // padding block: uniform over the block
    const int xb = x0 - 3, yb = y0 - 3;
    fill_slice(S.sx, g, X, xb);
    fill_slice(S.sy, g, Y, yb);
    // column values at levels L-2 .. L+1 (the last loaded by step L): v(x, y) of the own x-edge, u(x, y) of the
    // own y-edge, v or u of the halo side edge. Byte offsets at level 0; the z offset of level k is k times the
    // array's z stride (levels below 0 read level 0: their products are 0)
    auto lev = [](int k) { return k < 0 ? 0u : (unsigned)k; };
    F cv[4], cu[4], ch[4] = {F(0), F(0), F(0), F(0)};
    {
        const int t = threadIdx.x, x = x0 + (t & (LTX - 1)), y = y0 + t / LTX;
        const StageHalo h = stage_halo<LTX, LTY>(t, x0, y0);
        const unsigned ov = (unsigned)ix.v(x, y, 0) * SZ, ou = (unsigned)ix.u(x, y, 0) * SZ;
        const unsigned oh = h.ydir ? (unsigned)ix.v(x0 + h.sc, h.se, 0) * SZ : (unsigned)ix.u(h.se, y0 + h.sr, 0) * SZ;
        for (int k = 0; k < 3; k++) {
            cv[k] = fld<F>(B.v, ov + lev(z0 - 3 + k) * B.sv);
            cu[k] = fld<F>(B.u, ou + lev(z0 - 3 + k) * B.su);
            if (h.side) ch[k] = h.ydir ? fld<F>(B.v, oh + lev(z0 - 3 + k) * B.sv) : fld<F>(B.u, oh + lev(z0 - 3 + k) * B.su);
        }
    }
    // own narrow products at levels L-2 .. L, wide ones at L-3 .. L; own K of layers L-1, L
    F xy1[3] = {F(0), F(0), F(0)}, yx1[3] = {F(0), F(0), F(0)}, xy3[4] = {F(0), F(0), F(0), F(0)},
      yx3[4] = {F(0), F(0), F(0), F(0)};
    F kprev = F(0), kcur = F(0);
    // fused stage: own impulse of the u and v faces at layers L-3 .. L-1, of the w face at level L-1
    F mu[3] = {F(0), F(0), F(0)}, mv[3] = {F(0), F(0), F(0)}, mwp = F(0);
    // the impulse of the own w face at the current level (the first step's level z0 - 1)
    F mwc;
    {
        const int t = threadIdx.x, x = x0 + (t & (LTX - 1)), y = y0 + t / LTX;
        mwc = fld<F>(B.mz, (unsigned)ix.w(x, y, z0 - 1) * SZ);
    }
    __syncthreads();                                          // the table slices
    for (int L = z0 - 1, it = 0; L <= z1 + 1; L++, it++) {
        // lane positions re-derived at every step from an opaque copy of the thread id: fewer per-lane values
        // live across the loop
        int t = threadIdx.x;
        asm volatile("" : "+v"(t));
        const int lx = t & (LTX - 1), ly = t / LTX, x = x0 + lx, y = y0 + ly;
        const bool own = x < NX && y < NY;
        const StageHalo h = stage_halo<LTX, LTY>(t, x0, y0);
        const unsigned ov = (unsigned)ix.v(x, y, 0) * SZ, ou = (unsigned)ix.u(x, y, 0) * SZ;
        const unsigned oh = h.ydir ? (unsigned)ix.v(x0 + h.sc, h.se, 0) * SZ : (unsigned)ix.u(h.se, y0 + h.sr, 0) * SZ;
        const EdgeLevel<F> el = edge_level<F>(level_record(g, L), NZ, L);
        const unsigned wz = (unsigned)clampr(L, 0, NZ) * B.sw;     // w level of the side edges
        cv[3] = fld<F>(B.v, ov + lev(L + 1) * B.sv);
        cu[3] = fld<F>(B.u, ou + lev(L + 1) * B.su);
        if (h.side) ch[3] = h.ydir ? fld<F>(B.v, oh + lev(L + 1) * B.sv) : fld<F>(B.u, oh + lev(L + 1) * B.su);
        // w at the in-plane nodes e-2 .. e+1 of the own x-edge (wx) and y-edge (wy), read together (the outer
        // two also in order 2, where the products do not use them) so that both groups are one round trip
        F p[4], wa, wb, wc, wd, wx[4], wy[4];
        {
            const unsigned ox = (unsigned)ix.w(x, y - 2, 0) * SZ + wz, oy = (unsigned)ix.w(x - 2, y, 0) * SZ + wz;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                wx[k] = fld<F>(B.w, ox + k * B.rw);
                wy[k] = fld<F>(B.w, oy + k * SZ);
            }
        }
        side_edge_products<F>(o4, el, S.sy, yb, NY, y, true, cv, wx[0], wx[1], wx[2], wx[3], p);
        xy1[0] = xy1[1]; xy1[1] = xy1[2]; xy1[2] = p[0];
        xy3[0] = xy3[1]; xy3[1] = xy3[2]; xy3[2] = xy3[3]; xy3[3] = p[2];
        S.xz[0][ly + 1][lx] = p[1];
        S.xz[1][ly + 1][lx] = p[3];
        side_edge_products<F>(o4, el, S.sx, xb, NX, x, false, cu, wy[0], wy[1], wy[2], wy[3], p);
        yx1[0] = yx1[1]; yx1[1] = yx1[2]; yx1[2] = p[0];
        yx3[0] = yx3[1]; yx3[1] = yx3[2]; yx3[2] = yx3[3]; yx3[3] = p[2];
        S.yz[0][ly][lx + 1] = p[1];
        S.yz[1][ly][lx + 1] = p[3];
        if (h.side) {                                         // the halo x- or y-edge: E_XZ / E_YZ only
            const unsigned o = (h.ydir ? (unsigned)ix.w(x0 + h.sc, h.se - 2, 0) * SZ : (unsigned)ix.w(h.se - 2, y0 + h.sr, 0) * SZ) + wz;
            const unsigned d = h.ydir ? B.rw : SZ;
            wb = fld<F>(B.w, o + d);
            wc = fld<F>(B.w, o + 2 * d);
            wa = o4 ? fld<F>(B.w, o) : F(0);
            wd = o4 ? fld<F>(B.w, o + 3 * d) : F(0);
            side_edge_products<F>(o4, el, h.ydir ? S.sy : S.sx, h.ydir ? yb : xb, h.ydir ? NY : NX, h.se, h.ydir, ch,
                                  wa, wb, wc, wd, p);
            F* dst = h.ydir ? &S.xz[0][h.sr][h.sc] : &S.yz[0][h.sr][h.sc];
            dst[0] = p[1];
            dst[h.ydir ? LZY * LTX : LTY * LZX] = p[3];
        }
        // own impulse column: u and v faces of layer L, w faces of levels L and L + 1 (fused stage, and K); the w
        // face of level L was loaded as level L + 1 one step before
        const unsigned iuL = (unsigned)ix.u(x, y, L) * SZ, ivL = (unsigned)ix.v(x, y, L) * SZ, iwL = (unsigned)ix.w(x, y, L) * SZ;
        const F muL = fld<F>(B.mx, iuL), mvL = fld<F>(B.my, ivL), mwL = mwc, mwL1 = fld<F>(B.mz, iwL + B.sw);
        // w(x, y, L) is the middle node of the own x-edge (levels 0 .. nz; the K of level -1 is never used)
        const F wwL = wx[2];
        if (L < z1) {                                         // K of layer L (used by layers z0-1 .. z1-1)
            const int slot = it % 3;
            kprev = kcur;
            {   // backend.cu:218-222 with the impulse loaded above; u(x, y, L), v(x, y, L) are the column values
                const F u0 = cu[2], u1 = fld<F>(B.u, iuL + SZ), m1 = fld<F>(B.mx, iuL + SZ);
                const F v0 = cv[2], v1 = fld<F>(B.v, ivL + B.rw), n1 = fld<F>(B.my, ivL + B.rw);
                const F w1 = fld<F>(B.w, iwL + B.sw);
                kcur = F(0.5) * ((u0 * muL + u1 * m1) + (v0 * mvL + v1 * n1) + (wwL * mwL + w1 * mwL1));
            }
            S.k[slot][ly + 1][lx + 1] = kcur;
            if (t < STG_KHALO) S.k[slot][h.kr][h.kc] = kinetic<F>(g, B, x0 - 1 + h.kc, y0 - 1 + h.kr, L);
        }
        const int j = L - 2;
        {                                                     // impulse planes: u, v of layer j, w of level L
            P.p[0][ly + 1][lx + 1] = mu[1];
            P.p[1][ly + 1][lx + 1] = mv[1];
            P.p[2][ly + 1][lx + 1] = mwL;
            if (t < 3 * STG_PHALO) {
                const int c = t / STG_PHALO, k = t - c * STG_PHALO;
                int r, q;
                if (k < 2 * LTX) { r = k < LTX ? 0 : LTY + 1; q = 1 + (k & (LTX - 1)); }
                else { const int kk = k - 2 * LTX; q = kk < LTY ? 0 : LTX + 1; r = 1 + (kk & (LTY - 1)); }
                const int hx_ = x0 - 1 + q, hy_ = y0 - 1 + r;
                const F val = c == 0 ? fld<F>(B.mx, (unsigned)ix.u(hx_, hy_, j) * SZ)
                            : c == 1 ? fld<F>(B.my, (unsigned)ix.v(hx_, hy_, j) * SZ)
                                     : fld<F>(B.mz, (unsigned)ix.w(hx_, hy_, L) * SZ);
                P.p[c][r][q] = val;
            }
        }
        if (j >= z0) {                                        // z-edges of layer j
            zedge_products<true, F>(g, B, o4, S.sx, xb, S.sy, yb, x, y, j, -1, p, cv[0], cu[0]);
            for (int q = 0; q < 4; q++) S.z[q][ly + 1][lx + 1] = p[q];
            if (h.zon) {
                zedge_products<false, F>(g, B, o4, S.sx, xb, S.sy, yb, x0 - 1 + h.zc, y0 - 1 + h.zr, j, h.zx ? 1 : 0, p);
                S.z[h.zx ? 0 : 1][h.zr][h.zc] = p[0];
                S.z[h.zx ? 2 : 3][h.zr][h.zc] = p[2];
            }
        }
        // the gauge-reset impulse of this step's w face and of layer j's u and v faces (the RK base of stages 2 and
        // 3), read before the barrier so that its latency overlaps the wait (OOB reads 0 where a face is not
        // updated, or in stage 1)
        const bool based = a.has_base && own;
        const F bw0 = fld<F>(B.M0w.r, based && L >= z0 && L < z1 && L >= 1 ? iwL + B.M0w.off : OOB);
        const F bu0 = fld<F>(B.M0u.r, based && j >= z0 && x >= 1 ? (unsigned)ix.u(x, y, j) * SZ + B.M0u.off : OOB);
        const F bv0 = fld<F>(B.M0v.r, based && j >= z0 && y >= 1 ? (unsigned)ix.v(x, y, j) * SZ + B.M0v.off : OOB);
        __syncthreads();
        // the level again, opaque to the compiler: the z factors of the face updates are loaded here, after the
        // barrier, instead of at the top of the step, where they would stay live in scalar registers (which
        // then spill) through the edge products
        int Lq = L;
        asm volatile("" : "+s"(Lq));
        // likewise the stage coefficients and the viscosity, read again from the control block and the
        // descriptor (which no kernel writes while a stage runs; the same values as in a)
        const Ctl* cpq = ctlp;
        asm volatile("" : "+s"(cpq));
        const CNS Ctl& cq = *cnp(cpq);
        const F lid = (F)cq.lid, dt = (F)cq.dt, a0 = (F)cq.a0, bb = (F)cq.b;   // (float) in the FP32 fields
        const Mem* mpq = memp;
        asm volatile("" : "+s"(mpq));
        const F nu = (F)cnp(mpq)->nu;
        const CNS double* zq = level_record(g, Lq);        // the z factors of the face updates
        if (own && L >= z0 && L < z1 && L >= 1) {            // w face (x, y, L)
            F Lw = -F(0.5) * (S.xz[0][ly + 2][lx] + S.xz[0][ly + 1][lx]) + F(0.5) * (S.yz[0][ly][lx + 2] + S.yz[0][ly][lx + 1]);
            if (o4) {
                const F cz3 = (F)zq[ZR_C3];
                Lw -= cz3 * (S.xz[1][ly + 3][lx] + S.xz[1][ly][lx]);
                Lw += cz3 * (S.yz[1][ly][lx + 3] + S.yz[1][ly][lx]);
            }
            const F Iwv = Lw - dv<F>(kcur - kprev, (F)zq[ZR_RSP]);
            {                                                 // stage update of the w face (backend.cu:279-330)
                const F c = mwL, mzp = mwL1, mzm = mwp;
                const F mxp = x < NX - 1 ? P.p[2][ly + 1][lx + 2] : c, mxm = x > 0 ? P.p[2][ly + 1][lx] : c;
                const F myp = y < NY - 1 ? P.p[2][ly + 2][lx + 1] : c, mym = y > 0 ? P.p[2][ly][lx + 1] : c;
                const bool by_wall = x == 0 || x == NX - 1 || y == 0 || y == NY - 1;
                const F wall = by_wall ? wwL : F(0);
                const F base = a.has_base ? bw0 : c;
                const F rspx = S.sx.rsp[lx + 3], rwx = S.sx.rw[lx + 3], re = S.sx.rsp[lx + 4];
                const F rspy = S.sy.rsp[ly + 3], rwy = S.sy.rw[ly + 3], rn = S.sy.rsp[ly + 4];
                F lap = F(0);
                {
                    const F east = x < NX - 1 ? dv<F>(mxp - c, re) : dv<F>(-wall, re);
                    const F west = x > 0 ? dv<F>(c - mxm, rspx) : dv<F>(wall, rspx);
                    lap += dv<F>(east - west, rwx);
                }
                {
                    const F north = y < NY - 1 ? dv<F>(myp - c, rn) : dv<F>(F(0) - wall, rn);
                    const F south = y > 0 ? dv<F>(c - mym, rspy) : dv<F>(wall, rspy);
                    lap += dv<F>(north - south, rwy);
                }
                lap += dv<F>(dv<F>(mzp - c, (F)zq[ZR_RW]) - dv<F>(c - mzm, (F)zq[ZR_RWM1]), (F)zq[ZR_RSP]);
                const F rhs = nu * lap + F(0);
                F q = c + dt * (rhs + Iwv);
                q = a0 * base + bb * q;
                fst(B.Ow, iwL, q);
            }
        }
        if (own && ((L == 0 && z0 == 0) || (L == NZ && z1 == NZ))) fst(B.Ow, iwL, F(0));   // w walls
        if (own && j >= z0) {
            const int ks = (it + 1) % 3;                      // slot of layer j
            const F kc = S.k[ks][ly + 1][lx + 1];
            if (x >= 1) {                                     // u face (x, y, j)
                F Lu = F(0.5) * (S.z[0][ly + 2][lx + 1] + S.z[0][ly + 1][lx + 1]) - F(0.5) * (yx1[1] + yx1[0]);
                if (o4) {
                    const F cx3 = S.sx.c3[lx + 3];
                    Lu += cx3 * (S.z[2][ly + 3][lx + 1] + S.z[2][ly][lx + 1]);
                    Lu -= cx3 * (yx3[3] + yx3[0]);
                }
                const F Iuv = Lu - dv<F>(kc - S.k[ks][ly + 1][lx], S.sx.rsp[lx + 3]);
                {                                             // stage update of the u face
                    const unsigned iu = (unsigned)ix.u(x, y, j) * SZ;
                    const F c = mu[1], mzp = j < NZ - 1 ? mu[2] : c, mzm = j > 0 ? mu[0] : c;
                    const F mxp = P.p[0][ly + 1][lx + 2], mxm = P.p[0][ly + 1][lx];
                    const F myp = y < NY - 1 ? P.p[0][ly + 2][lx + 1] : c, mym = y > 0 ? P.p[0][ly][lx + 1] : c;
                    const bool by_wall = y == 0 || y == NY - 1 || j == 0 || j == NZ - 1;
                    const F wall = by_wall ? cu[0] : F(0);
                    const F base = a.has_base ? bu0 : c;
                    const F rspx = S.sx.rsp[lx + 3], rwx = S.sx.rw[lx + 3], rwxm = S.sx.rw[lx + 2];
                    const F rspy = S.sy.rsp[ly + 3], rwy = S.sy.rw[ly + 3], rn = S.sy.rsp[ly + 4];
                    F lap = F(0);
                    lap += dv<F>(dv<F>(mxp - c, rwx) - dv<F>(c - mxm, rwxm), rspx);
                    {
                        const F U = y == NY - 1 ? lid * (S.sx.lidf[lx + 3] * (F)zq[ZR_LIDCJ]) : F(0);
                        const F north = y < NY - 1 ? dv<F>(myp - c, rn) : dv<F>(U - wall, rn);
                        const F south = y > 0 ? dv<F>(c - mym, rspy) : dv<F>(wall, rspy);
                        lap += dv<F>(north - south, rwy);
                    }
                    {
                        const F rt = (F)zq[ZR_RSPT], rb = (F)zq[ZR_RSPJ];
                        const F top = j < NZ - 1 ? dv<F>(mzp - c, rt) : dv<F>(-wall, rt);
                        const F bottom = j > 0 ? dv<F>(c - mzm, rb) : dv<F>(wall, rb);
                        lap += dv<F>(top - bottom, (F)zq[ZR_RWJ]);
                    }
                    const F rhs = nu * lap + F(0);
                    F q = c + dt * (rhs + Iuv);
                    q = a0 * base + bb * q;
                    fst(B.Ou, iu, q);
                }
            }
            if (y >= 1) {                                     // v face (x, y, j)
                F Lv = -F(0.5) * (S.z[1][ly + 1][lx + 2] + S.z[1][ly + 1][lx + 1]) + F(0.5) * (xy1[1] + xy1[0]);
                if (o4) {
                    const F cy3 = S.sy.c3[ly + 3];
                    Lv -= cy3 * (S.z[3][ly + 1][lx + 3] + S.z[3][ly + 1][lx]);
                    Lv += cy3 * (xy3[3] + xy3[0]);
                }
                const F Ivv = Lv - dv<F>(kc - S.k[ks][ly][lx + 1], S.sy.rsp[ly + 3]);
                {                                             // stage update of the v face
                    const unsigned iv = (unsigned)ix.v(x, y, j) * SZ;
                    const F c = mv[1], mzp = j < NZ - 1 ? mv[2] : c, mzm = j > 0 ? mv[0] : c;
                    const F mxp = x < NX - 1 ? P.p[1][ly + 1][lx + 2] : c, mxm = x > 0 ? P.p[1][ly + 1][lx] : c;
                    const F myp = P.p[1][ly + 2][lx + 1], mym = P.p[1][ly][lx + 1];
                    const bool by_wall = x == 0 || x == NX - 1 || j == 0 || j == NZ - 1;
                    const F wall = by_wall ? cv[0] : F(0);
                    const F base = a.has_base ? bv0 : c;
                    const F rspx = S.sx.rsp[lx + 3], rwx = S.sx.rw[lx + 3], re = S.sx.rsp[lx + 4];
                    const F rspy = S.sy.rsp[ly + 3], rwy = S.sy.rw[ly + 3], rwym = S.sy.rw[ly + 2];
                    F lap = F(0);
                    {
                        const F east = x < NX - 1 ? dv<F>(mxp - c, re) : dv<F>(-wall, re);
                        const F west = x > 0 ? dv<F>(c - mxm, rspx) : dv<F>(wall, rspx);
                        lap += dv<F>(east - west, rwx);
                    }
                    lap += dv<F>(dv<F>(myp - c, rwy) - dv<F>(c - mym, rwym), rspy);
                    {
                        const F rt = (F)zq[ZR_RSPT], rb = (F)zq[ZR_RSPJ];
                        const F top = j < NZ - 1 ? dv<F>(mzp - c, rt) : dv<F>(-wall, rt);
                        const F bottom = j > 0 ? dv<F>(c - mzm, rb) : dv<F>(wall, rb);
                        lap += dv<F>(top - bottom, (F)zq[ZR_RWJ]);
                    }
                    const F rhs = nu * lap + F(0);
                    F q = c + dt * (rhs + Ivv);
                    q = a0 * base + bb * q;
                    fst(B.Ov, iv, q);
                }
            }
        }
        if (own && j >= z0) {                                 // wall faces of layer j
            if (x == 0) fst(B.Ou, (unsigned)ix.u(0, y, j) * SZ, F(0));
            if (x == NX - 1) fst(B.Ou, (unsigned)ix.u(NX, y, j) * SZ, F(0));
            if (y == 0) fst(B.Ov, (unsigned)ix.v(x, 0, j) * SZ, F(0));
            if (y == NY - 1) fst(B.Ov, (unsigned)ix.v(x, NY, j) * SZ, F(0));
        }
        __syncthreads();
        cv[0] = cv[1]; cv[1] = cv[2]; cv[2] = cv[3];
        cu[0] = cu[1]; cu[1] = cu[2]; cu[2] = cu[3];
        ch[0] = ch[1]; ch[1] = ch[2]; ch[2] = ch[3];
        mu[0] = mu[1]; mu[1] = mu[2]; mu[2] = muL;
        mv[0] = mv[1]; mv[1] = mv[2]; mv[2] = mvL;
        mwp = mwL;
        mwc = mwL1;
    }
}


// ------------------------------------------------------------------------------------------------ projection
// backend.cu:922-925 (+ guess_phi's lincomb/scaled, backend.cu:870-871, and add_to):
// base = [a0 phi0 +] b cur;  phi = [c1 h1 [+ c2 h2] +] base, into the stage's work potential (stages 1 and 3:
// CR_PS0, stage 2: CR_PS1; cur: phi0 for stage 1, else the previous stage's work potential). Stage 1's base is
// phi0 itself (a0 0, b 1), which its projection reads from CR_PHI0: only the guess is stored. Grid-stride over
// the cells: for stage 1 it usually exits at once, the previous step's last kernel having formed the guess
// (Ctl::g0_ready).
__global__ __launch_bounds__(TPB) void kb_base_guess(const Mem* const* __restrict__ D, const int* __restrict__ seg,
                                                     int nm, int st) {
    const Where w = locate(seg, nm);
    const GLB Mem& M = *gp(D[w.m]);
    const GLB Ctl& c = *gp(M.ctl);
    if (st == 0 && c.g0_ready) return;
    const int rot = c.rot;
    const GLB double* __restrict__ phi0 = cbuf(M, rot, CR_PHI0);
    const GLB double* cur = cbuf(M, rot, st == 0 ? CR_PHI0 : st == 1 ? CR_PS0 : CR_PS1);
    GLB double* phi = cbuf(M, rot, st == 1 ? CR_PS1 : CR_PS0);
    GLB double* base = cbuf(M, rot, CR_BASE);
    const double a0 = c.a0, b = c.b;
    for (i64 i = (i64)w.local * TPB + threadIdx.x; i < M.n_cell; i += (i64)w.G * TPB) {
        const double bs = a0 != 0.0 ? a0 * phi0[i] + b * cur[i] : b * cur[i];
        if (st != 0) base[i] = bs;      // (the first stage's base is phi0 itself, read from CR_PHI0)
        double gs = bs;
        if (c.gnh == 1) gs = c.gc1 * cbuf(M, rot, CR_H0 + c.gb1)[i] + bs;
        else if (c.gnh == 2) gs = (c.gc1 * cbuf(M, rot, CR_H0 + c.gb1)[i] + c.gc2 * cbuf(M, rot, CR_H0 + c.gb2)[i]) + bs;
        phi[i] = gs;
    }
}

// backend.cu:373-380: wall faces of the impulse in role mm set to 0, one thread per wall face
template <class F>
__global__ __launch_bounds__(TPB) void kb_zero_walls(const Mem* const* __restrict__ D, const int* __restrict__ seg,
                                                     int nm, int mm) {
    const Where w = locate(seg, nm);
    const GLB Mem& M = *gp(D[w.m]);
    const int par = gp(M.ctl)->parity;
    const int NX = M.g.nx, NY = M.g.ny, NZ = M.g.nz;
    GLB F* f0 = (GLB F*)fbuf(M, par, mm, 0), *f1 = (GLB F*)fbuf(M, par, mm, 1), *f2 = (GLB F*)fbuf(M, par, mm, 2);
    const int nu = 2 * NY * NZ, nv = 2 * NX * NZ, nw = 2 * NX * NY;
    int t = w.local * TPB + threadIdx.x;
    if (t < nu) {
        const int side = t % 2, r = t / 2, y = r % NY, z = r / NY;
        f0[(side ? NX : 0) + (i64)(NX + 1) * (y + (i64)NY * z)] = F(0);
        return;
    }
    t -= nu;
    if (t < nv) {
        const int side = t % 2, r = t / 2, x = r % NX, z = r / NX;
        f1[x + (i64)NX * ((side ? NY : 0) + (i64)(NY + 1) * z)] = F(0);
        return;
    }
    t -= nv;
    if (t < nw) {
        const int side = t % 2, r = t / 2, x = r % NX, y = r / NX;
        f2[x + (i64)NX * (y + (i64)NY * (side ? NZ : 0))] = F(0);
    }
}

// backend.cu:334-342: rhs = vol div(mm) with the rhs^2 partial sums (region P_MAIN) fused (z-march, w(z+1)
// carried to the next layer)
template <class F>
__global__ __launch_bounds__(TPB) void kb_divergence(const Mem* const* __restrict__ D, const int* __restrict__ seg,
                                                     int nm, int mm) {
    const Where w = locate(seg, nm);
    const GLB Mem& M = *gp(D[w.m]);
    const int par = gp(M.ctl)->parity;
    const int rot = gp(M.ctl)->rot;
    const Geo g = geo_of(M.g);
    const GLB F* __restrict__ u = (const GLB F*)fbuf(M, par, mm, 0), *__restrict__ v = (const GLB F*)fbuf(M, par, mm, 1),
               *__restrict__ wv = (const GLB F*)fbuf(M, par, mm, 2);
    GLB double* __restrict__ rhs = cbuf(M, rot, CR_RHS);
    const Ix ix{g.nx, g.ny, g.nz};
    const March Mh = march_of(g.nx, g.ny, g.nz);
    double acc = 0.0;
    int x, y, z0, z1;
    march_point(Mh, g.nz, w.local, w.G, x, y, z0, z1);
    if (x < g.nx && y < g.ny && z0 < z1) {
        const double dx = tb(g, X, T_W, x), dy = tb(g, Y, T_W, y);
        double w0 = (double)wv[ix.w(x, y, z0)];
#pragma unroll 2
        for (int z = z0; z < z1; z++) {
            const double dz = tb(g, Z, T_W, z), w1 = (double)wv[ix.w(x, y, z + 1)];
            const double r = ((double)u[ix.u(x + 1, y, z)] - (double)u[ix.u(x, y, z)]) * (dy * dz)
                           + ((double)v[ix.v(x, y + 1, z)] - (double)v[ix.v(x, y, z)]) * (dx * dz)
                           + (w1 - w0) * (dx * dy);
            rhs[ix.c(x, y, z)] = r;
            acc += r * r;
            w0 = w1;
        }
    }
    acc = block_reduce(1, acc);
    if (threadIdx.x == 0) gp(M.partial)[P_MAIN * RED_BLOCKS + w.local] = acc;
}

// backend.cu:345-359: r = rhs - L(phi) with the r^2 partial sums fused (z-march, phi(z-1), phi(z), phi(z+1) in
// registers); potential in role ph, partials into `region`. pred: only while the member's FDM loop is active.
// The partial sums of phi itself go to region P_PHI: the last residual a projection runs has seen its final
// potential, whose mean backend.cu:845-846 subtracts. store = 0: the residual of the last pass enqueued, which
// only its norm decides on (a projection that needs another pass stalls and its step is redone): r is not
// stored (a view of no elements drops the stores).
template <bool LOOP, bool RF>
__global__ __launch_bounds__(TPB) void kb_residual(const Mem* const* __restrict__ D, const int* __restrict__ seg,
                                                   int nm, int ph, int region, int pred, int store) {
    const Where w = locate(seg, nm);
    const GLB Mem& M = *gp(D[w.m]);
    const int rot = gp(M.ctl)->rot;
    if (pred && !gp(M.ctl)->active) return;
    const Geo g = geo_of(M.g);
    const int NX = g.nx, NY = g.ny, NZ = g.nz;
    const i64 n = (i64)NX * NY * NZ;
    const Rsrc rp = rsrc_of(cbuf(M, rot, ph), n), rh = rsrc_of(cbuf(M, rot, CR_RHS), n);
    const Rsrc rr = rsrc_of(cbuf(M, rot, CR_RES), store ? n : 0);
    constexpr bool rf = RF;                               // a float residual (the modes with float transforms)
    const unsigned sy = (unsigned)NX * 8u, sz = (unsigned)NX * NY * 8u;
    const March Mh = march_of(NX, NY, NZ);
    const int full = march_blocks(Mh);
    // logical blocks of the full march, strided when the launch has fewer (LOOP; same partials either way)
    for (int lb = w.local; lb < full; lb += LOOP ? w.G : full) {
        double acc = 0.0, sphi = 0.0;
        int x, y, z0, z1;
        march_point(Mh, NZ, lb, full, x, y, z0, z1);
        if (x < NX && y < NY && z0 < z1) {
            // Branch-free: an absent neighbour (wall) is read at the cell itself and its coefficient is 0,
            // so its term is +0 and the sum is bit-identical to skipping it. Two layers per trip: the loads
            // of both issue before either store (a buffer store may alias a later load, so the compiler keeps
            // that order), and the sums run in z order as with one layer per trip.
            const double dx = tb(g, X, T_W, x), dy = tb(g, Y, T_W, y);
            const double rx1 = x + 1 < NX ? tb(g, X, T_RSPD, x + 1) : 0.0, rx0 = x > 0 ? tb(g, X, T_RSPD, x) : 0.0;
            const double ry1 = y + 1 < NY ? tb(g, Y, T_RSPD, y + 1) : 0.0, ry0 = y > 0 ? tb(g, Y, T_RSPD, y) : 0.0;
            const unsigned oe = x + 1 < NX ? 8u : 0u, ow = x > 0 ? 8u : 0u, on = y + 1 < NY ? sy : 0u, os = y > 0 ? sy : 0u;
            unsigned i = ((unsigned)x + (unsigned)NX * ((unsigned)y + (unsigned)NY * (unsigned)z0)) * 8u;
            double prev = z0 > 0 ? bload(rp, i - sz) : 0.0, c = bload(rp, i);
            double rzm = z0 > 0 ? tb(g, Z, T_RSPD, z0) : 0.0;
            auto layer = [&](double pc, double pm, double pp, double pe, double pw, double pn, double ps, double rv,
                             double dz, double rzp, double rzm_) {
                double s = 0.0;
                s += (dy * dz) * (pe - pc) * rx1;
                s -= (dy * dz) * (pc - pw) * rx0;
                s += (dx * dz) * (pn - pc) * ry1;
                s -= (dx * dz) * (pc - ps) * ry0;
                s += (dx * dy) * (pp - pc) * rzp;
                s -= (dx * dy) * (pc - pm) * rzm_;
                return rv - s;
            };
            int z = z0;
            for (; z + 1 < z1; z += 2, i += 2 * sz) {
                const unsigned i2 = i + sz;                       // layer z + 1 (< z1 <= NZ)
                const bool up2 = z + 2 < NZ;
                const double n1 = bload(rp, i2), n2 = bload(rp, up2 ? i2 + sz : i2);
                const double pe1 = bload(rp, i + oe), pw1 = bload(rp, i - ow), pn1 = bload(rp, i + on), ps1 = bload(rp, i - os);
                const double pe2 = bload(rp, i2 + oe), pw2 = bload(rp, i2 - ow), pn2 = bload(rp, i2 + on), ps2 = bload(rp, i2 - os);
                const double rv1 = bload(rh, i), rv2 = bload(rh, i2);
                const double dz1 = tb(g, Z, T_W, z), rzp1 = tb(g, Z, T_RSPD, z + 1);
                const double dz2 = tb(g, Z, T_W, z + 1), rzp2 = up2 ? tb(g, Z, T_RSPD, z + 2) : 0.0;
                const double r1 = layer(c, prev, n1, pe1, pw1, pn1, ps1, rv1, dz1, rzp1, rzm);
                const double r2 = layer(n1, c, n2, pe2, pw2, pn2, ps2, rv2, dz2, rzp2, rzp1);
                rstore(rr, i, r1, rf);
                rstore(rr, i2, r2, rf);
                acc += r1 * r1;
                sphi += c;
                acc += r2 * r2;
                sphi += n1;
                prev = n1;
                c = n2;
                rzm = rzp2;
            }
            if (z < z1) {                                         // odd chunk length: the last layer
                const bool up = z + 1 < NZ;
                const double next = bload(rp, up ? i + sz : i);
                const double pe = bload(rp, i + oe), pw = bload(rp, i - ow), pn = bload(rp, i + on), ps = bload(rp, i - os);
                const double rv = bload(rh, i);
                const double dz = tb(g, Z, T_W, z), rzp = up ? tb(g, Z, T_RSPD, z + 1) : 0.0;
                const double ri = layer(c, prev, next, pe, pw, pn, ps, rv, dz, rzp, rzm);
                rstore(rr, i, ri, rf);
                acc += ri * ri;
                sphi += c;
            }
        }
        acc = block_reduce(1, acc);
        sphi = block_reduce(1, sphi);
        if (threadIdx.x == 0) {
            gp(M.partial)[region * RED_BLOCKS + lb] = acc;
            gp(M.partial)[P_PHI * RED_BLOCKS + lb] = sphi;
        }
    }
}

// The divergence of the impulse in role mm (backend.cu:334-342, kb_divergence) and the residual of the warm-start
// guess in role pw (backend.cu:345-359, kb_residual into region P_GUESS) in one z-march: each cell's rhs is formed
// from its six faces, stored, and used at once in r = rhs - L(phi). The same march, blocks and per-thread z order
// as the two kernels, so rhs, r and the partial sums (|rhs|^2 into P_MAIN, |r|^2 into P_GUESS, sum phi into
// P_PHI) are theirs bit for bit; the rhs is not read back from memory.
template <class F, bool RF>
__global__ __launch_bounds__(TPB) void kb_div_residual(const Mem* const* __restrict__ D, const int* __restrict__ seg,
                                                       int nm, int mm, int pw) {
    constexpr unsigned SZ = sizeof(F);
    const Where w = locate(seg, nm);
    const GLB Mem& M = *gp(D[w.m]);
    const int par = gp(M.ctl)->parity;
    const int rot = gp(M.ctl)->rot;
    const Geo g = geo_of(M.g);
    const int NX = g.nx, NY = g.ny, NZ = g.nz;
    const Ix ix{NX, NY, NZ};
    const i64 n = (i64)NX * NY * NZ;
    const i64 nu = (i64)(NX + 1) * NY * NZ, nv = (i64)NX * (NY + 1) * NZ, nw = (i64)NX * NY * (NZ + 1);
    const Rsrc rp = rsrc_of(cbuf(M, rot, pw), n), rh = rsrc_of(cbuf(M, rot, CR_RHS), n);
    const Rsrc rr = rsrc_of(cbuf(M, rot, CR_RES), n);
    constexpr bool rf = RF;                               // a float residual (the modes with float transforms)
    auto fview = [&](int k, i64 cnt) { return rsrc_of(fbuf(M, par, mm, k), (cnt * SZ + 7) / 8); };
    const Rsrc ru = fview(0, nu), rv = fview(1, nv), rw = fview(2, nw);
    const unsigned sy = (unsigned)NX * 8u, sz = (unsigned)NX * NY * 8u;
    const unsigned su = (unsigned)((NX + 1) * NY) * SZ, sv = (unsigned)(NX * (NY + 1)) * SZ, sw = (unsigned)(NX * NY) * SZ;
    const unsigned rowf = (unsigned)NX * SZ;
    const March Mh = march_of(NX, NY, NZ);
    double dacc = 0.0, acc = 0.0, sphi = 0.0;
    int x, y, z0, z1;
    march_point(Mh, NZ, w.local, w.G, x, y, z0, z1);
    if (x < NX && y < NY && z0 < z1) {
        // the residual as kb_residual (an absent neighbour is read at the cell itself with coefficient 0)
        const double dx = tb(g, X, T_W, x), dy = tb(g, Y, T_W, y);
        const double rx1 = x + 1 < NX ? tb(g, X, T_RSPD, x + 1) : 0.0, rx0 = x > 0 ? tb(g, X, T_RSPD, x) : 0.0;
        const double ry1 = y + 1 < NY ? tb(g, Y, T_RSPD, y + 1) : 0.0, ry0 = y > 0 ? tb(g, Y, T_RSPD, y) : 0.0;
        const unsigned oe = x + 1 < NX ? 8u : 0u, ow = x > 0 ? 8u : 0u, on = y + 1 < NY ? sy : 0u, os = y > 0 ? sy : 0u;
        unsigned i = ((unsigned)x + (unsigned)NX * ((unsigned)y + (unsigned)NY * (unsigned)z0)) * 8u;
        unsigned iu = (unsigned)ix.u(x, y, z0) * SZ, iv = (unsigned)ix.v(x, y, z0) * SZ, iw = (unsigned)ix.w(x, y, z0) * SZ;
        double prev = z0 > 0 ? bload(rp, i - sz) : 0.0, c = bload(rp, i);
        double rzm = z0 > 0 ? tb(g, Z, T_RSPD, z0) : 0.0;
        double w0 = (double)fld<F>(rw, iw);
        auto layer = [&](double pc, double pm, double pp, double pe, double pwv, double pn, double ps, double rv,
                         double dz, double rzp, double rzm_) {
            double s = 0.0;
            s += (dy * dz) * (pe - pc) * rx1;
            s -= (dy * dz) * (pc - pwv) * rx0;
            s += (dx * dz) * (pn - pc) * ry1;
            s -= (dx * dz) * (pc - ps) * ry0;
            s += (dx * dy) * (pp - pc) * rzp;
            s -= (dx * dy) * (pc - pm) * rzm_;
            return rv - s;
        };
        // backend.cu:334-342 from the face values of one layer
        auto div = [&](F u0, F u1, F v0, F v1, double wl, double wh, double dz) {
            return ((double)u1 - (double)u0) * (dy * dz) + ((double)v1 - (double)v0) * (dx * dz) + (wh - wl) * (dx * dy);
        };
        int z = z0;
        for (; z + 1 < z1; z += 2, i += 2 * sz, iu += 2 * su, iv += 2 * sv, iw += 2 * sw) {
            const unsigned i2 = i + sz;                       // layer z + 1 (< z1 <= NZ)
            const bool up2 = z + 2 < NZ;
            const double n1 = bload(rp, i2), n2 = bload(rp, up2 ? i2 + sz : i2);
            const double pe1 = bload(rp, i + oe), pw1 = bload(rp, i - ow), pn1 = bload(rp, i + on), ps1 = bload(rp, i - os);
            const double pe2 = bload(rp, i2 + oe), pw2 = bload(rp, i2 - ow), pn2 = bload(rp, i2 + on), ps2 = bload(rp, i2 - os);
            const F ua = fld<F>(ru, iu), ub = fld<F>(ru, iu + SZ), va = fld<F>(rv, iv), vb = fld<F>(rv, iv + rowf);
            const F uc = fld<F>(ru, iu + su), ud = fld<F>(ru, iu + su + SZ), vc = fld<F>(rv, iv + sv), vd = fld<F>(rv, iv + sv + rowf);
            const double w1 = (double)fld<F>(rw, iw + sw), w2 = (double)fld<F>(rw, iw + 2 * sw);
            const double dz1 = tb(g, Z, T_W, z), rzp1 = tb(g, Z, T_RSPD, z + 1);
            const double dz2 = tb(g, Z, T_W, z + 1), rzp2 = up2 ? tb(g, Z, T_RSPD, z + 2) : 0.0;
            const double d1 = div(ua, ub, va, vb, w0, w1, dz1), d2 = div(uc, ud, vc, vd, w1, w2, dz2);
            const double r1 = layer(c, prev, n1, pe1, pw1, pn1, ps1, d1, dz1, rzp1, rzm);
            const double r2 = layer(n1, c, n2, pe2, pw2, pn2, ps2, d2, dz2, rzp2, rzp1);
            bstore(rh, i, d1);
            bstore(rh, i2, d2);
            rstore(rr, i, r1, rf);
            rstore(rr, i2, r2, rf);
            dacc += d1 * d1;
            dacc += d2 * d2;
            acc += r1 * r1;
            sphi += c;
            acc += r2 * r2;
            sphi += n1;
            prev = n1;
            c = n2;
            rzm = rzp2;
            w0 = w2;
        }
        if (z < z1) {                                         // odd chunk length: the last layer
            const bool up = z + 1 < NZ;
            const double next = bload(rp, up ? i + sz : i);
            const double pe = bload(rp, i + oe), pwv = bload(rp, i - ow), pn = bload(rp, i + on), ps = bload(rp, i - os);
            const F ua = fld<F>(ru, iu), ub = fld<F>(ru, iu + SZ), va = fld<F>(rv, iv), vb = fld<F>(rv, iv + rowf);
            const double w1 = (double)fld<F>(rw, iw + sw);
            const double dz = tb(g, Z, T_W, z), rzp = up ? tb(g, Z, T_RSPD, z + 1) : 0.0;
            const double d = div(ua, ub, va, vb, w0, w1, dz);
            const double ri = layer(c, prev, next, pe, pwv, pn, ps, d, dz, rzp, rzm);
            bstore(rh, i, d);
            rstore(rr, i, ri, rf);
            dacc += d * d;
            acc += ri * ri;
            sphi += c;
        }
    }
    dacc = block_reduce(1, dacc);
    acc = block_reduce(1, acc);
    sphi = block_reduce(1, sphi);
    if (threadIdx.x == 0) {
        gp(M.partial)[P_MAIN * RED_BLOCKS + w.local] = dacc;
        gp(M.partial)[P_GUESS * RED_BLOCKS + w.local] = acc;
        gp(M.partial)[P_PHI * RED_BLOCKS + w.local] = sphi;
    }
}

// Start of solve_pressure (backend.cu:816-833): |rhs|, the guess residual, where the refinement starts.
__global__ __launch_bounds__(1024) void kb_proj_begin(const Mem* const* __restrict__ D,
                                                      const int* __restrict__ seg_march, int guess) {
    const GLB Mem& M = *gp(D[blockIdx.x]);
    const int count = seg_march[blockIdx.x + 1] - seg_march[blockIdx.x];
    const double rhs2 = fold1024(gp(M.partial) + P_MAIN * RED_BLOCKS, count, 1);
    const double res2 = guess ? fold1024(gp(M.partial) + P_GUESS * RED_BLOCKS, count, 1) : 0.0;
    if (threadIdx.x != 0) return;
    gp(M.scal)[S_RHS2] = rhs2;
    gp(M.scal)[S_RES2] = res2;
    GLB Ctl& c = *gp(M.ctl);
    c.passes = 0;
    c.zero_rhs = 0;
    c.zero_start = 0;
    c.r_rhs = 0;
    c.active = 0;
    c.run_xf = 0;
    if (c.err || c.stalled) return;
    const double rn = sqrt(rhs2);
    c.rhs_norm = rn;
    if (!isfinite(rn)) { c.err = ERR_RHS_NAN; return; }
    if (rn == 0.0) {
        c.zero_rhs = 1;
        c.zero_start = 1;
        c.rel = 0.0;
        return;
    }
    double r2 = rhs2, rel = 1.0;
    int r_rhs = 1, zs = 1;
    if (guess) {
        rel = sqrt(res2) / rn;
        r2 = res2;
        r_rhs = 0;
        zs = 0;
        if (rel >= 1.0) { zs = 1; r_rhs = 1; r2 = rhs2; rel = 1.0; }
    }
    c.rel = rel;
    c.r2 = r2;
    c.r_rhs = r_rhs;
    c.zero_start = zs;
    c.active = rel > c.rtol && 0 < c.max_passes;
    c.rms = sqrt(r2 / (double)M.n_cell);
    c.run_xf = c.active && c.rms > 0;
}

// phi = 0 where the refinement starts from zero (no guess, a guess worse than zero, or rhs = 0)
__global__ __launch_bounds__(TPB) void kb_zero_start(const Mem* const* __restrict__ D, const int* __restrict__ seg,
                                                     int nm, int ph) {
    const Where w = locate(seg, nm);
    const GLB Mem& M = *gp(D[w.m]);
    const GLB Ctl& c = *gp(M.ctl);
    const int rot = c.rot;
    if (!c.zero_start) return;
    GLB double* p = cbuf(M, rot, ph);
    // the float transforms take their first operand from the float residual array, the rhs on a cold start
    const bool cvt = c.r_rhs && M.mode != 0;
    const GLB double* __restrict__ rhs = cbuf(M, rot, CR_RHS);
    GLB float* __restrict__ rf = (GLB float*)cbuf(M, rot, CR_RES);
    for (i64 i = (i64)w.local * TPB + threadIdx.x; i < M.n_cell; i += (i64)w.G * TPB) {
        p[i] = 0.0;
        if (cvt) rf[i] = (float)rhs[i];
    }
}

// End of one refinement pass (backend.cu:834-840): the new residual norm, convergence, the next pass' scale
// (one thread, after the block has folded the pass' residual partials into r2)
__device__ void pass_end_member(const GLB Mem& M, double r2) {
    GLB Ctl& c = *gp(M.ctl);
    gp(M.scal)[S_RES2] = r2;
    c.passes++;
    c.r2 = r2;
    const double rel = sqrt(r2) / c.rhs_norm;
    c.rel = rel;
    if (!isfinite(rel)) {
        c.err = ERR_RESIDUAL_NAN;
        c.active = 0;
        c.run_xf = 0;
        return;
    }
    c.active = rel > c.rtol && c.passes < c.max_passes;
    c.rms = sqrt(r2 / (double)M.n_cell);
    c.run_xf = c.active && c.rms > 0;
}
__global__ __launch_bounds__(1024) void kb_pass_end(const Mem* const* __restrict__ D, const int* __restrict__ seg_march) {
    const GLB Mem& M = *gp(D[blockIdx.x]);
    if (!gp(M.ctl)->active) return;
    const double r2 = fold1024(gp(M.partial) + P_MAIN * RED_BLOCKS, seg_march[blockIdx.x + 1] - seg_march[blockIdx.x], 1);
    if (threadIdx.x == 0) pass_end_member(M, r2);
}

// End of the projection (backend.cu:841-847, 926-939): the phi sum for the mean (from the last residual's
// partials; phi = 0 when the right-hand side is 0), convergence failure, counters and the warm-start history
// entry. A member whose refinement would continue past the passes enqueued stops here (stalled): its later
// kernels in the step leave the state untouched and the host redoes the step.
// (one thread, after the block has folded the last residual's phi partials into sum)
__device__ void proj_end_member(const GLB Mem& M, double sum, int hist, int usual) {
    GLB Ctl& c = *gp(M.ctl);
    gp(M.scal)[S_PHISUM] = c.zero_rhs ? 0.0 : sum;
    c.fin_buf = hist ? c.out_buf : -1;
    if (c.err || c.stalled) return;
    if (!c.zero_rhs && c.active) {      // more passes allowed than were enqueued: the step is redone
        c.stalled = 1;
        return;
    }
    if (!c.zero_rhs) {
        c.projections++;
        c.pass_total += c.passes;
        c.phi_passes = c.passes;
        if (c.passes > usual) c.spare_mark = c.steps + 1;
        if (c.rel > c.rtol) {
            c.err = ERR_NOT_CONVERGED;
            c.err_val = c.rel;
            c.err_passes = c.passes;
            return;
        }
        c.phi_rel = c.rel;
    } else {
        c.phi_passes = 0;
        c.phi_rel = 0.0;
    }
    if (hist && c.out_buf >= 0) {
        if (c.hn == 3) {
            for (int k = 0; k < 2; k++) { c.ht[k] = c.ht[k + 1]; c.hs[k] = c.hs[k + 1]; c.hb[k] = c.hb[k + 1]; }
            c.hn = 2;
        }
        c.ht[c.hn] = c.ts;
        c.hs[c.hn] = c.b * c.dt;
        c.hb[c.hn] = c.out_buf;
        c.hn++;
    }
}
// The last pass' end, the projection's end and the next stage's control (next_stage 1, 2; 3: the next step's
// first-stage plan) of each member in one launch (block m: member m), in that order
__global__ __launch_bounds__(1024) void kb_proj_close(const Mem* const* __restrict__ D, const int* __restrict__ seg_march,
                                                      int pass_end, int hist, int usual, int next_stage) {
    const GLB Mem& M = *gp(D[blockIdx.x]);
    const int count = seg_march[blockIdx.x + 1] - seg_march[blockIdx.x];
    // (every thread reads the flag before thread 0 can change it: the fold's barriers come first)
    if (pass_end && gp(M.ctl)->active) {
        const double r2 = fold1024(gp(M.partial) + P_MAIN * RED_BLOCKS, count, 1);
        if (threadIdx.x == 0) pass_end_member(M, r2);
    }
    const double sum = fold1024(gp(M.partial) + P_PHI * RED_BLOCKS, count, 1);
    if (threadIdx.x != 0) return;
    // (the first stage has refined its guess in place: a retry after an error forms it again)
    if (next_stage == 1) gp(M.ctl)->g0_ready = 0;
    proj_end_member(M, sum, hist, usual);
    if (next_stage == 3) next_step_plan_member(M);
    else if (next_stage > 0) stage_control_member(M, next_stage);
}

// End of the step (backend.cu:948-954), after its last kernel: the clock, and the stage-3 buffers become
// the state. A member that failed or stalled keeps its clock and state.
__global__ void kb_step_end(const Mem* const* __restrict__ D, int nm) {
    for (int m = threadIdx.x; m < nm; m += blockDim.x) {
        GLB Ctl& c = *gp(gp(D[m])->ctl);
        if (c.err || c.stalled) continue;
        c.t += c.dt;
        c.steps++;
        c.last_dt = c.dt;
        c.last_cfl = c.dt * c.use_rate;
        c.parity ^= 1;
        c.rot = next_rot(c.rot);          // stage 3's potential becomes the state's
        c.g0_ready = c.nguess;            // and the next step's first base and guess are formed
        c.m0_ready = 1;                   // the last stage's gradient formed the new state's gauge reset
        c.cfl_ready = 1;                  // and its CFL rate
    }
}

// backend.cu:846 (subtract the mean), backend.cu:936 (dphi = phi - base into the stage's history buffer) and
// backend.cu:361-371 (out = mm - grad(phi) on every face, wall faces 0), in one z-march over the cell columns and
// the nz + 1 face levels. The refinement leaves its potential in role pw; this kernel reads it (never writes it:
// neighbours read by other threads stay as they were) and writes the mean-free potential into role ph and, with
// hist, dphi into the stage's history buffer. Every potential value it uses is p - mean, the value the
// reference's subtract_constant stores and its subtract_gradient reads, so the faces are those of the two
// separate kernels bit for bit. Column (x, y) writes its u(x), v(y) and w(z) faces; the last column and row also
// write the wall faces u(nx) and v(ny) (always 0), so no tile of the march is a lone lattice column or row.
// Branch-free: reads through buffer views (a neighbour that does not exist reads 0 or an unused value), a value
// that is not stored goes to an out-of-range offset, which the hardware drops.
// gz (the last stage of a step) adds the next step's start, from the state this projection completes, unless the
// step has stalled or failed (its state stays the old one, whose m0 is still in place):
//   - the gauge reset (backend.cu:880-882), m0 = r mm + (1 - r) out on every face and phi0 = r phi, the same
//     expressions on the same values as kb_gauge, which the next step then skips;
//   - the CFL rate (backend.cu:465-472): each cell's rate from its six faces (the three upper ones formed again
//     from the same values as the threads that store them), the block maximum (NaN wins), and the member's
//     maximum by an atomic max on the bits of a non-negative double (NaN above every number), which kb_dt then
//     uses instead of a kb_cfl pass over the velocity. A maximum does not depend on its order, so the rate is
//     kb_cfl's to the bit.
template <class F>
__global__ __launch_bounds__(TPB) void kb_gradient(const Mem* const* __restrict__ D, const int* __restrict__ seg,
                                                   int nm, int mm, int pw, int ph, int outr, int gz, int hist, int nxt,
                                                   int br) {
    constexpr unsigned SZ = sizeof(F);
    const Where w = locate(seg, nm);
    const GLB Mem& M = *gp(D[w.m]);
    const int par = gp(M.ctl)->parity;
    const int rot = gp(M.ctl)->rot;
    const Geo g = geo_of(M.g);
    const int NX = g.nx, NY = g.ny, NZ = g.nz;
    const Ix ix{NX, NY, NZ};
    const i64 nc = (i64)NX * NY * NZ;
    const i64 nw = (i64)NX * NY * (NZ + 1);
    const Rsrc rp = rsrc_of(cbuf(M, rot, pw), nc);
    // the three face arrays of a role through one buffer view (they follow one another in the arena, at the same
    // byte offsets in every role): fewer scalar registers than a view per array. Offsets of v and w faces add
    // offv, offw where they are valid, and an access that is to be dropped stays at OOB.
    const unsigned offv = (unsigned)((const GLB char*)fbuf(M, par, mm, 1) - (const GLB char*)fbuf(M, par, mm, 0));
    const unsigned offw = (unsigned)((const GLB char*)fbuf(M, par, mm, 2) - (const GLB char*)fbuf(M, par, mm, 0));
    const unsigned gbytes = offw + (unsigned)(nw * SZ);
    auto gview = [&](int role) {
        return __builtin_amdgcn_make_buffer_rsrc((void*)fbuf(M, par, role, 0), (short)0, (int)gbytes, 0x00020000);
    };
    const Rsrc gm = gview(mm), go = gview(outr);
    const GLB Ctl& cz = *gp(M.ctl);
    const int ob = hist ? cz.fin_buf : -1;
    const Rsrc rb = rsrc_of(cbuf(M, rot, br), nc);            // this stage's base (dphi = phi - base)
    const Rsrc rh = rsrc_of(cbuf(M, rot, CR_H0 + (ob >= 0 ? ob : 0)), ob >= 0 ? nc : 0);
    const bool wphi = ph >= 0;
    const Rsrc rf = rsrc_of(cbuf(M, rot, wphi ? ph : pw), wphi ? nc : 0);
    // the next stage's base and guess (kb_base_guess for stage nxt, whose control has run; nxt 3: the next step's
    // first stage, whose plan the projection's end formed (Ctl::nguess), when this step completes): its work
    // potential is the other stage buffer (nxt 3: the state's array, CR_PS0 once the step completes), and its
    // potential from the previous stage is this projection's (nxt 3: phi0 = r phi), in registers
    const bool g0 = gz && !cz.stalled && !cz.err;
    const bool nb = nxt == 3 ? g0 && cz.nguess : nxt > 0;
    const Rsrc rg = rsrc_of(cbuf(M, rot, nxt == 1 ? CR_PS1 : nxt == 2 ? CR_PS0 : CR_PHI), nb ? nc : 0);
    // (the next stage's base; a step's first stage has phi0 itself as its base, read from CR_PHI0, so none is
    // stored for it)
    const Rsrc rbn = rsrc_of(cbuf(M, rot, CR_BASE), nb && nxt != 3 ? nc : 0);
    const Rsrc r0 = rsrc_of(cbuf(M, rot, CR_PHI0), nc);        // (read here; the gauge reset below writes it)
    const double na0 = nxt == 3 ? 0.0 : cz.a0, nbb = nxt == 3 ? 1.0 : cz.b;     // (stage 1: a0 0, b 1)
    const double gc1 = nxt == 3 ? cz.ngc1 : cz.gc1, gc2 = nxt == 3 ? cz.ngc2 : cz.gc2;
    const int gnh = nb ? (nxt == 3 ? cz.ngnh : cz.gnh) : 0, gb1 = nxt == 3 ? cz.ngb1 : cz.gb1,
              gb2 = nxt == 3 ? cz.ngb2 : cz.gb2;
    const Rsrc rh1 = rsrc_of(cbuf(M, rot, CR_H0 + gb1), gnh >= 1 && gb1 != ob ? nc : 0);
    const Rsrc rh2 = rsrc_of(cbuf(M, rot, CR_H0 + gb2), gnh == 2 && gb2 != ob ? nc : 0);
    const double mean = gp(M.scal)[S_PHISUM] / (double)M.n_cell;
    const Rsrc gz0 = gview(FR_M0);
    const F r = (F)M.r, omr = (F)M.omr;                   // (float) in the FP32 fields (backend.cu:881)
    const March Mh = grad_march_of(NX, NY, NZ, M.tune[TU_GRAD_TARGET], M.tune[TU_GRAD_MINZ], M.tune[TU_GRAD_TILE]);
    int x, y, z0, z1;
    march_point_tiled(Mh, NZ + 1, w.local, w.G, x, y, z0, z1);
    if (x >= NX || y >= NY) { x = 0; y = 0; z1 = z0; }    // no column: an empty march (the block reduces below)
    const bool uwall = x == 0, vwall = y == 0;             // u(0), v(0); u(nx), v(ny) below
    const bool ue = x == NX - 1, ve = y == NY - 1;
    const double rx = tb(g, X, T_RSP, x), ry = tb(g, Y, T_RSP, y);
    const double rx1 = tb(g, X, T_RSP, x + 1), ry1 = tb(g, Y, T_RSP, y + 1);
    const F rwx = (F)tb(g, X, T_RW, x), rwy = (F)tb(g, Y, T_RW, y);
    const unsigned rowb = (unsigned)NX * 8u, rowf = (unsigned)NX * SZ;     // a row of cells, of v faces
    const unsigned lay8 = (unsigned)NX * NY * 8u;                                   // a cell layer
    // out = m - dphi / spacing: FP64 fields multiply by the reciprocal table; the FP32 fields divide in FP64 by the
    // float spacing and round the quotient to float (backend.cu:361-371)
    auto grad = [](F m, double dphi, double t) -> F {
        if constexpr (sizeof(F) == 8) return m - dphi * t;
        else return m - (F)(dphi / t);
    };
    double rate = -INFINITY;
    // (CFL rate: the pending cell's u and v part and |w| of its lower face; the lane whose x + 1 is not the next
    // lane of its wave)
    F uvp = F(0), awp = F(0);
    bool pend = false;
    const bool redge = (threadIdx.x & 63) == 63 || (threadIdx.x & (Mh.tw - 1)) == (unsigned)(Mh.tw - 1);
    double pprev = bload(rp, z0 > 0 && z0 < z1 ? (unsigned)ix.c(x, y, z0 - 1) * 8u : OOB) - mean;
    // one level of the march, compiled apart with and without the next step's start (the CFL rate's lane
    // exchange would keep the plain march from being unrolled; the march with it is unrolled by hand)
    auto level = [&](int z, auto gzc) {
        constexpr bool GZ = decltype(gzc)::value;
        const bool layer = z < NZ;
        const unsigned ic = (unsigned)ix.c(x, y, z) * 8u;
        const double pc = bload(rp, layer ? ic : OOB) - mean;
        const double pwst = bload(rp, ic - 8u) - mean, ps = bload(rp, ic - rowb) - mean;
        const unsigned iu = (unsigned)ix.u(x, y, z) * SZ, iv = (unsigned)ix.v(x, y, z) * SZ, iw = (unsigned)ix.w(x, y, z) * SZ;
        const F mu = fld<F>(gm, layer ? iu : OOB), mv = fld<F>(gm, layer ? iv + offv : OOB), mw = fld<F>(gm, iw + offw);
        const F gu = uwall ? F(0) : grad(mu, pc - pwst, rx);
        const F gv = vwall ? F(0) : grad(mv, pc - ps, ry);
        const F gw = (z == 0 || z == NZ) ? F(0) : grad(mw, pc - pprev, tb(g, Z, T_RSP, z));
        fst(go, layer ? iu : OOB, gu);
        fst(go, layer ? iv + offv : OOB, gv);
        fst(go, iw + offw, gw);
        fst(go, ue && layer ? iu + SZ : OOB, F(0));               // u(nx, y, z)
        fst(go, ve && layer ? iv + rowf + offv : OOB, F(0));      // v(x, ny, z)
        if (wphi) bstore(rf, layer ? ic : OOB, pc);                                                 // uniform
        const double dphi = 1.0 * pc + -1.0 * bload(rb, ob >= 0 && layer ? ic : OOB);
        if (ob >= 0) bstore(rh, layer ? ic : OOB, dphi);                                            // uniform
        if (nb) {                                                                                   // uniform
            // backend.cu:922-925 with guess_phi (backend.cu:870-871): base = a0 phi0 + b phi, then the guess;
            // a history entry this kernel has just written is taken from the register
            // (the previous stage's potential: this projection's pc, or for the next step's first stage the
            // phi0 = r pc formed below)
            const double p0 = nxt == 3 ? 0.0 : bload(r0, layer ? ic : OOB);
            const double h1 = gb1 == ob ? dphi : bload(rh1, layer ? ic : OOB);
            const double h2 = gb2 == ob ? dphi : bload(rh2, layer ? ic : OOB);
            const double cur = nxt == 3 ? M.r * pc : pc;
            const double bs = na0 != 0.0 ? na0 * p0 + nbb * cur : nbb * cur;
            double gs = bs;
            if (gnh == 1) gs = gc1 * h1 + bs;
            else if (gnh == 2) gs = (gc1 * h1 + gc2 * h2) + bs;
            if (nxt != 3) bstore(rbn, layer ? ic : OOB, bs);                                      // uniform
            bstore(rg, layer ? ic : OOB, gs);
        }
        if constexpr (GZ) {
            const F mue = fld<F>(gm, layer ? iu + SZ : OOB), mve = fld<F>(gm, layer ? iv + rowf + offv : OOB);
            fst(gz0, layer ? iu : OOB, r * mu + omr * gu);
            fst(gz0, layer ? iv + offv : OOB, r * mv + omr * gv);
            fst(gz0, iw + offw, r * mw + omr * gw);
            fst(gz0, ue && layer ? iu + SZ : OOB, r * mue + omr * F(0));
            fst(gz0, ve && layer ? iv + rowf + offv : OOB, r * mve + omr * F(0));
            bstore(r0, layer ? ic : OOB, M.r * pc);
            // the CFL rate of the cell below this level (its upper w face is this level's)
            if (pend) rate = comb(0, rate, (double)(uvp + dv<F>(fmax(awp, fabs(gw)), (F)tb(g, Z, T_RW, z - 1))));
            // this layer's cell: u(x + 1) from the next lane where it holds column x + 1 of the same row, else
            // formed here as that thread forms it; v(y + 1) formed here (0 on a wall)
            const F gun = __shfl_down(gu, 1, 64);
            const double pe = bload(rp, redge && !ue && layer ? ic + 8u : OOB) - mean;
            const double pn = bload(rp, ve || !layer ? OOB : ic + rowb) - mean;
            const F gu1 = ue ? F(0) : redge ? grad(mue, pe - pc, rx1) : gun;
            const F gv1 = ve ? F(0) : grad(mve, pn - pc, ry1);
            uvp = dv<F>(fmax(fabs(gu), fabs(gu1)), rwx) + dv<F>(fmax(fabs(gv), fabs(gv1)), rwy);
            awp = fabs(gw);
            pend = layer;
        }
        pprev = pc;
    };
    if (g0) {                                              // uniform over the member
        int z = z0;
        for (; z + 1 < z1; z += 2) {                       // (two levels per trip, written out: the lane exchange
            level(z, std::true_type{});                    // keeps the compiler from unrolling)
            level(z + 1, std::true_type{});
        }
        if (z < z1) level(z, std::true_type{});
    } else {
#pragma unroll 4
        for (int z = z0; z < z1; z++) level(z, std::false_type{});
    }
    if (g0) {
        if (pend) {                 // the chunk's last cell (below level z1 <= nz): its upper w face formed here
            const unsigned ic = (unsigned)ix.c(x, y, z1 - 1) * 8u;
            const bool wt = z1 == NZ;
            const double pu = bload(rp, wt ? OOB : ic + lay8) - mean;
            const F gw1 = wt ? F(0) : grad(fld<F>(gm, (unsigned)ix.w(x, y, z1) * SZ + offw), pu - pprev, tb(g, Z, T_RSP, z1));
            rate = comb(0, rate, (double)(uvp + dv<F>(fmax(awp, fabs(gw1)), (F)tb(g, Z, T_RW, z1 - 1))));
        }
        rate = block_reduce(0, rate);
        // the member's rate: every cell rate is >= +0 or NaN, so the bits order as unsigned integers do (a block
        // with no cells, at -inf, adds nothing)
        if (threadIdx.x == 0 && rate != -INFINITY)
            __hip_atomic_fetch_max(&M.ctl->cfl_bits, __builtin_bit_cast(unsigned long long, rate), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
    }
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
// 16-row matrix fragments per wave along the output index (a block covers 32 JF outputs). 4 doubles the MFMA
// work per staged chunk and halves the input re-reads, but measured no faster than 2 at 256^3 in FP64.
__host__ __device__ constexpr int xf_jf(int) { return 2; }
__host__ __device__ inline int pad_for_mode(int n, int) { return (n + XBK - 1) / XBK * XBK; }
// whether the chunked even/odd forward transform computes both halves in one block: the padded half fits the
// 16 xf_jf rows one wave column of the block covers, on the y and z axes (on the x axis, whose rows a thread
// stages as a run of PER values, staging both folds measured slower than reading the values twice)
__host__ __device__ inline bool sym_pair(bool xl, int hpad, int mode) { return !xl && hpad <= 16 * xf_jf(mode); }

__device__ __forceinline__ dbl4 mfma(double a, double b, dbl4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
__device__ __forceinline__ flt4 mfma(float a, float b, flt4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
__device__ __forceinline__ flt4 mfma(bf8 a, bf8 b, flt4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }

// Matrix fragments of one staged chunk of XBK contraction indices for a wave's two row tiles mt0, mt0 + 1
// (packed, contraction padded to npad). Loaded (mload) before the chunk's LDS staging so that their L2
// latency overlaps the barrier and the stores.
template <int MODE, class T> struct MFrag;
template <class T> struct MFrag<0, T> { T m[xf_jf(0)][XBK / 4]; };
template <class T> struct MFrag<1, T> { T m[xf_jf(1)][XBK / 4]; };
template <class T> struct MFrag<2, T> { bf8 hi[xf_jf(2)], lo[xf_jf(2)]; };

template <int MODE, class T>
__device__ __forceinline__ void mload(MFrag<MODE, T>& f, const void* mpk, const void* mpk_lo, int mt0,
                                      const bool (&jv)[xf_jf(MODE)], int npad, int k0, int lane) {
    constexpr int JF = xf_jf(MODE);
    if constexpr (MODE < 2) {
        const GLB T* mp = (const GLB T*)mpk;
        const int kts = npad / 4;
#pragma unroll
        for (int jf = 0; jf < JF; jf++)
#pragma unroll
            for (int ks = 0; ks < XBK / 4; ks++)
                f.m[jf][ks] = jv[jf] ? mp[((i64)(mt0 + jf) * kts + k0 / 4 + ks) * 64 + lane] : (T)0;
    } else {
        const GLB bf8* mhi = (const GLB bf8*)mpk;
        const GLB bf8* mlo = (const GLB bf8*)mpk_lo;
        const int kts = npad / 32, kt = k0 / 32;
#pragma unroll
        for (int jf = 0; jf < JF; jf++) {
            const i64 at = ((i64)(mt0 + jf) * kts + kt) * 64 + lane;
            if (jv[jf]) { f.hi[jf] = mhi[at]; f.lo[jf] = mlo[at]; }
        }
    }
}

// Accumulate one staged chunk into a wave's 2 x 2 fragment block: data from tile columns dc .. dc+31,
// matrix fragments f.
template <bool XL, int MODE, class T, class Acc>
__device__ __forceinline__ void mma_chunk(Acc (&acc)[2][xf_jf(MODE)], T (*tile)[64 + 2], int dc, int lane,
                                          const MFrag<MODE, T>& mf, const bool (&jv)[xf_jf(MODE)]) {
    constexpr int BK = XBK, JF = xf_jf(MODE);
    if constexpr (MODE < 2) {
#pragma unroll
        for (int ks = 0; ks < BK / 4; ks++) {
            const T d[2] = {tile[ks * 4 + lane / 16][dc + lane % 16], tile[ks * 4 + lane / 16][dc + 16 + lane % 16]};
#pragma unroll
            for (int jf = 0; jf < JF; jf++) {
                if (!jv[jf]) continue;
#pragma unroll
                for (int f = 0; f < 2; f++)
                    acc[f][jf] = XL ? mfma(d[f], mf.m[jf][ks], acc[f][jf]) : mfma(mf.m[jf][ks], d[f], acc[f][jf]);
            }
        }
    } else {
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
        for (int jf = 0; jf < JF; jf++) {
            if (!jv[jf]) continue;
#pragma unroll
            for (int f = 0; f < 2; f++) {
                if (XL) {
                    acc[f][jf] = mfma(dl[f], mf.hi[jf], acc[f][jf]);
                    acc[f][jf] = mfma(dh[f], mf.lo[jf], acc[f][jf]);
                    acc[f][jf] = mfma(dh[f], mf.hi[jf], acc[f][jf]);
                } else {
                    acc[f][jf] = mfma(mf.lo[jf], dh[f], acc[f][jf]);
                    acc[f][jf] = mfma(mf.hi[jf], dl[f], acc[f][jf]);
                    acc[f][jf] = mfma(mf.hi[jf], dh[f], acc[f][jf]);
                }
            }
        }
    }
}
// Element offset of index 0 of column col in the [A][n][B] layout (col = a B + b -> a n B + b). 32-bit: every
// array of a cavity holds fewer than 2^29 values.
__device__ __forceinline__ unsigned col_base(unsigned col, unsigned n, unsigned B) {
    const unsigned aa = col / B;
    return aa * n * B + (col - aa * B);
}

// result row within the 16 x 16 tile of register r: FP64 interleaves the lane groups (4 r + l/16), the
// FP32-accumulating instructions stack them (4 (l/16) + r)
template <int MODE> __device__ __forceinline__ int result_row(int r, int lane) {
    return MODE == 0 ? 4 * r + lane / 16 : 4 * (lane / 16) + r;
}

template <bool XL, int MODE, class TIN, class TOUT>
__device__ __forceinline__ void axis_tile(const AxisArgs& a, i64 ncols, int npad, int L) {
    using T = typename XT<MODE>::T;
    using Acc = typename XT<MODE>::Acc;
    constexpr int BK = XBK, PER = BK * 64 / 256;          // staged values per thread and tile
    constexpr int JF = xf_jf(MODE), JB = 32 * JF;          // output fragments per wave, outputs per block
    __shared__ T tile[BK][64 + 2];
    const GLB TIN* __restrict__ in = (const GLB TIN*)a.in;
    GLB TOUT* __restrict__ out = (GLB TOUT*)a.out;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // logical tile L = (column tile, j tile) with the j tiles of one column tile adjacent and on one XCD,
    // so the input tile they share is fetched once into that L2
    const int nby = (npad + JB - 1) / JB;
    const int by = L % nby;
    const i64 c0 = (i64)(L / nby) * 64;
    if (c0 >= ncols) return;
    const int wr = wave >> 1, wc = wave & 1;
    const int n = a.n;
    const int jt0 = (by * JB + wc * 16 * JF) / 16;
    bool jv[JF];
#pragma unroll
    for (int k = 0; k < JF; k++) jv[k] = 16 * (jt0 + k) < npad;
    Acc acc[2][JF];
    for (int f = 0; f < 2; f++)
        for (int jf = 0; jf < JF; jf++)
            for (int r = 0; r < 4; r++) acc[f][jf][r] = 0;

    // Staging map. XL: thread -> row c0 + tid/4, PER consecutive i from (tid%4)*PER.
    // [A][n][B]: thread -> 4 consecutive columns (tid%16)*4, rows tid/16 and tid/16 + 16.
    const int xcol = tid >> 2, xig = (tid & 3) * PER;
    const i64 xrow = c0 + xcol;
    const unsigned B = (unsigned)a.B, uc0 = (unsigned)c0, un = (unsigned)n;
    // [A][n][B] staging: one column per lane (tid % 64), rows tid / 64 + 4 k: each load instruction reads 64
    // consecutive values of one row
    const unsigned scol = uc0 + (tid & 63);
    const bool cok = !XL && scol < ncols;
    const unsigned cb = cok ? col_base(scol, un, B) : 0u;
    const int srow = tid >> 6, scl = tid & 63;
    T pre[PER];
    auto load = [&](int i0) {
        if (XL) {
#pragma unroll
            for (int q = 0; q < PER; q++) {
                const int i = i0 + xig + q;
                pre[q] = (xrow < ncols && i < n) ? (T)(in[(unsigned)xrow * un + i] * a.in_scale) : (T)0;
            }
        } else {
#pragma unroll
            for (int k = 0; k < PER; k++) {
                const int i = i0 + srow + 4 * k;
                pre[k] = (cok && i < n) ? (T)(in[cb + (unsigned)i * B] * a.in_scale) : (T)0;
            }
        }
    };
    auto store = [&]() {
        if (XL) {
#pragma unroll
            for (int q = 0; q < PER; q++) tile[xig + q][xcol] = pre[q];
        } else {
#pragma unroll
            for (int k = 0; k < PER; k++) tile[srow + 4 * k][scl] = pre[k];
        }
    };

    load(0);
    MFrag<MODE, T> mf;
    for (int i0 = 0; i0 < npad; i0 += BK) {
        mload<MODE, T>(mf, a.mpk, a.mpk_lo, jt0, jv, npad, i0, lane);
        __syncthreads();
        store();
        __syncthreads();
        if (i0 + BK < npad) load(i0 + BK);                  // next tile in flight during the MFMAs
        mma_chunk<XL, MODE, T, Acc>(acc, tile, wr * 32, lane, mf, jv);
    }
    const GLB T* il = (const GLB T*)a.il;
#pragma unroll
    for (int f = 0; f < 2; f++) {
        const unsigned col = uc0 + wr * 32 + f * 16 + lane % 16;          // [A][n][B]: this lane's column
        const unsigned cb = XL ? 0u : col_base(col < ncols ? col : 0u, un, B);
#pragma unroll
        for (int jf = 0; jf < JF; jf++) {
            if (!jv[jf]) continue;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                unsigned idx;
                const int rr = result_row<MODE>(r, lane);
                if (XL) {
                    const unsigned row = uc0 + wr * 32 + f * 16 + rr;
                    const int j = 16 * (jt0 + jf) + lane % 16;
                    if (row >= ncols || j >= n) continue;
                    idx = row * un + j;
                } else {
                    const int j = 16 * (jt0 + jf) + rr;
                    if (col >= ncols || j >= n) continue;
                    idx = cb + (unsigned)j * B;
                }
                const T val = acc[f][jf][r];
                if (a.epi == 0) out[idx] = (TOUT)val;
                else if (a.epi == 1) out[idx] = (TOUT)(val * il[idx]);
                else out[idx] = out[idx] + (TOUT)((double)val * a.out_scale);
            }
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
__device__ __forceinline__ void axis_symf_tile(const AxisArgs& a, i64 ncols, int hpad, int L,
                                               typename XT<MODE>::T (*tile)[64 + 2]) {
    using T = typename XT<MODE>::T;
    using Acc = typename XT<MODE>::Acc;
    constexpr int BK = XBK, PER = BK * 64 / 256, JF = xf_jf(MODE), JB = 32 * JF;
    const GLB TIN* __restrict__ in = (const GLB TIN*)a.in;
    GLB TOUT* __restrict__ out = (GLB TOUT*)a.out;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int n = a.n, h = n / 2, nbh = (hpad + JB - 1) / JB;
    // PAIR (hpad <= JB / 2): both halves in one block, wave column wc computing half wc, from one read of the
    // tile's values (both folds staged); else a block computes rows byh JB .. of one half
    const bool pair = sym_pair(XL, hpad, MODE);
    const int by = pair ? 0 : L % (2 * nbh), byh = pair ? 0 : by % nbh;
    const i64 c0 = (i64)(L / (pair ? 1 : 2 * nbh)) * 64;
    if (c0 >= ncols) return;
    const int wr = wave >> 1, wc = wave & 1;
    const int odd = pair ? wc : by / nbh;
    const int jt0 = pair ? 0 : (byh * JB + wc * 16 * JF) / 16;
    bool jv[JF];
#pragma unroll
    for (int k = 0; k < JF; k++) jv[k] = 16 * (jt0 + k) < hpad;
    const double sgn = odd ? -1.0 : 1.0;
    Acc acc[2][JF];
    for (int f = 0; f < 2; f++)
        for (int jf = 0; jf < JF; jf++)
            for (int r = 0; r < 4; r++) acc[f][jf][r] = 0;
    const int xcol = tid >> 2, xig = (tid & 3) * PER;
    const i64 xrow = c0 + xcol;
    const unsigned B = (unsigned)a.B, uc0 = (unsigned)c0, un = (unsigned)n;
    // [A][n][B] staging: one column per lane (tid % 64), rows tid / 64 + 4 k: each load instruction reads 64
    // consecutive values of one row
    const unsigned scol = uc0 + (tid & 63);
    const bool cok = !XL && scol < ncols;
    const unsigned cb = cok ? col_base(scol, un, B) : 0u;
    const int srow = tid >> 6, scl = tid & 63;
    // (r_i s + sgn (r_{n-1-i} s) with sgn = +-1 rounds exactly as the sum or difference of the two products: the
    // pair's even and odd folds are the single-half values)
    T pre[PER], pro[PER];
    auto fold = [&](double lo, double hi, T& e, T& o) {
        if (pair) { e = (T)(lo + hi); o = (T)(lo - hi); }
        else e = (T)(lo + sgn * hi);
    };
    auto load = [&](int i0) {
        if (XL) {
#pragma unroll
            for (int q = 0; q < PER; q++) {
                const int i = i0 + xig + q;
                T e = 0, o = 0;
                if (xrow < ncols && i < h)
                    fold(in[(unsigned)xrow * un + i] * a.in_scale, in[(unsigned)xrow * un + (n - 1 - i)] * a.in_scale, e, o);
                pre[q] = e;
                pro[q] = o;
            }
        } else {
#pragma unroll
            for (int k = 0; k < PER; k++) {
                const int i = i0 + srow + 4 * k;
                T e = 0, o = 0;
                if (cok && i < h)
                    fold(in[cb + (unsigned)i * B] * a.in_scale, in[cb + (unsigned)(n - 1 - i) * B] * a.in_scale, e, o);
                pre[k] = e;
                pro[k] = o;
            }
        }
    };
    auto store = [&]() {
        if (XL) {
#pragma unroll
            for (int q = 0; q < PER; q++) {
                tile[xig + q][xcol] = pre[q];
                if (pair) tile[XBK + xig + q][xcol] = pro[q];
            }
        } else {
#pragma unroll
            for (int k = 0; k < PER; k++) {
                tile[srow + 4 * k][scl] = pre[k];
                if (pair) tile[XBK + srow + 4 * k][scl] = pro[k];
            }
        }
    };
    const void* mp = odd ? a.mpk2 : a.mpk;
    const void* mpl = odd ? a.mpk2_lo : a.mpk_lo;
    load(0);
    MFrag<MODE, T> mf;
    for (int i0 = 0; i0 < hpad; i0 += BK) {
        mload<MODE, T>(mf, mp, mpl, jt0, jv, hpad, i0, lane);
        __syncthreads();
        store();
        __syncthreads();
        if (i0 + BK < hpad) load(i0 + BK);
        mma_chunk<XL, MODE, T, Acc>(acc, tile + (pair && odd ? XBK : 0), wr * 32, lane, mf, jv);
    }
    const GLB T* il = (const GLB T*)a.il;
#pragma unroll
    for (int f = 0; f < 2; f++) {
        const unsigned col = uc0 + wr * 32 + f * 16 + lane % 16;          // [A][n][B]: this lane's column
        const unsigned cb = XL ? 0u : col_base(col < ncols ? col : 0u, un, B);
#pragma unroll
        for (int jf = 0; jf < JF; jf++) {
            if (!jv[jf]) continue;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int rr = result_row<MODE>(r, lane);
                unsigned idx;
                if (XL) {
                    const unsigned row = uc0 + wr * 32 + f * 16 + rr;
                    const int j = 16 * (jt0 + jf) + lane % 16;
                    if (row >= ncols || j >= h) continue;
                    idx = row * un + odd * h + j;
                } else {
                    const int j = 16 * (jt0 + jf) + rr;
                    if (col >= ncols || j >= h) continue;
                    idx = cb + (unsigned)(odd * h + j) * B;
                }
                const T val = acc[f][jf][r];
                if (a.epi == 1) out[idx] = (TOUT)(val * il[idx]);
                else out[idx] = (TOUT)val;
            }
        }
    }
}

constexpr int symb_rows(int mode) { return XBK > 16 * xf_jf(mode) ? XBK : 16 * xf_jf(mode); }
// tile: [even | odd] spectral chunk [2][symb_rows][64 + 2], reused for the results
template <bool XL, int MODE, class TIN, class TOUT>
__device__ __forceinline__ void axis_symb_tile(const AxisArgs& a, i64 ncols, int hpad, int L,
                                               typename XT<MODE>::T (*tile)[symb_rows(MODE)][64 + 2]) {
    using T = typename XT<MODE>::T;
    using Acc = typename XT<MODE>::Acc;
    constexpr int BK = XBK, PER = BK * 64 / 256, JF = xf_jf(MODE), JR = 16 * JF;   // JR output rows per block
    constexpr int TR = symb_rows(MODE);
    const GLB TIN* __restrict__ in = (const GLB TIN*)a.in;
    GLB TOUT* __restrict__ out = (GLB TOUT*)a.out;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int n = a.n, h = n / 2, nbi = (hpad + JR - 1) / JR;
    const int by = L % nbi;
    const i64 c0 = (i64)(L / nbi) * 64;
    if (c0 >= ncols) return;
    const int part = wave >> 1, dsub = wave & 1;   // part 0 even, 1 odd; data columns dsub*32 .. +31
    const int mt0 = by * JF;                         // output rows i = JR by .. JR by + JR - 1
    bool jv[JF];
#pragma unroll
    for (int k = 0; k < JF; k++) jv[k] = 16 * (mt0 + k) < hpad;
    Acc acc[2][JF];
    for (int f = 0; f < 2; f++)
        for (int jf = 0; jf < JF; jf++)
            for (int r = 0; r < 4; r++) acc[f][jf][r] = 0;
    const int xcol = tid >> 2, xig = (tid & 3) * PER;
    const i64 xrow = c0 + xcol;
    const unsigned B = (unsigned)a.B, uc0 = (unsigned)c0, un = (unsigned)n;
    // [A][n][B] staging: one column per lane (tid % 64), rows tid / 64 + 4 k: each load instruction reads 64
    // consecutive values of one row
    const unsigned scol = uc0 + (tid & 63);
    const bool cok = !XL && scol < ncols;
    const unsigned cb = cok ? col_base(scol, un, B) : 0u;
    const int srow = tid >> 6, scl = tid & 63;
    T pre[2][PER];
    auto load = [&](int k0) {
#pragma unroll
        for (int p = 0; p < 2; p++) {
            if (XL) {
#pragma unroll
                for (int q = 0; q < PER; q++) {
                    const int k = k0 + xig + q;
                    pre[p][q] = (xrow < ncols && k < h) ? (T)in[(unsigned)xrow * un + p * h + k] : (T)0;
                }
            } else {
#pragma unroll
                for (int kk = 0; kk < PER; kk++) {
                    const int k = k0 + srow + 4 * kk;
                    pre[p][kk] = (cok && k < h) ? (T)in[cb + (unsigned)(p * h + k) * B] : (T)0;
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
                for (int kk = 0; kk < PER; kk++) tile[p][srow + 4 * kk][scl] = pre[p][kk];
            }
        }
    };
    const void* mp = part ? a.mpk2 : a.mpk;
    const void* mpl = part ? a.mpk2_lo : a.mpk_lo;
    load(0);
    MFrag<MODE, T> mf;
    for (int k0 = 0; k0 < hpad; k0 += BK) {
        mload<MODE, T>(mf, mp, mpl, mt0, jv, hpad, k0, lane);
        __syncthreads();
        store();
        __syncthreads();
        if (k0 + BK < hpad) load(k0 + BK);
        mma_chunk<XL, MODE, T, Acc>(acc, tile[part], dsub * 32, lane, mf, jv);
    }
    // exchange: res[part][row i within 32][data column within 64], aliased onto the staging tiles
    __syncthreads();
    T (*res)[TR][64 + 2] = tile;
#pragma unroll
    for (int f = 0; f < 2; f++)
#pragma unroll
        for (int jf = 0; jf < JF; jf++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int rr = result_row<MODE>(r, lane);
                if (XL) res[part][jf * 16 + lane % 16][dsub * 32 + f * 16 + rr] = acc[f][jf][r];
                else res[part][jf * 16 + rr][dsub * 32 + f * 16 + lane % 16] = acc[f][jf][r];
            }
    __syncthreads();
    // [A][n][B]: the column of element e is c0 + tid % 64 at every e (256 is a multiple of 64)
    const unsigned ecol = uc0 + (tid & 63);
    const unsigned ecb = XL ? 0u : col_base(ecol < ncols ? ecol : 0u, un, B);
    for (int e = tid; e < JR * 64; e += 256) {
        const int il_ = XL ? e % JR : e / 64, cl = XL ? e / JR : e % 64;
        const int i = by * JR + il_;
        const unsigned col = uc0 + cl;
        if (i >= h || col >= ncols) continue;
        const T ev = res[0][il_][cl], ov = res[1][il_][cl];
        const T vlo = ev + ov, vhi = ev - ov;
        unsigned lo, hi;
        if (XL) {
            lo = col * un + i;
            hi = col * un + (n - 1 - i);
        } else {
            lo = ecb + (unsigned)i * B;
            hi = ecb + (unsigned)(n - 1 - i) * B;
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

// ------------------------------------------------------------------------------------------------ symmetric transform by whole columns
// The even/odd transform for h = n/2 <= SC_H (n <= 256) by tiles of SC_BM columns, each computing every output
// of both halves: a tile's columns are read once (the forward folds r_lo +- rev r_hi while staging) and
// staged whole in LDS. Blocks of SC_T threads are persistent (sc_resident per CU): each wave holds its matrix
// fragments for the whole contraction in registers, loaded once, and loads the next tile's columns into
// registers while its MFMAs work on the current one, so HBM traffic and MFMA work overlap. Each wave owns two
// output slots of 16 rows: in the forward transform wave w contracts half w/4 for the output fragments
// w%4 and w%4 + 4; in the backward transform it contracts both halves for output fragment w and unfolds them
// in registers. Each output sums the same products in the same order as axis_symf_tile / axis_symb_tile, so
// the results are identical.
constexpr int SC_T = 512, SC_W = SC_T / 64, SC_NS = 2;      // threads and waves per block, slots per wave
constexpr int SC_BM = 32, SC_NF = SC_BM / 16, SC_H = 128, SC_LD = SC_BM + 2, SC_NP = SC_H * SC_BM / SC_T,
              SC_KS = SC_H / 4;
constexpr int SC_CUS = 256;                                   // compute units of an MI355X
__host__ __device__ inline bool sym_cols(int n, int mode) { return pad_for_mode(n / 2, mode) <= SC_H; }
// resident blocks per CU (two waves per SIMD: the FP64 matrix slots take 128 registers of a wave)
__host__ __device__ constexpr int sc_resident(int) { return 1; }

template <int MODE> struct SymColsTile {
    using T = typename XT<MODE>::T;
    T t[2][SC_H][SC_LD];               // [even | odd][contraction index][column]
};
// Split BF16: the staged values already split into bf16 hi + lo (once per value, not once per wave that reads
// it), [even | odd][column][contraction index], so a lane's eight consecutive indices of one MFMA operand are one
// 16-byte read. The row pad keeps the 16 columns one read instruction touches on distinct banks.
constexpr int SC_KLD = SC_H + 8;
template <> struct __align__(16) SymColsTile<2> {
    __bf16 hi[2][SC_BM][SC_KLD], lo[2][SC_BM][SC_KLD];
};
// this wave's matrix fragments for every contraction step, per output slot
template <int MODE> struct ScMat { typename XT<MODE>::T m[SC_NS][SC_KS]; };
template <> struct ScMat<2> { bf8 hi[SC_NS][SC_H / 32], lo[SC_NS][SC_H / 32]; };
// a tile's staged values in flight (q: value pair of this thread)
template <class TIN> struct ScPre { TIN lo[SC_NP], hi[SC_NP]; };

// Slot s of wave w: half `part`, output fragment `mt` (16 rows)
__device__ __forceinline__ void sc_slot(bool fwd, int wave, int s, int& part, int& mt) {
    if (fwd) { part = wave / (SC_W / 2); mt = wave % (SC_W / 2) + (SC_W / 2) * s; }
    else { part = s; mt = wave; }
}
// Pair q of this thread: contraction index k and tile column c. XL: wave w takes columns w + SC_W j, lane l
// indices l, l + 64; else thread tid owns column tid % SC_BM, rows tid / SC_BM + (SC_T / SC_BM) q (a load
// instruction reads 64 / SC_BM rows of SC_BM consecutive values).
// the smallest contraction index of pair q over the block (pairs from hpad on are skipped, uniformly)
template <bool XL> __host__ __device__ constexpr int sc_pair_k0(int q) {
    return XL ? (q % (SC_H / 64)) * 64 : (SC_T / SC_BM) * q;
}
template <bool XL> __device__ __forceinline__ void sc_pair(int q, int& k, int& c) {
    const int tid = threadIdx.x;
    if (XL) {
        c = (tid >> 6) + SC_W * (q / (SC_H / 64));
        k = (q % (SC_H / 64)) * 64 + (tid & 63);
    } else {
        c = tid & (SC_BM - 1);
        k = tid / SC_BM + (SC_T / SC_BM) * q;
    }
}
// Issue the loads of tile c0's values (FWD: r_k and r_{n-1-k}; else the spectral halves in[k], in[h + k]).
// Values that do not exist are read at index 0 and dropped by sc_store: no select on a loaded value, so
// the loads stay in flight.
// Buffer-view access to the transform arrays (fewer than 2^32 bytes each): 32-bit byte offsets, an offset
// at or past the end (OOB) loads 0 and stores nothing, so values that do not exist need no branches.
template <class T> __device__ __forceinline__ Rsrc xrsrc(const void* p, i64 n) {
    return __builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, (int)(n * (i64)sizeof(T)), 0x00020000);
}
template <class T> __device__ __forceinline__ T xload(Rsrc r, unsigned off);
template <> __device__ __forceinline__ double xload<double>(Rsrc r, unsigned off) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
template <> __device__ __forceinline__ float xload<float>(Rsrc r, unsigned off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
__device__ __forceinline__ void xstore(Rsrc r, unsigned off, double v) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned, v), r, off, 0, 0);
}
__device__ __forceinline__ void xstore(Rsrc r, unsigned off, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, off, 0, 0);
}

template <bool XL, class TIN, bool FWD>
__device__ __forceinline__ void sc_load(const AxisArgs& a, i64 ncols, Rsrc rin, unsigned c0, ScPre<TIN>& p) {
    const int n = a.n, h = n / 2;
    const unsigned B = (unsigned)a.B, un = (unsigned)n, sz = sizeof(TIN);
    unsigned cb = 0u;
    if (!XL) {                          // the thread's column is the same for all its pairs
        int k, c;
        sc_pair<XL>(0, k, c);
        cb = col_base(c0 + c < ncols ? c0 + c : 0u, un, B);
    }
#pragma unroll
    for (int q = 0; q < SC_NP; q++) {
        if (sc_pair_k0<XL>(q) >= h) continue;                               // uniform
        int k, c;
        sc_pair<XL>(q, k, c);
        const unsigned col = c0 + c;
        const bool ok = col < ncols && k < h;
        const unsigned k2 = FWD ? (unsigned)(n - 1 - k) : (unsigned)(h + k);
        const unsigned base = XL ? col * un : cb, st = XL ? 1u : B;
        p.lo[q] = xload<TIN>(rin, ok ? (base + (unsigned)k * st) * sz : OOB);
        p.hi[q] = xload<TIN>(rin, ok ? (base + k2 * st) * sz : OOB);
    }
}
// Fold (FWD: (r_k s +- r_{n-1-k} s)) and store the tile's values: tile[p][k][c] for k < hpad, 0 beyond h
template <bool XL, int MODE, class TIN, bool FWD>
__device__ __forceinline__ void sc_store(const AxisArgs& a, i64 ncols, int hpad, unsigned c0, const ScPre<TIN>& p,
                                         SymColsTile<MODE>& S) {
    using T = typename XT<MODE>::T;
    const int h = a.n / 2;
    const double s = a.in_scale;
#pragma unroll
    for (int q = 0; q < SC_NP; q++) {
        if (sc_pair_k0<XL>(q) >= hpad) continue;                            // uniform
        int k, c;
        sc_pair<XL>(q, k, c);
        if (k >= hpad) continue;
        const bool ok = c0 + c < ncols && k < h;
        T e = 0, o = 0;
        if (sc_pair_k0<XL>(q) < h) {    // uniform: sc_load filled this pair (ok selects its values)
            if (FWD) {
                const double l = (double)p.lo[q] * s, r = (double)p.hi[q] * s;
                e = (T)(l + r);
                o = (T)(l - r);
            } else {
                e = (T)p.lo[q];
                o = (T)p.hi[q];
            }
        }
        e = ok ? e : (T)0;
        o = ok ? o : (T)0;
        if constexpr (MODE == 2) {      // x = hi + lo: hi = bf16(x), lo = bf16(x - hi)
            const __bf16 eh = (__bf16)e, oh = (__bf16)o;
            S.hi[0][c][k] = eh;
            S.lo[0][c][k] = (__bf16)(e - (float)eh);
            S.hi[1][c][k] = oh;
            S.lo[1][c][k] = (__bf16)(o - (float)oh);
        } else {
            S.t[0][k][c] = e;
            S.t[1][k][c] = o;
        }
    }
}// Load this wave's matrix fragments for the contraction steps that exist (rows past hpad read row fragment 0
// and are never used). Through buffer loads: the compiler cannot prove these invariant, so it keeps them in
// registers rather than loading them again inside the tile loop (where such loads would queue behind the next
// tile's prefetch).
template <int MODE>
__device__ __forceinline__ void sc_matrix(const AxisArgs& a, int hpad, bool fwd, ScMat<MODE>& M) {
    // (the wave index is uniform over the wave: read as such, the matrix half it selects is a scalar buffer view
    // instead of a per-lane one that every load would loop over)
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
#pragma unroll
    for (int sl = 0; sl < SC_NS; sl++) {
        int part, mt;
        sc_slot(fwd, wave, sl, part, mt);
        const int mtc = 16 * mt < hpad ? mt : 0;
        if constexpr (MODE < 2) {
            using T = typename XT<MODE>::T;
            const int kts = hpad / 4;
            const Rsrc rm = xrsrc<T>(part ? a.mpk2 : a.mpk, (i64)(hpad / 16) * kts * 64);
#pragma unroll
            for (int ks = 0; ks < SC_KS; ks++)
                if (ks < kts) M.m[sl][ks] = xload<T>(rm, (unsigned)((mtc * kts + ks) * 64 + lane) * sizeof(T));   // uniform
        } else {
            const int kts = hpad / 32;
            const i64 nb = (i64)(hpad / 16) * kts * 64 * 16;             // bytes of one packed half (bf16 x 8)
            const Rsrc rh = __builtin_amdgcn_make_buffer_rsrc((void*)(part ? a.mpk2 : a.mpk), (short)0, (int)nb, 0x00020000);
            const Rsrc rl = __builtin_amdgcn_make_buffer_rsrc((void*)(part ? a.mpk2_lo : a.mpk_lo), (short)0, (int)nb, 0x00020000);
#pragma unroll
            for (int kt = 0; kt < SC_H / 32; kt++) {
                if (kt >= kts) continue;                                     // uniform
                const unsigned at = (unsigned)((mtc * kts + kt) * 64 + lane) * 16u;
                M.hi[sl][kt] = __builtin_bit_cast(bf8, __builtin_amdgcn_raw_buffer_load_b128(rh, at, 0, 0));
                M.lo[sl][kt] = __builtin_bit_cast(bf8, __builtin_amdgcn_raw_buffer_load_b128(rl, at, 0, 0));
            }
        }
    }
}
// Contract the staged tile: acc[f][sl] += M(slot sl) x data (tile columns 16 f .. 16 f + 15) over k < hpad
template <bool XL, int MODE, bool FWD, bool FULL>
__device__ __forceinline__ void sc_contract(int hpad, const ScMat<MODE>& M, const SymColsTile<MODE>& S,
                                            typename XT<MODE>::Acc (&acc)[SC_NF][SC_NS]) {
    using T = typename XT<MODE>::T;
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    bool jv[SC_NS];
#pragma unroll
    for (int sl = 0; sl < SC_NS; sl++) {
        int part, mt;
        sc_slot(FWD, wave, sl, part, mt);
        jv[sl] = FULL || 16 * mt < hpad;
    }
    if constexpr (MODE < 2) {
        const auto* t0 = S.t[FWD ? wave / (SC_W / 2) : 0];
        const auto* t1 = S.t[FWD ? wave / (SC_W / 2) : 1];
        const int kts = hpad / 4;
#pragma unroll
        for (int ks = 0; ks < SC_KS; ks++) {
            if (!FULL && ks >= kts) continue;         // uniform; a fixed trip count keeps M in registers
            const int k = 4 * ks + lane / 16;
            T d0[SC_NF], d1[SC_NF];
#pragma unroll
            for (int f = 0; f < SC_NF; f++) {
                d0[f] = t0[k][16 * f + lane % 16];
                d1[f] = FWD ? d0[f] : t1[k][16 * f + lane % 16];
            }
#pragma unroll
            for (int sl = 0; sl < SC_NS; sl++) {
                if (!jv[sl]) continue;
#pragma unroll
                for (int f = 0; f < SC_NF; f++) {
                    const T d = (FWD || sl == 0) ? d0[f] : d1[f];
                    acc[f][sl] = XL ? mfma(d, M.m[sl][ks], acc[f][sl]) : mfma(M.m[sl][ks], d, acc[f][sl]);
                }
            }
        }
    } else {
        const int kts = hpad / 32;
#pragma unroll
        for (int kt = 0; kt < SC_H / 32; kt++) {
            if (!FULL && kt >= kts) continue;
            bf8 dh[2][SC_NF], dl[2][SC_NF];          // [half][f]: indices 32 kt + 8 (lane / 16) + 0..7
#pragma unroll
            for (int hp = 0; hp < (FWD ? 1 : 2); hp++)
#pragma unroll
                for (int f = 0; f < SC_NF; f++) {
                    const int half = FWD ? wave / (SC_W / 2) : hp, c = f * 16 + lane % 16, k = 32 * kt + 8 * (lane / 16);
                    dh[hp][f] = *(const bf8*)&S.hi[half][c][k];
                    dl[hp][f] = *(const bf8*)&S.lo[half][c][k];
                }
#pragma unroll
            for (int sl = 0; sl < SC_NS; sl++) {
                if (!jv[sl]) continue;
                const int hp = (FWD || sl == 0) ? 0 : 1;
#pragma unroll
                for (int f = 0; f < SC_NF; f++) {
                    if (XL) {
                        acc[f][sl] = mfma(dl[hp][f], M.hi[sl][kt], acc[f][sl]);
                        acc[f][sl] = mfma(dh[hp][f], M.lo[sl][kt], acc[f][sl]);
                        acc[f][sl] = mfma(dh[hp][f], M.hi[sl][kt], acc[f][sl]);
                    } else {
                        acc[f][sl] = mfma(M.lo[sl][kt], dh[hp][f], acc[f][sl]);
                        acc[f][sl] = mfma(M.hi[sl][kt], dl[hp][f], acc[f][sl]);
                        acc[f][sl] = mfma(M.hi[sl][kt], dh[hp][f], acc[f][sl]);
                    }
                }
            }
        }
    }
}

// The epilogue operands of a tile in LDS: the eigenvalue factors of the z-forward transform (epi 1) or the
// potential the x-backward transform adds into (epi 2), for the tile's SC_BM columns at every position of the
// axis. They are copied from memory straight into LDS (buffer loads with the LDS flag, which take no
// registers) while the tile's MFMAs run, so that the epilogue does not wait on a memory round trip. XL:
// [column][position] (the tile's lines are one contiguous block); else [position][column] (at one position
// the tile's columns are contiguous in groups of 16 bytes: B, a multiple of nx, is even and a multiple of 4
// elements, so no group straddles the end of a row of B columns).
constexpr int SC_EPI_BYTES = SC_BM * 256 * 8;
struct __align__(16) ScEpi { unsigned char b[SC_EPI_BYTES]; };
#define LDSP __attribute__((address_space(3)))
template <bool XL, class E>
__device__ __forceinline__ void sc_epi_load(const AxisArgs& a, i64 ncols, Rsrc r, unsigned c0, ScEpi& S) {
    constexpr unsigned SZ = sizeof(E), EPL = 16 / SZ;
    const int n = a.n, lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (XL) {                           // SC_BM lines of n values from line c0 on: 1 KiB per wave and load
        const int chunks = (int)((SC_BM * (unsigned)n * SZ + 1023) / 1024);
        const unsigned base = c0 * (unsigned)n * SZ;
#pragma unroll 1
        for (int k = wave; k < chunks; k += SC_W)
            __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (LDSP void*)(S.b + k * 1024), 16, base + (unsigned)k * 1024u + lane * 16u,
                                                     0, 0, 0);
    } else {                            // per load PPI positions, LPP lanes of EPL columns each
        constexpr int LPP = SC_BM / EPL, PPI = 64 / LPP;
        const unsigned col = c0 + EPL * (lane % LPP), un = (unsigned)n, B = (unsigned)a.B;
        const unsigned cb = col < ncols ? col_base(col, un, B) : 0u;
#pragma unroll 1
        for (int p0 = PPI * wave; p0 < n; p0 += PPI * SC_W) {
            const int p = p0 + lane / LPP;
            const unsigned off = col < ncols && p < n ? (cb + (unsigned)p * B) * SZ : OOB;
            __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (LDSP void*)(S.b + p0 * SC_BM * SZ), 16, off, 0, 0, 0);
        }
    }
}

// Write tile c0's outputs, one 16-column half f at a time, branch-free through buffer views (an output that
// does not exist goes to OOB). FWD: slot sl holds half `part` rows 16 mt ..; the eigenvalue scaling (epi 1)
// loads its factors before the first store. Backward: slots 0 (even) and 1 (odd) hold the same rows, unfolded
// into phi_lo = e + o, rev phi_hi = e - o; the accumulation into phi (epi 2) loads every old value first.
// (STG: the operands are staged in E; else loaded here)
template <bool XL, int MODE, class TOUT, bool FWD, bool STG>
__device__ __forceinline__ void sc_epilogue(const AxisArgs& a, i64 ncols, Rsrc rout, Rsrc ril, unsigned c0,
                                            const typename XT<MODE>::Acc (&acc)[SC_NF][SC_NS], const ScEpi& E) {
    using T = typename XT<MODE>::T;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, n = a.n, h = n / 2;
    const unsigned B = (unsigned)a.B, un = (unsigned)n, so = sizeof(TOUT), st_ = sizeof(T);
    constexpr int NS = FWD ? SC_NS : 1;              // output slots (backward: even slot 0 with odd slot 1)
#pragma unroll
    for (int f = 0; f < SC_NF; f++) {
        const unsigned col = c0 + f * 16 + lane % 16;                // [A][n][B]: this lane's column
        const unsigned cb = XL ? 0u : col_base(col < ncols ? col : 0u, un, B);
        unsigned i0[NS][4], i1[NS][4];                              // element indices, OOB where absent
#pragma unroll
        for (int sl = 0; sl < NS; sl++) {
            int part, mt;
            sc_slot(FWD, wave, sl, part, mt);
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int rr = result_row<MODE>(r, lane);
                const unsigned row = XL ? c0 + f * 16 + rr : col;
                const int j = 16 * mt + (XL ? lane % 16 : rr);
                const bool ok = row < ncols && j < h;
                const unsigned base = XL ? row * un : cb, st = XL ? 1u : B;
                i0[sl][r] = ok ? base + (unsigned)((FWD ? part * h : 0) + j) * st : OOB;
                i1[sl][r] = ok ? base + (unsigned)(n - 1 - j) * st : OOB;
            }
        }
        auto at = [&](unsigned i, unsigned size) { return i == OOB ? OOB : i * size; };
        if (FWD) {
            if (a.epi == 1) {
                T fac[NS][4];
                if constexpr (!XL && STG) {     // [position][column]: position part h + j, column f 16 + lane % 16
                    const T* e = (const T*)E.b;
#pragma unroll
                    for (int sl = 0; sl < NS; sl++) {
                        int part, mt;
                        sc_slot(FWD, wave, sl, part, mt);
#pragma unroll
                        for (int r = 0; r < 4; r++)
                            fac[sl][r] = e[(part * h + 16 * mt + result_row<MODE>(r, lane)) * SC_BM + f * 16 + lane % 16];
                    }
                } else {
#pragma unroll
                    for (int sl = 0; sl < NS; sl++)
#pragma unroll
                        for (int r = 0; r < 4; r++) fac[sl][r] = xload<T>(ril, at(i0[sl][r], st_));
                }
#pragma unroll
                for (int sl = 0; sl < NS; sl++)
#pragma unroll
                    for (int r = 0; r < 4; r++) xstore(rout, at(i0[sl][r], so), (TOUT)(acc[f][sl][r] * fac[sl][r]));
            } else {
#pragma unroll
                for (int sl = 0; sl < NS; sl++)
#pragma unroll
                    for (int r = 0; r < 4; r++) xstore(rout, at(i0[sl][r], so), (TOUT)acc[f][sl][r]);
            }
            continue;
        }
        if (a.epi == 2) {
            TOUT o0[4], o1[4];
            if constexpr (XL && STG) {          // [column][position]: column f 16 + rr, positions j and n - 1 - j
                const TOUT* e = (const TOUT*)E.b;
                int part, mt;
                sc_slot(FWD, wave, 0, part, mt);
                const int j = 16 * mt + lane % 16, jj = j < h ? j : 0;
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int c = f * 16 + result_row<MODE>(r, lane);
                    o0[r] = e[c * n + jj];
                    o1[r] = e[c * n + n - 1 - jj];
                }
            } else {
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    o0[r] = xload<TOUT>(rout, at(i0[0][r], so));
                    o1[r] = xload<TOUT>(rout, at(i1[0][r], so));
                }
            }
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const T e = acc[f][0][r], o = acc[f][1][r];
                const T vlo = e + o, vhi = e - o;
                xstore(rout, at(i0[0][r], so), o0[r] + (TOUT)((double)vlo * a.out_scale));
                xstore(rout, at(i1[0][r], so), o1[r] + (TOUT)((double)vhi * a.out_scale));
            }
        } else {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const T e = acc[f][0][r], o = acc[f][1][r];
                xstore(rout, at(i0[0][r], so), (TOUT)(e + o));
                xstore(rout, at(i1[0][r], so), (TOUT)(e - o));
            }
        }
    }
}

// Tiles L0, L0 + Ls, ... (< T) of one member, by one persistent block. FULL: h = SC_H (every contraction step
// and output slot exists; no guards).
// STG: the z-forward (epi 1) or x-backward (epi 2) transform, whose epilogue operands are staged in Ep
template <bool XL, int MODE, class TIN, class TOUT, bool FWD, bool STG, bool FULL>
__device__ __forceinline__ void axis_sym_cols_run(const AxisArgs& a, i64 ncols, int hpad, i64 L0, i64 Ls, i64 T,
                                                  SymColsTile<MODE>& S, ScEpi& Ep) {
    using Acc = typename XT<MODE>::Acc;
    using TT = typename XT<MODE>::T;
    using EE = std::conditional_t<FWD, TT, TOUT>;      // the staged epilogue operand's type
    const i64 nc = ncols * a.n;
    const Rsrc rin = xrsrc<TIN>(a.in, nc), rout = xrsrc<TOUT>(a.out, nc), ril = xrsrc<TT>(a.il, nc);
    ScMat<MODE> M;
    sc_matrix<MODE>(a, hpad, FWD, M);
    ScPre<TIN> p;
    sc_load<XL, TIN, FWD>(a, ncols, rin, (unsigned)(L0 * SC_BM), p);
    // the matrix (and the first tile) in registers before the loop: with its loads known to be complete, the
    // loop's MFMAs carry no memory waits (the counter is in order, so such waits would also wait for the
    // previous tile's stores and the next tile's loads)
    __builtin_amdgcn_s_waitcnt(0);
    for (i64 L = L0; L < T; L += Ls) {
        const unsigned c0 = (unsigned)(L * SC_BM);
        sc_store<XL, MODE, TIN, FWD>(a, ncols, hpad, c0, p, S);
        __syncthreads();
        if (STG) sc_epi_load<XL, EE>(a, ncols, FWD ? ril : rout, c0, Ep);                       // this tile's epilogue operands
        if (L + Ls < T) sc_load<XL, TIN, FWD>(a, ncols, rin, (unsigned)((L + Ls) * SC_BM), p);   // next tile in flight
        asm volatile("" ::: "memory");
        Acc acc[SC_NF][SC_NS];
#pragma unroll
        for (int f = 0; f < SC_NF; f++)
#pragma unroll
            for (int sl = 0; sl < SC_NS; sl++)
#pragma unroll
                for (int r = 0; r < 4; r++) acc[f][sl][r] = 0;
        sc_contract<XL, MODE, FWD, FULL>(hpad, M, S, acc);
        if (STG || MODE == 2) {
            if (STG) __builtin_amdgcn_s_waitcnt(0);   // this wave's copies into LDS have landed (long since)
            __syncthreads();              // every wave's copies are visible, and every wave has read the tile
            sc_epilogue<XL, MODE, TOUT, FWD, STG>(a, ncols, rout, ril, c0, acc, Ep);
        } else {
            // The epilogue reads no LDS, so the barrier (every wave has read the tile: the next one may be stored)
            // follows it: the waves whose MFMAs finish first (the two waves on a SIMD do not finish together)
            // store their outputs while the others still contract. FP64 step 6.15 to 6.09 ms at 256^3; the
            // split-BF16 transforms, whose epilogue is shorter, do not gain and keep the barrier first.
            sc_epilogue<XL, MODE, TOUT, FWD, STG>(a, ncols, rout, ril, c0, acc, Ep);
            __syncthreads();
        }
    }
}
template <bool XL, int MODE, class TIN, class TOUT, bool FWD, bool STG>
__device__ __forceinline__ void axis_sym_cols(const AxisArgs& a, i64 ncols, int hpad, i64 L0, i64 Ls, i64 T,
                                              SymColsTile<MODE>& S, ScEpi& Ep) {
    if (L0 >= T) return;
    if (a.n / 2 == SC_H) axis_sym_cols_run<XL, MODE, TIN, TOUT, FWD, STG, true>(a, ncols, hpad, L0, Ls, T, S, Ep);
    else axis_sym_cols_run<XL, MODE, TIN, TOUT, FWD, STG, false>(a, ncols, hpad, L0, Ls, T, S, Ep);
}

// ------------------------------------------------------------------------------------------------ batched transforms
// The six transforms of one pressure pass (backend.cu:801-814), t = 0..5:
//   0: t2 = (r / rms) S_x          (r = rhs on the first pass of a cold start, else the residual; the float
//                                   transforms read both from the float residual array)
//   1: t1 = S_y^T t2      2: t2 = (S_z^T t1) * inv_lambda      3: t1 = S_z t2      4: t2 = S_y t1
//   5: phi += rms * (t2 S_x^T)
// One block per tile, the tiles of one XCD contiguous; a member whose refinement has converged (or whose
// residual is 0) skips the whole pass. Passes that are rarely needed run with few blocks looping over the
// tiles (LOOP): CAPPED_BLOCKS per launch, shared by the members, so their launches cost little when every
// member skips them.
constexpr int CAPPED_BLOCKS = 512;
inline int capped_per_member(int n) { return std::max(8, CAPPED_BLOCKS / std::max(1, n) / 8 * 8); }

__device__ __forceinline__ int xf_axis(int t) { return (t == 0 || t == 5) ? X : (t == 1 || t == 4) ? Y : Z; }

// One tile of a dense or chunked even/odd transform, with the kind's own staging tiles in LDS
template <bool XL, int MODE, class TIN, class TOUT, int KIND>
__device__ __forceinline__ void xf_chunk_tile(const AxisArgs& a, i64 ncols, int pad, int L) {
    using T = typename XT<MODE>::T;
    if constexpr (KIND == 0) {
        axis_tile<XL, MODE, TIN, TOUT>(a, ncols, pad, L);
    } else if constexpr (KIND == 1) {
        __shared__ T tl[2 * XBK][64 + 2];     // (both folds in PAIR blocks)
        axis_symf_tile<XL, MODE, TIN, TOUT>(a, ncols, pad, L, tl);
    } else {
        __shared__ T tl[2][symb_rows(MODE)][64 + 2];
        axis_symb_tile<XL, MODE, TIN, TOUT>(a, ncols, pad, L, tl);
    }
}

// transform layout of an axis: 0 dense, 1 even/odd by chunks, 2 even/odd by whole columns (axes of at most
// 2 SC_H cells)
// (whole columns need at least SC_MIN_TILES tiles per persistent block to amortise its start: resident matrix
// and pipeline fill; smaller cavities stay with the chunked kernel)
constexpr int SC_MIN_TILES = 2;
__host__ __device__ inline int xf_layout(bool sym, int n, i64 ncols, int mode) {
    if (!sym) return 0;
    const bool cols = sym_cols(n, mode) && (ncols + SC_BM - 1) / SC_BM >= (i64)SC_MIN_TILES * SC_CUS * sc_resident(mode);
    return cols ? 2 : 1;
}
// blocks of the whole-column kernel for one member: persistent, at most sc_resident per CU
__host__ __device__ inline i64 xc_blocks(i64 ncols, int mode) {
    const i64 t = (ncols + SC_BM - 1) / SC_BM, res = (i64)SC_CUS * sc_resident(mode);
    return t < res ? t : res;
}

__host__ __device__ inline i64 xf_tiles(bool xl, int n, i64 ncols, int kind, int mode) {
    const i64 ct = (ncols + 63) / 64;
    const int jb = 32 * xf_jf(mode), jr = 16 * xf_jf(mode);      // outputs per block (axis_tile / symf, symb)
    if (kind == 0) return ct * ((pad_for_mode(n, mode) + jb - 1) / jb);
    const int hpad = pad_for_mode(n / 2, mode);
    return kind == 1 ? ct * (sym_pair(xl, hpad, mode) ? 1 : 2 * ((hpad + jb - 1) / jb)) : ct * ((hpad + jr - 1) / jr);
}

// Arguments of transform t (0..5) of a pass for one member: axis, layout, buffers, packed matrices, scales
__device__ __forceinline__ void axis_args(const GLB Mem& M, const GLB Ctl& c, int t, int ph, int pass, int kind,
                                          AxisArgs& a) {
    const int axis = xf_axis(t), dir = t >= 3;
    const int nx = M.g.nx, ny = M.g.ny, nz = M.g.nz;
    a.n = axis == X ? nx : axis == Y ? ny : nz;
    a.A = axis == X ? (i64)ny * nz : axis == Y ? (i64)nz : 1;
    a.B = axis == X ? 1 : axis == Y ? (i64)nx : (i64)nx * ny;
    const void* t1 = M.C[CR_T1];
    const void* t2 = M.C[CR_T2];
    const int rot = c.rot;
    switch (t) {
        case 0: a.in = (pass == 0 && c.r_rhs && M.mode == 0) ? M.C[CR_RHS] : M.C[CR_RES]; a.out = (void*)t2; break;
        case 1: a.in = t2; a.out = (void*)t1; break;
        case 2: a.in = t1; a.out = (void*)t2; break;
        case 3: a.in = t2; a.out = (void*)t1; break;
        case 4: a.in = t1; a.out = (void*)t2; break;
        default: a.in = t2; a.out = M.C[cell_buffer(ph, rot)]; break;
    }
    a.mpk = M.mat[axis][dir][0];
    a.mpk_lo = M.mat[axis][dir][1];
    a.mpk2 = M.mat[axis][dir][2];
    a.mpk2_lo = M.mat[axis][dir][3];
    a.il = M.il;
    a.in_scale = t == 0 ? 1.0 / c.rms : 1.0;
    a.out_scale = t == 5 ? c.rms : 1.0;
    a.epi = t == 2 ? 1 : t == 5 ? 2 : 0;
    a.sym = kind;
}

// Whole-column even/odd transform (axes of at most 2 SC_H cells): persistent blocks striding over the
// member's column tiles. STG: the z-forward (t = 2) or x-backward (t = 5) transform, which stages its epilogue
// operands in LDS (for the z axis B = nx ny is a multiple of 4: nx and ny are even). Only the FP32 transforms do:
// the FP64 ones then spill 18-22 registers (and their z-forward moves 192 KB per tile, near its share of HBM
// bandwidth while the MFMAs run, so the hidden round trip gains nothing), and the split-BF16 ones do not gain.
template <bool XL, int MODE, class TIN, class TOUT, bool FWD, bool STG>
__global__ __launch_bounds__(SC_T) __attribute__((amdgpu_waves_per_eu(2 * sc_resident(MODE)))) void kb_axis_cols(
        const Mem* const* __restrict__ D, const int* __restrict__ seg, int nm, int t, int ph, int pass) {
    const Where w = locate(seg, nm);
    const GLB Mem& M = *gp(D[w.m]);
    const GLB Ctl& c = *gp(M.ctl);
    if (!c.run_xf) return;
    AxisArgs a;
    axis_args(M, c, t, ph, pass, FWD ? 1 : 2, a);
    const i64 ncols = M.n_cell / a.n;
    __shared__ SymColsTile<MODE> S;
    __shared__ ScEpi Ep;
    axis_sym_cols<XL, MODE, TIN, TOUT, FWD, STG>(a, ncols, pad_for_mode(a.n / 2, MODE), w.local, w.G,
                                           (ncols + SC_BM - 1) / SC_BM, S, Ep);
}

template <bool XL, int MODE, class TIN, class TOUT, int KIND, bool LOOP>
__global__ __launch_bounds__(256) void kb_axis(const Mem* const* __restrict__ D, const int* __restrict__ seg, int nm,
                                               int t, int ph, int pass) {
    const Where w = locate(seg, nm);
    const GLB Mem& M = *gp(D[w.m]);
    const GLB Ctl& c = *gp(M.ctl);
    if (!c.run_xf) return;
    AxisArgs a;
    axis_args(M, c, t, ph, pass, KIND, a);
    const i64 ncols = M.n_cell / a.n;
    const int pad = KIND == 0 ? pad_for_mode(a.n, MODE) : pad_for_mode(a.n / 2, MODE);
    const i64 T = xf_tiles(XL, a.n, ncols, KIND, MODE);
    const int per = (int)((T + 7) / 8), xcd = w.local % 8;
    auto tile = [&](int L) { xf_chunk_tile<XL, MODE, TIN, TOUT, KIND>(a, ncols, pad, L); };
    if (!LOOP) {                        // one tile per block
        const i64 L = (i64)xcd * per + w.local / 8;
        if (L >= T) return;
        tile((int)L);
        return;
    }
    for (int j = w.local / 8; j < per; j += w.G / 8) {
        const i64 L = (i64)xcd * per + j;
        if (L >= T) break;
        tile((int)L);
        __syncthreads();                // the next tile restages the shared tiles
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
// backend.cu:554-561: out = f + (field type)(noise s), m += a - b (in the field type)
template <class F>
__global__ __launch_bounds__(TPB) void k_add_noise(const F* __restrict__ f, const double* __restrict__ noise,
                                                   double s, F* __restrict__ out, i64 n) {
    const i64 i = (i64)blockIdx.x * TPB + threadIdx.x;
    if (i < n) out[i] = f[i] + (F)(noise[i] * s);
}
template <class F>
__global__ __launch_bounds__(TPB) void k_add_difference(F* __restrict__ m, const F* __restrict__ a,
                                                        const F* __restrict__ b, i64 n) {
    const i64 i = (i64)blockIdx.x * TPB + threadIdx.x;
    if (i < n) m[i] = m[i] + (a[i] - b[i]);
}
__global__ __launch_bounds__(TPB) void k_widen(const float* __restrict__ src, double* __restrict__ dst, i64 n) {
    for (i64 i = (i64)blockIdx.x * TPB + threadIdx.x; i < n; i += (i64)gridDim.x * TPB) dst[i] = (double)src[i];
}

}  // namespace

// Plain copy (16 B per lane per access): the calibration kernel for HBM byte counters.
__global__ __launch_bounds__(TPB) void k_copy1(const double* __restrict__ src, double* __restrict__ dst, i64 n) {
    for (i64 i = (i64)blockIdx.x * TPB + threadIdx.x; i < n; i += (i64)gridDim.x * TPB) dst[i] = src[i];
}
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
void copy_array(const Exec& ex, const double* src, double* dst, i64 n) {
    run(ex, F_KICK, k_copy1, dim3(red_blocks(n)), dim3(TPB), src, dst, n);
}

void fill(const Exec& ex, double* p, i64 n, double value) {
    run(ex, F_FILL, k_fill, dim3(red_blocks(n)), dim3(TPB), p, n, value);
}

// ------------------------------------------------------------------------------------------------ batches
int usual_passes(int mode) { return mode == 0 ? 1 : 2; }

int shape_blocks(const Mem& m, int shape, int members) {
    const int nx = m.g.nx, ny = m.g.ny, nz = m.g.nz;
    const i64 nc = m.n_cell;
    switch (shape) {
        case SH_GAUGE:
            // (grid-stride, so any count gives the same values; a batch shares the cap, since in most steps
            // every block exits at once and an empty block still costs its dispatch)
            return round8(std::min<i64>((i64)blocks_for((i64)(nx + 1) * ny * nz) + blocks_for((i64)nx * (ny + 1) * nz)
                                            + blocks_for((i64)nx * ny * (nz + 1)) + blocks_for(nc),
                                        std::max(256, GAUGE_BLOCKS / std::max(1, members))));
        case SH_MARCH: return march_blocks(march_of(nx, ny, nz));
        case SH_STAGE:
        case SH_STAGE16: {
            const int tile = m.tune[TU_STAGE_TILE] == 1 ? 1 : 0;
            if (tile != (shape == SH_STAGE16 ? 1 : 0)) return 0;
            return march_blocks(stage_march_of(nx, ny, nz, tile, m.tune[TU_STAGE_TARGET], m.tune[TU_STAGE_MINZ]));
        }
        case SH_CELLS: return round8(blocks_for(nc));
        case SH_WALLS: return round8(blocks_for(2 * ((i64)ny * nz + (i64)nx * nz + (i64)nx * ny)));
        case SH_RED: return round8(red_blocks(nc));
        case SH_GRAD:
            return march_blocks(grad_march_of(nx, ny, nz, m.tune[TU_GRAD_TARGET], m.tune[TU_GRAD_MINZ], m.tune[TU_GRAD_TILE]));
        default: break;
    }
    const int cap = capped_per_member(members);
    if (shape == SH_MARCH_CAP) return std::min(march_blocks(march_of(nx, ny, nz)), cap);
    const bool capped = shape >= SH_XF_CAP;
    const int k = shape - (capped ? SH_XF_CAP : SH_XF), axis = k / 6, dir = (k / 3) % 2, lay = k % 3;
    if (k < 0 || k >= 18) return 0;
    const int n = axis == X ? nx : axis == Y ? ny : nz;
    if (lay != transform_layout(m, axis)) return 0;
    const i64 T = lay == 2 ? xc_blocks(nc / n, m.mode) : xf_tiles(axis == X, n, nc / n, lay ? 1 + dir : 0, m.mode);
    if ((i64)round8(T) > INT_MAX / 2) fail("grid too large for one transform launch");
    return capped ? (int)std::min<i64>(round8(T), cap) : round8(T);
}

bool whole_columns_possible(const Mem& m, int axis) {
    const int n = axis == X ? m.g.nx : axis == Y ? m.g.ny : m.g.nz;
    return m.sym[axis] != 0 && sym_cols(n, m.mode);
}
int transform_layout(const Mem& m, int axis) {
    const int n = axis == X ? m.g.nx : axis == Y ? m.g.ny : m.g.nz;
    if (!m.sym[axis]) return 0;
    const int pick = (m.tune[TU_XF] >> (2 * axis)) & 3;
    if (pick == 1) return 1;
    if (pick == 2 && whole_columns_possible(m, axis)) return 2;
    return xf_layout(true, n, m.n_cell / n, m.mode);
}

void batch_segments(const std::vector<const Mem*>& members, std::vector<int>& seg, int* total) {
    seg.assign((size_t)SH_COUNT * (members.size() + 1), 0);
    batch_segments(members.data(), (int)members.size(), seg.data(), total);
}
void batch_segments(const Mem* const* members, int n, int* seg, int* total) {
    for (int s = 0; s < SH_COUNT; s++) {
        int acc = 0;
        for (int m = 0; m < n; m++) {
            seg[(size_t)s * (n + 1) + m] = acc;
            const int nb = shape_blocks(*members[m], s, n);
            if ((i64)acc + nb > INT_MAX) fail("batch too large for one launch");
            acc += nb;
        }
        seg[(size_t)s * (n + 1) + n] = acc;
        total[s] = acc;
    }
}

namespace {
inline const int* row(const Batch& b, int shape) { return b.d_seg + (size_t)shape * (b.n + 1); }
template <class K, class... Args>
void brun(const Batch& b, int family, int shape, K kernel, Args... args) {
    run(b.ex, family, kernel, dim3((unsigned)b.total[shape]), dim3(TPB), b.d_mem, row(b, shape), b.n, args...);
}
}  // namespace

// the FP32-field mode's face arrays hold floats; every other mode's hold doubles
inline bool f32_fields(const Batch& b) { return b.mode == CAV_PREC_FIELD_FP32; }

void b_gauge(const Batch& b) {
    if (f32_fields(b)) brun(b, F_GAUGE, SH_GAUGE, kb_gauge<float>);
    else brun(b, F_GAUGE, SH_GAUGE, kb_gauge<double>);
}

void b_cfl_dt(const Batch& b) {
    if (f32_fields(b)) brun(b, F_CFL, SH_MARCH, kb_cfl<float>);
    else brun(b, F_CFL, SH_MARCH, kb_cfl<double>);
    run(b.ex, F_CONTROL, kb_dt, dim3(b.n), dim3(1024), b.d_mem, row(b, SH_MARCH), 1);
}
void b_cfl_rate(const Batch& b) {
    if (f32_fields(b)) brun(b, F_CFL, SH_MARCH, kb_cfl<float>);
    else brun(b, F_CFL, SH_MARCH, kb_cfl<double>);
    run(b.ex, F_CONTROL, kb_dt, dim3(b.n), dim3(1024), b.d_mem, row(b, SH_MARCH), 0);
}


void b_stage(const Batch& b, int st) {
    if (f32_fields(b)) {
        brun(b, F_STAGE, SH_STAGE, kb_stage<STX, STY, float>, st);
        brun(b, F_STAGE, SH_STAGE16, kb_stage<16, 16, float>, st);
    } else {
        brun(b, F_STAGE, SH_STAGE, kb_stage<STX, STY, double>, st);
        brun(b, F_STAGE, SH_STAGE16, kb_stage<16, 16, double>, st);
    }
}

void b_base_guess(const Batch& b, int st) { brun(b, F_GUESS, SH_GAUGE, kb_base_guess, st); }

namespace {

void b_xform(const Batch& b, int t, int ph, int pass) {
    const int axis = t == 0 || t == 5 ? X : (t == 1 || t == 4) ? Y : Z, dir = t >= 3;
    const int family = axis == X ? F_XFORM_X : axis == Y ? F_XFORM_Y : F_XFORM_Z;
    const bool capped = pass >= usual_passes(b.mode);
    for (int lay = 0; lay < 3; lay++) {
        const int shape = (capped ? SH_XF_CAP : SH_XF) + 6 * axis + 3 * dir + lay;
        if (!b.total[shape]) continue;
        const int kind = lay ? 1 + dir : 0;
        const dim3 grid((unsigned)b.total[shape]);
#define CAV_XF1(XL, M, TI, TO, K)                                                                            \
    if (capped) brun(b, family, shape, kb_axis<XL, M, TI, TO, K, true>, t, ph, pass);                       \
    else brun(b, family, shape, kb_axis<XL, M, TI, TO, K, false>, t, ph, pass);
#define CAV_XC(XL, M, TI, TO)                                                                                 \
    if (kind == 1 && t == 2)                                                                                  \
        run(b.ex, family, kb_axis_cols<XL, M, TI, TO, true, !XL && M == 1>, grid, dim3(SC_T), b.d_mem, row(b, shape), b.n, \
            t, ph, pass);                                                                                      \
    else if (kind == 1)                                                                                        \
        run(b.ex, family, kb_axis_cols<XL, M, TI, TO, true, false>, grid, dim3(SC_T), b.d_mem, row(b, shape), b.n, \
            t, ph, pass);                                                                                      \
    else if (t == 5)                                                                                           \
        run(b.ex, family, kb_axis_cols<XL, M, TI, TO, false, XL && M == 1>, grid, dim3(SC_T), b.d_mem, row(b, shape), b.n, \
            t, ph, pass);                                                                                      \
    else run(b.ex, family, kb_axis_cols<XL, M, TI, TO, false, false>, grid, dim3(SC_T), b.d_mem, row(b, shape), \
             b.n, t, ph, pass);
#define CAV_XF(XL, M, TI, TO)                                                                                 \
    if (lay == 2) { CAV_XC(XL, M, TI, TO) }                                                                   \
    else if (kind == 0) { CAV_XF1(XL, M, TI, TO, 0) }                                                        \
    else if (kind == 1) { CAV_XF1(XL, M, TI, TO, 1) }                                                        \
    else { CAV_XF1(XL, M, TI, TO, 2) }
        if (b.mode == 0) {
            if (axis == X) { CAV_XF(true, 0, double, double) } else { CAV_XF(false, 0, double, double) }
        } else if (b.mode == 1 || b.mode == CAV_PREC_FIELD_FP32) {
            if (t == 0) { CAV_XF(true, 1, float, float) }
            else if (t == 5) { CAV_XF(true, 1, float, double) }
            else { CAV_XF(false, 1, float, float) }
        } else {
            if (t == 0) { CAV_XF(true, 2, float, float) }
            else if (t == 5) { CAV_XF(true, 2, float, double) }
            else { CAV_XF(false, 2, float, float) }
        }
#undef CAV_XF
#undef CAV_XC
#undef CAV_XF1
    }
}
}  // namespace

void b_step_end(const Batch& b) { run(b.ex, F_CONTROL, kb_step_end, dim3(1), dim3(64), b.d_mem, b.n); }

void b_project(const Batch& b, int mm, int pw, int ph, int out, bool guess, bool hist, int passes, bool gauge_next,
               bool walls, int next_stage, int base) {
    const bool f32 = f32_fields(b);
    if (walls) {
        if (f32) brun(b, F_FILL, SH_WALLS, kb_zero_walls<float>, mm);
        else brun(b, F_FILL, SH_WALLS, kb_zero_walls<double>, mm);
    }
    if (guess) {               // the divergence and the guess residual in one march
        if (f32) brun(b, F_DIVERGENCE, SH_MARCH, kb_div_residual<float, true>, mm, pw);
        else if (b.mode != 0) brun(b, F_DIVERGENCE, SH_MARCH, kb_div_residual<double, true>, mm, pw);
        else brun(b, F_DIVERGENCE, SH_MARCH, kb_div_residual<double, false>, mm, pw);
    } else {
        if (f32) brun(b, F_DIVERGENCE, SH_MARCH, kb_divergence<float>, mm);
        else brun(b, F_DIVERGENCE, SH_MARCH, kb_divergence<double>, mm);
    }
    run(b.ex, F_CONTROL, kb_proj_begin, dim3(b.n), dim3(1024), b.d_mem, row(b, SH_MARCH), guess ? 1 : 0);
    brun(b, F_FILL, SH_MARCH, kb_zero_start, pw);       // grid-stride: usually it only reads the flag
    for (int p = 0; p < passes; p++) {
        for (int t = 0; t < 6; t++) b_xform(b, t, pw, p);
        const int store = p + 1 < passes ? 1 : 0;     // the last pass' residual is only measured
        // (a float residual where the transforms take float operands; the last pass' residual stores nothing)
        const bool rf = b.mode != 0 && store;
        if (p >= usual_passes(b.mode)) {
            if (rf) brun(b, F_RESIDUAL, SH_MARCH_CAP, kb_residual<true, true>, pw, (int)P_MAIN, 1, store);
            else brun(b, F_RESIDUAL, SH_MARCH_CAP, kb_residual<true, false>, pw, (int)P_MAIN, 1, store);
        } else {
            if (rf) brun(b, F_RESIDUAL, SH_MARCH, kb_residual<false, true>, pw, (int)P_MAIN, 1, store);
            else brun(b, F_RESIDUAL, SH_MARCH, kb_residual<false, false>, pw, (int)P_MAIN, 1, store);
        }
        if (p + 1 < passes) run(b.ex, F_CONTROL, kb_pass_end, dim3(b.n), dim3(1024), b.d_mem, row(b, SH_MARCH));
    }
    // the last pass' end, the projection's end and the next stage's control in one launch
    run(b.ex, F_CONTROL, kb_proj_close, dim3(b.n), dim3(1024), b.d_mem, row(b, SH_MARCH), passes > 0 ? 1 : 0,
        hist ? 1 : 0, b.usual, next_stage);
    b_gradient(b, mm, pw, ph, out, gauge_next, hist, next_stage, base);
}

void b_gradient(const Batch& b, int mm, int pw, int ph, int out, bool gauge_next, bool hist, int next_stage,
                int base) {
    const int gz = gauge_next ? 1 : 0, hs = hist ? 1 : 0, nx = next_stage > 0 ? next_stage : 0;
    if (f32_fields(b)) brun(b, F_GRADIENT, SH_GRAD, kb_gradient<float>, mm, pw, ph, out, gz, hs, nx, base);
    else brun(b, F_GRADIENT, SH_GRAD, kb_gradient<double>, mm, pw, ph, out, gz, hs, nx, base);
}

void b_transforms(const Batch& b, int ph, int pass) {
    for (int t = 0; t < 6; t++) b_xform(b, t, ph, pass);
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
    run(ex, F_KICK, k_add_noise<double>, dim3(blocks_for(n)), dim3(TPB), f, noise, s, out, n);
}
void add_difference(const Exec& ex, double* m, const double* a, const double* b, i64 n) {
    run(ex, F_KICK, k_add_difference<double>, dim3(blocks_for(n)), dim3(TPB), m, a, b, n);
}
void add_noise(const Exec& ex, const float* f, const double* noise, double s, float* out, i64 n) {
    run(ex, F_KICK, k_add_noise<float>, dim3(blocks_for(n)), dim3(TPB), f, noise, s, out, n);
}
void add_difference(const Exec& ex, float* m, const float* a, const float* b, i64 n) {
    run(ex, F_KICK, k_add_difference<float>, dim3(blocks_for(n)), dim3(TPB), m, a, b, n);
}
void widen(const Exec& ex, const float* src, double* dst, i64 n) {
    run(ex, F_OUTPUT, k_widen, dim3(red_blocks(n)), dim3(TPB), src, dst, n);
}

}  // namespace cav
