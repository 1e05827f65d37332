// kernels.h — launch wrappers of the gfx950 kernels (kernels.hip). Every wrapper enqueues on exec.stream
// and reports its kernel family to the optional timer.
#pragma once

#include "cavity.h"

namespace cav {

enum Family {
    F_GAUGE, F_CFL, F_EDGES, F_STAGE, F_DIVERGENCE, F_GUESS, F_RESIDUAL, F_REDUCE, F_XFORM_X, F_XFORM_Y, F_XFORM_Z,
    F_FINALIZE, F_GRADIENT, F_FILL, F_OUTPUT, F_KICK, F_COUNT
};
const char* family_name(int f);

struct KernelTimer {
    virtual ~KernelTimer() = default;
    virtual void before(int family, hipStream_t s) = 0;
    virtual void after(int family, hipStream_t s) = 0;
};

struct Exec {
    hipStream_t stream = nullptr;
    KernelTimer* timer = nullptr;
};

// Reductions: kernels that reduce leave one partial per block in `partial`; reduce_final folds
// `count` partials in a fixed order into out[0]. op: 0 max (NaN wins), 1 sum, 2 max |x|, 3 sum x^2.
constexpr int RED_BLOCKS = 65536;      // partial slots per reduction
void reduce_array(const Exec& ex, const double* x, i64 n, int op, double* partial, double* out);
void reduce_final(const Exec& ex, const double* partial, int count, int op, double* out);

void fill(const Exec& ex, double* p, i64 n, double value);
// plain device-to-device copy of n doubles (counter calibration)
void copy_probe(const Exec& ex, const double* src, double* dst, i64 n_doubles);

// step pieces (backend.cu:877-955)
void gauge_reset(const Exec& ex, const Geo& g, const double* const m[3], const double* const u[3], const double* phi,
                 double r, double omr, double* const m0[3], double* phi0);
// per-cell CFL rate max over cells -> out (device scalar), NaN wins
void cfl_max(const Exec& ex, const Geo& g, const double* const u[3], double* partial, double* out);

// Edge products of the Lamb vector and the cell kinetic term, one array each (see k_edges):
//   E_ZX, E_ZY on z-edges (nx+1)(ny+1)nz, E_XY, E_XZ on x-edges nx(ny+1)(nz+1), E_YX, E_YZ on y-edges
//   (nx+1)ny(nz+1), each for the narrow (q = 1) and wide (q = 3) stencils; E_K on cells.
enum EdgeArray { E_ZX1, E_ZY1, E_XY1, E_XZ1, E_YX1, E_YZ1, E_ZX3, E_ZY3, E_XY3, E_XZ3, E_YX3, E_YZ3, E_K, E_COUNT };
i64 edge_array_size(int nx, int ny, int nz, int which);

struct StageArgs {
    const double* u[3];     // velocity the inviscid term and the wall fluxes use
    const double* m[3];     // impulse being advanced
    const double* m0[3];    // gauge-reset impulse (base of stages 2, 3)
    double* out[3];
    double* edge[E_COUNT];  // scratch for the edge products
    double nu, lid, dt, a0, b;
    int has_base, o4;
};
void stage(const Exec& ex, const Geo& g, const StageArgs& a);

// rhs = vol div(f); partial sums of rhs^2 -> partial (count returned)
int divergence(const Exec& ex, const Geo& g, const double* const f[3], double* rhs, double* partial);
void zero_walls(const Exec& ex, const Geo& g, double* const f[3]);

// base = [a0 phi0 +] b cur;  phi = [c1 h1 [+ c2 h2] +] base
void base_guess(const Exec& ex, i64 n, const double* phi0, const double* cur, double a0, double b, int has_a0,
                const double* h1, double c1, const double* h2, double c2, int nh, double* base, double* phi);
// r = rhs - L(phi); partial sums of r^2
int residual(const Exec& ex, const Geo& g, const double* phi, const double* rhs, double* r, double* partial);

// One axis of the FDM transform: out[.., j, ..] = sum_i M[j][i] in[.., i, ..] with M packed for the mode
// (pack_fragments / pack_bf16_fragments). layout_x: rows of length n contiguous (x axis); else [A][n][B]
// with B the inner stride. mode 0 FP64, 1 FP32, 2 split-BF16 x3 (in/out FP32 unless in_f64/out_f64).
// epi: 0 store, 1 store * il[idx] (il in the mode's type), 2 out[idx] += acc * out_scale (out FP64).
// sym: 0 dense; 1 symmetric forward, 2 symmetric backward (mpk / mpk2 the even / odd half matrices, the
// spectral index ordered [even | odd]).
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
int pad_for_mode(int n, int mode);
void axis_transform(const Exec& ex, bool layout_x, int mode, bool in_f64, bool out_f64, const AxisArgs& a,
                    int family);

// phi -= sum/n (sum at dsum, device); if hist: dphi = phi - base
void finalize_phi(const Exec& ex, i64 n, double* phi, const double* dsum, const double* base, double* dphi);
// out_k = m_k - grad(phi), walls 0
void subtract_gradient(const Exec& ex, const Geo& g, const double* const m[3], const double* phi, double* const out[3]);

// outputs (backend.cu:445-540)
void cell_diagnostics(const Exec& ex, const Geo& g, const double* const u[3], double lid, double* partial3,
                      double* out3);
void max_abs_diff(const Exec& ex, const double* a, const double* b, i64 n, double* partial, double* out);
void first_index_of_abs_diff(const Exec& ex, const double* a, const double* b, i64 n, const double* value, int* first);
void probe_values(const Exec& ex, const Geo& g, const double* u, const double* v, const int* cells, int n,
                  float* out);
void add_statistics(const Exec& ex, const Geo& g, const double* const u[3], double weight, double* sums);
void centreline_sample(const Exec& ex, const Geo& g, const double* const u[3], double* out);
void midplane(const Exec& ex, const Geo& g, const double* const u[3], double lid, float* out);
void q_criterion(const Exec& ex, const Geo& g, const double* const u[3], double lid, double qscale, double inv_U,
                 int colour_w, float* q, float* colour);

// kick (backend.cu:542-561)
void smooth(const Exec& ex, const double* in, double* out, int axis, int sx, int sy, int sz);
void add_noise(const Exec& ex, const double* f, const double* noise, double s, double* out, i64 n);
void add_difference(const Exec& ex, double* m, const double* a, const double* b, i64 n);

}  // namespace cav
