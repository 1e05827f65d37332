// kernels.h — launch wrappers of the gfx950 kernels (kernels.hip). Every wrapper enqueues on exec.stream
// and reports its kernel family to the optional timer.
//
// The time-step kernels are batched: one launch advances any number of cavities (members), each with its
// own sizes, grid, coefficients and precision mode. The blocks of a launch are split into one contiguous
// segment per member (block offsets in a device table, segments a multiple of 8 blocks so that the
// XCD-aware block remap works inside each segment); a block finds its member by binary search and reads the
// member's descriptor (Mem). A single cavity is a batch of one.
#pragma once

#include "cavity.h"

namespace cav {

enum Family {
    F_GAUGE, F_CFL, F_STAGE, F_DIVERGENCE, F_GUESS, F_RESIDUAL, F_REDUCE, F_XFORM_X, F_XFORM_Y, F_XFORM_Z,
    F_FINALIZE, F_GRADIENT, F_FILL, F_OUTPUT, F_KICK, F_CONTROL, F_COUNT
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

// ------------------------------------------------------------------------------------------------ batches
// Launch shapes: how many blocks one member takes in a launch of that kind
enum Shape {
    SH_GAUGE,      // gauge reset: faces of the three components and the cells, 256 per block
    SH_MARCH,      // z-march over the cells (CFL, divergence, residual)
    SH_STAGE,      // the RK stage: z-march of 32 x 8 column tiles over the cells
    SH_CELLS,      // one cell per thread
    SH_WALLS,      // one wall face per thread
    SH_RED,        // grid-stride reduction over the cells
    SH_GRAD,       // z-march over the cell columns and the nz + 1 face levels
    SH_MARCH_CAP,  // the z-march of the residual on a capped grid (rarely needed passes)
    SH_STAGE16,    // the RK stage of cavities tuned to 16 x 16 column tiles
    SH_XF,         // FDM transforms: SH_XF + 6 axis + 3 backward + layout (0 dense, 1 even/odd by chunks,
                   // 2 even/odd by whole columns)
    SH_XF_CAP = SH_XF + 18,   // the same on a capped grid, blocks looping over the tiles
    SH_COUNT = SH_XF_CAP + 18
};
// Blocks of `shape` for one member of a batch of `members` (host), with the member's mode and layout for the
// transforms
int shape_blocks(const Mem& m, int shape, int members);
// Transform layout (0 dense, 1 even/odd by chunks, 2 even/odd by whole columns) of an axis of a member, and
// whether the whole-column layout is possible there (the autotuner's choice)
int transform_layout(const Mem& m, int axis);
bool whole_columns_possible(const Mem& m, int axis);

struct Batch {
    const Mem* const* d_mem = nullptr;   // device array of n member descriptors
    const int* d_seg = nullptr;          // device: SH_COUNT rows of n + 1 block offsets
    int n = 0;
    int total[SH_COUNT] = {};            // blocks per shape over all members
    int mode = 0;                        // precision mode shared by all members
    int usual = 1;                       // refinement passes a projection usually needs (a spare comes on top)
    Exec ex;
};
// Fill seg (SH_COUNT x (n+1) ints) and total from the host copies of the member descriptors
void batch_segments(const std::vector<const Mem*>& members, std::vector<int>& seg, int* total);
// The same into a caller's array of SH_COUNT x (n + 1) ints (no allocation)
void batch_segments(const Mem* const* members, int n, int* seg, int* total);

// Reductions: kernels that reduce leave one partial per block; the final fold runs in a fixed order.
// op: 0 max (NaN wins), 1 sum, 2 max |x|, 3 sum x^2.
void reduce_array(const Exec& ex, const double* x, i64 n, int op, double* partial, double* out);
void reduce_final(const Exec& ex, const double* partial, int count, int op, double* out);

void fill(const Exec& ex, double* p, i64 n, double value);
void copy_array(const Exec& ex, const double* src, double* dst, i64 n);
// plain device-to-device copy of n doubles (counter calibration)
void copy_probe(const Exec& ex, const double* src, double* dst, i64 n_doubles);

// ---- the step (backend.cu:877-955), batched. Stage st = 0, 1, 2 reads and writes the roles
//   st 0: u FR_U,   m FR_M0,  out FR_NM0, projection -> FR_NU0, potential CR_PS0, previous potential CR_PHI0
//   st 1: u FR_NU0, m FR_NM0, out FR_NM1, projection -> FR_NU1, potential CR_PS1, previous potential CR_PS0
//   st 2: u FR_NU1, m FR_NM1, out FR_NM0, projection -> FR_NU0, potential CR_PS0 (mean-free into CR_PS1),
//         previous potential CR_PS1
// (a stage refines its guess in place; the previous stage's potential is dead once the base and guess are
// formed from it). When the step completes the state's faces swap with the stage buffers (Ctl::parity) and the
// potentials rotate (Ctl::rot): CR_PS1 becomes the state's, the state's array CR_PS0, where the step's last
// gradient has formed the next step's first guess (next_stage 3), and CR_PS0's array CR_PS1.
void b_gauge(const Batch& b);                       // m0 = r m + (1-r) u, phi0 = r phi
void b_cfl_dt(const Batch& b);                      // CFL rate, then dt and the first stage's control (stage time,
                                                    // lid, coefficients, warm-start plan)
void b_stage(const Batch& b, int st);               // inviscid term, viscous term and RK update of every face
void b_base_guess(const Batch& b, int st);          // base and warm-start guess into the stage's work potential
// Refinement passes a warm-started projection usually needs in a precision mode (FP64 transforms: one,
// FP32 / split-BF16: two). Passes beyond them are enqueued on capped grids.
int usual_passes(int mode);
// Projection of the impulse in role mm into velocity role out (backend.cu:816-854). The refinement works on the
// potential in role pw; the last kernel writes the mean-free potential into role ph (pw != ph; ph < 0: not
// kept, when only the next stage's base uses it).
// guess: start from the potential already in pw; hist: store dphi into the warm-start history;
// passes: FDM passes enqueued (each skipped by a member whose refinement has converged).
// gauge_next: the projection completes a step's state; its gradient also forms the next step's gauge reset and
// CFL rate. walls: zero the wall faces of mm first (backend.cu:850; the stage kernel writes them as 0 itself).
// next_stage (1 or 2): run that stage's control after the projection's, and form its base and guess in the
// gradient (in place of kb_base_guess); 3 (with gauge_next): plan the next step's first stage when its time step
// is known already (fixed, or lagged) and form its guess in the gradient, which that stage then keeps
// (Ctl::g0_ready). base: the role holding the base dphi is taken against (a step's first stage: CR_PHI0).
void b_project(const Batch& b, int mm, int pw, int ph, int out, bool guess, bool hist, int passes,
               bool gauge_next = false, bool walls = true, int next_stage = 0, int base = CR_BASE);
// ph = pw - mean (ph >= 0), dphi into the history (hist), out = mm - grad(phi) on every face, the next stage's
// base and guess (next_stage > 0): the projection's last kernel
void b_gradient(const Batch& b, int mm, int pw, int ph, int out, bool gauge_next = false, bool hist = false,
                int next_stage = 0, int base = CR_BASE);
// the six FDM transforms of refinement pass `pass`, the last one adding into the potential in role ph (run by a
// member only while its control block has run_xf set)
void b_transforms(const Batch& b, int ph, int pass);
// after the step's last kernel: advance the clock and make the stage-3 buffers the state
void b_step_end(const Batch& b);
// CFL rate of the state velocity into scal[S_CFL] (diagnostics)
void b_cfl_rate(const Batch& b);

// ---- outputs and kick (single cavity, backend.cu:445-561)
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
void smooth(const Exec& ex, const double* in, double* out, int axis, int sx, int sy, int sz);
void add_noise(const Exec& ex, const double* f, const double* noise, double s, double* out, i64 n);
void add_difference(const Exec& ex, double* m, const double* a, const double* b, i64 n);
// the same on the FP32-field mode's float face arrays
void add_noise(const Exec& ex, const float* f, const double* noise, double s, float* out, i64 n);
void add_difference(const Exec& ex, float* m, const float* a, const float* b, i64 n);
// dst = (double)src, n values (the FP32-field mode's outputs are computed from widened copies)
void widen(const Exec& ex, const float* src, double* dst, i64 n);

}  // namespace cav
