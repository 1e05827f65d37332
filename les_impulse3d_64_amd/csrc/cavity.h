// cavity.h — internal declarations shared by the host set-up, the HIP kernels and the C-ABI layer of
// libcavity_solver.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/cavity_solver.h"

namespace cav {

using i64 = long long;
enum { X = 0, Y = 1, Z = 2 };

[[noreturn]] inline void fail(const std::string& msg) { throw std::runtime_error(msg); }
inline void check(hipError_t e, const char* what) {
    if (e != hipSuccess) fail(std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
}
#define HIP_CHECK(call) ::cav::check((call), #call)

// ------------------------------------------------------------------------------------------------ host grid
struct Axis {
    int n = 0;
    double length = 0;
    std::vector<double> face, centre, width, spacing;
};
// backend.cu:672-699 in FP64; round_float rounds every value through float as the reference does
// (backend.cu:695-697; the FP32-field mode), otherwise no rounding
Axis make_axis(const CavConfig& c, int n, double length, bool round_float = false);

// 1-D FDM eigensystem of one axis (backend.cu:750-777): S[i*n + j] = eigenvector j at cell i,
// W-orthonormal; lambda with the constant mode (argmax) set to 0.
struct Eigen {
    int n = 0, zero_mode = 0;
    std::vector<double> S, lambda;
};
// Cached process-wide by the axis' widths and spacings (identical input, identical result).
const Eigen& axis_eigen(const Axis& a);

// MFMA fragment packing of an n x n matrix M (row-major, M[j*n + i]) for the 16x16x4 operand layout:
// packed[((jt*(n32/4) + kt)*64 + l] = M[16 jt + l%16][4 kt + l/16], zero outside n x n; n32 = n rounded
// up to 32 (the transforms stage 32 contraction indices at a time).
int pad32(int n);
template <class T>
std::vector<T> pack_fragments(const std::vector<double>& M, int n);
// Even/odd split of the FDM eigensystem of an axis whose grid mirrors about its centre (n even, widths and
// spacings equal under i -> n-1-i to 1e-12): the even and odd eigenvectors come from the two folded
// h x h problems (h = n/2). Spectral order [even | odd]; lambda in that order with the constant mode
// (index zero) set to 0. Half matrices (row-major h x h), S the W-orthonormal eigenvectors:
//   fe[j][i] = S[i][even j], fo[j][i] = S[i][odd j] for i < h   (forward: fold r_i +/- r_{n-1-i})
//   be = fe^T, bo = fo^T                                         (backward: unfold e +/- o)
// ok = false: use the dense transform.
struct SymSplit {
    bool ok = false;
    int h = 0, zero = 0;
    std::vector<double> lambda, fe, fo, be, bo;
};
SymSplit symmetric_split(const Axis& a);
// Same for the 16x16x32 bf16 operand layout (lane l: row l%16, k = 8 (l/16) + 0..7, 8 values per lane),
// n rounded up to 32; M split as bf16 hi + bf16 lo (the rounded remainder).
void pack_bf16_fragments(const std::vector<double>& M, int n, std::vector<uint16_t>& hi, std::vector<uint16_t>& lo);

// ------------------------------------------------------------------------------------------------ device view
// Grid tables on the device (FP64) plus sizes; passed by value to every kernel.
struct Geo {
    int nx, ny, nz, lid_shaped;
    const double* face[3];
    const double* centre[3];
    const double* width[3];
    const double* spacing[3];
    double length[3];
    // Per-axis coefficient tables, TAB_KINDS rows of n+1 entries each (see axis_tables)
    const double* tab[3];
};

// Kinds of per-axis coefficient tables. Index e is a cell (W, RW, CE) or a face/edge position (others).
enum TabKind {
    T_W = 0,      // width[e]
    T_RW,         // 1 / width[e]
    T_SP,         // spacing[e]
    T_RSP,        // 1 / spacing[e]
    T_CE,         // centre[e]
    T_C3,         // wide-stencil face weight: edge width / (2 spacing[e]) (backend.cu:124-128, 234)
    T_RAMP,       // 4th-order wall ramp clamp((min(e, n-e) - 1) / K, 0, 1) (backend.cu:130-133)
    T_RGAP3,      // 1 / (centre[e+1] - centre[e-2]): wide node distance (backend.cu:114-117)
    T_RWS3,       // 1 / (width[e-2] + width[e+1]): wide edge interpolation (backend.cu:119-122)
    T_R2SP,       // 1 / (2 spacing[e]): narrow edge interpolation
    T_LIDF,       // lid profile factor at face e: lid_edge(2 face[e] / length - 1), 1 on a flat lid (backend.cu:58-65)
    T_LIDC,       // the same at cell centre e
    T_RSPD,       // 1 / spacing[e] in FP64 in every mode (the FP64 residual of the potential)
    TAB_KINDS
};
// f32: the FP32-field mode's tables (float values; the reciprocal kinds hold the divisors themselves)
std::vector<double> axis_tables(const Axis& a, double K, bool lid_shaped, bool f32 = false);
// The z-axis factors the stage kernel uses at march level L, gathered from the z tables into one record of
// ZREC values per level (L = -1 .. nz + 2, record L + 1), so a level's factors are one contiguous read. Indices
// are clamped into the tables as the kernel clamps them; entries a level never uses hold the clamped values.
enum ZRec {
    ZR_RAMP = 0, ZR_RSP, ZR_RGAP3, ZR_R2SP, ZR_RWS3,   // x- and y-edges of level L (at L clamped to [0, nz])
    ZR_WM2, ZR_WM1, ZR_W0, ZR_WP1,                     // T_W at the cells L-2 .. L+1 (clamped to [0, nz-1])
    ZR_C3, ZR_RW, ZR_RWM1,                             // the w face of level L: T_C3[L], T_RW[L], T_RW[L-1]
    ZR_LIDCJ, ZR_RSPT, ZR_RSPJ, ZR_RWJ,                // the u, v faces of layer j = L-2: T_LIDC[j], T_RSP[top of j],
                                                       // T_RSP[j], T_RW[j]
    ZREC
};
std::vector<double> level_records(const std::vector<double>& ztab, int nz);

// ------------------------------------------------------------------------------------------------ member state
// Face-array roles (three components each): state u, m; gauge-reset m0; the impulse and velocity of RK
// stages 1 and 2 (stage 3 writes into the stage-1 buffers, which become the state when the step ends). The
// state and the stage-1 buffers swap roles after every completed step (Ctl::parity), so a step that has to be
// redone still finds the state it started from. Between steps only u and m hold state (the kick uses the
// others as scratch).
enum FaceRole { FR_U, FR_M, FR_M0, FR_NM0, FR_NM1, FR_NU0, FR_NU1, FR_COUNT };
// Cell-array roles: state phi, gauge-reset phi0, the two work potentials of the stages (stages 1 and 3 refine
// CR_PS0, stage 2 CR_PS1; stage 3 writes its mean-free potential into CR_PS1), base of the warm start,
// divergence rhs, residual (FP64; float in the modes with FP32 or split-BF16 transforms, which alone read it, the
// rhs converted into it on a cold start), the history buffers of dphi, two FDM transform temporaries.
constexpr int HIST_BUFFERS = 6;     // 3 history entries kept from the previous step + 3 written by this one
enum CellRole { CR_PHI, CR_PHI0, CR_PS0, CR_PS1, CR_BASE, CR_RHS, CR_RES, CR_H0, CR_T1 = CR_H0 + HIST_BUFFERS, CR_T2,
                CR_COUNT };
// Physical buffer of a role under a parity: the state and the stage-1 buffers trade places
__host__ __device__ inline int face_buffer(int role, int parity) {
    if (!parity) return role;
    return role == FR_U ? FR_NU0 : role == FR_NU0 ? FR_U : role == FR_M ? FR_NM0 : role == FR_NM0 ? FR_M : role;
}
// The state potential and the two work potentials rotate through three arrays (Ctl::rot): when a step completes,
// stage 3's potential (CR_PS1) becomes the state's, the state's array becomes CR_PS0 (where the step's last
// kernel has formed the next step's first guess) and CR_PS0's array becomes CR_PS1.
__host__ __device__ inline int cell_buffer(int role, int rot) {
    const int k = role == CR_PHI ? 0 : role == CR_PS0 ? 1 : role == CR_PS1 ? 2 : -1;
    if (k < 0 || rot == 0) return role;
    const int j = (k + rot) % 3;
    return j == 0 ? CR_PHI : j == 1 ? CR_PS0 : CR_PS1;
}
// the rotation after a completed step
__host__ __device__ inline int next_rot(int rot) { return (rot + 2) % 3; }

// Device scalar slots (Mem::scal) and partial-sum regions (Mem::partial, RED_BLOCKS doubles each)
enum Slot { S_CFL = 0, S_RHS2, S_RES2, S_PHISUM, S_ZERO, S_KE, S_ENS, S_DIV2, S_UMAX, S_GAUGE = S_UMAX + 3,
            S_KICK = S_GAUGE + 3, S_COUNT = 32 };
constexpr int RED_BLOCKS = 65536;
enum PartialRegion { P_MAIN = 0, P_GUESS = 1, P_PHI = 2, P_AUX = 3, P_COUNT = 4 };

// Error codes a step leaves in Ctl::err (the first one wins)
enum StepError { ERR_NONE = 0, ERR_VELOCITY_NAN, ERR_RHS_NAN, ERR_RESIDUAL_NAN, ERR_NOT_CONVERGED };

// Control scalars of one cavity, resident on the device: the step's time-step choice, stage
// coefficients, warm-start history and FDM refinement loop run in single-thread control kernels
// (backend.cu:884-899, 816-847, 857-873, 919-939), so a whole step needs no host round trip.
struct Ctl {
    // configuration
    double dt_fixed, dt_max, cfl_target, diffusion_limit, viscous_rate, rtol, u_lid, lid_ramp_t;
    int adaptive, dt_lag, max_passes, extrapolate;
    // clock (backend.cu:599-640)
    double t, last_dt, last_cfl, prev_rate[2];
    long long steps;
    int have_prev, dt_reason;
    double dt, use_rate;
    // current RK stage
    double ts, lid, a0, b;
    // warm-start history: entries oldest first (t, b dt, buffer), hn <= 3
    double ht[3], hs[3];
    int hb[3], hn;
    int gnh, gb1, gb2, out_buf;
    double gc1, gc2;
    // projection
    double rhs_norm, rel, r2, rms, phi_rel;
    int passes, active, run_xf, r_rhs, zero_start, zero_rhs, phi_passes;
    long long projections, pass_total;
    // error; a projection that needed more refinement passes than the step enqueued (the step is redone)
    int err, err_passes;
    double err_val;
    int stalled, parity;
    // m0 and phi0 hold the gauge reset of the current state, and cfl_bits its CFL rate (the bits of a double): the
    // last stage of the previous step formed them with its velocity (kb_gradient); the next step's gauge reset
    // and CFL pass are then skipped. Cleared by every host-side change of the state.
    int m0_ready, cfl_ready;
    unsigned long long cfl_bits;
    // history buffer of the projection that just ended (its dphi; -1: none), set at the projection's end
    // (kb_proj_close): the next stage's control may pick another out_buf before the projection's last kernel
    // writes this one
    int fin_buf, pad1;
    // 1 + the step index of the last projection that ran more refinement passes than its batch usually enqueues
    // (0: none); the host keeps a spare pass enqueued for a while after it (kb_proj_close)
    long long spare_mark;
    // rotation of the potentials (cell_buffer); the next step's first guess has been formed in CR_PS0 (its base
    // is phi0) by this step's last kernel: the next stage 1 skips kb_base_guess
    int rot, g0_ready;
    // the next step's first-stage warm-start plan (as its stage control will form it), when the next dt is known
    // before the step's last kernel (lagged or fixed dt): nguess set
    double ngc1, ngc2;
    int ngnh, ngb1, ngb2, nguess;
};
static_assert(sizeof(Ctl) % sizeof(long long) == 0, "the control block is copied as long longs");

// Launch choices the autotuner makes per cavity (Mem::tune). None changes a result: the stage and gradient
// kernels compute every face the same way whatever the tile and z chunks, and both transform layouts of a
// mirror-symmetric axis sum the same products in the same order.
enum TuneKnob {
    TU_STAGE_TILE = 0,   // kb_stage column tile: 0 32 x 8, 1 16 x 16
    TU_STAGE_TARGET,     // kb_stage march: blocks aimed at (0: 768)
    TU_STAGE_MINZ,       // kb_stage march: shortest z chunk (0: 8 levels)
    TU_GRAD_TARGET,      // kb_gradient march: blocks aimed at (0: 1792)
    TU_GRAD_MINZ,        // kb_gradient march: shortest z chunk (0: 8 levels)
    TU_XF,               // transform layout per axis, 2 bits each (axis a at bits 2a): 0 the built-in rule,
                         // 1 even/odd by chunks, 2 even/odd by whole columns (where the axis allows it)
    TU_GRAD_TILE,        // kb_gradient march tile: 0 32 x 8, 1 64 x 4, 2 128 x 2, 3 256 x 1
    TU_COUNT
};

// Everything a batched kernel needs to know about one cavity: grid tables, the device control block,
// buffers by role and the FDM operators for its precision mode and transform layout. Immutable while
// the cavity exists (a mode or layout change rewrites it between steps).
struct Mem {
    Geo g;
    Ctl* ctl;
    Ctl* ctl_bak;                      // the control block as it was when the current step began
    double* scal;
    double* partial;                   // P_COUNT regions of RED_BLOCKS
    double* F[FR_COUNT][3];
    double* C[CR_COUNT];
    // [axis][0 forward, 1 backward][0 matrix or even half, 1 its bf16 lo part, 2 odd half, 3 its lo part]
    const void* mat[3][2][4];
    const void* il;                    // 1 / eigenvalue sums in the mode's type and the layout's order
    int sym[3];
    int tune[8];                       // launch choices of the autotuner (TuneKnob); 0 = the built-in default
    int mode, o4, max_passes;
    double nu, r, omr;
    i64 n_cell;
};

}  // namespace cav
