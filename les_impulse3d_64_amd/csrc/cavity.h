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
// backend.cu:672-699 in FP64 (no rounding through float)
Axis make_axis(const CavConfig& c, int n, double length);

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
    TAB_KINDS
};
std::vector<double> axis_tables(const Axis& a, double K);

}  // namespace cav
