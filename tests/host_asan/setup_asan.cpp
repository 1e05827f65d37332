// Host set-up of libcavity_solver.so (csrc/setup.cpp) under AddressSanitizer: every routine that builds what the
// kernels read (grid axes in FP64 and float-rounded, Jacobi eigensystems, the even/odd split, MFMA fragment
// packing in FP64, FP32 and split bf16, the coefficient tables of both field types and the per-level z records)
// on the smallest accepted grids, odd and even sizes, uniform and stretched, both lid shapes. Built with g++
// (host code only) by tests/test_host_asan.py; it prints "setup asan ok" when every check passes.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "cavity.h"

namespace {
int failures = 0;
void expect(bool ok, const char* what, int n) {
    if (!ok) {
        std::printf("FAIL %s (n = %d)\n", what, n);
        failures++;
    }
}
CavConfig config(int stretch, double aspect) {
    CavConfig c{};
    c.nx = c.ny = 8;
    c.nz = 6;
    c.grid_stretch = stretch;
    c.grid_dx_wall = 1e-4;
    c.grid_max_aspect = aspect;
    c.length = c.lz = 1.0;
    c.u_lid = 1.0;
    c.impulse_order = 4;
    c.impulse_order4_ramp = 2.0;
    return c;
}
}  // namespace

int main() {
    using namespace cav;
    const int sizes[] = {5, 6, 7, 9, 10, 16, 33, 48, 64};
    for (int stretch = 0; stretch <= 1; stretch++)
        for (int n : sizes)
            for (int rf = 0; rf <= 1; rf++) {
                const CavConfig c = config(stretch, 1.7);
                const Axis a = make_axis(c, n, 1.0, rf != 0);
                expect((int)a.face.size() == n + 1 && (int)a.width.size() == n && (int)a.spacing.size() == n + 1,
                       "axis sizes", n);
                expect(std::fabs(a.face[0]) < 1e-15 && std::fabs(a.face[n] - 1.0) < 1e-6, "axis ends", n);
                const Eigen& e = axis_eigen(a);
                expect(e.n == n && (int)e.S.size() == n * n && (int)e.lambda.size() == n, "eigen sizes", n);
                if (n % 2 == 0) {
                    const SymSplit sp = symmetric_split(a);
                    if (sp.ok) {
                        const int h = n / 2;
                        expect((int)sp.fe.size() == h * h && (int)sp.bo.size() == h * h, "split sizes", n);
                        const std::vector<double> pd = pack_fragments<double>(sp.fe, h);
                        const std::vector<float> pf = pack_fragments<float>(sp.fo, h);
                        std::vector<uint16_t> hi, lo;
                        pack_bf16_fragments(sp.be, h, hi, lo);
                        expect(!pd.empty() && !pf.empty() && hi.size() == lo.size(), "split packing", n);
                    }
                }
                const std::vector<double> pd = pack_fragments<double>(e.S, n);
                std::vector<uint16_t> hi, lo;
                pack_bf16_fragments(e.S, n, hi, lo);
                expect(pd.size() >= (size_t)n * n && hi.size() == lo.size(), "dense packing", n);
                for (int f32 = 0; f32 <= 1; f32++)
                    for (int lid = 0; lid <= 1; lid++) {
                        const std::vector<double> t = axis_tables(a, 2.0, lid != 0, f32 != 0);
                        expect(t.size() == (size_t)TAB_KINDS * (n + 1), "table size", n);
                        const std::vector<double> r = level_records(t, n);
                        expect(r.size() == (size_t)(n + 4) * ZREC, "level record size", n);
                        bool finite = true;
                        for (double v : r) finite &= std::isfinite(v);
                        expect(finite, "level records finite", n);
                    }
            }
    if (failures) return 1;
    std::printf("setup asan ok\n");
    return 0;
}
