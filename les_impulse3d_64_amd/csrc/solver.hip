// solver.hip — host orchestration of one cavity solve on a gfx950 device and the C-ABI of
// libcavity_solver.so (include/cavity_solver.h).
//
// One Cavity owns a HIP stream and one device arena holding the FP64 state (u, m, phi), the work arrays,
// the packed FDM matrices and the reduction scratch. The step follows the reference's sequence
// (backend.cu:877-955): gauge reset, CFL/dt, three SSP-RK3 stages of fused inviscid+viscous update,
// warm-started projection with the FDM refinement loop, history of the potential. Host round trips are
// limited to the scalars that steer control flow (CFL rate, residual norms); the potential's mean stays on
// the device.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>

#include "kernels.h"

namespace cav {
namespace {

thread_local std::string g_error;

// device scalar slots
enum Slot { S_CFL = 0, S_RHS2, S_RES2, S_PHISUM, S_ZERO, S_KE, S_ENS, S_DIV2, S_UMAX, S_GAUGE = S_UMAX + 3,
            S_KICK = S_GAUGE + 3, S_COUNT = 32 };

struct EventTimer final : KernelTimer {
    struct Pair { hipEvent_t a, b; int family; };
    std::vector<Pair> pool;
    size_t used = 0;
    double ms[F_COUNT] = {};
    long long launches[F_COUNT] = {};
    ~EventTimer() override {
        for (auto& p : pool) { (void)hipEventDestroy(p.a); (void)hipEventDestroy(p.b); }
    }
    void before(int family, hipStream_t s) override {
        if (used == pool.size()) {
            if (pool.size() >= 8192) flush();
            else {
                Pair p{};
                HIP_CHECK(hipEventCreate(&p.a));
                HIP_CHECK(hipEventCreate(&p.b));
                pool.push_back(p);
            }
        }
        pool[used].family = family;
        HIP_CHECK(hipEventRecord(pool[used].a, s));
    }
    void after(int, hipStream_t s) override {
        HIP_CHECK(hipEventRecord(pool[used].b, s));
        used++;
    }
    void flush() {
        if (!used) return;
        HIP_CHECK(hipEventSynchronize(pool[used - 1].b));
        for (size_t k = 0; k < used; k++) {
            float t = 0;
            HIP_CHECK(hipEventElapsedTime(&t, pool[k].a, pool[k].b));
            ms[pool[k].family] += t;
            launches[pool[k].family]++;
        }
        used = 0;
    }
};

struct History { double t, scale; int buf; };

struct Layout {              // byte offsets of every device array in the arena
    size_t total = 0;
    size_t take(size_t bytes) {
        const size_t off = total;
        total += (bytes + 255) / 256 * 256;
        return off;
    }
};

}  // namespace
}  // namespace cav

struct cav_solver {
    CavConfig c{};
    int nx = 0, ny = 0, nz = 0;
    cav::i64 n_cell = 0, size[3] = {0, 0, 0};
    cav::Axis ax[3];
    double volume = 0, viscous_rate = 0;
    int mode = CAV_PREC_FP64;

    hipStream_t stream = nullptr;
    cav::Exec ex;
    std::unique_ptr<cav::EventTimer> timer;
    bool timing = false;

    char* arena = nullptr;
    size_t arena_bytes = 0;
    cav::Geo geo{};
    double *u[3]{}, *m[3]{}, *phi = nullptr;
    double *m0[3]{}, *new_m[2][3]{}, *new_u[2][3]{}, *tmp[3]{};
    double *phi0 = nullptr, *phi_stage[2]{}, *base = nullptr, *rhs = nullptr, *res = nullptr, *hist_buf[4]{};
    double *t1 = nullptr, *t2 = nullptr, *noise[2]{};
    double* edge[cav::E_COUNT]{};  // Lamb edge products and K (stage scratch)
    double* mpk[3][2]{};     // packed FDM matrices per axis: [0] forward (S^T), [1] backward (S)
    float* mpk32[3][2]{};    // the same in FP32 (mixed FDM)
    uint16_t* mpkbf[3][2][2]{};   // the same split as bf16 hi [..][0] and lo [..][1] (split-BF16 FDM)
    double* inv_lambda = nullptr;
    float* il_f = nullptr;   // 1/eigenvalue sums in FP32 (mixed FDM)
    // even/odd split per axis (symmetric grids): half matrices fe, fo, be, bo in the three precisions, and
    // 1/eigenvalue sums in the [even | odd] spectral order of the split axes
    bool sym_ok[3] = {false, false, false};
    bool use_sym = true;
    double* spk[3][4]{};
    float* spk32[3][4]{};
    uint16_t* spkbf[3][4][2]{};
    double* il_sym = nullptr;
    float* il_sym_f = nullptr;
    double *partial = nullptr, *scal = nullptr, *stats = nullptr, *line = nullptr;
    float *probe_out = nullptr, *plane = nullptr;
    int *probe_cells = nullptr, *first = nullptr;
    double* h_scal = nullptr;                               // pinned host mirror of scal

    double t = 0, last_dt = 0, last_cfl = 0, phi_rel = 0, prev_rates[2] = {0, 0};
    long long steps = 0, projections = 0, pass_total = 0;
    int dt_reason = 0, phi_passes = 0;
    bool have_prev_rates = false;
    std::vector<cav::History> hist;
};

namespace cav {
namespace {

using Solver = cav_solver;

template <class T> T* carve(Solver& s, size_t off) { return (T*)(s.arena + off); }

// Device arrays of one solver (offsets into the arena). Also the footprint reported for admission.
struct Plan {
    Layout L;
    size_t u[3], m[3], phi, m0[3], new_m[2][3], new_u[2][3], tmp[3], phi0, phi_stage[2], base, rhs, res, hist[4],
        t1, t2, noise[2], mpk[3][2], mpk32[3][2], mpkbf[3][2][2], inv_lambda, il_f, spk[3][4], spk32[3][4],
        spkbf[3][4][2], il_sym, il_sym_f, partial, scal, stats, line, probe_out, plane, probe_cells, first,
        geo[3][4], tab[3], edge[E_COUNT];
    Plan(int nx, int ny, int nz) {
        const i64 nc = (i64)nx * ny * nz;
        const i64 sz[3] = {(i64)(nx + 1) * ny * nz, (i64)nx * (ny + 1) * nz, (i64)nx * ny * (nz + 1)};
        const i64 big = std::max(sz[0], std::max(sz[1], sz[2]));
        const int n[3] = {nx, ny, nz};
        const size_t D = sizeof(double);
        for (int k = 0; k < 3; k++) {
            u[k] = L.take(sz[k] * D); m[k] = L.take(sz[k] * D); m0[k] = L.take(sz[k] * D); tmp[k] = L.take(sz[k] * D);
            for (int j = 0; j < 2; j++) { new_m[j][k] = L.take(sz[k] * D); new_u[j][k] = L.take(sz[k] * D); }
        }
        phi = L.take(nc * D); phi0 = L.take(nc * D); phi_stage[0] = L.take(nc * D); phi_stage[1] = L.take(nc * D);
        base = L.take(nc * D); rhs = L.take(nc * D); res = L.take(nc * D);
        for (int k = 0; k < 4; k++) hist[k] = L.take(nc * D);
        t1 = L.take(nc * D); t2 = L.take(nc * D);
        noise[0] = L.take(big * D); noise[1] = L.take(big * D);
        for (int k = 0; k < E_COUNT; k++) edge[k] = L.take(edge_array_size(nx, ny, nz, k) * D);
        for (int a = 0; a < 3; a++) {
            const int n32 = pad32(n[a]);
            for (int d = 0; d < 2; d++) {
                mpk[a][d] = L.take((size_t)n32 * n32 * D);
                mpk32[a][d] = L.take((size_t)n32 * n32 * sizeof(float));
                for (int h = 0; h < 2; h++) mpkbf[a][d][h] = L.take((size_t)n32 * n32 * sizeof(uint16_t));
            }
            const int h32 = pad32(n[a] / 2);
            for (int d = 0; d < 4; d++) {
                spk[a][d] = L.take((size_t)h32 * h32 * D);
                spk32[a][d] = L.take((size_t)h32 * h32 * sizeof(float));
                for (int h = 0; h < 2; h++) spkbf[a][d][h] = L.take((size_t)h32 * h32 * sizeof(uint16_t));
            }
            geo[a][0] = L.take((n[a] + 1) * D);   // face
            geo[a][1] = L.take(n[a] * D);         // centre
            geo[a][2] = L.take(n[a] * D);         // width
            geo[a][3] = L.take((n[a] + 1) * D);   // spacing
            tab[a] = L.take((size_t)TAB_KINDS * (n[a] + 1) * D);
        }
        inv_lambda = L.take(nc * D);
        il_f = L.take(nc * sizeof(float));
        il_sym = L.take(nc * D);
        il_sym_f = L.take(nc * sizeof(float));
        partial = L.take(4 * RED_BLOCKS * D);
        scal = L.take(S_COUNT * D);
        stats = L.take(6 * (i64)nx * ny * D);
        line = L.take(5 * (i64)nx * D);
        probe_out = L.take(16 * sizeof(float));
        plane = L.take(4 * (i64)nx * ny * sizeof(float));
        probe_cells = L.take(24 * sizeof(int));
        first = L.take(sizeof(int));
    }
};

void validate(const CavConfig& c) {                                     // backend.cu:1016-1020
    if (c.nx % 2 || c.ny % 2 || c.nx < 6 || c.ny < 6) fail("nx and ny must be even and >= 6");
    if (c.nz < 5) fail("need nz >= 5");
    if (std::max(c.nx, std::max(c.ny, c.nz)) > 1024) fail("at most 1024 cells per axis");
    if (c.impulse_order != 2 && c.impulse_order != 4) fail("IMPULSE_ORDER must be 2 or 4");
    if ((double)(c.nx + 1) * c.ny * (c.nz + 1) > 2e9) fail("grid too large");
    if (c.n_probes < 0 || c.n_probes > 8) fail("n_probes must be 0..8");
}

double lid_speed(const Solver& s, double time) {                       // backend.cu:623-626
    if (s.c.lid_ramp_t > 0) return s.c.u_lid * std::min(1.0, std::max(time, 0.0) / s.c.lid_ramp_t);
    return s.c.u_lid;
}

void sync_read(Solver& s, int first, int count) {
    HIP_CHECK(hipMemcpyAsync(s.h_scal + first, s.scal + first, count * sizeof(double), hipMemcpyDeviceToHost, s.stream));
    HIP_CHECK(hipStreamSynchronize(s.stream));
}

void h2d(Solver& s, void* dst, const void* src, size_t bytes) {
    HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s.stream));
    HIP_CHECK(hipStreamSynchronize(s.stream));
}
void d2h(Solver& s, void* dst, const void* src, size_t bytes) {
    HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s.stream));
    HIP_CHECK(hipStreamSynchronize(s.stream));
}

void create(Solver& s, const CavConfig& cfg) {                         // backend.cu:1009-1073
    validate(cfg);
    s.c = cfg;
    s.nx = cfg.nx; s.ny = cfg.ny; s.nz = cfg.nz;
    s.n_cell = (i64)s.nx * s.ny * s.nz;
    s.size[0] = (i64)(s.nx + 1) * s.ny * s.nz;
    s.size[1] = (i64)s.nx * (s.ny + 1) * s.nz;
    s.size[2] = (i64)s.nx * s.ny * (s.nz + 1);
    s.ax[X] = make_axis(s.c, s.nx, s.c.length);
    s.ax[Y] = make_axis(s.c, s.ny, s.c.length);
    s.ax[Z] = make_axis(s.c, s.nz, s.nz * (s.c.lz / s.nz));
    s.volume = s.ax[X].length * s.ax[Y].length * s.ax[Z].length;
    // viscous rate (backend.cu:1030-1034): nu * max over cells of sum over axes of 4/(width * smaller
    // spacing). Rounded addition and multiplication are monotone, so the max over cells of the sum equals
    // the sum of the per-axis maxima, bit for bit.
    double gmax[3];
    for (int a = 0; a < 3; a++) {
        gmax[a] = 0.0;
        const Axis& g = s.ax[a];
        for (int i = 0; i < g.n; i++)
            gmax[a] = std::max(gmax[a], 4.0 / (g.width[i] * std::min(g.spacing[i], g.spacing[i + 1])));
    }
    s.viscous_rate = std::max(0.0, s.c.nu * (gmax[X] + gmax[Y] + gmax[Z]));

    HIP_CHECK(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
    s.ex.stream = s.stream;
    Plan P(s.nx, s.ny, s.nz);
    s.arena_bytes = P.L.total;
    HIP_CHECK(hipMalloc((void**)&s.arena, s.arena_bytes));
    HIP_CHECK(hipMemsetAsync(s.arena, 0, s.arena_bytes, s.stream));
    HIP_CHECK(hipHostMalloc((void**)&s.h_scal, S_COUNT * sizeof(double), hipHostMallocDefault));
    for (int k = 0; k < 3; k++) {
        s.u[k] = carve<double>(s, P.u[k]); s.m[k] = carve<double>(s, P.m[k]); s.m0[k] = carve<double>(s, P.m0[k]);
        s.tmp[k] = carve<double>(s, P.tmp[k]);
        for (int j = 0; j < 2; j++) { s.new_m[j][k] = carve<double>(s, P.new_m[j][k]); s.new_u[j][k] = carve<double>(s, P.new_u[j][k]); }
    }
    s.phi = carve<double>(s, P.phi); s.phi0 = carve<double>(s, P.phi0);
    s.phi_stage[0] = carve<double>(s, P.phi_stage[0]); s.phi_stage[1] = carve<double>(s, P.phi_stage[1]);
    s.base = carve<double>(s, P.base); s.rhs = carve<double>(s, P.rhs); s.res = carve<double>(s, P.res);
    for (int k = 0; k < 4; k++) s.hist_buf[k] = carve<double>(s, P.hist[k]);
    s.t1 = carve<double>(s, P.t1); s.t2 = carve<double>(s, P.t2);
    s.noise[0] = carve<double>(s, P.noise[0]); s.noise[1] = carve<double>(s, P.noise[1]);
    for (int k = 0; k < E_COUNT; k++) s.edge[k] = carve<double>(s, P.edge[k]);
    s.inv_lambda = carve<double>(s, P.inv_lambda);
    s.partial = carve<double>(s, P.partial); s.scal = carve<double>(s, P.scal);
    s.stats = carve<double>(s, P.stats); s.line = carve<double>(s, P.line);
    s.probe_out = carve<float>(s, P.probe_out); s.plane = carve<float>(s, P.plane);
    s.probe_cells = carve<int>(s, P.probe_cells); s.first = carve<int>(s, P.first);

    // grid tables
    s.geo.nx = s.nx; s.geo.ny = s.ny; s.geo.nz = s.nz; s.geo.lid_shaped = s.c.lid_shaped;
    for (int a = 0; a < 3; a++) {
        const Axis& g = s.ax[a];
        double* p[4] = {carve<double>(s, P.geo[a][0]), carve<double>(s, P.geo[a][1]), carve<double>(s, P.geo[a][2]),
                        carve<double>(s, P.geo[a][3])};
        h2d(s, p[0], g.face.data(), g.face.size() * sizeof(double));
        h2d(s, p[1], g.centre.data(), g.centre.size() * sizeof(double));
        h2d(s, p[2], g.width.data(), g.width.size() * sizeof(double));
        h2d(s, p[3], g.spacing.data(), g.spacing.size() * sizeof(double));
        s.geo.face[a] = p[0]; s.geo.centre[a] = p[1]; s.geo.width[a] = p[2]; s.geo.spacing[a] = p[3];
        const std::vector<double> tabs = axis_tables(g, std::max(s.c.impulse_order4_ramp, 1.0));
        double* tp = carve<double>(s, P.tab[a]);
        h2d(s, tp, tabs.data(), tabs.size() * sizeof(double));
        s.geo.tab[a] = tp;
        s.geo.length[a] = g.length;
    }
    // FDM eigensystems (backend.cu:750-787)
    const Eigen* e[3];
    for (int a = 0; a < 3; a++) {
        e[a] = &axis_eigen(s.ax[a]);
        const int n = s.ax[a].n;
        std::vector<double> ST((size_t)n * n);
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) ST[(size_t)j * n + i] = e[a]->S[(size_t)i * n + j];
        std::vector<double> pf = pack_fragments<double>(ST, n), pb = pack_fragments<double>(e[a]->S, n);
        s.mpk[a][0] = carve<double>(s, P.mpk[a][0]);
        s.mpk[a][1] = carve<double>(s, P.mpk[a][1]);
        h2d(s, s.mpk[a][0], pf.data(), pf.size() * sizeof(double));
        h2d(s, s.mpk[a][1], pb.data(), pb.size() * sizeof(double));
        const std::vector<double>* mats[2] = {&ST, &e[a]->S};
        for (int d = 0; d < 2; d++) {
            std::vector<float> f32 = pack_fragments<float>(*mats[d], n);
            s.mpk32[a][d] = carve<float>(s, P.mpk32[a][d]);
            h2d(s, s.mpk32[a][d], f32.data(), f32.size() * sizeof(float));
            std::vector<uint16_t> hi, lo;
            pack_bf16_fragments(*mats[d], n, hi, lo);
            s.mpkbf[a][d][0] = carve<uint16_t>(s, P.mpkbf[a][d][0]);
            s.mpkbf[a][d][1] = carve<uint16_t>(s, P.mpkbf[a][d][1]);
            h2d(s, s.mpkbf[a][d][0], hi.data(), hi.size() * sizeof(uint16_t));
            h2d(s, s.mpkbf[a][d][1], lo.data(), lo.size() * sizeof(uint16_t));
        }
    }
    {
        std::vector<double> inv(s.n_cell);
        for (int z = 0; z < s.nz; z++)
            for (int y = 0; y < s.ny; y++)
                for (int x = 0; x < s.nx; x++) {
                    const bool zero = x == e[X]->zero_mode && y == e[Y]->zero_mode && z == e[Z]->zero_mode;
                    inv[x + (i64)s.nx * (y + (i64)s.ny * z)] =
                        zero ? 0.0 : 1.0 / (e[Z]->lambda[z] + e[Y]->lambda[y] + e[X]->lambda[x]);
                }
        h2d(s, s.inv_lambda, inv.data(), inv.size() * sizeof(double));
        std::vector<float> invf(inv.size());
        for (size_t i = 0; i < inv.size(); i++) invf[i] = (float)inv[i];
        s.il_f = carve<float>(s, P.il_f);
        h2d(s, s.il_f, invf.data(), invf.size() * sizeof(float));
        // even/odd split where the axis allows it; spectral order [even | odd] on those axes
        const std::vector<double>* lam[3];
        int zero[3];
        SymSplit sps[3];
        for (int a = 0; a < 3; a++) {
            sps[a] = symmetric_split(s.ax[a]);
            const SymSplit& sp = sps[a];
            s.sym_ok[a] = sp.ok;
            lam[a] = sp.ok ? &sp.lambda : &e[a]->lambda;
            zero[a] = sp.ok ? sp.zero : e[a]->zero_mode;
            const std::vector<double>* mats[4] = {&sp.fe, &sp.fo, &sp.be, &sp.bo};
            for (int d = 0; d < 4; d++) {
                s.spk[a][d] = carve<double>(s, P.spk[a][d]);
                s.spk32[a][d] = carve<float>(s, P.spk32[a][d]);
                s.spkbf[a][d][0] = carve<uint16_t>(s, P.spkbf[a][d][0]);
                s.spkbf[a][d][1] = carve<uint16_t>(s, P.spkbf[a][d][1]);
                if (!sp.ok) continue;
                const std::vector<double> f64 = pack_fragments<double>(*mats[d], sp.h);
                const std::vector<float> f32 = pack_fragments<float>(*mats[d], sp.h);
                std::vector<uint16_t> hi, lo;
                pack_bf16_fragments(*mats[d], sp.h, hi, lo);
                h2d(s, s.spk[a][d], f64.data(), f64.size() * sizeof(double));
                h2d(s, s.spk32[a][d], f32.data(), f32.size() * sizeof(float));
                h2d(s, s.spkbf[a][d][0], hi.data(), hi.size() * sizeof(uint16_t));
                h2d(s, s.spkbf[a][d][1], lo.data(), lo.size() * sizeof(uint16_t));
            }
        }
        std::vector<double> invs(s.n_cell);
        std::vector<float> invsf(s.n_cell);
        for (int z = 0; z < s.nz; z++)
            for (int y = 0; y < s.ny; y++)
                for (int x = 0; x < s.nx; x++) {
                    const bool zm = x == zero[X] && y == zero[Y] && z == zero[Z];
                    const double v = zm ? 0.0 : 1.0 / ((*lam[Z])[z] + (*lam[Y])[y] + (*lam[X])[x]);
                    const i64 to = x + (i64)s.nx * (y + (i64)s.ny * z);
                    invs[to] = v;
                    invsf[to] = (float)v;
                }
        s.il_sym = carve<double>(s, P.il_sym);
        s.il_sym_f = carve<float>(s, P.il_sym_f);
        h2d(s, s.il_sym, invs.data(), invs.size() * sizeof(double));
        h2d(s, s.il_sym_f, invsf.data(), invsf.size() * sizeof(float));
    }
    // probes: the cell containing (px L, py L) on the mid-span layer (backend.cu:1058-1070)
    std::vector<int> cells;
    for (int p = 0; p < s.c.n_probes; p++) {
        auto cell_of = [](const Axis& g, double frac) {
            const int k = (int)(std::upper_bound(g.face.begin(), g.face.end(), frac * g.length) - g.face.begin()) - 1;
            return std::min(g.n - 1, std::max(0, k));
        };
        cells.push_back(cell_of(s.ax[X], s.c.probe_x[p]));
        cells.push_back(cell_of(s.ax[Y], s.c.probe_y[p]));
        cells.push_back(s.nz / 2);
    }
    if (!cells.empty()) h2d(s, s.probe_cells, cells.data(), cells.size() * sizeof(int));
    HIP_CHECK(hipStreamSynchronize(s.stream));
}

void destroy(Solver* s) {
    if (!s) return;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    s->timer.reset();
    if (s->arena) (void)hipFree(s->arena);
    if (s->h_scal) (void)hipHostFree(s->h_scal);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

// ------------------------------------------------------------------------------------------------ pressure
// backend.cu:801-814: phi += rms * S (Lambda^-1 o S^T (r / rms)), r2 = sum r^2. In FP64 mode every
// transform is FP64; in the mixed modes the scaled residual is rounded to FP32 (to_float_scaled,
// backend.cu:403-406), the six transforms run in FP32 or split-BF16 x3, the eigenvalue division in FP32
// (backend.cu:407-410) and the update is added to phi in FP64 (backend.cu:411-414).
void pressure_pass(Solver& s, const double* r, double r2, double* ph) {
    const double rms = std::sqrt(r2 / s.n_cell);
    if (!(rms > 0)) return;
    const int mode = s.mode;
    const bool wide = mode == CAV_PREC_FP64;
    const i64 plane = (i64)s.nx * s.ny;
    bool sym[3];
    for (int k = 0; k < 3; k++) sym[k] = s.use_sym && s.sym_ok[k];
    // matrices of one axis transform: dense S^T / S, or the even/odd halves (forward fe, fo; backward be, bo)
    auto mat = [&](int axis, int dir, AxisArgs& a) {
        a.sym = sym[axis] ? 1 + dir : 0;
        if (!sym[axis]) {
            a.mpk2 = a.mpk2_lo = nullptr;
            if (mode == 0) { a.mpk = s.mpk[axis][dir]; a.mpk_lo = nullptr; }
            else if (mode == 1) { a.mpk = s.mpk32[axis][dir]; a.mpk_lo = nullptr; }
            else { a.mpk = s.mpkbf[axis][dir][0]; a.mpk_lo = s.mpkbf[axis][dir][1]; }
            return;
        }
        const int de = 2 * dir, dd = 2 * dir + 1;
        if (mode == 0) { a.mpk = s.spk[axis][de]; a.mpk2 = s.spk[axis][dd]; a.mpk_lo = a.mpk2_lo = nullptr; }
        else if (mode == 1) { a.mpk = s.spk32[axis][de]; a.mpk2 = s.spk32[axis][dd]; a.mpk_lo = a.mpk2_lo = nullptr; }
        else {
            a.mpk = s.spkbf[axis][de][0]; a.mpk_lo = s.spkbf[axis][de][1];
            a.mpk2 = s.spkbf[axis][dd][0]; a.mpk2_lo = s.spkbf[axis][dd][1];
        }
    };
    const bool any_sym = sym[X] || sym[Y] || sym[Z];
    AxisArgs a{};
    a.in_scale = 1.0;
    a.out_scale = 1.0;
    // forward x: t2 = (r/rms) S_x
    a.in = r; a.out = s.t2; mat(X, 0, a); a.n = s.nx; a.A = (i64)s.ny * s.nz; a.B = 1;
    a.in_scale = 1.0 / rms; a.epi = 0;
    axis_transform(s.ex, true, mode, true, wide, a, F_XFORM_X);
    a.in_scale = 1.0;
    // forward y (each z layer): t1 = S_y^T t2
    a.in = s.t2; a.out = s.t1; mat(Y, 0, a); a.n = s.ny; a.A = s.nz; a.B = s.nx;
    axis_transform(s.ex, false, mode, wide, wide, a, F_XFORM_Y);
    // forward z, divided by the eigenvalues: t2 = (S_z^T t1) * inv_lambda
    a.in = s.t1; a.out = s.t2; mat(Z, 0, a); a.n = s.nz; a.A = 1; a.B = plane; a.epi = 1;
    if (wide) a.il = any_sym ? (const void*)s.il_sym : (const void*)s.inv_lambda;
    else a.il = any_sym ? (const void*)s.il_sym_f : (const void*)s.il_f;
    axis_transform(s.ex, false, mode, wide, wide, a, F_XFORM_Z);
    // back z: t1 = S_z t2
    a.in = s.t2; a.out = s.t1; mat(Z, 1, a); a.epi = 0; a.il = nullptr;
    axis_transform(s.ex, false, mode, wide, wide, a, F_XFORM_Z);
    // back y: t2 = S_y t1
    a.in = s.t1; a.out = s.t2; mat(Y, 1, a); a.n = s.ny; a.A = s.nz; a.B = s.nx;
    axis_transform(s.ex, false, mode, wide, wide, a, F_XFORM_Y);
    // back x: phi += rms * (t2 S_x^T)
    a.in = s.t2; a.out = ph; mat(X, 1, a); a.n = s.nx; a.A = (i64)s.ny * s.nz; a.B = 1; a.epi = 2;
    a.out_scale = rms;
    axis_transform(s.ex, true, mode, wide, true, a, F_XFORM_X);
}

// backend.cu:816-847 and backend.cu:849-854. mm's wall faces are zeroed in place; with a guess, phi
// already holds it on entry. With hist_out, dphi = phi - base is written after the mean is removed.
void project(Solver& s, double* const mm[3], bool guess, double* ph, double* const out[3], double* hist_out) {
    const i64 n = s.n_cell;
    zero_walls(s.ex, s.geo, mm);
    const double* mc[3] = {mm[0], mm[1], mm[2]};
    const int nb = divergence(s.ex, s.geo, mc, s.rhs, s.partial);
    reduce_final(s.ex, s.partial, nb, 3, s.scal + S_RHS2);
    if (guess) {
        const int nr = residual(s.ex, s.geo, ph, s.rhs, s.res, s.partial + RED_BLOCKS);
        reduce_final(s.ex, s.partial + RED_BLOCKS, nr, 3, s.scal + S_RES2);
    }
    sync_read(s, S_RHS2, 2);
    const double rhs2 = s.h_scal[S_RHS2];
    const double rhs_norm = std::sqrt(rhs2);
    s.phi_passes = 0;
    if (!std::isfinite(rhs_norm)) fail("pressure solve: right-hand side is NaN/inf (the run has blown up)");
    if (rhs_norm == 0.0) {
        fill(s.ex, ph, n, 0.0);
        s.phi_rel = 0;
        if (hist_out) finalize_phi(s.ex, n, ph, s.scal + S_ZERO, s.base, hist_out);
    } else {
        const double* r = s.rhs;
        double r2 = rhs2, rel = 1.0;
        if (guess) {
            const double g2 = s.h_scal[S_RES2];
            rel = std::sqrt(g2) / rhs_norm;
            r = s.res;
            r2 = g2;
            if (rel >= 1.0) { fill(s.ex, ph, n, 0.0); r = s.rhs; r2 = rhs2; rel = 1.0; }
        } else {
            fill(s.ex, ph, n, 0.0);
        }
        while (rel > s.c.fdm_rtol && s.phi_passes < s.c.fdm_max_passes) {
            pressure_pass(s, r, r2, ph);
            s.phi_passes++;
            const int nr = residual(s.ex, s.geo, ph, s.rhs, s.res, s.partial);
            reduce_final(s.ex, s.partial, nr, 3, s.scal + S_RES2);
            sync_read(s, S_RES2, 1);
            r = s.res;
            r2 = s.h_scal[S_RES2];
            rel = std::sqrt(r2) / rhs_norm;
            if (!std::isfinite(rel)) fail("pressure solve: residual is NaN/inf");
        }
        s.projections++;
        s.pass_total += s.phi_passes;
        if (rel > s.c.fdm_rtol) {
            char msg[200];
            std::snprintf(msg, sizeof msg, "pressure solve: residual %.2e > FDM_RTOL after %d passes", rel, s.phi_passes);
            fail(msg);
        }
        s.phi_rel = rel;
        reduce_array(s.ex, ph, n, 1, s.partial + 2 * RED_BLOCKS, s.scal + S_PHISUM);
        finalize_phi(s.ex, n, ph, s.scal + S_PHISUM, s.base, hist_out);
    }
    const double* mc2[3] = {mm[0], mm[1], mm[2]};
    subtract_gradient(s.ex, s.geo, mc2, ph, out);
}

// backend.cu:857-873: which history entries the warm start combines, and with which weights
struct Guess { int nh = 0; int b1 = 0, b2 = 0; double c1 = 0, c2 = 0; };
Guess guess_plan(const Solver& s, double tau, double scale) {
    Guess gp;
    if (!s.c.phi_extrapolate || s.hist.empty()) return gp;
    std::vector<int> order(s.hist.size());
    for (size_t k = 0; k < order.size(); k++) order[k] = (int)k;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        const double da = std::fabs(s.hist[a].t - tau), db = std::fabs(s.hist[b].t - tau);
        return da != db ? da < db : a > b;
    });
    const History& h1 = s.hist[order[0]];
    const History* h2 = nullptr;
    for (size_t k = 1; k < order.size() && !h2; k++)
        if (std::fabs(s.hist[order[k]].t - h1.t) > 1e-12 * std::max(1.0, std::fabs(h1.t))) h2 = &s.hist[order[k]];
    const double w2 = h2 ? (tau - h1.t) / (h2->t - h1.t) : 0.0, w1 = 1.0 - w2;
    gp.nh = h2 ? 2 : 1;
    gp.b1 = h1.buf;
    gp.c1 = scale * w1 / h1.scale;
    if (h2) { gp.b2 = h2->buf; gp.c2 = scale * w2 / h2->scale; }
    return gp;
}

// backend.cu:877-955
void step(Solver& s) {
    const double* mc[3] = {s.m[0], s.m[1], s.m[2]};
    const double* uc[3] = {s.u[0], s.u[1], s.u[2]};
    gauge_reset(s.ex, s.geo, mc, uc, s.phi, s.c.reinit_r, 1.0 - s.c.reinit_r, s.m0, s.phi0);
    cfl_max(s.ex, s.geo, uc, s.partial, s.scal + S_CFL);
    sync_read(s, S_CFL, 1);
    const double rates[2] = {s.h_scal[S_CFL], s.viscous_rate};
    double use[2] = {rates[0], rates[1]};
    if (s.c.dt_lag && s.have_prev_rates) { use[0] = s.prev_rates[0]; use[1] = s.prev_rates[1]; }
    s.prev_rates[0] = rates[0]; s.prev_rates[1] = rates[1]; s.have_prev_rates = true;
    if (!std::isfinite(use[0])) fail("velocity is NaN/inf: the solution has blown up");
    double dt;
    if (!s.c.adaptive_dt) { dt = s.c.dt_fixed; s.dt_reason = 0; }
    else {
        const double caps[3] = {s.c.dt_max, s.c.cfl_target / std::max(use[0], 1e-300),
                                s.c.diffusion_limit / std::max(use[1], 1e-300)};
        int best = 0;
        for (int k = 1; k < 3; k++) if (caps[k] < caps[best]) best = k;
        dt = caps[best];
        s.dt_reason = best + 1;
    }
    if (s.c.dt_lag && s.last_dt > 0) dt = std::min(dt, 1.1 * s.last_dt);

    const double a0s[3] = {0.0, 0.75, 1.0 / 3.0}, bs[3] = {1.0, 0.25, 2.0 / 3.0};
    const double times[3] = {s.t, s.t + dt, s.t + 0.5 * dt};
    double* cur_u[3] = {s.u[0], s.u[1], s.u[2]};
    double* cur_m[3] = {s.m0[0], s.m0[1], s.m0[2]};
    double* cur_phi = s.phi0;
    for (int st = 0; st < 3; st++) {
        const double a0 = a0s[st], b = bs[st], ts = times[st];
        double** nm = s.new_m[st % 2];
        double** nu = s.new_u[st % 2];
        StageArgs sa{};
        for (int k = 0; k < 3; k++) { sa.u[k] = cur_u[k]; sa.m[k] = cur_m[k]; sa.m0[k] = s.m0[k]; sa.out[k] = nm[k]; }
        for (int k = 0; k < E_COUNT; k++) sa.edge[k] = s.edge[k];
        sa.nu = s.c.nu; sa.lid = lid_speed(s, ts); sa.dt = dt; sa.a0 = a0; sa.b = b;
        sa.has_base = st > 0 ? 1 : 0;
        sa.o4 = s.c.impulse_order == 4;
        stage(s.ex, s.geo, sa);
        double* new_phi = s.phi_stage[st % 2];
        const Guess gp = guess_plan(s, ts, b * dt);
        base_guess(s.ex, s.n_cell, s.phi0, cur_phi, a0, b, a0 != 0.0, gp.nh ? s.hist_buf[gp.b1] : nullptr, gp.c1,
                   gp.nh == 2 ? s.hist_buf[gp.b2] : nullptr, gp.c2, gp.nh, s.base, new_phi);
        int buf = -1;
        if (s.c.phi_extrapolate) {
            for (int k = 0; k < 4 && buf < 0; k++) {
                bool used = false;
                for (const History& h : s.hist) used |= (h.buf == k);
                if (!used) buf = k;
            }
        }
        project(s, nm, true, new_phi, nu, buf >= 0 ? s.hist_buf[buf] : nullptr);
        if (buf >= 0) {
            s.hist.push_back({ts, b * dt, buf});
            if (s.hist.size() > 3) s.hist.erase(s.hist.begin());
        }
        for (int k = 0; k < 3; k++) { cur_u[k] = nu[k]; cur_m[k] = nm[k]; }
        cur_phi = new_phi;
    }
    for (int k = 0; k < 3; k++) {
        std::swap(s.u[k], s.new_u[0][k]);
        std::swap(s.m[k], s.new_m[0][k]);
    }
    std::swap(s.phi, s.phi_stage[0]);
    s.t += dt;
    s.steps++;
    s.last_dt = dt;
    s.last_cfl = dt * use[0];
}

// backend.cu:958-978
void kick(Solver& s, const double* const nz_[3], double amp, int passes) {
    for (int k = 0; k < 3; k++) if (!nz_[k]) fail("kick: noise arrays must not be NULL");
    for (int k = 0; k < 3; k++) {
        const i64 n = s.size[k];
        const int sx = s.nx + (k == 0), sy = s.ny + (k == 1), sz = s.nz + (k == 2);
        h2d(s, s.noise[0], nz_[k], n * sizeof(double));
        double *a = s.noise[0], *b = s.noise[1];
        for (int p = 0; p < std::max(1, passes); p++)
            for (int axis = 2; axis >= 0; axis--) {
                smooth(s.ex, a, b, axis, sx, sy, sz);
                std::swap(a, b);
            }
        reduce_array(s.ex, a, n, 3, s.partial, s.scal + S_KICK);
        sync_read(s, S_KICK, 1);
        const double rms = std::sqrt(s.h_scal[S_KICK] / n);
        add_noise(s.ex, s.u[k], a, amp * s.c.u_lid / std::max(rms, 1e-30), s.tmp[k], n);
    }
    project(s, s.tmp, false, s.phi_stage[0], s.new_u[0], nullptr);
    for (int k = 0; k < 3; k++) {
        add_difference(s.ex, s.m[k], s.new_u[0][k], s.u[k], s.size[k]);
        std::swap(s.u[k], s.new_u[0][k]);
    }
    HIP_CHECK(hipStreamSynchronize(s.stream));
}

void initialize(Solver& s, const double* vn, const double* wn) {       // backend.cu:1075-1088
    const double* src[3] = {nullptr, vn, wn};
    for (int k = 0; k < 3; k++) {
        if (src[k]) h2d(s, s.m[k], src[k], s.size[k] * sizeof(double));
        else fill(s.ex, s.m[k], s.size[k], 0.0);
    }
    project(s, s.m, false, s.phi, s.u, nullptr);
    s.t = 0; s.steps = 0; s.last_dt = 0; s.hist.clear(); s.have_prev_rates = false;
    HIP_CHECK(hipStreamSynchronize(s.stream));
}

void fill_report(const Solver& s, CavReport* r) {                      // backend.cu:995-999
    if (!r) return;
    r->steps = s.steps; r->time = s.t; r->dt = s.last_dt; r->cfl = s.last_cfl; r->phi_rel = s.phi_rel;
    r->dt_reason = s.dt_reason; r->phi_passes = s.phi_passes;
}

void diagnostics(Solver& s, CavDiagnostics* d) {                       // backend.cu:1096-1129
    const double* uc[3] = {s.u[0], s.u[1], s.u[2]};
    cell_diagnostics(s.ex, s.geo, uc, lid_speed(s, s.t), s.partial, s.scal + S_KE);
    for (int k = 0; k < 3; k++) reduce_array(s.ex, s.u[k], s.size[k], 2, s.partial + 3 * RED_BLOCKS, s.scal + S_UMAX + k);
    cfl_max(s.ex, s.geo, uc, s.partial + 3 * RED_BLOCKS, s.scal + S_CFL);
    for (int k = 0; k < 3; k++) max_abs_diff(s.ex, s.m[k], s.u[k], s.size[k], s.partial + 3 * RED_BLOCKS, s.scal + S_GAUGE + k);
    sync_read(s, 0, S_KICK);
    const double* h = s.h_scal;
    d->kinetic_energy = h[S_KE] / s.volume;
    d->enstrophy = 0.5 * h[S_ENS] / s.volume;
    d->div_rms = std::sqrt(h[S_DIV2] / s.volume);
    double umax = 0;
    for (int k = 0; k < 3; k++) umax = std::max(umax, h[S_UMAX + k]);
    d->umax = umax;
    d->cfl_rate = h[S_CFL];
    double best = -1;
    int best_comp = 0, best_i = 0;
    for (int k = 0; k < 3; k++) {
        const double val = h[S_GAUGE + k];
        if (val > best) { best = val; best_comp = k; }
    }
    if (best >= 0) {
        first_index_of_abs_diff(s.ex, s.m[best_comp], s.u[best_comp], s.size[best_comp], s.scal + S_GAUGE + best_comp, s.first);
        d2h(s, &best_i, s.first, sizeof(int));
    }
    const int sx = s.nx + (best_comp == 0), sy = s.ny + (best_comp == 1);
    const int ix = best_i % sx, iy = (best_i / sx) % sy;
    d->gauge_ratio = best / std::max(umax, 1e-30);
    d->gauge_x = (best_comp == 0 ? s.ax[X].face[ix] : s.ax[X].centre[std::min(ix, s.nx - 1)]) / s.c.length;
    d->gauge_y = (best_comp == 1 ? s.ax[Y].face[iy] : s.ax[Y].centre[std::min(iy, s.ny - 1)]) / s.c.length;
}

double* field_ptr(Solver& s, int f) {
    if (f < 0 || f > 6) fail("field must be 0..6");
    return f < 3 ? s.u[f] : f < 6 ? s.m[f - 3] : s.phi;
}
i64 field_size(const Solver& s, int f) { return f < 3 ? s.size[f] : f < 6 ? s.size[f - 3] : s.n_cell; }

void get_field(Solver& s, int f, void* out, bool wide) {
    double* p = field_ptr(s, f);
    const i64 n = field_size(s, f);
    if (wide || f == 6) { d2h(s, out, p, n * sizeof(double)); return; }
    std::vector<double> tmpv(n);
    d2h(s, tmpv.data(), p, n * sizeof(double));
    float* o = (float*)out;
    for (i64 i = 0; i < n; i++) o[i] = (float)tmpv[i];
}

void set_field(Solver& s, int f, const void* in, bool wide) {          // backend.cu:1147-1155
    double* p = field_ptr(s, f);
    const i64 n = field_size(s, f);
    if (wide || f == 6) h2d(s, p, in, n * sizeof(double));
    else {
        std::vector<double> tmpv(n);
        const float* src = (const float*)in;
        for (i64 i = 0; i < n; i++) tmpv[i] = (double)src[i];
        h2d(s, p, tmpv.data(), n * sizeof(double));
    }
    s.hist.clear();
    s.have_prev_rates = false;
}

template <class F> int guard(F f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        g_error = e.what();
        return 1;
    } catch (...) {
        g_error = "unknown error";
        return 1;
    }
}

Solver& need(cav_solver* s) {
    if (!s) fail("no solver: call solver_create first");
    return *s;
}

cav_solver* g_singleton = nullptr;

}  // namespace
}  // namespace cav

using namespace cav;

// ================================================================================================ handle API
extern "C" {

int cav_create(const CavConfig* config, cav_solver** out) {
    return guard([&] {
        if (!config || !out) fail("cav_create: NULL argument");
        *out = nullptr;
        auto* s = new cav_solver();
        try {
            create(*s, *config);
        } catch (...) {
            destroy(s);
            throw;
        }
        *out = s;
    });
}
int cav_destroy(cav_solver* s) { return guard([&] { destroy(s); }); }
int cav_initialize(cav_solver* s, const double* v, const double* w) { return guard([&] { initialize(need(s), v, w); }); }
int cav_step(cav_solver* s, CavReport* r) { return guard([&] { step(need(s)); fill_report(*s, r); }); }
int cav_steps(cav_solver* s, int n, CavReport* r) {
    return guard([&] {
        for (int k = 0; k < n; k++) step(need(s));
        fill_report(need(s), r);
    });
}
int cav_kick(cav_solver* s, const double* nu, const double* nv, const double* nw, double amp, int passes) {
    return guard([&] { const double* n[3] = {nu, nv, nw}; kick(need(s), n, amp, passes); });
}
int cav_report(cav_solver* s, CavReport* r) { return guard([&] { fill_report(need(s), r); }); }
int cav_diagnostics(cav_solver* s, CavDiagnostics* d) { return guard([&] { diagnostics(need(s), d); }); }
int cav_grid(cav_solver* s, int axis, double* faces, double* centres) {
    return guard([&] {
        if (axis < 0 || axis > 2) fail("axis must be 0, 1 or 2");
        const Axis& g = need(s).ax[axis];
        if (faces) std::copy(g.face.begin(), g.face.end(), faces);
        if (centres) std::copy(g.centre.begin(), g.centre.end(), centres);
    });
}
int cav_get_field(cav_solver* s, int f, void* out) { return guard([&] { get_field(need(s), f, out, false); }); }
int cav_set_field(cav_solver* s, int f, const void* in) { return guard([&] { set_field(need(s), f, in, false); }); }
int cav_get_field64(cav_solver* s, int f, double* out) { return guard([&] { get_field(need(s), f, out, true); }); }
int cav_set_field64(cav_solver* s, int f, const double* in) { return guard([&] { set_field(need(s), f, in, true); }); }
int cav_set_clock(cav_solver* s, double t, long long steps, double last_dt) {
    return guard([&] { Solver& q = need(s); q.t = t; q.steps = steps; q.last_dt = last_dt; });
}
int cav_probes(cav_solver* s, float* out) {                            // backend.cu:1159-1166
    return guard([&] {
        Solver& q = need(s);
        if (q.c.n_probes == 0) return;
        probe_values(q.ex, q.geo, q.u[0], q.u[1], q.probe_cells, q.c.n_probes, q.probe_out);
        d2h(q, out, q.probe_out, 2 * q.c.n_probes * sizeof(float));
    });
}
int cav_add_statistics(cav_solver* s, double weight, double* centreline) {   // backend.cu:1168-1175
    return guard([&] {
        Solver& q = need(s);
        const double* uc[3] = {q.u[0], q.u[1], q.u[2]};
        if (centreline && q.nx != q.ny) fail("centreline sample needs nx == ny");
        add_statistics(q.ex, q.geo, uc, weight, q.stats);
        if (centreline) {
            centreline_sample(q.ex, q.geo, uc, q.line);
            d2h(q, centreline, q.line, 5 * (size_t)q.nx * sizeof(double));
        } else {
            HIP_CHECK(hipStreamSynchronize(q.stream));
        }
    });
}
int cav_get_statistics(cav_solver* s, double* sums) {
    return guard([&] { Solver& q = need(s); d2h(q, sums, q.stats, 6 * (size_t)q.nx * q.ny * sizeof(double)); });
}
int cav_set_statistics(cav_solver* s, const double* sums) {
    return guard([&] { Solver& q = need(s); h2d(q, q.stats, sums, 6 * (size_t)q.nx * q.ny * sizeof(double)); });
}
int cav_midplane(cav_solver* s, float* out) {                           // backend.cu:1178-1184
    return guard([&] {
        Solver& q = need(s);
        const double* uc[3] = {q.u[0], q.u[1], q.u[2]};
        midplane(q.ex, q.geo, uc, lid_speed(q, q.t), q.plane);
        d2h(q, out, q.plane, 4 * (size_t)q.nx * q.ny * sizeof(float));
    });
}
int cav_q_volume(cav_solver* s, int colour_w, float* qv, float* colour) {    // backend.cu:1185-1194
    return guard([&] {
        Solver& q = need(s);
        const double* uc[3] = {q.u[0], q.u[1], q.u[2]};
        const double LU = q.c.length / q.c.u_lid;
        float* dq = (float*)q.t1;                      // scratch: two float cell arrays fit in one double one
        float* dc = dq + q.n_cell;
        q_criterion(q.ex, q.geo, uc, lid_speed(q, q.t), LU * LU, 1.0 / q.c.u_lid, colour_w, dq, dc);
        d2h(q, qv, dq, q.n_cell * sizeof(float));
        d2h(q, colour, dc, q.n_cell * sizeof(float));
    });
}
int cav_set_precision(cav_solver* s, int mode) {
    return guard([&] {
        Solver& q = need(s);
        if (mode != CAV_PREC_FP64 && mode != CAV_PREC_FDM_FP32 && mode != CAV_PREC_FDM_BF16X3)
            fail("precision mode must be 0 (FP64), 1 (FP32 FDM) or 2 (split-BF16 FDM)");
        q.mode = mode;
    });
}
int cav_set_fdm_layout(cav_solver* s, int symmetric) {
    return guard([&] { need(s).use_sym = symmetric != 0; });
}
int cav_fdm_layout(cav_solver* s, int* axes) {
    return guard([&] {
        Solver& q = need(s);
        *axes = (q.use_sym && q.sym_ok[X] ? 1 : 0) | (q.use_sym && q.sym_ok[Y] ? 2 : 0) | (q.use_sym && q.sym_ok[Z] ? 4 : 0);
    });
}
int cav_pass_counters(cav_solver* s, long long* projections, long long* passes) {
    return guard([&] { Solver& q = need(s); *projections = q.projections; *passes = q.pass_total; });
}
int cav_footprint(const CavConfig* config, int mode, long long* bytes) {
    return guard([&] {
        if (!config || !bytes) fail("cav_footprint: NULL argument");
        validate(*config);
        (void)mode;
        Plan P(config->nx, config->ny, config->nz);
        *bytes = (long long)P.L.total;
    });
}
const char* cav_last_error(void) { return g_error.c_str(); }

int cav_time_steps(cav_solver* s, int n, double* ms) {
    return guard([&] {
        Solver& q = need(s);
        hipEvent_t a, b;
        HIP_CHECK(hipEventCreate(&a));
        HIP_CHECK(hipEventCreate(&b));
        HIP_CHECK(hipDeviceSynchronize());
        HIP_CHECK(hipEventRecord(a, q.stream));
        for (int k = 0; k < n; k++) step(q);
        HIP_CHECK(hipEventRecord(b, q.stream));
        HIP_CHECK(hipEventSynchronize(b));
        HIP_CHECK(hipDeviceSynchronize());
        float t = 0;
        HIP_CHECK(hipEventElapsedTime(&t, a, b));
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        *ms = t;
    });
}
int cav_copy_probe(long long bytes, int repeats, double* ms) {
    return guard([&] {
        if (bytes < 16 || repeats < 1) fail("cav_copy_probe: bytes >= 16 and repeats >= 1");
        double *a = nullptr, *b = nullptr;
        hipStream_t st;
        HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        HIP_CHECK(hipMalloc((void**)&a, bytes));
        HIP_CHECK(hipMalloc((void**)&b, bytes));
        HIP_CHECK(hipMemsetAsync(a, 0, bytes, st));
        Exec ex;
        ex.stream = st;
        hipEvent_t e0, e1;
        HIP_CHECK(hipEventCreate(&e0));
        HIP_CHECK(hipEventCreate(&e1));
        copy_probe(ex, a, b, bytes / 8);
        HIP_CHECK(hipEventRecord(e0, st));
        for (int r = 0; r < repeats; r++) copy_probe(ex, a, b, bytes / 8);
        HIP_CHECK(hipEventRecord(e1, st));
        HIP_CHECK(hipEventSynchronize(e1));
        float t = 0;
        HIP_CHECK(hipEventElapsedTime(&t, e0, e1));
        *ms = t / repeats;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipFree(a);
        (void)hipFree(b);
        (void)hipStreamDestroy(st);
    });
}

int cav_set_kernel_timing(cav_solver* s, int on) {
    return guard([&] {
        Solver& q = need(s);
        if (on) {
            if (!q.timer) q.timer.reset(new EventTimer());
            q.ex.timer = q.timer.get();
        } else {
            if (q.timer) q.timer->flush();
            q.ex.timer = nullptr;
        }
    });
}
int cav_kernel_times(cav_solver* s, int max, const char** names, double* ms, long long* launches, int* count) {
    return guard([&] {
        Solver& q = need(s);
        int k = 0;
        if (q.timer) q.timer->flush();
        for (int f = 0; f < F_COUNT && k < max; f++, k++) {
            names[k] = family_name(f);
            ms[k] = q.timer ? q.timer->ms[f] : 0.0;
            launches[k] = q.timer ? q.timer->launches[f] : 0;
        }
        *count = k;
    });
}

// ================================================================================================ reference ABI
// backend.cu:980-1195: one process-wide solver.
int solver_destroy(void) {
    return guard([&] {
        destroy(g_singleton);
        g_singleton = nullptr;
    });
}
int solver_create(const CavConfig* config) {
    return guard([&] {
        destroy(g_singleton);
        g_singleton = nullptr;
        if (!config) fail("solver_create: NULL config");
        auto* s = new cav_solver();
        try {
            create(*s, *config);
        } catch (...) {
            destroy(s);
            throw;
        }
        g_singleton = s;
    });
}
int solver_initialize(const double* v, const double* w) { return cav_initialize(g_singleton, v, w); }
int solver_step(CavReport* r) { return cav_step(g_singleton, r); }
int solver_kick(const double* nu, const double* nv, const double* nw, double amp, int passes) {
    return cav_kick(g_singleton, nu, nv, nw, amp, passes);
}
int solver_report(CavReport* r) { return cav_report(g_singleton, r); }
int solver_diagnostics(CavDiagnostics* d) { return cav_diagnostics(g_singleton, d); }
int solver_grid(int axis, double* faces, double* centres) { return cav_grid(g_singleton, axis, faces, centres); }
int solver_get_field(int field, void* out) { return cav_get_field(g_singleton, field, out); }
int solver_set_field(int field, const void* in) { return cav_set_field(g_singleton, field, in); }
int solver_set_clock(double t, long long steps, double last_dt) { return cav_set_clock(g_singleton, t, steps, last_dt); }
int solver_probes(float* out) { return cav_probes(g_singleton, out); }
int solver_add_statistics(double weight, double* centreline) { return cav_add_statistics(g_singleton, weight, centreline); }
int solver_get_statistics(double* sums) { return cav_get_statistics(g_singleton, sums); }
int solver_set_statistics(const double* sums) { return cav_set_statistics(g_singleton, sums); }
int solver_midplane(float* out) { return cav_midplane(g_singleton, out); }
int solver_q_volume(int colour_w, float* q, float* colour) { return cav_q_volume(g_singleton, colour_w, q, colour); }
const char* solver_last_error(void) { return g_error.c_str(); }
int solver_get_field64(int field, double* out) { return cav_get_field64(g_singleton, field, out); }
int solver_set_field64(int field, const double* in) { return cav_set_field64(g_singleton, field, in); }
int solver_set_precision(int mode) { return cav_set_precision(g_singleton, mode); }
int solver_pass_counters(long long* projections, long long* passes) {
    return cav_pass_counters(g_singleton, projections, passes);
}

}  // extern "C"
