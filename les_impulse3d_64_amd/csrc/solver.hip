// solver.hip — host side of libcavity_solver.so (include/cavity_solver.h): cavity set-up, batches of cavities
// stepped together, HIP graphs, and the C-ABI.
//
// A cavity owns a HIP stream and one device arena: the FP64 state (u, m, phi), work arrays, packed FDM
// operators, reduction scratch, its control block (Ctl) and its descriptor (Mem). The step follows the
// reference's sequence (backend.cu:877-955) with every decision (time step, warm start, FDM convergence)
// taken by control kernels on the device, so a step is a fixed sequence of launches with no host round trip:
// it is captured once into a HIP graph and replayed. The FDM refinement loop is enqueued for the passes the
// precision mode normally needs (plus a spare in split BF16, and in a batch once one of its cavities has
// stalled), passes after convergence exiting at their first instruction; a cavity whose refinement would
// need more stops (stalled) with its state untouched, and the host redoes that step from the control block
// it began with, with all fdm_max_passes passes enqueued.
//
// A batch is a list of cavities advanced by the same launches (kernels.h); one cavity alone is a batch of
// one, an ensemble (cav_ensemble_*) is a batch of many with its own stream and graph.
#include <algorithm>
#include <array>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>

#include "kernels.h"

// Launch plan of cav_batch_steps_noalloc (member descriptor pointers, then the block offsets of every launch
// shape, then a one-member plan for redone steps) in a fixed device array: at most 256 cavities per call
namespace cav {
constexpr int NOALLOC_MAX_MEMBERS = 256;
constexpr size_t NOALLOC_PTRS = (NOALLOC_MAX_MEMBERS * sizeof(Mem*) + 255) / 256 * 256;
constexpr size_t NOALLOC_SEG = ((size_t)SH_COUNT * (NOALLOC_MAX_MEMBERS + 1) * sizeof(int) + 255) / 256 * 256;
constexpr size_t NOALLOC_ONE = 256 + ((size_t)SH_COUNT * 2 * sizeof(int) + 255) / 256 * 256;
constexpr size_t NOALLOC_SCRATCH_BYTES = NOALLOC_PTRS + NOALLOC_SEG + NOALLOC_ONE;
static_assert(NOALLOC_SCRATCH_BYTES <= 64 * 1024, "the no-allocation plan array stays within 64 KiB");
}  // namespace cav
// One copy per device, held until the process ends (file scope, device memory)
static __device__ __align__(256) char scratch[cav::NOALLOC_SCRATCH_BYTES];

namespace cav {
namespace {

thread_local std::string g_error;

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

struct Layout {              // byte offsets of every device array in the arena
    size_t total = 0;
    size_t take(size_t bytes) {
        const size_t off = total;
        total += (bytes + 255) / 256 * 256;
        return off;
    }
};

// Launch plan of a batch: member descriptors and segment tables on the device, the captured step graph.
struct BatchRt {
    std::vector<cav_solver*> members;
    std::vector<unsigned> versions;     // member descriptor versions the plan was built for
    Batch b;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    char* dev = nullptr;                // member pointer array + segment table
    size_t dev_bytes = 0;               // its allocated size (a rebuild that fits reuses it)
    int passes_fast = 0, passes_full = 0;   // refinement passes enqueued per projection: usual / all allowed
    int passes_set = 0;                     // > 0: the usual count replaced (testing the redo path)
    bool spare_on = false;                  // the plan enqueues one pass beyond the usual count (a member wants it)
    bool use_graph = true;
    hipGraph_t graph[2] = {nullptr, nullptr};          // [0] usual passes, [1] all passes
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    std::unique_ptr<EventTimer> timer;
    ~BatchRt();
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
    bool use_sym = true;
    bool sym_ok[3] = {false, false, false};

    hipStream_t stream = nullptr;
    cav::Exec ex;                                   // the cavity's own launches (outputs, kick)
    char* arena = nullptr;
    size_t arena_bytes = 0;
    cav::Mem mem{};                                 // host copy of the descriptor
    cav::Mem* d_mem = nullptr;                      // the descriptor on the device
    cav::Ctl* h_ctl = nullptr;                      // pinned host mirror of the control block
    unsigned version = 0;                           // bumped whenever the descriptor changes
    int in_ensembles = 0;
    long long redone = 0;                           // steps redone with every refinement pass enqueued
    long long spare_until = 0;                      // until its clock reaches this step, every batch stepping this
                                                    // cavity enqueues one refinement pass beyond the usual count

    // FDM operators: dense (per axis, [0] forward S^T, [1] backward S) and even/odd halves
    // ([0] fe, [1] fo, [2] be, [3] bo), in FP64, FP32 and split BF16 (hi, lo)
    double* mpk[3][2]{};
    float* mpk32[3][2]{};
    uint16_t* mpkbf[3][2][2]{};
    double* spk[3][4]{};
    float* spk32[3][4]{};
    uint16_t* spkbf[3][4][2]{};
    double* inv_lambda = nullptr;
    float* il_f = nullptr;
    double* il_sym = nullptr;
    float* il_sym_f = nullptr;
    double *stats = nullptr, *line = nullptr;
    float *probe_out = nullptr, *plane = nullptr;
    int *probe_cells = nullptr, *first = nullptr;
    double* h_scal = nullptr;                       // pinned host mirror of the scalar slots

    std::unique_ptr<cav::BatchRt> solo;             // this cavity as a batch of one
};

struct cav_ensemble {
    std::unique_ptr<cav::BatchRt> rt;
};

namespace cav {
namespace {

using Solver = cav_solver;

template <class T> T* carve(Solver& s, size_t off) { return (T*)(s.arena + off); }

// Device arrays of one cavity (offsets into the arena). Also the footprint reported for admission.
struct Plan {
    Layout L;
    size_t F[FR_COUNT][3], C[CR_COUNT], mpk[3][2], mpk32[3][2], mpkbf[3][2][2], inv_lambda, il_f, spk[3][4],
        spk32[3][4], spkbf[3][4][2], il_sym, il_sym_f, partial, scal, stats, line, probe_out, plane, probe_cells, first,
        geo[3][4], tab[3], ctl, ctl_bak, mem;
    Plan(int nx, int ny, int nz) {
        const i64 nc = (i64)nx * ny * nz;
        const i64 sz[3] = {(i64)(nx + 1) * ny * nz, (i64)nx * (ny + 1) * nz, (i64)nx * ny * (nz + 1)};
        const int n[3] = {nx, ny, nz};
        const size_t D = sizeof(double);
        // the u, v, w arrays of a role follow one another, alike in every role: the stage kernel views the
        // three as one buffer with the same v and w offsets for all roles
        for (int r = 0; r < FR_COUNT; r++)
            for (int k = 0; k < 3; k++) F[r][k] = L.take(sz[k] * D);
        for (int r = 0; r < CR_COUNT; r++) C[r] = L.take(nc * D);
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
            tab[a] = L.take((size_t)(TAB_KINDS * (n[a] + 1) + (a == Z ? (n[a] + 4) * ZREC : 0)) * D);
        }
        inv_lambda = L.take(nc * D);
        il_f = L.take(nc * sizeof(float));
        il_sym = L.take(nc * D);
        il_sym_f = L.take(nc * sizeof(float));
        partial = L.take((size_t)P_COUNT * RED_BLOCKS * D);
        scal = L.take(S_COUNT * D);
        stats = L.take(6 * (i64)nx * ny * D);
        line = L.take(5 * (i64)nx * D);
        probe_out = L.take(16 * sizeof(float));
        plane = L.take(4 * (i64)nx * ny * sizeof(float));
        probe_cells = L.take(24 * sizeof(int));
        first = L.take(sizeof(int));
        ctl = L.take(sizeof(Ctl));
        ctl_bak = L.take(sizeof(Ctl));
        mem = L.take(sizeof(Mem));
    }
};

void validate(const CavConfig& c) {                                     // backend.cu:1016-1020
    if (c.nx % 2 || c.ny % 2 || c.nx < 6 || c.ny < 6) fail("nx and ny must be even and >= 6");
    if (c.nz < 5) fail("need nz >= 5");
    if (std::max(c.nx, std::max(c.ny, c.nz)) > 1024) fail("at most 1024 cells per axis");
    if (c.impulse_order != 2 && c.impulse_order != 4) fail("IMPULSE_ORDER must be 2 or 4");
    // the stage kernels address every face array with 32-bit byte offsets
    if ((double)(c.nx + 1) * (c.ny + 1) * (c.nz + 1) * 8.0 >= 4294967296.0) fail("grid too large: (nx+1)(ny+1)(nz+1) must stay below 2^29");
    // ... and view the u, v and w arrays of a field together (256-byte aligned, one after another)
    const double faces = (double)(c.nx + 1) * c.ny * c.nz + (double)c.nx * (c.ny + 1) * c.nz + (double)c.nx * c.ny * (c.nz + 1);
    if (faces * 8.0 + 512.0 >= 4294967296.0) fail("grid too large: the face arrays of a field must stay below 4 GiB together");
    if (c.n_probes < 0 || c.n_probes > 8) fail("n_probes must be 0..8");
}

double lid_speed(const Solver& s, double time) {                       // backend.cu:623-626
    if (s.c.lid_ramp_t > 0) return s.c.u_lid * std::min(1.0, std::max(time, 0.0) / s.c.lid_ramp_t);
    return s.c.u_lid;
}

void sync_read(Solver& s, int first, int count) {
    HIP_CHECK(hipMemcpyAsync(s.h_scal + first, s.mem.scal + first, count * sizeof(double), hipMemcpyDeviceToHost,
                             s.stream));
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
void pull_ctl(Solver& s) { d2h(s, s.h_ctl, s.mem.ctl, sizeof(Ctl)); }
void push_ctl(Solver& s) { h2d(s, s.mem.ctl, s.h_ctl, sizeof(Ctl)); }
// a buffer by role under the current parity (faces) and rotation (potentials); the mirror is current between calls
double* face(Solver& s, int role, int k) { return s.mem.F[face_buffer(role, s.h_ctl->parity)][k]; }
double* cell(Solver& s, int role) { return s.mem.C[cell_buffer(role, s.h_ctl->rot)]; }
// The FP32-field mode: u, m and the stage arrays hold float values (in the first half of each face array)
bool f32_fields(const Solver& s) { return s.mode == CAV_PREC_FIELD_FP32; }
size_t field_bytes(const Solver& s) { return f32_fields(s) ? sizeof(float) : sizeof(double); }
float* face32(Solver& s, int role, int k) { return (float*)face(s, role, k); }
// u (and m) as FP64 arrays for the output kernels: the state itself, or in the FP32-field mode its values
// widened into the RK stage-2 arrays, which hold nothing between steps
void wide_state(Solver& s, const double* uc[3], const double* mc[3] = nullptr) {
    for (int k = 0; k < 3; k++) {
        if (!f32_fields(s)) {
            uc[k] = face(s, FR_U, k);
            if (mc) mc[k] = face(s, FR_M, k);
            continue;
        }
        widen(s.ex, face32(s, FR_U, k), face(s, FR_NM1, k), s.size[k]);
        uc[k] = face(s, FR_NM1, k);
        if (mc) {
            widen(s.ex, face32(s, FR_M, k), face(s, FR_NU1, k), s.size[k]);
            mc[k] = face(s, FR_NU1, k);
        }
    }
}
double* field_ptr(Solver& s, int f) {
    if (f < 0 || f > 6) fail("field must be 0..6");
    return f < 3 ? face(s, FR_U, f) : f < 6 ? face(s, FR_M, f - 3) : cell(s, CR_PHI);
}
i64 field_size(const Solver& s, int f) { return f < 3 ? s.size[f] : f < 6 ? s.size[f - 3] : s.n_cell; }

// u and m as FP64 (from the float fields, the values widened) into host memory
void read_fields(Solver& s, int f, double* out) {
    const i64 n = field_size(s, f);
    if (f == 6 || !f32_fields(s)) { d2h(s, out, field_ptr(s, f), n * sizeof(double)); return; }
    std::vector<float> tmpv(n);
    d2h(s, tmpv.data(), field_ptr(s, f), n * sizeof(float));
    for (i64 i = 0; i < n; i++) out[i] = (double)tmpv[i];
}
void write_fields(Solver& s, int f, const double* in) {
    const i64 n = field_size(s, f);
    if (f == 6 || !f32_fields(s)) { h2d(s, field_ptr(s, f), in, n * sizeof(double)); return; }
    std::vector<float> tmpv(n);
    for (i64 i = 0; i < n; i++) tmpv[i] = (float)in[i];
    h2d(s, field_ptr(s, f), tmpv.data(), n * sizeof(float));
}


// Message of a device-side step error (the reference's wording, backend.cu:829-840, 894)
std::string error_message(const Ctl& c) {
    switch (c.err) {
        case ERR_VELOCITY_NAN: return "velocity is NaN/inf: the solution has blown up";
        case ERR_RHS_NAN: return "pressure solve: right-hand side is NaN/inf (the run has blown up)";
        case ERR_RESIDUAL_NAN: return "pressure solve: residual is NaN/inf";
        case ERR_NOT_CONVERGED: {
            char msg[200];
            std::snprintf(msg, sizeof msg, "pressure solve: residual %.2e > FDM_RTOL after %d passes", c.err_val,
                          c.err_passes);
            return msg;
        }
        default: return "";
    }
}
// After device work: refresh the mirror and raise the first error (which is then cleared on the device)
void settle(Solver& s) {
    pull_ctl(s);
    if (s.h_ctl->err) {
        const std::string msg = error_message(*s.h_ctl);
        s.h_ctl->err = ERR_NONE;
        push_ctl(s);
        fail(msg);
    }
}

// Descriptor of the cavity for its current precision mode and transform layout
void write_mem(Solver& s) {
    Mem& M = s.mem;
    M.mode = s.mode;
    // FP32 transforms in the FP32-FDM and FP32-field modes
    const bool fp32_xf = s.mode == CAV_PREC_FDM_FP32 || s.mode == CAV_PREC_FIELD_FP32;
    bool any_sym = false;
    for (int a = 0; a < 3; a++) {
        M.sym[a] = s.use_sym && s.sym_ok[a];
        any_sym |= M.sym[a] != 0;
        for (int d = 0; d < 2; d++) {
            const void** m = M.mat[a][d];
            for (int k = 0; k < 4; k++) m[k] = nullptr;
            if (!M.sym[a]) {
                if (s.mode == 0) m[0] = s.mpk[a][d];
                else if (fp32_xf) m[0] = s.mpk32[a][d];
                else { m[0] = s.mpkbf[a][d][0]; m[1] = s.mpkbf[a][d][1]; }
            } else {
                const int de = 2 * d, dd = 2 * d + 1;
                if (s.mode == 0) { m[0] = s.spk[a][de]; m[2] = s.spk[a][dd]; }
                else if (fp32_xf) { m[0] = s.spk32[a][de]; m[2] = s.spk32[a][dd]; }
                else {
                    m[0] = s.spkbf[a][de][0]; m[1] = s.spkbf[a][de][1];
                    m[2] = s.spkbf[a][dd][0]; m[3] = s.spkbf[a][dd][1];
                }
            }
        }
    }
    if (s.mode == 0) M.il = any_sym ? (const void*)s.il_sym : (const void*)s.inv_lambda;
    else M.il = any_sym ? (const void*)s.il_sym_f : (const void*)s.il_f;
    h2d(s, s.d_mem, &M, sizeof(Mem));
    s.version++;
}

constexpr long long SPARE_KEEP = 32;     // steps a spare pass stays enqueued after a projection needed it
bool wants_spare(const Solver& s) { return s.spare_until > s.h_ctl->steps; }

void drop_graphs(BatchRt& rt) {
    for (int k = 0; k < 2; k++) {
        if (rt.exec[k]) (void)hipGraphExecDestroy(rt.exec[k]);
        if (rt.graph[k]) (void)hipGraphDestroy(rt.graph[k]);
        rt.exec[k] = nullptr;
        rt.graph[k] = nullptr;
    }
}

BatchRt::~BatchRt() {
    drop_graphs(*this);
    if (dev) (void)hipFree(dev);
    if (own_stream && stream) (void)hipStreamDestroy(stream);
}

// (Re)build the launch plan of a batch for its members' current descriptors
void build(BatchRt& rt) {
    const int n = (int)rt.members.size();
    if (n < 1) fail("a batch needs at least one cavity");
    std::vector<const Mem*> hm(n);
    std::vector<const Mem*> dm(n);
    rt.passes_full = 0;
    bool spare = false;
    for (int m = 0; m < n; m++) {
        hm[m] = &rt.members[m]->mem;
        dm[m] = rt.members[m]->d_mem;
        if (rt.members[m]->mode != rt.members[0]->mode) fail("cavities stepped together must share the precision mode");
        rt.passes_full = std::max(rt.passes_full, std::max(0, rt.members[m]->c.fdm_max_passes));
        spare |= wants_spare(*rt.members[m]);
    }
    // FP64 (an exact inverse) converges in one pass, FP32 and split BF16 in two. A batch holding a cavity that
    // has stalled or used a spare recently (in this batch or in any other) enqueues one pass more.
    const int mode = rt.members[0]->mode, usual = rt.passes_set > 0 ? rt.passes_set : usual_passes(mode);
    rt.spare_on = spare;
    rt.passes_fast = std::min(rt.passes_full, usual + (spare ? 1 : 0));
    rt.b.usual = usual;
    std::vector<int> seg;
    batch_segments(hm, seg, rt.b.total);
    const size_t pbytes = (n * sizeof(Mem*) + 255) / 256 * 256, need_bytes = pbytes + seg.size() * sizeof(int);
    // The plan is read only by launches on the batch's stream, so the copies below, queued behind them, may
    // overwrite it in place. A larger plan gets a new allocation: the old one is freed once that stream is idle
    // (the graphs that read it are dropped first).
    drop_graphs(rt);
    if (need_bytes > rt.dev_bytes) {
        if (rt.dev) {
            HIP_CHECK(hipStreamSynchronize(rt.stream));
            (void)hipFree(rt.dev);
        }
        rt.dev = nullptr;
        rt.dev_bytes = 0;
        HIP_CHECK(hipMalloc((void**)&rt.dev, need_bytes));
        rt.dev_bytes = need_bytes;
    }
    HIP_CHECK(hipMemcpyAsync(rt.dev, dm.data(), n * sizeof(Mem*), hipMemcpyHostToDevice, rt.stream));
    HIP_CHECK(hipMemcpyAsync(rt.dev + pbytes, seg.data(), seg.size() * sizeof(int), hipMemcpyHostToDevice, rt.stream));
    HIP_CHECK(hipStreamSynchronize(rt.stream));
    rt.b.d_mem = (const Mem* const*)rt.dev;
    rt.b.d_seg = (const int*)(rt.dev + pbytes);
    rt.b.n = n;
    rt.b.mode = rt.members[0]->mode;
    rt.b.ex.stream = rt.stream;
    rt.versions.resize(n);
    for (int m = 0; m < n; m++) rt.versions[m] = rt.members[m]->version;
}
bool stale(const BatchRt& rt) {
    for (size_t m = 0; m < rt.members.size(); m++)
        if (rt.versions[m] != rt.members[m]->version) return true;
    return rt.b.d_mem == nullptr;
}
// After a chunk of steps: a cavity whose projections have stalled or run more passes than usual keeps a spare
// pass enqueued for SPARE_KEEP steps after the last of them, so projections that need it run in the replay
// instead of in a redone step (a cold start of split BF16 needs a third pass). A batch re-plans (and captures
// its graphs again) when the spare its members want changes.
void note_spare(Solver& s) {
    const long long mark = s.h_ctl->spare_mark;
    if (mark > 0) s.spare_until = std::max(s.spare_until, mark - 1 + SPARE_KEEP);
}
void update_spare(BatchRt& rt) {
    bool want = false;
    for (Solver* m : rt.members) {
        note_spare(*m);
        want |= wants_spare(*m);
    }
    if (want != rt.spare_on) rt.b.d_mem = nullptr;
}

// backend.cu:877-955: gauge reset, CFL/dt, three SSP-RK3 stages each followed by a warm-started projection
// (the control, base and guess of stages 2 and 3 run at the end of the previous stage's projection)
void enqueue_step(const Batch& b, int passes) {
    b_gauge(b);
    b_cfl_dt(b);                                     // (and the first stage's control)
    for (int st = 0; st < 3; st++) {
        b_stage(b, st);
        if (st == 0) b_base_guess(b, 0);
        const int mm = st == 1 ? FR_NM1 : FR_NM0;
        const int pw = st == 1 ? CR_PS1 : CR_PS0;
        const int out = st == 1 ? FR_NU1 : FR_NU0;
        // (the last projection also plans the next step's first stage and forms its base and guess)
        b_project(b, mm, pw, st == 2 ? CR_PS1 : -1, out, true, true, passes, st == 2, false, st < 2 ? st + 1 : 3,
                  st == 0 ? CR_PHI0 : CR_BASE);
    }
    b_step_end(b);
}

// Enqueue n steps of the batch on its stream: graph replays, or plain launches while kernels are timed.
// full: every allowed refinement pass enqueued (redone steps).
// Rebuild a stale plan and capture the step graph a replay will need (nothing to do when both are current),
// so that a caller timing the step can do this set-up outside its events
void prepare(BatchRt& rt, bool full = false) {
    if (stale(rt)) build(rt);
    const bool timed = rt.b.ex.timer != nullptr;
    const int g = full ? 1 : 0, passes = full ? rt.passes_full : rt.passes_fast;
    if (rt.use_graph && !timed && !rt.exec[g]) {
        HIP_CHECK(hipStreamBeginCapture(rt.stream, hipStreamCaptureModeThreadLocal));
        try {
            enqueue_step(rt.b, passes);
        } catch (...) {
            hipGraph_t gr = nullptr;
            (void)hipStreamEndCapture(rt.stream, &gr);
            if (gr) (void)hipGraphDestroy(gr);
            throw;
        }
        HIP_CHECK(hipStreamEndCapture(rt.stream, &rt.graph[g]));
        HIP_CHECK(hipGraphInstantiate(&rt.exec[g], rt.graph[g], nullptr, nullptr, 0));
    }
}
void enqueue_steps(BatchRt& rt, int n, hipEvent_t* marks = nullptr, bool full = false) {
    prepare(rt, full);
    const bool timed = rt.b.ex.timer != nullptr;
    const int g = full ? 1 : 0, passes = full ? rt.passes_full : rt.passes_fast;
    for (int k = 0; k < n; k++) {
        if (marks) HIP_CHECK(hipEventRecord(marks[k], rt.stream));
        if (rt.use_graph && !timed) HIP_CHECK(hipGraphLaunch(rt.exec[g], rt.stream));
        else enqueue_step(rt.b, passes);
    }
    if (marks) HIP_CHECK(hipEventRecord(marks[n], rt.stream));
}

// Clear an error left on the device by an earlier failed call
void clear_error(Solver& s) {
    if (s.h_ctl->err) {
        s.h_ctl->err = ERR_NONE;
        push_ctl(s);
    }
}

extern bool g_autotune;
void autotune(BatchRt& rt);

// The grid and everything computed from it: the axes (rounded through float for the FP32 fields), the viscous
// rate, the grid arrays and coefficient tables, the FDM operators of both layouts and the probe cells. Written
// into the cavity's arena at create and again when the precision mode enters or leaves the FP32 fields.
void setup_grid(Solver& s, bool f32) {                                 // backend.cu:1022-1070
    s.ax[X] = make_axis(s.c, s.nx, s.c.length, f32);
    s.ax[Y] = make_axis(s.c, s.ny, s.c.length, f32);
    s.ax[Z] = make_axis(s.c, s.nz, s.nz * (s.c.lz / s.nz), f32);
    s.volume = s.ax[X].length * s.ax[Y].length * s.ax[Z].length;
    // viscous rate (backend.cu:1030-1034): nu * max over cells of sum over axes of 4/(width * smaller
    // spacing), in the field type. Rounded addition and multiplication are monotone, so the max over cells of
    // the sum equals the sum of the per-axis maxima, bit for bit.
    if (f32) {
        float gmax[3];
        for (int a = 0; a < 3; a++) {
            gmax[a] = 0.0f;
            const Axis& g = s.ax[a];
            for (int i = 0; i < g.n; i++)
                gmax[a] = std::max(gmax[a], 4.0f / ((float)g.width[i] * std::min((float)g.spacing[i], (float)g.spacing[i + 1])));
        }
        s.viscous_rate = (double)std::max(0.0f, (float)s.c.nu * (gmax[X] + gmax[Y] + gmax[Z]));
    } else {
        double gmax[3];
        for (int a = 0; a < 3; a++) {
            gmax[a] = 0.0;
            const Axis& g = s.ax[a];
            for (int i = 0; i < g.n; i++)
                gmax[a] = std::max(gmax[a], 4.0 / (g.width[i] * std::min(g.spacing[i], g.spacing[i + 1])));
        }
        s.viscous_rate = std::max(0.0, s.c.nu * (gmax[X] + gmax[Y] + gmax[Z]));
    }
    const Plan P(s.nx, s.ny, s.nz);
    Mem& M = s.mem;

    // grid tables
    Geo& geo = M.g;
    geo.nx = s.nx; geo.ny = s.ny; geo.nz = s.nz; geo.lid_shaped = s.c.lid_shaped;
    for (int a = 0; a < 3; a++) {
        const Axis& g = s.ax[a];
        double* p[4] = {carve<double>(s, P.geo[a][0]), carve<double>(s, P.geo[a][1]), carve<double>(s, P.geo[a][2]),
                        carve<double>(s, P.geo[a][3])};
        h2d(s, p[0], g.face.data(), g.face.size() * sizeof(double));
        h2d(s, p[1], g.centre.data(), g.centre.size() * sizeof(double));
        h2d(s, p[2], g.width.data(), g.width.size() * sizeof(double));
        h2d(s, p[3], g.spacing.data(), g.spacing.size() * sizeof(double));
        geo.face[a] = p[0]; geo.centre[a] = p[1]; geo.width[a] = p[2]; geo.spacing[a] = p[3];
        std::vector<double> tabs = axis_tables(g, std::max(s.c.impulse_order4_ramp, 1.0), s.c.lid_shaped != 0, f32);
        if (a == Z) {                                   // the stage kernel's per-level records follow the z tables
            const std::vector<double> rec = level_records(tabs, g.n);
            tabs.insert(tabs.end(), rec.begin(), rec.end());
        }
        double* tp = carve<double>(s, P.tab[a]);
        h2d(s, tp, tabs.data(), tabs.size() * sizeof(double));
        geo.tab[a] = tp;
        geo.length[a] = g.length;
    }
    // FDM eigensystems (backend.cu:750-787), dense
    const Eigen* e[3];
    for (int a = 0; a < 3; a++) {
        e[a] = &axis_eigen(s.ax[a]);
        const int n = s.ax[a].n;
        std::vector<double> ST((size_t)n * n);
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) ST[(size_t)j * n + i] = e[a]->S[(size_t)i * n + j];
        const std::vector<double>* mats[2] = {&ST, &e[a]->S};
        for (int d = 0; d < 2; d++) {
            s.mpk[a][d] = carve<double>(s, P.mpk[a][d]);
            s.mpk32[a][d] = carve<float>(s, P.mpk32[a][d]);
            s.mpkbf[a][d][0] = carve<uint16_t>(s, P.mpkbf[a][d][0]);
            s.mpkbf[a][d][1] = carve<uint16_t>(s, P.mpkbf[a][d][1]);
            const std::vector<double> f64 = pack_fragments<double>(*mats[d], n);
            const std::vector<float> f32 = pack_fragments<float>(*mats[d], n);
            std::vector<uint16_t> hi, lo;
            pack_bf16_fragments(*mats[d], n, hi, lo);
            h2d(s, s.mpk[a][d], f64.data(), f64.size() * sizeof(double));
            h2d(s, s.mpk32[a][d], f32.data(), f32.size() * sizeof(float));
            h2d(s, s.mpkbf[a][d][0], hi.data(), hi.size() * sizeof(uint16_t));
            h2d(s, s.mpkbf[a][d][1], lo.data(), lo.size() * sizeof(uint16_t));
        }
    }
    {
        std::vector<double> inv(s.n_cell);
        std::vector<float> invf(s.n_cell);
        for (int z = 0; z < s.nz; z++)
            for (int y = 0; y < s.ny; y++)
                for (int x = 0; x < s.nx; x++) {
                    const bool zero = x == e[X]->zero_mode && y == e[Y]->zero_mode && z == e[Z]->zero_mode;
                    const i64 i = x + (i64)s.nx * (y + (i64)s.ny * z);
                    inv[i] = zero ? 0.0 : 1.0 / (e[Z]->lambda[z] + e[Y]->lambda[y] + e[X]->lambda[x]);
                    invf[i] = (float)inv[i];
                }
        h2d(s, s.inv_lambda, inv.data(), inv.size() * sizeof(double));
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
        for (int z = 0; z < s.nz; z++)
            for (int y = 0; y < s.ny; y++)
                for (int x = 0; x < s.nx; x++) {
                    const bool zm = x == zero[X] && y == zero[Y] && z == zero[Z];
                    const double v = zm ? 0.0 : 1.0 / ((*lam[Z])[z] + (*lam[Y])[y] + (*lam[X])[x]);
                    const i64 i = x + (i64)s.nx * (y + (i64)s.ny * z);
                    inv[i] = v;
                    invf[i] = (float)v;
                }
        h2d(s, s.il_sym, inv.data(), inv.size() * sizeof(double));
        h2d(s, s.il_sym_f, invf.data(), invf.size() * sizeof(float));
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

}

void create(Solver& s, const CavConfig& cfg) {                         // backend.cu:1009-1073
    validate(cfg);
    s.c = cfg;
    s.nx = cfg.nx; s.ny = cfg.ny; s.nz = cfg.nz;
    s.n_cell = (i64)s.nx * s.ny * s.nz;
    s.size[0] = (i64)(s.nx + 1) * s.ny * s.nz;
    s.size[1] = (i64)s.nx * (s.ny + 1) * s.nz;
    s.size[2] = (i64)s.nx * s.ny * (s.nz + 1);
    HIP_CHECK(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
    const Plan P(s.nx, s.ny, s.nz);
    s.ex.stream = s.stream;
    s.arena_bytes = P.L.total;
    HIP_CHECK(hipMalloc((void**)&s.arena, s.arena_bytes));
    HIP_CHECK(hipMemsetAsync(s.arena, 0, s.arena_bytes, s.stream));
    HIP_CHECK(hipHostMalloc((void**)&s.h_scal, S_COUNT * sizeof(double), hipHostMallocDefault));
    HIP_CHECK(hipHostMalloc((void**)&s.h_ctl, sizeof(Ctl), hipHostMallocDefault));
    std::memset(s.h_ctl, 0, sizeof(Ctl));
    Mem& M = s.mem;
    for (int r = 0; r < FR_COUNT; r++)
        for (int k = 0; k < 3; k++) M.F[r][k] = carve<double>(s, P.F[r][k]);
    for (int r = 0; r < CR_COUNT; r++) M.C[r] = carve<double>(s, P.C[r]);
    M.partial = carve<double>(s, P.partial);
    M.scal = carve<double>(s, P.scal);
    M.ctl = carve<Ctl>(s, P.ctl);
    M.ctl_bak = carve<Ctl>(s, P.ctl_bak);
    s.d_mem = carve<Mem>(s, P.mem);
    s.inv_lambda = carve<double>(s, P.inv_lambda);
    s.il_f = carve<float>(s, P.il_f);
    s.il_sym = carve<double>(s, P.il_sym);
    s.il_sym_f = carve<float>(s, P.il_sym_f);
    s.stats = carve<double>(s, P.stats);
    s.line = carve<double>(s, P.line);
    s.probe_out = carve<float>(s, P.probe_out);
    s.plane = carve<float>(s, P.plane);
    s.probe_cells = carve<int>(s, P.probe_cells);
    s.first = carve<int>(s, P.first);
    setup_grid(s, false);

    // control block: configuration, clock at 0 (backend.cu:1036-1054)
    Ctl& c = *s.h_ctl;
    c.dt_fixed = s.c.dt_fixed; c.dt_max = s.c.dt_max; c.cfl_target = s.c.cfl_target;
    c.diffusion_limit = s.c.diffusion_limit; c.viscous_rate = s.viscous_rate; c.rtol = s.c.fdm_rtol;
    c.u_lid = s.c.u_lid; c.lid_ramp_t = s.c.lid_ramp_t;
    c.adaptive = s.c.adaptive_dt; c.dt_lag = s.c.dt_lag; c.max_passes = s.c.fdm_max_passes;
    c.extrapolate = s.c.phi_extrapolate;
    push_ctl(s);
    M.o4 = s.c.impulse_order == 4;
    M.max_passes = s.c.fdm_max_passes;
    M.nu = s.c.nu;
    M.r = s.c.reinit_r;
    M.omr = 1.0 - s.c.reinit_r;
    M.n_cell = s.n_cell;
    write_mem(s);

    s.solo.reset(new BatchRt());
    s.solo->members = {&s};
    s.solo->stream = s.stream;
    build(*s.solo);
    HIP_CHECK(hipStreamSynchronize(s.stream));
    if (g_autotune) autotune(*s.solo);
    HIP_CHECK(hipStreamSynchronize(s.stream));
}

void destroy(Solver* s) {
    if (!s) return;
    if (s->in_ensembles) fail("cavity is still part of an ensemble: destroy the ensemble first");
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    s->solo.reset();
    if (s->arena) (void)hipFree(s->arena);
    if (s->h_scal) (void)hipHostFree(s->h_scal);
    if (s->h_ctl) (void)hipHostFree(s->h_ctl);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

// Redo a stalled step from the control block it began with (its state is untouched), all passes enqueued
void redo(Solver& s) {
    HIP_CHECK(hipMemcpyAsync(s.mem.ctl, s.mem.ctl_bak, sizeof(Ctl), hipMemcpyDeviceToDevice, s.stream));
    enqueue_steps(*s.solo, 1, nullptr, true);
    pull_ctl(s);
    s.redone++;
    s.spare_until = std::max(s.spare_until, s.h_ctl->steps + SPARE_KEEP);
    update_spare(*s.solo);
}
// Step the cavity alone until its clock shows `target` steps (or it fails), redoing stalled steps
void advance(Solver& s, long long target) {
    constexpr int CHUNK = 16;         // steps between two looks at the control block
    for (;;) {
        if (s.h_ctl->err) return;
        if (s.h_ctl->stalled) { redo(s); continue; }
        const long long left = target - s.h_ctl->steps;
        if (left <= 0) return;
        enqueue_steps(*s.solo, (int)std::min<long long>(left, CHUNK));
        pull_ctl(s);
        update_spare(*s.solo);
    }
}

void steps(Solver& s, int n) {
    clear_error(s);
    advance(s, s.h_ctl->steps + n);
    settle(s);
}

// backend.cu:958-978. Scratch between steps: the noise is smoothed in the RK stage-2 buffers and u + noise
// formed in the gauge-reset impulse m0; a step rewrites all three before it reads them.
void kick(Solver& s, const double* const nz_[3], double amp, int passes) {
    for (int k = 0; k < 3; k++) if (!nz_[k]) fail("kick: noise arrays must not be NULL");
    clear_error(s);
    if (stale(*s.solo)) build(*s.solo);
    Mem& M = s.mem;
    for (int k = 0; k < 3; k++) {
        const i64 n = s.size[k];
        const int sx = s.nx + (k == 0), sy = s.ny + (k == 1), sz = s.nz + (k == 2);
        double *a = face(s, FR_NM1, k), *b = face(s, FR_NU1, k);
        h2d(s, a, nz_[k], n * sizeof(double));
        for (int p = 0; p < std::max(1, passes); p++)
            for (int axis = 2; axis >= 0; axis--) {
                smooth(s.ex, a, b, axis, sx, sy, sz);
                std::swap(a, b);
            }
        reduce_array(s.ex, a, n, 3, M.partial + P_AUX * RED_BLOCKS, M.scal + S_KICK);
        sync_read(s, S_KICK, 1);
        const double rms = std::sqrt(s.h_scal[S_KICK] / n);
        const double scale = amp * s.c.u_lid / std::max(rms, 1e-30);
        if (f32_fields(s)) add_noise(s.ex, face32(s, FR_U, k), a, scale, face32(s, FR_M0, k), n);
        else add_noise(s.ex, face(s, FR_U, k), a, scale, face(s, FR_M0, k), n);
    }
    b_project(s.solo->b, FR_M0, CR_PS1, -1, FR_NU0, false, false, s.solo->passes_full);
    settle(s);
    for (int k = 0; k < 3; k++) {
        if (f32_fields(s)) add_difference(s.ex, face32(s, FR_M, k), face32(s, FR_NU0, k), face32(s, FR_U, k), s.size[k]);
        else add_difference(s.ex, face(s, FR_M, k), face(s, FR_NU0, k), face(s, FR_U, k), s.size[k]);
        // face arrays hold an even number of values: the float ones are whole doubles
        copy_array(s.ex, face(s, FR_NU0, k), face(s, FR_U, k), s.size[k] * (i64)field_bytes(s) / 8);
    }
    HIP_CHECK(hipStreamSynchronize(s.stream));
    s.h_ctl->m0_ready = 0;                               // the state changed (and m0 held the kick's scratch)
    s.h_ctl->cfl_ready = 0;
    s.h_ctl->g0_ready = 0;
    push_ctl(s);
}

// ------------------------------------------------------------------------------------------------ autotuner
// Launch choices per cavity (Mem::tune, TuneKnob): the stage kernel's column tile, z-chunk target and shortest
// chunk, the gradient's z-chunk target and shortest chunk, the transform layout of each mirror-symmetric axis.
// Every choice gives the same results bit for bit; only the time differs. The tuner times the candidates on
// the cavity itself (the stage kernel into the stage-1 impulse buffers, the gradient into the stage-1 velocity
// buffers, the warm-start base and the stage-1 potential, the transforms into the warm-start base: all rewritten
// by a step before it reads them), picks the
// fastest of each knob group in turn and caches the result per configuration (sizes, mode, layout), so a
// shape is timed once per process. A batch of cavities is tuned as one: its members' z-chunk targets are
// shares of a total in proportion to their cells.
bool g_autotune = true;                  // tune every cavity as it is created (cav_set_autotune)
std::mutex g_tune_lock;
std::map<std::string, std::vector<int>> g_tune_cache;    // key -> TU_COUNT knobs per member

std::string tune_key(const BatchRt& rt) {
    std::string k = "m" + std::to_string(rt.members[0]->mode);
    for (const Solver* m : rt.members) {
        k += ":" + std::to_string(m->nx) + "x" + std::to_string(m->ny) + "x" + std::to_string(m->nz) + "s";
        for (int a = 0; a < 3; a++) k += m->mem.sym[a] ? "1" : "0";
    }
    return k;
}

void set_knobs(BatchRt& rt, const std::vector<int>& knobs) {
    for (size_t m = 0; m < rt.members.size(); m++) {
        Solver& s = *rt.members[m];
        bool diff = false;
        for (int k = 0; k < TU_COUNT; k++) diff |= s.mem.tune[k] != knobs[m * TU_COUNT + k];
        if (!diff) continue;
        for (int k = 0; k < TU_COUNT; k++) s.mem.tune[k] = knobs[m * TU_COUNT + k];
        write_mem(s);
    }
    if (stale(rt)) build(rt);
}
// Milliseconds per call of launch() on the batch's stream: one untimed call, then enough timed calls for
// about a millisecond in all (3 to 30)
template <class F> double time_calls(BatchRt& rt, F launch) {
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    auto span = [&](int reps) {
        HIP_CHECK(hipEventRecord(e0, rt.stream));
        for (int r = 0; r < reps; r++) launch();
        HIP_CHECK(hipEventRecord(e1, rt.stream));
        HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0;
        HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        return (double)ms / reps;
    };
    double ms = 0;
    try {
        launch();
        const double one = span(1);
        const int reps = std::max(3, std::min(30, (int)std::ceil(1.0 / std::max(one, 1e-3))));
        ms = span(reps);
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ms;
}

// Tune a batch (a cavity alone is a batch of one). Leaves the fastest knobs applied and cached.
void autotune(BatchRt& rt) {
    const size_t n = rt.members.size();
    const std::string key = tune_key(rt);
    {
        std::lock_guard<std::mutex> lock(g_tune_lock);
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) {
            set_knobs(rt, it->second);
            return;
        }
    }
    // the trials overwrite the warm-start base: a first guess a previous step formed is formed again
    for (size_t m = 0; m < n; m++) {
        Solver& s = *rt.members[m];
        pull_ctl(s);
        if (s.h_ctl->g0_ready) {
            s.h_ctl->g0_ready = 0;
            push_ctl(s);
        }
    }
    const bool alone = n == 1;
    double cells = 0;
    for (const Solver* m : rt.members) cells += (double)m->n_cell;
    std::vector<int> best(n * TU_COUNT, 0), trial;
    // a z-chunk target for member m: absolute alone; a share of the batch's total in proportion to its cells
    auto target_of = [&](size_t m, int total) {
        if (total == 0) return 0;
        if (alone) return total;
        return std::max(1, (int)std::lround(total * (double)rt.members[m]->n_cell / cells));
    };
    auto time_with = [&](const std::vector<int>& kn, int what) {
        set_knobs(rt, kn);
        if (what == 0) return time_calls(rt, [&] { b_stage(rt.b, 0); });
        // (the gradient as in a step's first two stages: velocity and the next stage's base and guess, into the
        // stage-1 velocity, the base and the stage-1 potential; the history entries are only read)
        if (what == 1) return time_calls(rt, [&] { b_gradient(rt.b, FR_M, CR_PHI, -1, FR_NU0, false, false, 1); });
        return time_calls(rt, [&] { b_transforms(rt.b, CR_BASE, 0); });
    };
    // 1. the stage kernel: tile x chunk target x shortest chunk
    {
        // whole rounds of two or three resident blocks per CU (the FP64 and the float kernel)
        const int totals_alone[] = {384, 512, 768, 1024, 1536}, totals_batch[] = {0, 512, 768, 1536, 3072},
                  minz[] = {8, 16, 32};
        const int* totals = alone ? totals_alone : totals_batch;
        const int nt = 5;
        double fastest = 1e300;
        for (int tile = 0; tile < 2; tile++)
            for (int ti = 0; ti < nt; ti++)
                for (int mz : minz) {
                    trial = best;
                    for (size_t m = 0; m < n; m++) {
                        trial[m * TU_COUNT + TU_STAGE_TILE] = tile;
                        trial[m * TU_COUNT + TU_STAGE_TARGET] = target_of(m, totals[ti]);
                        trial[m * TU_COUNT + TU_STAGE_MINZ] = mz;
                    }
                    const double ms = time_with(trial, 0);
                    if (ms < fastest) { fastest = ms; best = trial; }
                }
    }
    // 2. the gradient: column tile x chunk target x shortest chunk (a tile wider than a member's x extent
    //    leaves threads idle: the narrowest member bounds the tile a batch tries)
    {
        const int totals_alone[] = {896, 1792, 3584}, totals_batch[] = {0, 1792, 3584, 7168}, minz[] = {8, 16};
        const int* totals = alone ? totals_alone : totals_batch;
        const int nt = alone ? 3 : 4;
        int widest = 0;
        {
            int nxmin = INT_MAX;
            for (const Solver* m : rt.members) nxmin = std::min(nxmin, m->nx);
            widest = nxmin >= 256 ? 3 : nxmin >= 128 ? 2 : nxmin >= 64 ? 1 : 0;
        }
        double fastest = 1e300;
        std::vector<int> base = best;
        for (int tile = 0; tile <= widest; tile++)
            for (int ti = 0; ti < nt; ti++)
                for (int mz : minz) {
                    trial = base;
                    for (size_t m = 0; m < n; m++) {
                        trial[m * TU_COUNT + TU_GRAD_TILE] = tile;
                        trial[m * TU_COUNT + TU_GRAD_TARGET] = target_of(m, totals[ti]);
                        trial[m * TU_COUNT + TU_GRAD_MINZ] = mz;
                    }
                    const double ms = time_with(trial, 1);
                    if (ms < fastest) { fastest = ms; best = trial; }
                }
    }
    // 3. transform layouts: per axis, even/odd by chunks or by whole columns, where some member allows both
    {
        int axes = 0;
        for (int a = 0; a < 3; a++)
            for (const Solver* m : rt.members)
                if (whole_columns_possible(m->mem, a)) axes |= 1 << a;
        if (axes) {
            // the transforms run only with run_xf set: set it for the timing, then put the control blocks back
            std::vector<Ctl> saved(n);
            for (size_t m = 0; m < n; m++) {
                Solver& s = *rt.members[m];
                pull_ctl(s);
                saved[m] = *s.h_ctl;
                s.h_ctl->run_xf = 1;
                s.h_ctl->rms = 1.0;
                s.h_ctl->r_rhs = 0;
                push_ctl(s);
            }
            try {
                double fastest = 1e300;
                std::vector<int> base = best;
                for (int combo = 0; combo < 8; combo++) {
                    if (combo & ~axes) continue;
                    int xf = 0;
                    for (int a = 0; a < 3; a++) xf |= ((combo >> a) & 1 ? 2 : 1) << (2 * a);
                    trial = base;
                    for (size_t m = 0; m < n; m++) trial[m * TU_COUNT + TU_XF] = xf;
                    const double ms = time_with(trial, 2);
                    if (ms < fastest) { fastest = ms; best = trial; }
                }
            } catch (...) {
                for (size_t m = 0; m < n; m++) { *rt.members[m]->h_ctl = saved[m]; push_ctl(*rt.members[m]); }
                throw;
            }
            for (size_t m = 0; m < n; m++) { *rt.members[m]->h_ctl = saved[m]; push_ctl(*rt.members[m]); }
        }
    }
    set_knobs(rt, best);
    std::lock_guard<std::mutex> lock(g_tune_lock);
    g_tune_cache[key] = best;
}

void initialize(Solver& s, const double* vn, const double* wn) {       // backend.cu:1075-1088
    clear_error(s);
    const double* src[3] = {nullptr, vn, wn};
    for (int k = 0; k < 3; k++) {
        if (src[k]) write_fields(s, 3 + k, src[k]);
        else fill(s.ex, face(s, FR_M, k), s.size[k], 0.0);     // covers the float array too
    }
    if (stale(*s.solo)) build(*s.solo);
    b_project(s.solo->b, FR_M, CR_PS0, CR_PHI, FR_U, false, false, s.solo->passes_full);
    settle(s);
    Ctl& c = *s.h_ctl;
    c.t = 0; c.steps = 0; c.last_dt = 0; c.hn = 0; c.have_prev = 0;
    c.m0_ready = 0;
    c.cfl_ready = 0;
    c.g0_ready = 0;
    push_ctl(s);
}

// Entering or leaving the FP32-field mode changes the grid (rounded through float there) and the field type:
// the set-up is redone on the new grid, u and m carried over (rounded to float, or widened), phi and the clock
// kept; the warm-start history and the lagged dt rates are cleared, as after solver_set_field.
void set_precision(Solver& s, int mode) {
    if (f32_fields(s) == (mode == CAV_PREC_FIELD_FP32)) {
        s.mode = mode;
        write_mem(s);
        return;
    }
    HIP_CHECK(hipStreamSynchronize(s.stream));
    pull_ctl(s);
    std::vector<double> keep[6];
    for (int f = 0; f < 6; f++) {
        keep[f].resize(field_size(s, f));
        read_fields(s, f, keep[f].data());
    }
    s.mode = mode;
    setup_grid(s, f32_fields(s));
    for (int f = 0; f < 6; f++) write_fields(s, f, keep[f].data());
    Ctl& c = *s.h_ctl;
    c.viscous_rate = s.viscous_rate;
    c.hn = 0;
    c.have_prev = 0;
    c.m0_ready = 0;
    c.cfl_ready = 0;
    c.g0_ready = 0;
    push_ctl(s);
    write_mem(s);
}

void fill_report(const Solver& s, CavReport* r) {                      // backend.cu:995-999
    if (!r) return;
    const Ctl& c = *s.h_ctl;
    r->steps = c.steps; r->time = c.t; r->dt = c.last_dt; r->cfl = c.last_cfl; r->phi_rel = c.phi_rel;
    r->dt_reason = c.dt_reason; r->phi_passes = c.phi_passes;
}

void diagnostics(Solver& s, CavDiagnostics* d) {                       // backend.cu:1096-1129
    Mem& M = s.mem;
    const double *uc[3], *mc[3];
    wide_state(s, uc, mc);
    double* aux = M.partial + P_AUX * RED_BLOCKS;
    cell_diagnostics(s.ex, M.g, uc, lid_speed(s, s.h_ctl->t), M.partial, M.scal + S_KE);
    for (int k = 0; k < 3; k++) reduce_array(s.ex, uc[k], s.size[k], 2, aux, M.scal + S_UMAX + k);
    if (stale(*s.solo)) build(*s.solo);
    b_cfl_rate(s.solo->b);
    for (int k = 0; k < 3; k++) max_abs_diff(s.ex, mc[k], uc[k], s.size[k], aux, M.scal + S_GAUGE + k);
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
        first_index_of_abs_diff(s.ex, mc[best_comp], uc[best_comp], s.size[best_comp], M.scal + S_GAUGE + best_comp,
                                s.first);
        d2h(s, &best_i, s.first, sizeof(int));
    }
    const int sx = s.nx + (best_comp == 0), sy = s.ny + (best_comp == 1);
    const int ix = best_i % sx, iy = (best_i / sx) % sy;
    d->gauge_ratio = best / std::max(umax, 1e-30);
    d->gauge_x = (best_comp == 0 ? s.ax[X].face[ix] : s.ax[X].centre[std::min(ix, s.nx - 1)]) / s.c.length;
    d->gauge_y = (best_comp == 1 ? s.ax[Y].face[iy] : s.ax[Y].centre[std::min(iy, s.ny - 1)]) / s.c.length;
}

void get_field(Solver& s, int f, void* out, bool wide) {
    const i64 n = field_size(s, f);
    if (wide || f == 6) { read_fields(s, f, (double*)out); return; }
    if (f32_fields(s)) { d2h(s, out, field_ptr(s, f), n * sizeof(float)); return; }
    std::vector<double> tmpv(n);
    d2h(s, tmpv.data(), field_ptr(s, f), n * sizeof(double));
    float* o = (float*)out;
    for (i64 i = 0; i < n; i++) o[i] = (float)tmpv[i];
}

void set_field(Solver& s, int f, const void* in, bool wide) {          // backend.cu:1147-1155
    const i64 n = field_size(s, f);
    if (wide || f == 6) write_fields(s, f, (const double*)in);
    else if (f32_fields(s)) h2d(s, field_ptr(s, f), in, n * sizeof(float));
    else {
        std::vector<double> tmpv(n);
        const float* src = (const float*)in;
        for (i64 i = 0; i < n; i++) tmpv[i] = (double)src[i];
        h2d(s, field_ptr(s, f), tmpv.data(), n * sizeof(double));
    }
    s.h_ctl->hn = 0;
    s.h_ctl->have_prev = 0;
    s.h_ctl->m0_ready = 0;
    s.h_ctl->cfl_ready = 0;
    s.h_ctl->g0_ready = 0;
    push_ctl(s);
}

// ------------------------------------------------------------------------------------------------ snapshot
// Whole state of a cavity between calls, enough to continue bit for bit: the control block (clock, lagged dt
// rates, warm-start history entries, counters, buffer parity), u, m and phi, the six history buffers of dphi
// and the mid-span statistics sums. Everything else a step reads (m0, phi0, the stage buffers, the warm-start
// base) is rewritten by the step before it is read. Layout: header, control block, arrays, packed (every
// section a whole number of doubles, no gaps), so equal states give equal bytes. The destination may be device
// or host memory (hipMemcpyDefault).
struct SnapHeader {
    char magic[8];
    int version, mode;
    long long bytes;
    CavConfig config;                  // the configuration of the solver the snapshot was taken from
};
constexpr char SNAP_MAGIC[8] = {'C', 'A', 'V', 'S', 'N', 'A', 'P', '1'};
struct SnapLayout {
    size_t ctl = 0, u[3]{}, m[3]{}, phi = 0, hist[HIST_BUFFERS]{}, stats = 0, total = 0;
};
static_assert(sizeof(SnapHeader) % 8 == 0 && sizeof(Ctl) % 8 == 0, "snapshot sections are whole doubles");
SnapLayout snap_layout(const Solver& s) {
    SnapLayout o;
    size_t at = sizeof(SnapHeader);
    auto take = [&](size_t bytes) { const size_t off = at; at += bytes; return off; };
    o.ctl = take(sizeof(Ctl));
    // u and m in the field type (face arrays hold an even number of values: whole doubles either way)
    for (int k = 0; k < 3; k++) o.u[k] = take(s.size[k] * field_bytes(s));
    for (int k = 0; k < 3; k++) o.m[k] = take(s.size[k] * field_bytes(s));
    o.phi = take(s.n_cell * sizeof(double));
    for (int k = 0; k < HIST_BUFFERS; k++) o.hist[k] = take(s.n_cell * sizeof(double));
    o.stats = take(6 * (size_t)s.nx * s.ny * sizeof(double));
    o.total = at;
    return o;
}
// The configurations two solvers were created from agree in every field (padding ignored)
bool same_config(const CavConfig& a, const CavConfig& b) {
    const int ia[] = {a.nx, a.ny, a.nz, a.lid_shaped, a.impulse_order, a.adaptive_dt, a.dt_lag, a.phi_extrapolate,
                      a.fdm_max_passes, a.grid_stretch, a.n_probes};
    const int ib[] = {b.nx, b.ny, b.nz, b.lid_shaped, b.impulse_order, b.adaptive_dt, b.dt_lag, b.phi_extrapolate,
                      b.fdm_max_passes, b.grid_stretch, b.n_probes};
    const double da[] = {a.length, a.u_lid, a.nu, a.lz, a.lid_ramp_t, a.reinit_r, a.impulse_order4_ramp,
                         a.grid_dx_wall, a.grid_max_aspect, a.dt_fixed, a.dt_max, a.cfl_target, a.diffusion_limit,
                         a.fdm_rtol};
    const double db[] = {b.length, b.u_lid, b.nu, b.lz, b.lid_ramp_t, b.reinit_r, b.impulse_order4_ramp,
                         b.grid_dx_wall, b.grid_max_aspect, b.dt_fixed, b.dt_max, b.cfl_target, b.diffusion_limit,
                         b.fdm_rtol};
    return std::memcmp(ia, ib, sizeof ia) == 0 && std::memcmp(da, db, sizeof da) == 0 &&
           std::memcmp(a.probe_x, b.probe_x, sizeof a.probe_x) == 0 &&
           std::memcmp(a.probe_y, b.probe_y, sizeof a.probe_y) == 0;
}
void copy_any(Solver& s, void* dst, const void* src, size_t bytes) {
    HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, s.stream));
}
void snapshot(Solver& s, void* dst, long long bytes) {
    const SnapLayout o = snap_layout(s);
    if (!dst) fail("snapshot: NULL destination");
    if (bytes < (long long)o.total) fail("snapshot: the destination holds " + std::to_string(bytes) +
                                          " bytes, the snapshot needs " + std::to_string(o.total));
    pull_ctl(s);
    SnapHeader h;
    std::memset(&h, 0, sizeof h);
    std::memcpy(h.magic, SNAP_MAGIC, sizeof h.magic);
    h.version = 1;
    h.mode = s.mode;
    h.bytes = (long long)o.total;
    h.config = s.c;
    char* d = (char*)dst;
    copy_any(s, d, &h, sizeof h);
    copy_any(s, d + o.ctl, s.h_ctl, sizeof(Ctl));
    for (int k = 0; k < 3; k++) {
        copy_any(s, d + o.u[k], face(s, FR_U, k), s.size[k] * field_bytes(s));
        copy_any(s, d + o.m[k], face(s, FR_M, k), s.size[k] * field_bytes(s));
    }
    copy_any(s, d + o.phi, cell(s, CR_PHI), s.n_cell * sizeof(double));
    for (int k = 0; k < HIST_BUFFERS; k++) copy_any(s, d + o.hist[k], s.mem.C[CR_H0 + k], s.n_cell * sizeof(double));
    copy_any(s, d + o.stats, s.stats, 6 * (size_t)s.nx * s.ny * sizeof(double));
    HIP_CHECK(hipStreamSynchronize(s.stream));          // the header copy reads a stack object
}
void restore(Solver& s, const void* src, long long bytes) {
    const SnapLayout o = snap_layout(s);
    if (!src) fail("restore: NULL source");
    if (bytes < (long long)sizeof(SnapHeader)) fail("restore: not a cavity snapshot (too short)");
    SnapHeader h;
    copy_any(s, &h, src, sizeof h);
    HIP_CHECK(hipStreamSynchronize(s.stream));
    if (std::memcmp(h.magic, SNAP_MAGIC, sizeof h.magic) != 0 || h.version != 1)
        fail("restore: not a cavity snapshot");
    if (!same_config(h.config, s.c)) fail("restore: the snapshot was taken from a solver of another configuration");
    if (h.mode != s.mode) fail("restore: the snapshot was taken in precision mode " + std::to_string(h.mode) +
                               ", the solver runs mode " + std::to_string(s.mode));
    if (h.bytes != (long long)o.total || bytes < h.bytes) fail("restore: snapshot size mismatch");
    const char* d = (const char*)src;
    copy_any(s, s.h_ctl, d + o.ctl, sizeof(Ctl));
    HIP_CHECK(hipStreamSynchronize(s.stream));           // the parity of the restored state picks the buffers
    s.h_ctl->err = ERR_NONE;
    s.h_ctl->stalled = 0;
    s.h_ctl->m0_ready = 0;                               // m0 and phi0 are not part of a snapshot
    s.h_ctl->cfl_ready = 0;
    s.h_ctl->g0_ready = 0;                               // nor the next first guess
    for (int k = 0; k < 3; k++) {
        copy_any(s, face(s, FR_U, k), d + o.u[k], s.size[k] * field_bytes(s));
        copy_any(s, face(s, FR_M, k), d + o.m[k], s.size[k] * field_bytes(s));
    }
    copy_any(s, cell(s, CR_PHI), d + o.phi, s.n_cell * sizeof(double));
    for (int k = 0; k < HIST_BUFFERS; k++) copy_any(s, s.mem.C[CR_H0 + k], d + o.hist[k], s.n_cell * sizeof(double));
    copy_any(s, s.stats, d + o.stats, 6 * (size_t)s.nx * s.ny * sizeof(double));
    push_ctl(s);
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
cav_ensemble& need(cav_ensemble* e) {
    if (!e) fail("no ensemble");
    return *e;
}

// n steps of a batch, replayed in chunks; after each chunk a member that stalled is redone and brought level
// on its own. ev: optional 2n events bracketing each replay on the batch's stream.
void run_batch(BatchRt& rt, int n, hipEvent_t* ev) {
    constexpr int CHUNK = 16;         // replays between two looks at the control blocks
    if (stale(rt)) build(rt);
    std::vector<long long> start(rt.members.size());
    for (size_t m = 0; m < rt.members.size(); m++) start[m] = rt.members[m]->h_ctl->steps;
    for (int done = 0; done < n;) {
        const int k = std::min(CHUNK, n - done);
        for (int j = 0; j < k; j++) {
            prepare(rt);                  // a rebuild after a stall runs before the step's start event
            if (ev) HIP_CHECK(hipEventRecord(ev[2 * (done + j)], rt.stream));
            enqueue_steps(rt, 1);
            if (ev) HIP_CHECK(hipEventRecord(ev[2 * (done + j) + 1], rt.stream));
        }
        HIP_CHECK(hipStreamSynchronize(rt.stream));
        done += k;
        for (size_t m = 0; m < rt.members.size(); m++) {
            Solver& s = *rt.members[m];
            pull_ctl(s);
            if (s.h_ctl->stalled && !s.h_ctl->err) advance(s, start[m] + done);
        }
        update_spare(rt);
    }
}

// Times of n steps of a batch: total from the first replay's start to the last one's end (redone steps
// included), per step the replay alone; device synchronised around
void time_batch(BatchRt& rt, int n, double* total_ms, double* per_step) {
    std::vector<hipEvent_t> ev(2 * (size_t)n);
    for (auto& e : ev) HIP_CHECK(hipEventCreate(&e));
    HIP_CHECK(hipDeviceSynchronize());
    try {
        run_batch(rt, n, ev.data());
        HIP_CHECK(hipDeviceSynchronize());
        float t = 0;
        HIP_CHECK(hipEventElapsedTime(&t, ev[0], ev[2 * (size_t)n - 1]));
        *total_ms = t;
        if (per_step)
            for (int k = 0; k < n; k++) {
                HIP_CHECK(hipEventElapsedTime(&t, ev[2 * k], ev[2 * k + 1]));
                per_step[k] = t;
            }
    } catch (...) {
        for (auto& e : ev) (void)hipEventDestroy(e);
        throw;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
}

void set_timing(BatchRt& rt, bool on) {
    if (on) {
        if (!rt.timer) rt.timer.reset(new EventTimer());
        rt.b.ex.timer = rt.timer.get();
    } else {
        if (rt.timer) rt.timer->flush();
        rt.b.ex.timer = nullptr;
    }
}
void kernel_times(BatchRt& rt, int max, const char** names, double* ms, long long* launches, int* count) {
    int k = 0;
    if (rt.timer) rt.timer->flush();
    for (int f = 0; f < F_COUNT && k < max; f++, k++) {
        names[k] = family_name(f);
        ms[k] = rt.timer ? rt.timer->ms[f] : 0.0;
        launches[k] = rt.timer ? rt.timer->launches[f] : 0;
    }
    *count = k;
}

// After ensemble steps: refresh every member's mirror; status[m] = its error (cleared on the device)
void settle_members(BatchRt& rt, int* status, char* messages, int msg_stride) {
    for (size_t m = 0; m < rt.members.size(); m++) {
        Solver& s = *rt.members[m];
        pull_ctl(s);
        if (status) status[m] = s.h_ctl->err;
        if (messages && msg_stride > 0) {
            const std::string msg = error_message(*s.h_ctl);
            std::snprintf(messages + m * msg_stride, msg_stride, "%s", msg.c_str());
        }
        if (s.h_ctl->err) {
            s.h_ctl->err = ERR_NONE;
            push_ctl(s);
        }
    }
}

cav_solver* g_singleton = nullptr;

// ------------------------------------------------------------------------------------------------ no allocation
// cav_batch_steps_noalloc steps a list of cavities together without allocating: the only memory beyond the
// cavities' own arenas that the batched kernels read is the launch plan (member descriptor pointers and the
// per-shape block offsets), which an ensemble keeps in hipMalloc'ed memory and this path keeps in the
// file-scope array below. Its device address is taken once per device with hipGetSymbolAddress and used for
// every copy and kernel argument. The array is shared by every call on that device and holds the last call's
// plan: each call writes the whole plan it uses before its first launch, and calls that overlap on one device
// race.
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_lock;
char* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// Device address of this device's copy of the plan array, taken on the first call on the device. Host code
// uses only this address, never the array's name.
char* scratch_on_device() {
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) fail("cav_batch_steps_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_lock);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) fail("cav_batch_steps_noalloc: no device address for the plan array");
        g_scratch_ptr[dev] = (char*)p;
    }
    return g_scratch_ptr[dev];
}

// Host staging of a plan (per thread: no heap memory per call)
struct PlanHost {
    const Mem* ptr[NOALLOC_MAX_MEMBERS];
    const Mem* host[NOALLOC_MAX_MEMBERS];
    int seg[SH_COUNT * (NOALLOC_MAX_MEMBERS + 1)];
};
thread_local PlanHost t_plan;

// Write the plan of members[0 .. n) to the device at ptrs / segs on `stream` and describe it in b
void write_plan(cav_solver* const* members, int n, char* ptrs, char* segs, hipStream_t stream, Batch& b) {
    PlanHost& h = t_plan;
    for (int m = 0; m < n; m++) {
        h.ptr[m] = members[m]->d_mem;
        h.host[m] = &members[m]->mem;
    }
    batch_segments(h.host, n, h.seg, b.total);
    HIP_CHECK(hipMemcpyAsync(ptrs, h.ptr, n * sizeof(Mem*), hipMemcpyHostToDevice, stream));
    HIP_CHECK(hipMemcpyAsync(segs, h.seg, (size_t)SH_COUNT * (n + 1) * sizeof(int), hipMemcpyHostToDevice, stream));
    HIP_CHECK(hipStreamSynchronize(stream));        // the staging copy is per thread
    b.d_mem = (const Mem* const*)ptrs;
    b.d_seg = (const int*)segs;
    b.n = n;
    b.mode = members[0]->mode;
    b.ex.stream = stream;
    b.ex.timer = nullptr;
}

void pull_after(Solver& s, hipStream_t stream) {
    HIP_CHECK(hipStreamSynchronize(stream));
    pull_ctl(s);
}

// Bring a stalled member level (its clock at `target` steps) as a batch of one: the stalled step redone from
// the control block it began with, with every pass enqueued, then plain steps (the same as advance() does
// through the cavity's own plan, here through the one-member plan in the array)
void advance_noalloc(Solver& s, long long target, const Batch& one) {
    constexpr int CHUNK = 16;
    const int full = std::max(0, s.c.fdm_max_passes);
    for (;;) {
        if (s.h_ctl->err) return;
        if (s.h_ctl->stalled) {
            HIP_CHECK(hipMemcpyAsync(s.mem.ctl, s.mem.ctl_bak, sizeof(Ctl), hipMemcpyDeviceToDevice, one.ex.stream));
            enqueue_step(one, full);
            pull_after(s, one.ex.stream);
            s.redone++;
            s.spare_until = std::max(s.spare_until, s.h_ctl->steps + SPARE_KEEP);
            continue;
        }
        const long long left = target - s.h_ctl->steps;
        if (left <= 0) return;
        note_spare(s);
        const int fast = std::min(full, one.usual + (wants_spare(s) ? 1 : 0));
        for (int k = 0; k < (int)std::min<long long>(left, CHUNK); k++) enqueue_step(one, fast);
        pull_after(s, one.ex.stream);
    }
}

void batch_steps_noalloc(cav_solver* const* members, int n, int steps, CavReport* reports, int* status) {
    if (n < 1 || n > NOALLOC_MAX_MEMBERS)
        fail("cav_batch_steps_noalloc: between 1 and " + std::to_string(NOALLOC_MAX_MEMBERS) + " cavities per call (got " +
             std::to_string(n) + ")");
    if (steps < 0) fail("cav_batch_steps_noalloc: steps must be >= 0");
    if (!members) fail("cav_batch_steps_noalloc: NULL member list");
    int full = 0, passes_set = 0;
    for (int m = 0; m < n; m++) {
        if (!members[m]) fail("cav_batch_steps_noalloc: NULL cavity");
        for (int k = 0; k < m; k++)
            if (members[k] == members[m]) fail("cav_batch_steps_noalloc: a cavity appears twice");
        if (members[m]->mode != members[0]->mode) fail("cavities stepped together must share the precision mode");
        full = std::max(full, std::max(0, members[m]->c.fdm_max_passes));
        passes_set = std::max(passes_set, members[m]->solo->passes_set);
    }
    char* base = scratch_on_device();
    const hipStream_t stream = members[0]->stream;
    for (int m = 0; m < n; m++) {
        HIP_CHECK(hipStreamSynchronize(members[m]->stream));
        clear_error(*members[m]);
    }
    Batch b;
    write_plan(members, n, base, base + NOALLOC_PTRS, stream, b);
    // the passes an ensemble of these members would enqueue (build()): the usual count of the mode, plus a
    // spare while a member wants one
    const int mode = members[0]->mode, usual = passes_set > 0 ? passes_set : usual_passes(mode);
    b.usual = usual;
    auto fast_now = [&] {
        bool spare = false;
        for (int m = 0; m < n; m++) {
            note_spare(*members[m]);
            spare |= wants_spare(*members[m]);
        }
        return std::min(full, usual + (spare ? 1 : 0));
    };
    int fast = fast_now();
    constexpr int CHUNK = 16;
    long long start[NOALLOC_MAX_MEMBERS];
    for (int m = 0; m < n; m++) start[m] = members[m]->h_ctl->steps;
    for (int done = 0; done < steps;) {
        const int k = std::min(CHUNK, steps - done);
        for (int j = 0; j < k; j++) enqueue_step(b, fast);
        HIP_CHECK(hipStreamSynchronize(stream));
        done += k;
        for (int m = 0; m < n; m++) {
            Solver& s = *members[m];
            pull_ctl(s);
            if (s.h_ctl->stalled && !s.h_ctl->err) {
                Batch one;
                char* o = base + NOALLOC_PTRS + NOALLOC_SEG;
                write_plan(&members[m], 1, o, o + 256, stream, one);
                one.usual = usual;
                advance_noalloc(s, start[m] + done, one);
            }
        }
        fast = fast_now();
    }
    for (int m = 0; m < n; m++) {
        Solver& s = *members[m];
        pull_ctl(s);
        if (status) status[m] = s.h_ctl->err;
        if (s.h_ctl->err) {
            s.h_ctl->err = ERR_NONE;
            push_ctl(s);
        }
        fill_report(s, reports ? reports + m : nullptr);
    }
}

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
int cav_step(cav_solver* s, CavReport* r) { return guard([&] { steps(need(s), 1); fill_report(*s, r); }); }
int cav_steps(cav_solver* s, int n, CavReport* r) {
    return guard([&] {
        if (n < 0) fail("cav_steps: n must be >= 0");
        steps(need(s), n);
        fill_report(*s, r);
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
    return guard([&] {
        Solver& q = need(s);
        q.h_ctl->t = t;
        q.h_ctl->steps = steps;
        q.h_ctl->last_dt = last_dt;
        push_ctl(q);
    });
}
int cav_probes(cav_solver* s, float* out) {                            // backend.cu:1159-1166
    return guard([&] {
        Solver& q = need(s);
        if (q.c.n_probes == 0) return;
        const double* uc[3];
        wide_state(q, uc);
        probe_values(q.ex, q.mem.g, uc[0], uc[1], q.probe_cells, q.c.n_probes, q.probe_out);
        d2h(q, out, q.probe_out, 2 * q.c.n_probes * sizeof(float));
    });
}
int cav_add_statistics(cav_solver* s, double weight, double* centreline) {   // backend.cu:1168-1175
    return guard([&] {
        Solver& q = need(s);
        if (centreline && q.nx != q.ny) fail("centreline sample needs nx == ny");
        const double* uc[3];
        wide_state(q, uc);
        add_statistics(q.ex, q.mem.g, uc, weight, q.stats);
        if (centreline) {
            centreline_sample(q.ex, q.mem.g, uc, q.line);
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
        const double* uc[3];
        wide_state(q, uc);
        midplane(q.ex, q.mem.g, uc, lid_speed(q, q.h_ctl->t), q.plane);
        d2h(q, out, q.plane, 4 * (size_t)q.nx * q.ny * sizeof(float));
    });
}
int cav_q_volume(cav_solver* s, int colour_w, float* qv, float* colour) {    // backend.cu:1185-1194
    return guard([&] {
        Solver& q = need(s);
        const double* uc[3];
        wide_state(q, uc);
        const double LU = q.c.length / q.c.u_lid;
        float* dq = (float*)q.mem.C[CR_T1];            // scratch: two float cell arrays fit in one double one
        float* dc = dq + q.n_cell;
        q_criterion(q.ex, q.mem.g, uc, lid_speed(q, q.h_ctl->t), LU * LU, 1.0 / q.c.u_lid, colour_w, dq, dc);
        d2h(q, qv, dq, q.n_cell * sizeof(float));
        d2h(q, colour, dc, q.n_cell * sizeof(float));
    });
}
int cav_set_precision(cav_solver* s, int mode) {
    return guard([&] {
        Solver& q = need(s);
        if (mode != CAV_PREC_FP64 && mode != CAV_PREC_FDM_FP32 && mode != CAV_PREC_FDM_BF16X3 &&
            mode != CAV_PREC_FIELD_FP32)
            fail("precision mode must be 0 (FP64), 1 (FP32 FDM), 2 (split-BF16 FDM) or 3 (FP32 fields)");
        if (q.mode == mode) return;
        set_precision(q, mode);
    });
}
int cav_set_fdm_layout(cav_solver* s, int symmetric) {
    return guard([&] {
        Solver& q = need(s);
        if (q.use_sym == (symmetric != 0)) return;
        q.use_sym = symmetric != 0;
        write_mem(q);
    });
}
int cav_fdm_layout(cav_solver* s, int* axes) {
    return guard([&] {
        Solver& q = need(s);
        *axes = (q.mem.sym[X] ? 1 : 0) | (q.mem.sym[Y] ? 2 : 0) | (q.mem.sym[Z] ? 4 : 0);
    });
}
int cav_set_graph(cav_solver* s, int on) {
    return guard([&] { need(s).solo->use_graph = on != 0; });
}
int cav_pass_counters(cav_solver* s, long long* projections, long long* passes) {
    return guard([&] { Solver& q = need(s); *projections = q.h_ctl->projections; *passes = q.h_ctl->pass_total; });
}
int cav_redone_steps(cav_solver* s, long long* steps) {
    return guard([&] { *steps = need(s).redone; });
}
int cav_set_enqueued_passes(cav_solver* s, int passes) {
    return guard([&] {
        BatchRt& rt = *need(s).solo;
        rt.passes_set = std::max(0, passes);
        rt.b.d_mem = nullptr;         // rebuilt (and re-captured) at the next step
    });
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
int cav_snapshot_bytes(cav_solver* s, long long* bytes) {
    return guard([&] {
        if (!bytes) fail("cav_snapshot_bytes: NULL argument");
        *bytes = (long long)snap_layout(need(s)).total;
    });
}
int cav_snapshot(cav_solver* s, void* dst, long long bytes) { return guard([&] { snapshot(need(s), dst, bytes); }); }
int cav_device_alloc(long long bytes, void** out) {
    return guard([&] {
        if (!out || bytes <= 0) fail("cav_device_alloc: NULL argument or no bytes");
        *out = nullptr;
        HIP_CHECK(hipMalloc(out, (size_t)bytes));
    });
}
int cav_device_free(void* p) { return guard([&] { if (p) HIP_CHECK(hipFree(p)); }); }
int cav_memcpy(void* dst, const void* src, long long bytes) {
    return guard([&] {
        if (!dst || !src || bytes < 0) fail("cav_memcpy: NULL argument or negative size");
        HIP_CHECK(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDefault));
    });
}
int cav_restore(cav_solver* s, const void* src, long long bytes) { return guard([&] { restore(need(s), src, bytes); }); }
const char* cav_last_error(void) { return g_error.c_str(); }

int cav_time_steps(cav_solver* s, int n, double* ms) {
    return guard([&] {
        Solver& q = need(s);
        if (n < 1) fail("cav_time_steps: n must be >= 1");
        clear_error(q);
        time_batch(*q.solo, n, ms, nullptr);
        settle(q);
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
        Exec ex;
        ex.stream = st;
        fill(ex, a, bytes / 8, 1.0);
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
        set_timing(*q.solo, on != 0);
        q.ex.timer = q.solo->b.ex.timer;
    });
}
int cav_kernel_times(cav_solver* s, int max, const char** names, double* ms, long long* launches, int* count) {
    return guard([&] { kernel_times(*need(s).solo, max, names, ms, launches, count); });
}

// ------------------------------------------------------------------------------------------------ ensembles
int cav_ensemble_create(cav_solver* const* members, int n, cav_ensemble** out) {
    return guard([&] {
        if (!members || !out || n < 1) fail("cav_ensemble_create: need at least one cavity");
        *out = nullptr;
        std::unique_ptr<cav_ensemble> e(new cav_ensemble());
        e->rt.reset(new BatchRt());
        BatchRt& rt = *e->rt;
        for (int m = 0; m < n; m++) {
            if (!members[m]) fail("cav_ensemble_create: NULL cavity");
            for (int k = 0; k < m; k++)
                if (members[k] == members[m]) fail("cav_ensemble_create: a cavity appears twice");
            rt.members.push_back(members[m]);
        }
        HIP_CHECK(hipStreamCreateWithFlags(&rt.stream, hipStreamNonBlocking));
        rt.own_stream = true;
        build(rt);
        for (auto* s : rt.members) s->in_ensembles++;
        *out = e.release();
    });
}
int cav_ensemble_destroy(cav_ensemble* e) {
    return guard([&] {
        if (!e) return;
        if (e->rt) {
            if (e->rt->stream) (void)hipStreamSynchronize(e->rt->stream);
            for (auto* s : e->rt->members) s->in_ensembles--;
        }
        delete e;
    });
}
int cav_ensemble_size(cav_ensemble* e, int* n) {
    return guard([&] { *n = (int)need(e).rt->members.size(); });
}
int cav_ensemble_steps(cav_ensemble* e, int n, CavReport* reports, int* status) {
    return guard([&] {
        BatchRt& rt = *need(e).rt;
        if (n < 0) fail("cav_ensemble_steps: n must be >= 0");
        for (auto* s : rt.members) {
            HIP_CHECK(hipStreamSynchronize(s->stream));
            clear_error(*s);
        }
        run_batch(rt, n, nullptr);
        settle_members(rt, status, nullptr, 0);
        if (reports)
            for (size_t m = 0; m < rt.members.size(); m++) fill_report(*rt.members[m], reports + m);
    });
}
int cav_set_autotune(int on) {
    return guard([&] { g_autotune = on != 0; });
}
int cav_autotune(cav_solver* s) {
    return guard([&] {
        Solver& q = need(s);
        HIP_CHECK(hipStreamSynchronize(q.stream));
        autotune(*q.solo);
        HIP_CHECK(hipStreamSynchronize(q.stream));
    });
}
int cav_ensemble_autotune(cav_ensemble* e) {
    return guard([&] {
        BatchRt& rt = *need(e).rt;
        for (auto* s : rt.members) HIP_CHECK(hipStreamSynchronize(s->stream));
        autotune(rt);
        HIP_CHECK(hipStreamSynchronize(rt.stream));
    });
}
int cav_get_tuning(cav_solver* s, int* knobs, int count) {
    return guard([&] {
        Solver& q = need(s);
        if (!knobs || count < 1) fail("cav_get_tuning: need an array");
        for (int k = 0; k < std::min(count, (int)TU_COUNT); k++) knobs[k] = q.mem.tune[k];
    });
}
int cav_set_tuning(cav_solver* s, const int* knobs, int count) {
    return guard([&] {
        Solver& q = need(s);
        if (!knobs || count != (int)TU_COUNT) fail("cav_set_tuning: need " + std::to_string((int)TU_COUNT) + " knobs");
        if (knobs[TU_STAGE_TILE] < 0 || knobs[TU_STAGE_TILE] > 1) fail("cav_set_tuning: stage tile must be 0 or 1");
        for (int k = 1; k < TU_COUNT; k++)
            if (knobs[k] < 0) fail("cav_set_tuning: knobs must be >= 0");
        if ((knobs[TU_STAGE_MINZ] && knobs[TU_STAGE_MINZ] > 1024) || (knobs[TU_GRAD_MINZ] && knobs[TU_GRAD_MINZ] > 1025))
            fail("cav_set_tuning: shortest z chunk out of range");
        if (knobs[TU_XF] >= 64) fail("cav_set_tuning: transform layout bits out of range");
        if (knobs[TU_GRAD_TILE] > 3) fail("cav_set_tuning: gradient tile must be 0, 1, 2 or 3");
        HIP_CHECK(hipStreamSynchronize(q.stream));
        for (int k = 0; k < TU_COUNT; k++) q.mem.tune[k] = knobs[k];
        write_mem(q);
    });
}
int cav_batch_steps_noalloc(cav_solver* const* members, int n, int steps, CavReport* reports, int* status) {
    return guard([&] { batch_steps_noalloc(members, n, steps, reports, status); });
}
int cav_ensemble_time_steps(cav_ensemble* e, int n, double* total_ms, double* per_step_ms, int* status) {
    return guard([&] {
        BatchRt& rt = *need(e).rt;
        if (n < 1) fail("cav_ensemble_time_steps: n must be >= 1");
        for (auto* s : rt.members) {
            HIP_CHECK(hipStreamSynchronize(s->stream));
            clear_error(*s);
        }
        time_batch(rt, n, total_ms, per_step_ms);
        settle_members(rt, status, nullptr, 0);
    });
}
int cav_ensemble_set_enqueued_passes(cav_ensemble* e, int passes) {
    return guard([&] {
        BatchRt& rt = *need(e).rt;
        rt.passes_set = std::max(0, passes);
        rt.b.d_mem = nullptr;
    });
}
int cav_ensemble_set_graph(cav_ensemble* e, int on) {
    return guard([&] { need(e).rt->use_graph = on != 0; });
}
int cav_ensemble_set_kernel_timing(cav_ensemble* e, int on) {
    return guard([&] { set_timing(*need(e).rt, on != 0); });
}
int cav_ensemble_kernel_times(cav_ensemble* e, int max, const char** names, double* ms, long long* launches,
                              int* count) {
    return guard([&] { kernel_times(*need(e).rt, max, names, ms, launches, count); });
}
const char* cav_error_text(int code) {
    switch (code) {
        case ERR_NONE: return "";
        case ERR_VELOCITY_NAN: return "velocity is NaN/inf: the solution has blown up";
        case ERR_RHS_NAN: return "pressure solve: right-hand side is NaN/inf (the run has blown up)";
        case ERR_RESIDUAL_NAN: return "pressure solve: residual is NaN/inf";
        case ERR_NOT_CONVERGED: return "pressure solve: residual > FDM_RTOL after fdm_max_passes passes";
        default: return "unknown error";
    }
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
