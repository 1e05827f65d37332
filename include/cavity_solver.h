/* cavity_solver.h — C-ABI of libcavity_solver.so, the MI355X (gfx950) lid-driven cavity solver.
 *
 * The first block is the drop-in boundary: the 18 `solver_*` entry points and the three records that the
 * reference ctypes binding loads (python_api.py:11-28 records, python_api.py:43-161 calls; implemented by
 * backend.cu:980-1195). Same names, argument meaning, return convention (0 ok, 1 error, message through
 * solver_last_error) and byte layouts, so python_api.py, main.py and validate_gpu.py drive this library
 * unchanged. One difference is deliberate: nx != ny is accepted (nx, ny even, >= 6); the centreline sample
 * of solver_add_statistics still needs nx == ny.
 *
 * The second block are additions with no counterpart in the reference: FP64 field I/O, precision modes,
 * pass counters, timing hooks, and a handle API (cav_*) that the singleton API is a thin wrapper over.
 *
 * Precision: the default mode computes every field (u, m, phi) and every stencil in FP64. The float32
 * field I/O of solver_get_field/solver_set_field converts at the boundary, exactly as the reference's
 * float fields would read.  Fields are x-fastest flat arrays, NumPy order [z, y, x]:
 *   u, mx: (nz, ny, nx+1)   v, my: (nz, ny+1, nx)   w, mz: (nz+1, ny, nx)   phi: (nz, ny, nx)
 */
#ifndef CAVITY_SOLVER_H
#define CAVITY_SOLVER_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- records (byte layouts are the ABI) */

/* Config — backend.cu:584-589, python_api.py:11-18. 288 bytes: 11 int32, 4 bytes padding, 14 float64,
 * probe_x[8], probe_y[8]. Offsets: nx 0, n_probes 40, length 48, fdm_rtol 152, probe_x 160, probe_y 224. */
typedef struct CavConfig {
    int nx, ny, nz, lid_shaped, impulse_order, adaptive_dt, dt_lag, phi_extrapolate, fdm_max_passes,
        grid_stretch, n_probes;
    double length, u_lid, nu, lz, lid_ramp_t, reinit_r, impulse_order4_ramp, grid_dx_wall, grid_max_aspect;
    double dt_fixed, dt_max, cfl_target, diffusion_limit, fdm_rtol;
    double probe_x[8], probe_y[8];
} CavConfig;

/* Report — backend.cu:590-594, python_api.py:21-23. 48 bytes. dt_reason: 0 fixed, 1 DT_MAX, 2 CFL,
 * 3 viscous. phi_rel is the relative residual (= relative vol-weighted divergence) of the last projection,
 * phi_passes its FDM pass count. */
typedef struct CavReport {
    long long steps;
    double time, dt, cfl, phi_rel;
    int dt_reason, phi_passes;
} CavReport;

/* Diagnostics — backend.cu:595-597, python_api.py:26-28. 64 bytes. */
typedef struct CavDiagnostics {
    double kinetic_energy, enstrophy, umax, div_rms, cfl_rate, gauge_ratio, gauge_x, gauge_y;
} CavDiagnostics;

/* ---------------------------------------------------------------- the reference boundary (singleton) */

/* backend.cu:1009-1073. Destroys any previous solver, validates, builds the grid and the FDM
 * eigensystems, allocates device state. */
int solver_create(const CavConfig* config);
/* backend.cu:1000-1008. */
int solver_destroy(void);
/* backend.cu:1075-1088. m = (0, v_noise, w_noise) (NULL = zero), u = its projection, clock reset.
 * v_noise has (nz, ny+1, nx) entries, w_noise (nz+1, ny, nx). */
int solver_initialize(const double* v_noise, const double* w_noise);
/* backend.cu:1089-1091 -> step() backend.cu:877-955: one SSP-RK3 step. */
int solver_step(CavReport* report);
/* backend.cu:1092-1094 -> kick() backend.cu:958-978. Three face-shaped noise arrays (u, v, w shapes). */
int solver_kick(const double* noise_u, const double* noise_v, const double* noise_w, double amp, int passes);
/* backend.cu:1095. */
int solver_report(CavReport* report);
/* backend.cu:1096-1129. */
int solver_diagnostics(CavDiagnostics* d);
/* backend.cu:1131-1137. axis 0 x, 1 y, 2 z: faces (n+1) and centres (n). */
int solver_grid(int axis, double* faces, double* centres);
/* backend.cu:1139-1146. field 0..5 = u, v, w, mx, my, mz as float32; 6 = phi as float64. */
int solver_get_field(int field, void* out);
/* backend.cu:1147-1155. Also clears the projection warm-start history and the lagged dt rates. */
int solver_set_field(int field, const void* in);
/* backend.cu:1156-1158. */
int solver_set_clock(double t, long long steps, double last_dt);
/* backend.cu:1159-1166. u, v at the centre of each probe cell, 2*n_probes float32. */
int solver_probes(float* out);
/* backend.cu:1168-1175. sums += weight*(u, v, uu, vv, ww, uv) on the mid-span plane; centreline gets
 * 5*nx float64 (u(y) at x = L/2, v(x) at y = L/2, uu, vv, ww). centreline may be NULL when nx != ny. */
int solver_add_statistics(double weight, double* centreline);
/* backend.cu:1176-1177. 6*nx*ny float64. */
int solver_get_statistics(double* sums);
int solver_set_statistics(const double* sums);
/* backend.cu:1178-1184. [u | v | w | |omega|] on the mid-span plane, 4*nx*ny float32. */
int solver_midplane(float* out);
/* backend.cu:1185-1194. Q L^2/U^2 and colour (w/U or |u|/U) per cell, float32. */
int solver_q_volume(int colour_w, float* q, float* colour);
/* backend.cu:1195. Valid until the next failing call on this thread. */
const char* solver_last_error(void);

/* ---------------------------------------------------------------- additions (no reference counterpart) */

enum {
    CAV_PREC_FP64 = 0,        /* every field, stencil and FDM transform in FP64 (default) */
    CAV_PREC_FDM_FP32 = 1,    /* FP64 fields/stencils/residuals; FDM transforms FP32 MFMA, FP64 refinement */
    CAV_PREC_FDM_BF16X3 = 2,  /* as 1, transforms as split-BF16 x3 MFMA products with FP32 accumulation */
    CAV_PREC_FIELD_FP32 = 3   /* the reference's own precision: float velocity and impulse, grid rounded through
                                 float, float stencil arithmetic, FP64 potential and residuals, FP32 transforms
                                 with FP64 refinement. Entering or leaving this mode redoes the set-up on the
                                 new grid: u and m are carried over (rounded to float, or widened), phi and the
                                 clock kept, the warm-start history and lagged dt rates cleared. */
};

/* Full-precision field I/O for the FP64 state (fields 0..6, all float64). */
int solver_get_field64(int field, double* out);
int solver_set_field64(int field, const double* in);
/* Precision of the pressure solve (CAV_PREC_*); may change between steps. */
int solver_set_precision(int mode);
/* FDM pass counters since create: projections done and passes summed over them. */
int solver_pass_counters(long long* projections, long long* passes);

/* Handle API. Every solver_* call above runs on one process-wide handle; these run on any number of
 * independent solvers on the current device. Not thread-safe per handle. */
typedef struct cav_solver cav_solver;
int cav_create(const CavConfig* config, cav_solver** out);
int cav_destroy(cav_solver* s);
int cav_initialize(cav_solver* s, const double* v_noise, const double* w_noise);
int cav_step(cav_solver* s, CavReport* report);
/* n steps back to back; report after the last one. */
int cav_steps(cav_solver* s, int n, CavReport* report);
int cav_kick(cav_solver* s, const double* nu, const double* nv, const double* nw, double amp, int passes);
int cav_report(cav_solver* s, CavReport* report);
int cav_diagnostics(cav_solver* s, CavDiagnostics* d);
int cav_grid(cav_solver* s, int axis, double* faces, double* centres);
int cav_get_field(cav_solver* s, int field, void* out);
int cav_set_field(cav_solver* s, int field, const void* in);
int cav_get_field64(cav_solver* s, int field, double* out);
int cav_set_field64(cav_solver* s, int field, const double* in);
int cav_set_clock(cav_solver* s, double t, long long steps, double last_dt);
int cav_probes(cav_solver* s, float* out);
int cav_add_statistics(cav_solver* s, double weight, double* centreline);
int cav_get_statistics(cav_solver* s, double* sums);
int cav_set_statistics(cav_solver* s, const double* sums);
int cav_midplane(cav_solver* s, float* out);
int cav_q_volume(cav_solver* s, int colour_w, float* q, float* colour);
int cav_set_precision(cav_solver* s, int mode);
int cav_pass_counters(cav_solver* s, long long* projections, long long* passes);
/* Steps redone since create because a projection needed more refinement passes than a step enqueued: by
 * default the usual count of the precision mode (FP64 1, FP32 2, split BF16 2 + 1), plus one after the first
 * stall of the cavity (or of the ensemble stepping it); the redone step enqueues all fdm_max_passes. */
int cav_redone_steps(cav_solver* s, long long* steps);
/* Refinement passes a step enqueues per projection before a stall (0: the default), the spare after a first
 * stall still added; a smaller count makes steps stall and be redone, with the same results. */
int cav_set_enqueued_passes(cav_solver* s, int passes);
/* FDM transform layout: symmetric != 0 (default) uses the even/odd half transforms on every axis whose grid
 * mirrors about its centre (half the MFMA work); 0 forces the dense n x n transforms. cav_fdm_layout
 * reports the axes using the split as a bit mask (1 x, 2 y, 4 z). */
int cav_set_fdm_layout(cav_solver* s, int symmetric);
int cav_fdm_layout(cav_solver* s, int* axes);
/* Steps run as one captured HIP graph per solver (default on); 0 launches the kernels one by one. */
int cav_set_graph(cav_solver* s, int on);
/* Snapshot and restore of a solver's whole state between calls (clock, lagged dt rates, warm-start history,
 * u, m, phi, statistics sums): a restored solver continues bit for bit as the one the snapshot was taken
 * from. For preempting a serving slot (snapshot to host memory, destroy, later create + restore) or keeping
 * a device-resident copy; dst/src may be device or host memory. The solver restored into must have been
 * created from the same configuration and run the same precision mode. No counterpart in the reference
 * (its checkpoints, main.py:159-191, keep only the fields and the clock). */
int cav_snapshot_bytes(cav_solver* s, long long* bytes);
int cav_snapshot(cav_solver* s, void* dst, long long bytes);
int cav_restore(cav_solver* s, const void* src, long long bytes);
/* Device memory of this library's HIP runtime, for snapshots kept on the device (a process may hold another
 * HIP runtime, e.g. one bundled with a framework, whose allocations this library cannot rely on), and a copy
 * between any two buffers, host or device. No counterpart in the reference. */
int cav_device_alloc(long long bytes, void** out);
int cav_device_free(void* p);
int cav_memcpy(void* dst, const void* src, long long bytes);
/* Device bytes held by one solver of this configuration (state + work arrays + FDM tables). */
int cav_footprint(const CavConfig* config, int mode, long long* bytes);
/* Message of the last failing cav_* or solver_* call on this thread. */
const char* cav_last_error(void);

/* Timing hooks. cav_time_steps runs n steps between two events on the solver's stream (device
 * synchronised before and after) and returns the elapsed milliseconds. With kernel timing on, every
 * kernel launch of the solver is bracketed by events, accumulated per kernel family. */
int cav_time_steps(cav_solver* s, int n, double* ms);
int cav_set_kernel_timing(cav_solver* s, int on);
/* Families in a fixed order; names[k] is a static string. Returns the number of families in *count. */
int cav_kernel_times(cav_solver* s, int max, const char** names, double* ms, long long* launches, int* count);
/* Plain device copy of `bytes` (two fresh buffers), averaged over `repeats` launches: the reference
 * kernel for calibrating HBM byte counters and the achievable copy bandwidth. */
int cav_copy_probe(long long bytes, int repeats, double* ms);

/* Ensembles: any number of solvers (sizes, grids, Reynolds numbers may differ; the precision mode must be
 * the same) advanced together, every kernel of the step launched once for all of them and the whole step
 * replayed as one HIP graph. Each member gives the same fields as when stepped alone. Members stay usable
 * on their own between ensemble calls (field I/O, kicks, outputs); a member cannot be destroyed while an
 * ensemble holds it. reports[n] and status[n] (0 ok, else the error code of cav_error_text) are optional;
 * a member that fails stops its clock and its fields are undefined until it is re-initialised or its
 * fields are set. */
typedef struct cav_ensemble cav_ensemble;
int cav_ensemble_create(cav_solver* const* members, int n, cav_ensemble** out);
int cav_ensemble_destroy(cav_ensemble* e);
int cav_ensemble_size(cav_ensemble* e, int* n);
int cav_ensemble_steps(cav_ensemble* e, int steps, CavReport* reports, int* status);
/* Steps between HIP events: total and (optional, steps entries) per-step milliseconds */
int cav_ensemble_time_steps(cav_ensemble* e, int steps, double* total_ms, double* per_step_ms, int* status);
int cav_ensemble_set_graph(cav_ensemble* e, int on);
int cav_ensemble_set_enqueued_passes(cav_ensemble* e, int passes);
int cav_ensemble_set_kernel_timing(cav_ensemble* e, int on);
int cav_ensemble_kernel_times(cav_ensemble* e, int max, const char** names, double* ms, long long* launches,
                              int* count);
/* Steps `steps` times the cavities members[0 .. n) together, as cav_ensemble_create + cav_ensemble_steps +
 * cav_ensemble_destroy would, with the same results bit for bit, but allocating nothing: no hipMalloc, no
 * hipFree, no stream, event or graph. The launch plan (member descriptor pointers and per-launch block
 * offsets) lives in a fixed 64 KiB device array of the library, one per device and held until the process
 * ends; the steps are plain kernel launches on members[0]'s stream. At most 256 cavities per call (more is
 * refused with a message); all must share the precision mode and live on the current device. Every call on a
 * device shares that device's array and rewrites it before its first launch, so calls that overlap on one
 * device (other threads) race. reports[n] and status[n] (0 ok, else the code of cav_error_text) are
 * optional. No counterpart in the reference: for callers that cannot allocate device memory (tight serving
 * loops, code that must not call hipMalloc). */
int cav_batch_steps_noalloc(cav_solver* const* members, int n, int steps, CavReport* reports, int* status);
/* Autotuner. Each cavity's launch choices (the stage kernel's column tile 32 x 8 or 16 x 16, its z-chunk
 * target and shortest chunk, the gradient's z chunks, the FDM transform layout of each mirror-symmetric
 * axis) are timed on the device and the fastest kept; no choice changes a result bit. By default every
 * cavity is tuned as cav_create makes it, once per configuration and process (cached). cav_set_autotune(0)
 * turns that off for cavities created later (the built-in defaults then apply). cav_autotune tunes one cavity
 * now; cav_ensemble_autotune tunes an ensemble's members for stepping together (their z-chunk targets become
 * shares of a total). Knobs (7 ints, 0 = default): tile (0 32x8, 1 16x16), stage chunk target, stage shortest
 * chunk, gradient chunk target, gradient shortest chunk, transform layout (2 bits per axis: 1 even/odd by
 * chunks, 2 by whole columns), gradient tile (0 32x8, 1 64x4, 2 128x2, 3 256x1). No counterpart in the reference (one
 * launch shape, backend.cu:23). */
int cav_set_autotune(int on);
int cav_autotune(cav_solver* s);
int cav_ensemble_autotune(cav_ensemble* e);
int cav_get_tuning(cav_solver* s, int* knobs, int count);
int cav_set_tuning(cav_solver* s, const int* knobs, int count);

/* Text of a step error code */
const char* cav_error_text(int code);

#ifdef __cplusplus
}
#endif
#endif
