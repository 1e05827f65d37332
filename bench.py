"""Throughput of the cavity time step on one MI355X: Mcell-steps/s of the plain FP64 SSP-RK3 step on the
256^3 tanh-stretched cavity (stretch factor 2.0, Re 3200, regularised lid, CFL 0.5, FDM rtol 1e-10), with the
mixed-precision FDM modes, the 64-cavity ensemble (p50/p99 step ms) and a serving window beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--size N] [--no-cpu-baseline] [--traffic 0|1]
                    [--ensemble-steps S] [--serve-seconds T]

A step is one full solver step (gauge reset, CFL/dt, three RK stages with projections and their FDM
refinement loops) through the library's C-ABI on inputs resident in HBM, replayed as a HIP graph. The timed
region is K steps between two HIP events on the solver's stream, device-synchronised on both sides. A second
region of K steps launches kernel by kernel with an event pair around every launch, giving each kernel
family's average duration; the dominant family's algorithmic bytes per launch over that duration is the
roofline figure, and two rocprofv3 counter runs (FETCH_SIZE, WRITE_SIZE, scaled by a copy kernel of known
size) give its HBM traffic per launch. The CPU baseline is the FP64 host oracle (OpenMP) on a bounded sample
of the same workload.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = 'Mcell-steps/s at 256^3 FP64 vs mixed-precision FDM; ensemble p50/p99 step ms'
HBM_PEAK_GBS = 8000.0
MIXED_EXTRA_WARMUP = 40          # untimed steps the mixed-precision modes add to --warmup (their steady state)
FP64_MFMA_PEAK_TFLOPS = 78.6


def algorithmic_work(family, nx, ny, nz, launches_per_step, sym_axes=0, FB=8):
    """(bytes, flops) per launch of a kernel family: each array read or written once, 8 B per value (FB per
    face value: 4 in the FP32-field mode, whose potential and cell arrays stay FP64). Averages over the
    launches one step makes (e.g. the three RK stages). An axis transform is a dense n x n contraction (2n
    flops per value), or on a mirror-symmetric axis (bit of sym_axes set) two h x h contractions of the folded
    halves (n flops per value)."""
    n = nx * ny * nz
    faces = (nx + 1) * ny * nz + nx * (ny + 1) * nz + nx * ny * (nz + 1)
    D = 8
    for bit, (name, length) in enumerate((('xform_x', nx), ('xform_y', ny), ('xform_z', nz))):
        if family == name:
            return 2 * n * D, (1 if sym_axes >> bit & 1 else 2) * n * length
    if family == 'stage':                      # reads u, m, m0 in stages 2 and 3; writes the stage's m
        return (3 * faces + faces * 2 / 3) * FB, 0
    if family == 'residual':                   # the FP64 step's one pass per projection: its residual is only
        return 2 * n * D, 0                    # measured (reads phi and rhs; r is not stored)
    if family == 'divergence':                 # fused with the guess residual: reads m and phi, writes rhs and r
        return faces * FB + 3 * n * D, 0
    if family == 'subtract_gradient':
        # every launch reads m, the work potential and the base (the first stage's base is phi0) and writes the
        # velocity and dphi. The first two of a step's three also form the next stage's base and guess: they read
        # the one history entry not in registers and (the second) phi0, and write both. The last writes the
        # state's potential and the next step's m0, phi0 and first guess (reading one history entry); its CFL
        # rate re-reads only values the launch has just touched.
        return 2 * faces * FB + 3 * n * D + (3 * n * D + 4 * n * D + faces * FB + 4 * n * D) / 3, 0
    if family == 'gauge_reset':
        return 3 * faces * FB + 2 * n * D, 0
    if family == 'cfl_max':
        return faces * FB, 0
    if family == 'base_guess':                 # the first stage's guess when the previous step could not form it
        return 4 * n * D, 0                    # (reads phi0 and two history entries, writes the guess)
    return None, None


def cpu_model():
    try:
        for line in open('/proc/cpuinfo'):
            if line.startswith('model name'):
                return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return 'unknown'


def cpu_baseline(settings, noise, steps, warmup=1):
    """The FP64 OpenMP oracle on the GPU box's host cores: `warmup` untimed steps, then `steps` timed steps of
    the same workload from the same initial state; the time inside its FDM transforms (a cache-blocked,
    fixed-order FP64 GEMM standing in for the reference's cuBLAS SGEMM) is reported apart from the rest
    (stencils, residuals, reductions)."""
    from oracle.oracle import Oracle
    o = Oracle(settings)
    o.initialize(*noise)
    if warmup:
        o.steps(warmup)
    x0, threads = o.timing()
    t0 = time.perf_counter()
    o.steps(steps)
    dt = time.perf_counter() - t0
    x1, _ = o.timing()
    n = settings['nx'] * settings['ny'] * settings['nz']
    o.close()
    xform = x1 - x0
    return {'value': n * steps / dt / 1e6, 'unit': 'Mcell-steps/s', 'cores': threads, 'kind': 'port',
            'cpu_model': cpu_model(), 'omp_proc_bind': os.environ.get('OMP_PROC_BIND', 'unset'),
            'host_cpus_visible': os.cpu_count(),
            'seconds_per_step': dt / steps, 'transform_seconds_per_step': xform / steps,
            'stencil_and_rest_seconds_per_step': (dt - xform) / steps,
            'sample': f'{warmup} warm-up + {steps} timed step(s) of the same {settings["nx"]}x{settings["ny"]}x'
                      f'{settings["nz"]} workload from the same initial state, FP64 OpenMP oracle on {threads} '
                      f'thread(s), {dt:.1f} s timed; FDM transforms as a cache-blocked fixed-order FP64 GEMM'}


def run_ensemble(args):
    """The 64-cavity ensemble (sizes from {32, 48, 64, 96} per axis, stretch U[1, 2], Re log-U[100, 5000],
    FP64 stencils, mixed FDM): every kernel launched once for all members, the step replayed as one HIP
    graph. Per-step times from HIP events give p50/p99; the same cavities stepped one at a time (each its
    own graph) give the sequential throughput."""
    import numpy as np
    from les_impulse3d_64_amd import workloads as W
    from les_impulse3d_64_amd.python_api import Ensemble, Solver
    members = []
    print(f'[bench] ensemble: creating {args.ensemble_members} cavities', file=sys.stderr, flush=True)
    for i in range(args.ensemble_members):
        st, noise = W.ensemble_member(i)
        g = Solver(st, handle=True, precision=args.ensemble_precision)
        g.initialize(*noise)
        members.append(g)
    cells = np.array([g.nx * g.ny * g.nz for g in members], dtype=np.float64)
    zero = {k: 0 for k in Solver.TUNE_KNOBS}
    for g in members:                        # the built-in launch choices first
        g.set_tuning(**zero)
    e = Ensemble(members)
    e.steps(10)
    fixed_steps = max(20, args.ensemble_steps // 5)
    print(f'[bench] ensemble: built-in launch choices, timed {fixed_steps} steps', file=sys.stderr, flush=True)
    _, per_fixed, _ = e.time_steps(fixed_steps)
    t0 = time.perf_counter()
    e.autotune()
    tune_s = time.perf_counter() - t0
    e.steps(5)
    print(f'[bench] ensemble: tuned ({tune_s:.2f} s), timed {args.ensemble_steps} steps', file=sys.stderr, flush=True)
    total, per, errors = e.time_steps(args.ensemble_steps)
    reports, _ = e.steps(1)
    tuned = [g.tuning() for g in members]
    e.close()
    failed = [k for k, t in enumerate(errors) if t]
    batched = cells.sum() * args.ensemble_steps / (total / 1e3) / 1e6
    sequential = seq_per_cavity = None
    if args.sequential_steps > 0:
        print(f'[bench] ensemble: sequential {args.sequential_steps} steps per cavity', file=sys.stderr, flush=True)
        seq_ms = 0.0
        for g in members:
            g.autotune()                     # its own (cached) choices for stepping alone
            g.steps(2)
            seq_ms += g.time_steps(args.sequential_steps)
        sequential = cells.sum() * args.sequential_steps / (seq_ms / 1e3) / 1e6
        seq_per_cavity = seq_ms / args.sequential_steps / len(members)
    for g in members:
        g.close()
    return {'members': len(members), 'cells_total': int(cells.sum()), 'steps': args.ensemble_steps,
            'precision': args.ensemble_precision, 'step_ms_p50': float(np.percentile(per, 50)),
            'step_ms_p99': float(np.percentile(per, 99)), 'step_ms_mean': float(per.mean()),
            'fixed_choices_step_ms_p50': float(np.percentile(per_fixed, 50)),
            'fixed_choices_step_ms_p99': float(np.percentile(per_fixed, 99)), 'fixed_choices_steps': fixed_steps,
            'autotune_seconds': tune_s,
            'tuned_stage_tile_16x16_members': sum(1 for t in tuned if t['stage_tile'] == 1),
            'mcell_steps_per_s': batched, 'sequential_mcell_steps_per_s': sequential,
            'sequential_step_ms_per_cavity_mean': seq_per_cavity,
            'speedup_vs_sequential': batched / sequential if sequential else None,
            'phi_rel_max': max(r['phi_rel'] for r in reports), 'failed_members': failed}


COPY_BYTES = 1 << 30
# dispatches of a kernel in the profiling child's two kernel-by-kernel FP64 steps
PMC_DISPATCHES = {'kb_stage': 6, 'kb_residual': 12}


def pmc_child(args):
    """The program the counter runs profile: a plain copy of COPY_BYTES (the calibration kernel), then two
    kernel-by-kernel steps of the workload."""
    import ctypes as ct
    from les_impulse3d_64_amd import workloads as W
    from les_impulse3d_64_amd.python_api import Solver, load_library
    lib = load_library()
    ms = ct.c_double()
    lib.cav_copy_probe.argtypes = [ct.c_longlong, ct.c_int, ct.c_void_p]
    if lib.cav_copy_probe(COPY_BYTES, 3, ct.byref(ms)):
        raise RuntimeError(lib.cav_last_error().decode())
    settings, noise, _ = W.c2_mixed(args.size)
    s = Solver(settings, handle=True)
    s.initialize(*noise)
    s.set_graph(False)
    s.steps(2)
    s.close()


def measure_traffic(args, kernel):
    """HBM bytes per launch of `kernel` (a kernel-name fragment): FETCH_SIZE and WRITE_SIZE (KiB per
    dispatch) from two rocprofv3 counter runs of the profiling child, each scaled by what the same counter
    reads for the copy kernel of COPY_BYTES read + COPY_BYTES written."""
    import csv
    import glob
    import shutil
    import subprocess
    import tempfile
    exe = shutil.which('rocprofv3')
    if not exe:
        return None, 'rocprofv3 not found'
    out = {}
    work = tempfile.mkdtemp(prefix='cav_pmc_', dir='/tmp')
    env = dict(os.environ, TMPDIR='/tmp')
    for counter in ('FETCH_SIZE', 'WRITE_SIZE'):
        d = os.path.join(work, counter)
        cmd = [exe, '--pmc', counter, '--output-format', 'csv', '-d', d, '-o', 'run', '--', sys.executable,
               str(ROOT / 'bench.py'), '--pmc-child', '--size', str(args.size)]
        print(f'[bench] counters: {counter}', file=sys.stderr, flush=True)
        try:
            subprocess.run(cmd, cwd='/tmp', env=env, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                           timeout=240, check=False)
        except subprocess.TimeoutExpired:
            return None, f'rocprofv3 {counter} run timed out'
        files = glob.glob(os.path.join(d, '**', '*counter_collection.csv'), recursive=True)
        if not files:
            return None, f'no {counter} counter file'
        vals = {}
        with open(files[0]) as f:
            for row in csv.DictReader(f):
                if row.get('Counter_Name') != counter:
                    continue
                name = row['Kernel_Name']
                # the kernel itself (its name followed by its arguments or template arguments), not another
                # kernel whose name starts the same (kb_residual / kb_residual_x)
                mine = kernel + '(' in name or kernel + '<' in name
                key = 'copy' if 'k_copy(' in name else kernel if mine else None
                if key:
                    order = int(row['Dispatch_Id']) if row.get('Dispatch_Id') else len(vals.get(key, []))
                    vals.setdefault(key, []).append((order, float(row['Counter_Value'])))
        if 'copy' not in vals or kernel not in vals:
            return None, f'{counter}: kernel not found in the counter file'
        # the dispatches of the child's two kernel-by-kernel steps: the last ones (earlier ones are the
        # autotuner's trials of other launch choices at creation)
        mine = [v for _, v in sorted(vals[kernel], key=lambda ov: ov[0])][-PMC_DISPATCHES.get(kernel, 6):]
        copy = [v for _, v in vals['copy']]
        factor = COPY_BYTES / (sum(copy) / len(copy) * 1024.0)
        out[counter] = (sum(mine) / len(mine) * 1024.0 * factor, factor, len(mine))
    shutil.rmtree(work, ignore_errors=True)
    fetch, write = out['FETCH_SIZE'], out['WRITE_SIZE']
    return {'bytes': fetch[0] + write[0], 'read_bytes': fetch[0], 'write_bytes': write[0],
            'fetch_scale': fetch[1], 'write_scale': write[1], 'dispatches': fetch[2]}, None


def run_serving(args):
    """Continuous batching under the serving workload (Poisson arrivals at 20/s, sizes 32^3..192^3 with
    anisotropic variants, Re log-U[100, 1e4], 100..2000 steps, FDM FP32 at rtol 1e-10) for a bounded window,
    once per admission policy: admission stops at the end of the window; solves still waiting or running then
    are unfinished, and their latency is reported censored at the window end beside the completed ones'."""
    from les_impulse3d_64_amd.serving import Server, poisson_requests
    out = {}
    for policy in args.serve_policies.split(','):
        reqs = poisson_requests(args.serve_seconds, rate=20.0, seed=4)
        print(f'[bench] serving ({policy}): {len(reqs)} requests over {args.serve_seconds:.0f} s', file=sys.stderr,
              flush=True)
        srv = Server(budget_bytes=int(args.serve_budget_gib * 2 ** 30), max_members=args.serve_members, chunk=20,
                     precision='fdm_fp32', policy=policy, prepare_workers=1)
        def report(now, done, waiting, running):
            print(f'[bench] serving ({policy}) t={now:.0f} s: {done} completed, {waiting} waiting, {running} running',
                  file=sys.stderr, flush=True)
        stats = srv.run(reqs, duration=args.serve_seconds, drain=False, progress=report)
        summ = stats.summary()
        rule = ('in arrival order, stopping at the first that does not fit' if policy == 'fifo' else
                'smallest remaining work (cells x steps) first, passing over what does not fit')
        summ.update(offered_total=len(reqs), window_s=args.serve_seconds, arrival_rate_per_s=20.0,
                    policy=f'{policy}: admit {rule}, while <= {args.serve_members} solves and their footprints fit '
                           f'in {args.serve_budget_gib:g} GiB; solvers created on a worker thread; chunks of <= 20 '
                           f'steps; no drain at the window end',
                    queue_depth=[(round(t, 2), w, a) for t, w, a in
                                 stats.queue_depth[::max(1, len(stats.queue_depth) // 20)]])
        out[policy] = summ
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=5)
    ap.add_argument('--size', type=int, default=256)
    ap.add_argument('--cpu-steps', type=int, default=3)
    ap.add_argument('--cpu-warmup', type=int, default=1)
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--fp64-only', action='store_true')
    ap.add_argument('--modes', default='fp64,fdm_fp32,fdm_bf16x3,field_fp32',
                    help='precision modes measured beside FP64 (FP64 always runs)')
    ap.add_argument('--copy-probe', type=int, default=0, help='only run a plain copy of this many bytes and print its bandwidth')
    ap.add_argument('--ensemble-steps', type=int, default=500, help='steps of the 64-cavity ensemble (0: skip it)')
    ap.add_argument('--ensemble-members', type=int, default=64)
    ap.add_argument('--sequential-steps', type=int, default=20, help='steps per cavity of the one-at-a-time run (0: skip it)')
    ap.add_argument('--ensemble-precision', default='fdm_fp32')
    ap.add_argument('--serve-seconds', type=float, default=45.0, help='serving window per policy (0: skip it)')
    ap.add_argument('--serve-policies', default='fifo,smallest')
    ap.add_argument('--serve-members', type=int, default=32)
    ap.add_argument('--serve-budget-gib', type=float, default=160.0)
    ap.add_argument('--traffic', type=int, default=1, help='measure HBM bytes per launch with rocprofv3 counters')
    ap.add_argument('--pmc-child', action='store_true', help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f'bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); each MI355X box here has '
              f'one GPU, and a cavity is not sharded across GPUs in this build.', file=sys.stderr)
        sys.exit(2)

    from les_impulse3d_64_amd import workloads as W
    from les_impulse3d_64_amd.python_api import Solver, load_library

    if args.pmc_child:
        pmc_child(args)
        return
    if args.copy_probe:
        import ctypes as ct
        lib = load_library()
        ms = ct.c_double()
        lib.cav_copy_probe.argtypes = [ct.c_longlong, ct.c_int, ct.c_void_p]
        if lib.cav_copy_probe(args.copy_probe, 20, ct.byref(ms)):
            raise RuntimeError(lib.cav_last_error().decode())
        print(json.dumps({'copy_bytes': args.copy_probe, 'ms': ms.value,
                          'GBs_read_plus_write': 2 * args.copy_probe / (ms.value / 1e3) / 1e9}), flush=True)
        return

    settings, noise, _ = W.c2_mixed(args.size)
    nx, ny, nz = settings['nx'], settings['ny'], settings['nz']
    ncell = nx * ny * nz
    runs = {}
    solvers = {}
    wanted = {m.strip() for m in args.modes.split(',') if m.strip()}
    for mode in ('fp64', 'fdm_fp32', 'fdm_bf16x3', 'field_fp32'):
        if mode != 'fp64' and (args.fp64_only or mode not in wanted):
            continue
        print(f'[bench] workload {nx}x{ny}x{nz} {mode}: create', file=sys.stderr, flush=True)
        s = Solver(settings, handle=True, precision=mode)
        s.initialize(*noise)
        # the mixed modes' cold start can need a third refinement pass, which keeps a spare pass enqueued for
        # a while (the solver drops it after SPARE_KEEP quiet steps): they warm up past that to their steady state
        warm = args.warmup + (MIXED_EXTRA_WARMUP if mode != 'fp64' else 0)
        print(f'[bench] {mode}: warmup {warm}', file=sys.stderr, flush=True)
        if warm:
            s.steps(warm)
        p0 = s.pass_counters()
        print(f'[bench] {mode}: timed {args.steps}', file=sys.stderr, flush=True)
        ms = s.time_steps(args.steps)
        p1 = s.pass_counters()
        rep = s.report()
        runs[mode] = {'value': ncell * args.steps / (ms / 1e3) / 1e6, 'ms_per_step': ms / args.steps,
                      'fdm_passes_mean': (p1[1] - p0[1]) / max(1, p1[0] - p0[0]), 'phi_rel_last': rep['phi_rel'],
                      'fdm_rtol': settings['fdm_rtol']}
        solvers[mode] = s
    base = runs['fp64']
    if len(runs) > 1:
        solvers['fp64'].steps(MIXED_EXTRA_WARMUP)      # level with the mixed modes' longer warm-up
    for mode, r in runs.items():
        if mode == 'fp64':
            continue
        r['speedup_vs_fp64'] = r['value'] / base['value']
        r['warmup'] = args.warmup + MIXED_EXTRA_WARMUP
        # both solvers advanced the same number of steps from the same state
        err = 0.0
        for c in ('u', 'v', 'w'):
            a, b = solvers[mode].field64(c), solvers['fp64'].field64(c)
            err = max(err, float(abs(a - b).max() / max(abs(b).max(), 1e-300)))
        r['velocity_rel_err_vs_fp64'] = err
    reference_precision = None
    if 'field_fp32' in solvers:
        # the reference's own precision (float fields): its stage kernel against the float face bytes
        f = solvers['field_fp32']
        f.steps(2)
        f.set_kernel_timing(True)
        f.time_steps(args.steps)
        f.set_kernel_timing(False)
        kt = f.kernel_times()
        reference_precision = dict(runs.pop('field_fp32'))
        kms, launches = kt.get('stage', (0.0, 0))
        if launches:
            b, _ = algorithmic_work('stage', nx, ny, nz, launches / args.steps, f.fdm_layout(), FB=4)
            avg = kms / launches
            reference_precision['stage'] = {'avg_launch_ms': avg, 'algorithmic_bytes': b,
                                            'achieved_GBs': b / (avg / 1e3) / 1e9,
                                            'frac': b / (avg / 1e3) / 1e9 / HBM_PEAK_GBS}
        reference_precision['kernel_ms_per_step'] = {k: v[0] / args.steps for k, v in kt.items() if v[1]}
    for mode in list(solvers):
        if mode != 'fp64':
            solvers.pop(mode).close()

    s = solvers['fp64']
    tuning = s.tuning()
    s.set_tuning(**{k: 0 for k in tuning})
    s.steps(2)
    untuned_ms = s.time_steps(args.steps) / args.steps
    s.set_tuning(**tuning)
    s.steps(2)
    print('[bench] kernel-timed region', file=sys.stderr, flush=True)
    q0 = s.pass_counters()
    s.set_kernel_timing(True)
    s.time_steps(args.steps)
    s.set_kernel_timing(False)
    q1 = s.pass_counters()
    projections, passes = q1[0] - q0[0], q1[1] - q0[1]
    sym_axes = s.fdm_layout()
    times = s.kernel_times()
    kernels = {}
    for fam, (kms, launches) in times.items():
        if launches == 0:
            continue
        # The FDM loop is enqueued fdm_max_passes times per projection; passes after convergence exit at
        # once. Their (small) time is charged to the passes that ran.
        if fam.startswith('xform_'):
            active = 2 * passes
        elif fam == 'residual':                # the residual after each pass (the guess residual runs with the divergence)
            active = passes
        else:
            active = launches
        per_step = active / args.steps
        avg_ms = kms / max(active, 1)
        b, f = algorithmic_work(fam, nx, ny, nz, per_step, sym_axes)
        entry = {'ms_per_step': kms / args.steps, 'launches_per_step': launches / args.steps,
                 'active_launches_per_step': per_step, 'avg_launch_ms': avg_ms}
        skippable = fam in ('gauge_reset', 'cfl_max', 'base_guess')
        if b is not None and skippable and avg_ms < b / (HBM_PEAK_GBS * 1e9) * 1e3:
            # faster than its bytes can move: the kernel exited (the previous step's last gradient did its work)
            entry['skipped'] = True
            b, f = None, None
        if b is not None:
            entry['algorithmic_bytes'] = b
            entry['achieved_GBs'] = b / (avg_ms / 1e3) / 1e9
            if f:
                entry['flops'] = f
                entry['achieved_TFLOPs'] = f / (avg_ms / 1e3) / 1e12
        kernels[fam] = entry
    # the whole step against its own bound: every kernel at max(bytes / HBM peak, flops / FP64 MFMA peak)
    t_lb = 0.0
    for e in kernels.values():
        if e.get('algorithmic_bytes') is None:
            continue
        per = max(e['algorithmic_bytes'] / (HBM_PEAK_GBS * 1e9), e.get('flops', 0) / (FP64_MFMA_PEAK_TFLOPS * 1e12))
        t_lb += per * e['active_launches_per_step']
    dom = max(kernels, key=lambda k: kernels[k]['ms_per_step'])
    d = kernels[dom]
    if d.get('flops') and d['flops'] / d['algorithmic_bytes'] > FP64_MFMA_PEAK_TFLOPS * 1e12 / (HBM_PEAK_GBS * 1e9):
        roof = {'kernel': dom, 'bound': 'mfma', 'achieved': d['achieved_TFLOPs'], 'peak': FP64_MFMA_PEAK_TFLOPS,
                'unit': 'TFLOP/s', 'frac': d['achieved_TFLOPs'] / FP64_MFMA_PEAK_TFLOPS, 'traffic': None}
    else:
        roof = {'kernel': dom, 'bound': 'hbm', 'achieved': d.get('achieved_GBs'), 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                'frac': (d.get('achieved_GBs') or 0) / HBM_PEAK_GBS, 'traffic': None}
    roof['algorithmic_bytes'] = d.get('algorithmic_bytes')
    s.close()
    value, ms = base['value'], base['ms_per_step'] * args.steps
    passes_mean = base['fdm_passes_mean']
    rep = {'phi_rel': base['phi_rel_last']}

    if args.traffic:
        kname = {'stage': 'kb_stage', 'residual': 'kb_residual'}.get(dom)
        if kname:
            traffic, why = measure_traffic(args, kname)
            roof['traffic'] = traffic['bytes'] if traffic else None
            roof['traffic_detail'] = traffic if traffic else why
    ensemble = run_ensemble(args) if args.ensemble_steps > 0 else None
    serving = run_serving(args) if args.serve_seconds > 0 else None

    cpu = None
    if not args.no_cpu_baseline:
        print('[bench] cpu baseline', file=sys.stderr, flush=True)
        cpu = cpu_baseline(settings, noise, args.cpu_steps, args.cpu_warmup)

    out = {
        'metric': METRIC, 'value': value, 'unit': 'Mcell-steps/s', 'n_gpus': 1, 'steps': args.steps,
        'warmup': args.warmup, 'ms_per_step': ms / args.steps, 'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': None, 'dtype': 'f64', 'data': 'synthetic',
        'config': {'workload': f'{nx}^3 lid-driven cavity, tanh-stretched grid (largest/smallest cell 2.0), Re 3200, '
                               f'regularised lid, adaptive dt CFL 0.5, FDM rtol 1e-10, plain FP64 step',
                   'nx': nx, 'ny': ny, 'nz': nz, 'precision': 'fp64', 'fdm_passes_mean': passes_mean,
                   'phi_rel_last': rep['phi_rel'], 'launch_choices': tuning,
                   'built_in_choices_ms_per_step': untuned_ms},
        'roofline': roof,
        'step_roofline': {'lower_bound_ms': t_lb * 1e3, 'ms_per_step': base['ms_per_step'],
                          'frac': t_lb * 1e3 / base['ms_per_step'],
                          'rule': 'sum over the step kernels of max(algorithmic bytes / 8 TB/s, useful flops / '
                                  '78.6 TFLOP/s) per launch, against the graph-replayed step'},
        'cpu_baseline': cpu,
        'kernels': kernels,
        'mixed_precision_fdm': {k: v for k, v in runs.items() if k != 'fp64'},
        'reference_precision_field_fp32': reference_precision,
        'ensemble': ensemble,
        'serving': serving,
    }
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
