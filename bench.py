#!/usr/bin/env python
"""Benchmark: FP64 classic 2D Euler shock-bubble step at 4096x1024 on one GPU.

A "step" is one unsplit classic time step (Roe solver with entropy fix, MC
limiter on all five waves, transverse correction waves, inflow / wall /
extrapolation boundaries) over the whole grid: one launch of the fused HIP
kernel.  Inputs are the shock-bubble initial data, resident in HBM before the
timed region; dt is fixed at the value giving Courant number 0.8 on the
initial data (the kernel still computes and max-combines the Courant number
of every step; it is reported).

The headline runs the FP64 kernels built with fused multiply-add (precision
CLAW_F64_FMA); the same line reports the reference-arithmetic build's speed
and the per-field difference between the two after one step and after the
whole run.  --precision f64 times the bit-exact build instead.

Prints one JSON line (see the keys below).  --gpus N>1 is not supported: this
build targets one GPU.
"""
import argparse
import ctypes
import json
import os
import subprocess
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL = 80          # read q (5 x f64) + write q (5 x f64)
# HBM bytes per launch of the Euler 4096x1024 FP64 kernels measured with
# rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE at the given launch shape, each scaled
# by the same counter's reading for a 512 MiB device copy (profiles/
# classic_unsplit_euler_4096x1024.md); reported only when the run uses that shape.
MEASURED_TRAFFIC = {('f64', 4096, 1024, 0, 2, 1): 398687027, ('f64fma', 4096, 1024, 0, 2, 1): 397183898,
                    ('f64', 4096, 1024, 18, 2, 1): 408753562, ('f64fma', 4096, 1024, 18, 2, 1): 407951360}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=300)
    ap.add_argument('--warmup', type=int, default=10)
    ap.add_argument('--mx', type=int, default=4096)
    ap.add_argument('--my', type=int, default=1024)
    ap.add_argument('--precision', default=None, choices=['f64', 'f64fma', 'f32'],
                    help="f64fma: FP64 with fused multiply-add and shared reciprocals (default of the euler "
                         "workload; within rounding of the reference arithmetic per step); f64: the reference "
                         "arithmetic bit for bit (default of the other workloads); f32")
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--no-autotune', action='store_true', help="keep the default row-band height")
    ap.add_argument('--launch', default=None,
                    help="ROWS,WAVES[,WPB]: set the Euler kernel's launch shape (row-band height, waves per "
                         "SIMD build, waves per block) instead of autotuning; the "
                         "4096x1024 FP64 lines default to 18,2,1 (as profiled)")
    ap.add_argument('--autotune', action='store_true', help="autotune the launch shape even where a default is pinned")
    ap.add_argument('--cpu-sample-steps', type=int, default=3)
    ap.add_argument('--workload', default='euler',
                    choices=['euler', 'batched', 'sharpclaw', 'sharpclaw_app', 'acoustics', 'acoustics_app', 'adjoint',
                             'copy'],
                    help="euler: the headline line; the others measure the other configurations")
    ap.add_argument('--instances', type=int, default=512)
    ap.add_argument('--adjoint-steps', type=int, default=2000)
    ap.add_argument('--headline-only', action='store_true',
                    help="euler workload: skip the rest of the metric (batched speedup, adjoint/forward, "
                         "SharpClaw, the run to t=0.2) in the line's extra")
    return ap.parse_args()


def cpu_baseline(q0, dx, dy, dt, mx, my, nsteps):
    """Single-thread C restatement (oracle/, the reference's step2 arithmetic)
    on the same grid for `nsteps` steps."""
    import oracle as O
    from pyclaw_amd.problems import shock_state
    O.set_params(O.RP_EULER, dict(gamma=1.4, gamma1=0.4))
    meth = O.method_array(2, 2, False)
    lim = O.mthlim_array(4, 5)
    qbc = np.zeros((5, mx + 4, my + 4), order='F')
    qbc[:, 2:-2, 2:-2] = q0
    rinf, vinf, einf = shock_state()
    t0 = time.perf_counter()
    for _ in range(nsteps):
        qbc[:, :2, :] = np.array([rinf, rinf * vinf, 0., einf, 0.])[:, None, None]
        O.fill_bc(qbc, 2, [99, 3], [1, 1], True)
        qn, _ = O.step2(O.RP_EULER, qbc, None, dx, dy, dt, meth, lim, 5)
        qbc = qn
    el = time.perf_counter() - t0
    return mx * my * nsteps / el / 1e9, el


def host_info():
    """The host the CPU baseline ran on: core counts, CPU model, the oracle's compiler and flags."""
    import oracle as O
    info = {"nproc": os.cpu_count(), "cores_usable": O.host_threads(cap=1 << 16)}
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    info["cpu_model"] = line.split(':', 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        cc = subprocess.run(['gcc', '--version'], stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True)
        info["compiler"] = cc.stdout.splitlines()[0] if cc.stdout else None
    except OSError:
        info["compiler"] = None
    mk = open(os.path.join(HERE, 'oracle', 'Makefile')).read()
    info["flags"] = [ln.split('=', 1)[1].strip() for ln in mk.splitlines() if ln.startswith('CFLAGS')][0]
    return info


def cpu_baseline_cores(q0, dx, dy, dt, mx, my, nsteps, threads):
    """The C restatement on `threads` host cores (OpenMP over row and column
    bands, bit-identical to one thread): Gcell-updates/s and seconds."""
    import oracle as O
    old = O.set_threads(threads)
    try:
        return cpu_baseline(q0, dx, dy, dt, mx, my, nsteps)
    finally:
        O.set_threads(old)


def measure_euler_run(args, q0, dx, dy, tfinal=0.2):
    """The shock-bubble at the headline size as an application: Controller.run
    to t = 0.2 through the drop-in API (ClawSolver2D, variable dt at CFL 0.8,
    cfl_max 1.0, the inflow as a static custom boundary, the step loop on the
    device), in the headline's arithmetic.  Wall time of the run over accepted
    steps: Gcell-updates/s of the whole application."""
    import torch
    import pyclaw_amd as pc
    from pyclaw_amd.problems import GAMMA, GAMMA1, euler_max_speed, shock_state
    from oracle.shockbubble_ref import shockbubble_inflow_numpy
    mx, my = q0.shape[1], q0.shape[2]
    dt0 = 0.8 * min(dx, dy) / euler_max_speed(q0)

    def run():
        x = pc.Dimension('x', 0.0, 2.0, mx)
        y = pc.Dimension('y', 0.0, 0.5, my)
        dom = pc.Domain([x, y])
        st = pc.State(dom, 5)
        st.problem_data.update(gamma=GAMMA, gamma1=GAMMA1)
        st.q[...] = q0
        s = pc.ClawSolver2D(pc.riemann.rp2_euler_5wave)
        s.dimensional_split = False
        s.transverse_waves = 2
        s.limiters = pc.limiters.tvd.MC
        s.cfl_max, s.cfl_desired = 1.0, 0.8
        s.dt_initial = dt0
        s.bc_lower = [pc.BC.custom, pc.BC.wall]
        s.bc_upper = [pc.BC.extrap, pc.BC.extrap]
        s.user_bc_lower = shockbubble_inflow_numpy
        s.custom_bc_static = True
        s.fma = args.precision == 'f64fma'
        claw = pc.Controller()
        claw.solution = pc.Solution(st, dom)
        claw.solver = s
        claw.output_format = None
        claw.keep_copy = True
        claw.tfinal = tfinal
        claw.num_output_times = 1
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        claw.run()
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        qf = claw.frames[-1].state.q
        return el, s.status, bool(np.isfinite(qf).all()), claw.solution.t
    run()                                     # warm-up (kernel loads, buffers)
    el, status, finite, tf = run()
    n = status['numsteps']
    return {"gcell_updates_per_s": mx * my * n / el / 1e9, "wall_s": el, "accepted_steps": n,
            "ms_per_step": 1e3 * el / n, "t_final": tf, "cflmax": status['cflmax'], "finite": finite,
            "arithmetic": args.precision,
            "how": "Controller.run to t=%g, one output interval, step loop on the device (ClawSolver2D, "
                   "device_loop default), wall clock including host waits" % tfinal}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: multi-GPU runs (domain decomposition with halo "
              "exchange, or sharded batches) are not part of this implementation", file=sys.stderr)
        sys.exit(2)
    if args.precision is None:
        args.precision = 'f64fma' if args.workload == 'euler' else 'f64'
    if args.workload != 'euler':
        return OTHER[args.workload](args)
    import torch
    from pyclaw_amd import _lib
    from pyclaw_amd.problems import euler_max_speed, shock_state, shockbubble

    mx, my = args.mx, args.my
    q0, dx, dy = shockbubble(mx, my)
    smax = euler_max_speed(q0)
    dt = 0.8 * min(dx, dy) / smax
    dev = torch.device('cuda', 0)
    dtype = torch.float32 if args.precision == 'f32' else torch.float64
    ng = 3
    nxp, nyp = mx + 2 * ng, my + 2 * ng
    buf = torch.zeros((5, nyp, nxp), dtype=torch.float64)
    buf[:, ng:ng + my, ng:ng + mx] = torch.from_numpy(np.ascontiguousarray(q0.transpose(0, 2, 1)))
    rinf, vinf, einf = shock_state()
    buf[:, :, :ng] = torch.tensor([rinf, rinf * vinf, 0., einf, 0.], dtype=torch.float64)[:, None, None]
    qa = buf.to(dev, dtype)
    qb = qa.clone()           # ghost cells (the constant inflow state) present in both buffers

    grid = _lib.ClawGrid(mx, my, ng, 0, nxp, nxp * nyp, dx, dy)
    m = _lib.ClawMethod()
    m.order, m.transverse, m.split, m.fwave, m.num_waves = 2, 2, 0, 0, 5
    for k in range(5):
        m.mthlim[k] = 4
    r = _lib.ClawRiemann()
    r.kind, r.efix = _lib.RP_EULER, 1
    r.p[0], r.p[1] = 1.4, 0.4
    bc = _lib.ClawBC()
    bc.lower[0], bc.lower[1] = _lib.BC_STORED, _lib.BC_WALL     # inflow (stored constant state), wall
    bc.upper[0], bc.upper[1] = _lib.BC_EXTRAP, _lib.BC_EXTRAP
    for s in range(2):
        bc.aux_lower[s] = bc.aux_upper[s] = _lib.BC_EXTRAP
    ctx = _lib.Context(0)
    L = _lib.lib()
    dtd = torch.tensor(dt, dtype=torch.float64, device=dev)
    cfl = torch.zeros((), dtype=torch.float64, device=dev)
    prec = {'f64': _lib.F64, 'f64fma': _lib.F64_FMA, 'f32': _lib.F32}[args.precision]
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    bufs = [qa, qb]

    def step(k):
        rc = L.claw2d_classic_step_ex(ctx.handle, grid, m, r, bc, prec, bufs[k % 2].data_ptr(),
                                      bufs[(k + 1) % 2].data_ptr(), None, dtd.data_ptr(), cfl.data_ptr(),
                                      _lib.STEP_ACCUMULATE_CFL, sh)
        _lib.check(rc, 'claw2d_classic_step_ex')

    band = None
    if args.launch is None and not args.autotune and args.precision != 'f32' and (mx, my) == (4096, 1024):
        # the paired fill decomposition (one-wavefront blocks of the two-per-SIMD build, one per
        # wavefront slot, each SIMD's older wavefront taking two units of rows for the younger
        # one's one): the library's default for this grid, pinned so the line does not depend on
        # the tuner, which ranks shapes on the initial data, where the per-step cost differs
        # (--launch 18,2,1: the bands in about two rounds of wavefronts before it)
        args.launch = '0,2,1'
    if args.launch:
        shape = [int(v) for v in args.launch.split(',')]
        rows_, waves_, wpb_ = shape[0], shape[1], (shape[2] if len(shape) > 2 else 4)
        for pr_ in {prec, _lib.F64}:
            _lib.check(L.claw2d_classic_set_launch_ex(ctx.handle, grid, m, r, pr_, rows_, waves_, wpb_),
                       'claw2d_classic_set_launch_ex')
    elif not args.no_autotune:
        # pick the kernel's launch shape for this grid (answers do not depend on it beyond the last
        # bits of the fused build)
        hb = ctypes.c_int(0)
        _lib.check(L.claw2d_classic_autotune(ctx.handle, grid, m, r, bc, prec, qa.data_ptr(), qb.data_ptr(), None,
                                             dtd.data_ptr(), cfl.data_ptr(), sh, ctypes.byref(hb)),
                   'claw2d_classic_autotune')
    sb, sw, sp = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
    _lib.check(L.claw2d_classic_get_launch(ctx.handle, grid, m, r, prec, ctypes.byref(sb), ctypes.byref(sw),
                                           ctypes.byref(sp)), 'claw2d_classic_get_launch')
    band = {"rows": sb.value, "waves_per_simd": sw.value, "waves_per_block": sp.value}
    for k in range(args.warmup):
        step(k)
    torch.cuda.synchronize()
    cfl.zero_()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    for k in range(args.steps):
        step(args.warmup + k)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = e0.elapsed_time(e1)
    ms_per_step = gpu_ms / args.steps
    cells = mx * my
    value = cells * args.steps / (gpu_ms * 1e-3) / 1e9
    bytes_per_cell = BYTES_PER_CELL // 2 if args.precision == 'f32' else BYTES_PER_CELL
    achieved = bytes_per_cell * cells / (ms_per_step * 1e-3) / 1e9
    cfl_max = float(cfl.item())
    qfin = bufs[(args.warmup + args.steps) % 2]
    finite = bool(torch.isfinite(qfin).all().item())

    f32_err = None
    if args.precision == 'f32':
        # error of the FP32 variant against the FP64 kernels over the same steps (untimed)
        ref = [buf.to(dev, torch.float64), None]
        ref[1] = ref[0].clone()
        for k in range(args.warmup + args.steps):
            _lib.check(L.claw2d_classic_step_ex(ctx.handle, grid, m, r, bc, _lib.F64, ref[k % 2].data_ptr(),
                                                ref[(k + 1) % 2].data_ptr(), None, dtd.data_ptr(), cfl.data_ptr(),
                                                _lib.STEP_ACCUMULATE_CFL, sh), 'f64 reference step')
        r64 = ref[(args.warmup + args.steps) % 2][:, ng:ng + my, ng:ng + mx]
        r32 = qfin[:, ng:ng + my, ng:ng + mx].to(torch.float64)
        # per field, relative to the field's scale (both momenta share the momentum scale)
        scale = [r64[0].abs().max(), r64[1:3].abs().max(), r64[1:3].abs().max(), r64[3].abs().max(),
                 max(float(r64[4].abs().max()), 1e-300)]
        f32_err = [float(((r32[c] - r64[c]).abs().max() / scale[c]).item()) for c in range(5)]

    fma_err = None
    if args.precision == 'f64fma':
        # one step from the final state through both FP64 builds (untimed), and
        # the whole run repeated with the reference-arithmetic build
        fin = qfin.clone()
        te, tf = torch.zeros_like(fin), torch.zeros_like(fin)
        for out_, pr_ in ((te, _lib.F64), (tf, _lib.F64_FMA)):
            out_.copy_(fin)
            _lib.check(L.claw2d_classic_step_ex(ctx.handle, grid, m, r, bc, pr_, fin.data_ptr(), out_.data_ptr(),
                                                None, dtd.data_ptr(), cfl.data_ptr(), _lib.STEP_ACCUMULATE_CFL, sh),
                       'comparison step')
        # the same run through the reference-arithmetic build, its timed steps
        # bracketed like the headline's
        ref = [buf.to(dev, torch.float64), None]
        ref[1] = ref[0].clone()
        f0 = torch.cuda.Event(enable_timing=True)
        f1 = torch.cuda.Event(enable_timing=True)
        for k in range(args.warmup + args.steps):
            if k == args.warmup:
                torch.cuda.synchronize()
                f0.record(stream)
            _lib.check(L.claw2d_classic_step_ex(ctx.handle, grid, m, r, bc, _lib.F64, ref[k % 2].data_ptr(),
                                                ref[(k + 1) % 2].data_ptr(), None, dtd.data_ptr(), cfl.data_ptr(),
                                                _lib.STEP_ACCUMULATE_CFL, sh), 'f64 reference step')
        f1.record(stream)
        torch.cuda.synchronize()
        ref_ms = f0.elapsed_time(f1) / args.steps
        rfin = ref[(args.warmup + args.steps) % 2]
        sl = (slice(None), slice(ng, ng + my), slice(ng, ng + mx))

        def per_field(a, b):
            # per field, relative to the field's scale (both momenta share the momentum scale)
            a, b = a[sl], b[sl]
            scale = [b[0].abs().max(), b[1:3].abs().max(), b[1:3].abs().max(), b[3].abs().max(), b[4].abs().max()]
            return [float(((a[c] - b[c]).abs().max() / scale[c]).item()) for c in range(5)]
        fma_err = {"reference_arithmetic_kernel_gcell_s": cells / (ref_ms * 1e-3) / 1e9,
                   "reference_arithmetic_kernel_ms_per_step": ref_ms,
                   "one_step_rel_linf_per_field": per_field(tf, te),
                   "after_%d_steps_rel_linf_per_field" % (args.warmup + args.steps): per_field(qfin, rfin)}

    cpu = None
    if not args.no_cpu_baseline:
        import oracle as O
        nt = O.host_threads()
        v1, el1 = cpu_baseline(q0, dx, dy, dt, mx, my, args.cpu_sample_steps)
        # all cores: the same restatement on OpenMP row/column bands, a sample
        # of about 10 s of that work
        nall = max(3, min(40, int(10.0 / max(el1 / args.cpu_sample_steps / nt, 1e-3))))
        vn, eln = cpu_baseline_cores(q0, dx, dy, dt, mx, my, nall, nt)
        cpu = {"value": vn, "unit": "Gcell-updates/s", "cores": nt, "kind": "port",
               "sample": "%d steps of the same 4096x1024 FP64 unsplit step (%.1f s) on %d host threads: "
                         "the C restatement of step2/flux2/limiter + Euler Roe solver (oracle/claw_oracle.c) "
                         "with OpenMP over row bands then column bands, each band recomputing one slice on "
                         "either side (bit-identical to one thread); it stands in for the reference's "
                         "f2py Fortran + petclaw MPI path, which cannot run here, and the reference Euler "
                         "solver cannot run this grid (maxm2=1800)" % (nall, eln, nt),
               "single_core": {"value": v1, "steps": args.cpu_sample_steps, "seconds": el1},
               "host": host_info()}

    out = {
        "metric": "Gcell-updates/s",
        "value": value,
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": 'f32' if args.precision == 'f32' else 'f64',
        "data": "synthetic (shock-bubble initial data, generated in-process)",
        "config": {"workload": "classic 2D Euler shock-bubble %dx%d, unsplit, Roe+efix, MC on 5 waves, "
                               "transverse correction waves, fixed dt at CFL 0.8%s" % (
                                   mx, my, ", FP64 with fused multiply-add" if args.precision == 'f64fma' else ""),
                   "grid": [mx, my], "global_batch": 1, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": MEASURED_TRAFFIC.get((args.precision, mx, my, band["rows"], band["waves_per_simd"],
                                                      band["waves_per_block"])),
                     "bytes_per_cell_update": bytes_per_cell,
                     "kernel": "%sclassic_unsplit_kernel%s<Euler5<%s>>" % (
                         'fused::' if args.precision == 'f64fma' else '',
                         '_w2' if band["waves_per_simd"] == 2 else '',
                         'float' if args.precision == 'f32' else 'double')},
        "cpu_baseline": cpu,
        "extra": {"cfl_max_over_timed_steps": cfl_max, "finite": finite, "dt": dt, "wall_s": wall,
                  "hbm_pct": 100.0 * achieved / HBM_PEAK_GBS, "launch_shape": band,
                  "f32_rel_linf_vs_f64_per_field": f32_err, "fma_vs_reference_arithmetic": fma_err},
    }
    if not args.headline_only and (mx, my) == (4096, 1024) and args.precision != 'f32':
        # the rest of the metric, each part on its own configuration
        del qa, qb, bufs
        torch.cuda.empty_cache()
        rest = {}
        for key, fn in (("euler_4096x1024_run_to_t0.2", lambda: measure_euler_run(args, q0, dx, dy)),
                        ("batched_vs_plain_speedup", lambda: measure_batched(args, nsteps=200)),
                        ("adjoint_over_forward", lambda: measure_adjoint(args)),
                        # launch='': the library's default shape for this grid, not the Euler line's
                        ("sharpclaw_shallow_1024", lambda: measure_sharpclaw(args, steps=500, warmup=5,
                                                                             precision='f64', launch=''))):
            t0 = time.perf_counter()
            try:
                r_ = fn()
            except Exception as e:      # recorded in the line, the headline still printed
                r_ = {"error": "%s: %s" % (type(e).__name__, e)}
            r_["bench_seconds"] = time.perf_counter() - t0
            rest[key] = r_
            torch.cuda.empty_cache()
        out["extra"]["rest_of_metric"] = rest
    print(json.dumps(out), flush=True)


def measure_batched(args, nsteps=None):
    """512 independent layered vc-acoustics grids (sizes from {64,...,512}^2),
    nsteps steps each (--steps, 200 in the default line) with its own CFL-0.9
    dt: one hipGraph-captured sequence of persistent work-queue launches vs the
    same grids stepped one at a time through the plain single-grid entry point.
    Both legs have their buffers uploaded before their timer starts; each times
    only the stepping (the batched leg one graph replay, the other leg the
    launches of every instance's steps)."""
    import torch
    import pyclaw_amd as pc
    from pyclaw_amd.batch import BatchedClawSolver2D, OneByOne
    from pyclaw_amd.problems import batched_sizes, layered_vc_acoustics
    nsteps = args.steps if nsteps is None else nsteps
    rng = np.random.default_rng(2024)
    sizes = batched_sizes(rng, args.instances)
    inst = []
    for (mx, my) in sizes:
        q, aux = layered_vc_acoustics(rng, mx, my)
        dx, dy = 1.0 / mx, 1.0 / my
        inst.append(dict(q=q, aux=aux, dx=dx, dy=dy, dt_initial=0.9 * min(dx, dy) / aux[1].max()))
    cells = sum(mx * my for mx, my in sizes)
    rp = pc.riemann.rp2_vc_acoustics

    b = BatchedClawSolver2D(rp, {}, inst)
    shape = None
    if not args.no_autotune:
        shape = b.autotune()               # tile rows x wavefronts per block (results do not depend on it)
    b.run(nsteps, use_graph=True)          # graph capture + first replay (warm-up)
    torch.cuda.synchronize()
    for k in range(b.n):                   # back to the initial data, then the timed replay
        b._put(k, inst[k]['q'], None)
    b.dt.copy_(torch.tensor([i['dt_initial'] for i in inst], dtype=torch.float64))
    b.t.zero_(); b.parity.zero_(); b.naccepted.zero_(); b.nrejected.zero_(); b.cflmax.zero_()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.run(nsteps, use_graph=True)
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    acc = b.naccepted.cpu().numpy()
    warm = OneByOne(rp, {}, inst[:8])       # warm-up of the plain path, and the equality check
    warm.run(nsteps)
    ref = warm.results()
    same = all(np.array_equal(b.q(k), ref[k]) for k in range(8))
    del warm
    ob = OneByOne(rp, {}, inst)             # uploads (untimed)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ob.run(nsteps)
    torch.cuda.synchronize()
    t_loop = time.perf_counter() - t0
    ob.results()                            # no step rejected (the ping-pong assumption)
    del ob
    gcell = cells * nsteps / t_batch / 1e9
    achieved = 64.0 * cells * nsteps / t_batch / 1e9
    return {"metric": "batched vs plain speedup", "value": t_loop / t_batch, "unit": "x", "n_gpus": 1,
            "steps": nsteps, "warmup": nsteps, "ms_per_step": 1e3 * t_batch / nsteps, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded layered media)",
            "config": {"workload": "%d vc-acoustics instances, sizes {64..512}^2, %d steps each, classic unsplit "
                                   "MC trans_inc, CFL 0.9 per instance" % (args.instances, nsteps),
                       "global_batch": args.instances, "parallelism": "single GPU"},
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "traffic": None, "bytes_per_cell_update": 64,
                         "kernel": "batch_step_kernel (whole step, update kernel included)"},
            "extra": {"batched_s": t_batch, "one_by_one_s": t_loop, "cells_per_step": cells,
                      "batched_gcell_updates_per_s": gcell,
                      "one_by_one_gcell_updates_per_s": cells * nsteps / t_loop / 1e9,
                      "timed": "stepping only in both legs (buffers uploaded beforehand)",
                      "bitwise_equal_first8": same, "all_accepted": bool((acc == nsteps).all()),
                      "work_items_per_step": b.work_items,
                      "tile_shape": None if shape is None else {"rows": shape[0], "waves_per_block": shape[1]}}}


def bench_batched(args):
    print(json.dumps(measure_batched(args)), flush=True)


def _grid_bufs(q0, ng, dtype, dev):
    import torch
    ne, mx, my = q0.shape
    buf = torch.zeros((ne, my + 2 * ng, mx + 2 * ng), dtype=torch.float64)
    buf[:, ng:ng + my, ng:ng + mx] = torch.from_numpy(np.ascontiguousarray(q0.transpose(0, 2, 1)))
    return buf.to(dev, dtype)


def measure_sharpclaw(args, steps=None, warmup=None, precision=None, launch=None):
    """SharpClaw shallow water 1024x1024 periodic, WENO5 + SSP104, FP64: one
    step = 10 fused stage launches (one C call).  The reference-arithmetic
    build by default (bit-exact); --precision f64fma times the fused build,
    with the reference build's speed and the difference between the two
    reported beside it (on this smooth data WENO5's weights amplify rounding
    differences: one-ulp input noise moves a step by ~1e-9)."""
    import torch
    from pyclaw_amd import _lib
    from pyclaw_amd.problems import shallow_hump
    steps = args.steps if steps is None else steps
    warmup = args.warmup if warmup is None else warmup
    precision = args.precision if precision is None else precision
    launch = args.launch if launch is None else launch
    mx = my = 1024
    rng = np.random.default_rng(7)
    x0 = 0.25 + 0.5 * rng.random(2)
    q0 = shallow_hump(mx, my, x0)
    dx = dy = 1.0 / mx
    smax = float((np.sqrt(q0[0])).max())
    dt = 0.5 * dx / smax
    dev = torch.device('cuda', 0)
    ng = 3
    q = _grid_bufs(q0, ng, torch.float64, dev)
    qn, s1a, s1b, s2 = (torch.zeros_like(q) for _ in range(4))
    g = _lib.ClawGrid(mx, my, ng, 0, mx + 2 * ng, (mx + 2 * ng) * (my + 2 * ng), dx, dy)
    r = _lib.ClawRiemann(); r.kind = _lib.RP_SHALLOW; r.efix = 1; r.p[0] = 1.0
    bc = _lib.ClawBC()
    for s_ in range(2):
        bc.lower[s_] = bc.upper[s_] = _lib.BC_PERIODIC
        bc.aux_lower[s_] = bc.aux_upper[s_] = _lib.BC_PERIODIC
    L = _lib.lib()
    dtd = torch.tensor(dt, dtype=torch.float64, device=dev)
    cfl = torch.zeros(10, dtype=torch.float64, device=dev)
    sh = torch.cuda.current_stream().cuda_stream
    pair = [q, qn]

    ctx = _lib.Context(0)
    if precision == 'f32':
        raise SystemExit("the sharpclaw workload runs FP64 (--precision f64fma or f64)")
    prec = _lib.F64_FMA if precision == 'f64fma' else _lib.F64
    band = None
    # without --launch or --autotune: the library's default shape for this grid (the fill
    # decomposition, one wavefront per slot of the default build)
    if launch:
        shape = [int(v) for v in launch.split(',')]
        for pr_ in {prec, _lib.F64}:
            _lib.check(L.claw2d_sharpclaw_set_launch_ex(ctx.handle, g, r, pr_, shape[0], shape[1],
                                                        shape[2] if len(shape) > 2 else 4),
                       'claw2d_sharpclaw_set_launch_ex')
    elif args.autotune:
        for pr_ in ([prec, _lib.F64] if prec != _lib.F64 else [prec]):
            hb = ctypes.c_int(0)
            _lib.check(L.claw2d_sharpclaw_autotune(ctx.handle, g, r, bc, pr_, q.data_ptr(), qn.data_ptr(), None,
                                                   dtd.data_ptr(), cfl.data_ptr(), sh, ctypes.byref(hb)),
                       'claw2d_sharpclaw_autotune')

    sb, sw, sp = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
    _lib.check(L.claw2d_sharpclaw_get_launch(ctx.handle, g, r, prec, ctypes.byref(sb), ctypes.byref(sw),
                                             ctypes.byref(sp)), 'claw2d_sharpclaw_get_launch')
    band = {"rows": sb.value, "waves_per_simd": sw.value, "waves_per_block": sp.value}

    def step(k, pr_=prec, bufs=pair):
        rc = L.claw2d_sharpclaw_ssp104_step(ctx.handle, g, r, bc, pr_, bufs[k % 2].data_ptr(),
                                            bufs[(k + 1) % 2].data_ptr(), s1a.data_ptr(), s1b.data_ptr(),
                                            s2.data_ptr(), None, dtd.data_ptr(), cfl.data_ptr(), sh)
        _lib.check(rc, 'ssp104')
    for k in range(warmup):
        step(k)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True); e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for k in range(steps):
        step(warmup + k)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    fin = pair[(warmup + steps) % 2][:, ng:ng + my, ng:ng + mx].to('cpu').numpy()
    import math
    tot0 = np.array([math.fsum(q0[m].ravel()) for m in range(3)])
    tot1 = np.array([math.fsum(fin[m].ravel()) for m in range(3)])
    h_rel = abs(tot1[0] - tot0[0]) / tot0[0]
    mom_rel = np.abs(tot1[1:] - tot0[1:]) / tot0[0]
    cmp = None
    if prec == _lib.F64_FMA:
        # the same run through the reference-arithmetic build (timed the same
        # way), then one step from the fused build's final state through both
        sl = (slice(None), slice(ng, ng + my), slice(ng, ng + mx))

        def per_field(a, b):
            a, b = a[sl], b[sl]
            scale = [b[0].abs().max(), b[1:3].abs().max(), b[1:3].abs().max()]
            return [float(((a[c] - b[c]).abs().max() / scale[c]).item()) for c in range(3)]
        ref = [_grid_bufs(q0, ng, torch.float64, dev), None]
        ref[1] = torch.zeros_like(ref[0])
        f0 = torch.cuda.Event(enable_timing=True); f1 = torch.cuda.Event(enable_timing=True)
        for k in range(warmup + steps):
            if k == warmup:
                torch.cuda.synchronize()
                f0.record()
            step(k, _lib.F64, ref)
        f1.record()
        torch.cuda.synchronize()
        ref_ms = f0.elapsed_time(f1) / steps
        qf = pair[(warmup + steps) % 2]
        after = per_field(qf, ref[(warmup + steps) % 2])
        te, tf = [qf.clone(), torch.zeros_like(qf)], [qf.clone(), torch.zeros_like(qf)]
        step(0, _lib.F64, te)
        step(0, _lib.F64_FMA, tf)
        cmp = {"reference_arithmetic_gcell_s": mx * my / (ref_ms * 1e-3) / 1e9,
               "reference_arithmetic_ms_per_step": ref_ms,
               "one_step_rel_linf_per_field": per_field(tf[1], te[1]),
               "after_%d_steps_rel_linf_per_field" % (warmup + steps): after}
    bytes_step = 552
    achieved = bytes_step * mx * my / (ms * 1e-3) / 1e9
    out = {"metric": "Gcell-updates/s", "value": mx * my / (ms * 1e-3) / 1e9, "unit": "Gcell-updates/s",
           "n_gpus": 1, "steps": steps, "warmup": warmup, "ms_per_step": ms, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded Gaussian hump)",
           "config": {"workload": "SharpClaw shallow water 1024x1024 periodic, WENO5 comp-wise + SSP104 (10 fused "
                                  "stage launches per step), CFL 0.5%s" % (
                                      ", FP64 with fused multiply-add" if prec == _lib.F64_FMA else ""),
                      "global_batch": 1, "parallelism": "single GPU"},
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS, "traffic": None, "bytes_per_cell_update": bytes_step},
           "extra": {"mass_rel_change": float(h_rel), "momentum_rel_change_over_mass": mom_rel.tolist(),
                     "totals": "compensated sums (math.fsum); momenta relative to total h",
                     "cfl_last_stage": float(cfl[9].item()), "band_rows": band,
                     "fma_vs_reference_arithmetic": cmp}}
    return out


def bench_sharpclaw(args):
    print(json.dumps(measure_sharpclaw(args)), flush=True)


def bench_acoustics(args):
    """Config-1 homogeneous acoustics 200x200 (cache resident): microseconds per
    unsplit step, trans_cor, MC, walls low / extrapolation high."""
    import torch
    from pyclaw_amd import _lib
    from pyclaw_amd.problems import acoustics_pulse
    mx = my = 200
    rng = np.random.default_rng(1)
    c0 = tuple(-0.1 + 0.2 * rng.random(2))
    q0 = acoustics_pulse(mx, my, center=c0)
    dx = dy = 2.0 / mx
    dt = 0.9 * dx / 2.0
    dev = torch.device('cuda', 0)
    ng = 3
    qa = _grid_bufs(q0, ng, torch.float64, dev); qb = qa.clone()
    g = _lib.ClawGrid(mx, my, ng, 0, mx + 2 * ng, (mx + 2 * ng) * (my + 2 * ng), dx, dy)
    m = _lib.ClawMethod(); m.order, m.transverse, m.split, m.fwave, m.num_waves = 2, 2, 0, 0, 2
    m.mthlim[0] = m.mthlim[1] = 4
    r = _lib.ClawRiemann(); r.kind = _lib.RP_ACOUSTICS; r.p[0], r.p[1], r.p[2], r.p[3] = 1.0, 4.0, 2.0, 2.0
    bc = _lib.ClawBC(); bc.lower[0] = bc.lower[1] = _lib.BC_WALL; bc.upper[0] = bc.upper[1] = _lib.BC_EXTRAP
    L = _lib.lib(); ctx = _lib.Context(0)
    dtd = torch.tensor(dt, dtype=torch.float64, device=dev); cfl = torch.zeros((), dtype=torch.float64, device=dev)
    sh = torch.cuda.current_stream().cuda_stream
    pair = [qa, qb]

    def step(k):
        _lib.check(L.claw2d_classic_step_ex(ctx.handle, g, m, r, bc, _lib.F64, pair[k % 2].data_ptr(),
                                            pair[(k + 1) % 2].data_ptr(), None, dtd.data_ptr(), cfl.data_ptr(),
                                            _lib.STEP_ACCUMULATE_CFL, sh), 'step')
    for k in range(args.warmup):
        step(k)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True); e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for k in range(args.steps):
        step(k)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    out = {"metric": "us per step", "value": ms * 1e3, "unit": "us/step", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": False, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "homogeneous acoustics 200x200, unsplit MC trans_cor (cache resident)",
                      "global_batch": 1, "parallelism": "single GPU"},
           "extra": {"gcell_updates_per_s": mx * my / (ms * 1e-3) / 1e9}}
    print(json.dumps(out), flush=True)


def bench_acoustics_app(args):
    """Config-1 homogeneous acoustics 200x200 as an application: Controller.run
    to t=0.6 with 10 output frames kept in memory, through the drop-in API
    (evolve_to_time's step loop, first dt 0.1 rejected as in the reference's
    defaults).  Wall-clock milliseconds per accepted step with the step loop's
    bookkeeping on the device (the default) and on the host (one host
    synchronisation per step).  --steps: the number of timed runs."""
    import time
    import torch
    import pyclaw_amd as pc
    from pyclaw_amd.problems import acoustics_pulse
    mx = my = 200
    rng = np.random.default_rng(1)
    c0 = tuple(-0.1 + 0.2 * rng.random(2))
    q0 = acoustics_pulse(mx, my, center=c0)

    def run(device_loop):
        x = pc.Dimension('x', -1., 1., mx)
        y = pc.Dimension('y', -1., 1., my)
        dom = pc.Domain([x, y])
        st = pc.State(dom, 3)
        st.problem_data.update(rho=1.0, bulk=4.0, cc=2.0, zz=2.0)
        st.q[...] = q0
        s = pc.ClawSolver2D(pc.riemann.rp2_acoustics)
        s.dimensional_split = False
        s.transverse_waves = 2
        s.limiters = pc.limiters.tvd.MC
        s.cfl_max, s.cfl_desired = 1.0, 0.9
        s.bc_lower = [pc.BC.wall, pc.BC.wall]
        s.bc_upper = [pc.BC.extrap, pc.BC.extrap]
        s.device_loop = device_loop
        claw = pc.Controller()
        claw.solution = pc.Solution(st, dom)
        claw.solver = s
        claw.output_format = None
        claw.keep_copy = True
        claw.tfinal = 0.6
        claw.num_output_times = 10
        nsteps = [0]
        evolve = s.evolve_to_time

        def counted(*a, **k):
            r = evolve(*a, **k)
            nsteps[0] += r['numsteps']
            return r
        s.evolve_to_time = counted
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        claw.run()
        torch.cuda.synchronize()
        return time.perf_counter() - t0, nsteps[0], claw.frames[-1].state.q

    reps = max(1, min(args.steps, 20))
    for w in range(max(1, min(args.warmup, 3))):
        run(True), run(False)
    res = {}
    for mode in (True, False):
        best = None
        for _ in range(reps):
            el, n, qf = run(mode)
            best = el if best is None else min(best, el)
        res[mode] = (best, n, qf)
    same = bool(np.array_equal(res[True][2], res[False][2]))
    ms = res[True][0] / res[True][1] * 1e3
    cpu = None
    if not args.no_cpu_baseline:
        # the same application on the host cores: the oracle's evolve_to_time
        # loop around its single-thread C restatement of step2/flux2
        import time as _time
        import oracle as O
        from oracle import driver
        P = driver.Problem(O.RP_ACOUSTICS, 2, dict(rho=1.0, bulk=4.0, cc=2.0, zz=2.0), 2.0 / mx, 2.0 / my,
                           (3, 3), (1, 1), kind='classic', dimensional_split=False, transverse=2, limiters=4,
                           cfl_max=1.0, cfl_desired=0.9, dt_initial=0.1)
        run = driver.Run(P, q0)
        t0 = _time.perf_counter()
        frames = run.run(0.6, 10)
        el = _time.perf_counter() - t0
        cpu = {"value": el / run.nsteps * 1e3, "unit": "ms/step", "cores": 1, "kind": "port",
               "sample": "the whole application (%d accepted steps, %.1f s): the oracle's evolve_to_time loop "
                         "around its C step (oracle/claw_oracle.c)" % (run.nsteps, el),
               "final_frame_equal_to_gpu": bool(np.array_equal(frames[-1], res[True][2]))}
    out = {"metric": "ms per accepted step (application loop)", "value": ms, "unit": "ms/step", "n_gpus": 1,
           "steps": res[True][1], "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": False,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "homogeneous acoustics 200x200 application, Controller.run to t=0.6, 10 frames",
                      "global_batch": 1, "parallelism": "single GPU"},
           "extra": {"host_loop_ms_per_step": res[False][0] / res[False][1] * 1e3,
                     "device_loop_speedup": res[False][0] / res[True][0], "frames_identical": same,
                     "runs_timed": reps, "best_of": "min wall time"}}
    if cpu is not None:
        out["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)


def bench_sharpclaw_app(args):
    """SharpClaw shallow water 1024x1024 periodic (WENO5 + SSP104, Roe, g=1)
    as an application: one Controller.run output interval of about 500 steps at
    Courant number 0.5 through the drop-in API, FP64.  Milliseconds per
    accepted step with the step loop on the device (the default) and on the
    host, and the relative change of total h, hu, hv."""
    import time
    import torch
    import pyclaw_amd as pc
    from pyclaw_amd.problems import shallow_hump
    mx = my = 1024
    rng = np.random.default_rng(7)
    x0 = 0.25 + 0.5 * rng.random(2)
    q0 = shallow_hump(mx, my, x0)
    smax = float(np.max(np.sqrt(q0[0])))
    tfinal = 500 * 0.5 * (1.0 / mx) / smax

    def run(device_loop):
        x = pc.Dimension('x', 0., 1., mx)
        y = pc.Dimension('y', 0., 1., my)
        dom = pc.Domain([x, y])
        st = pc.State(dom, 3)
        st.problem_data['grav'] = 1.0
        st.q[...] = q0
        s = pc.SharpClawSolver2D(pc.riemann.rp2_shallow_roe_with_efix)
        s.cfl_max, s.cfl_desired = 0.6, 0.5
        s.dt_initial = 0.5 * (1.0 / mx) / smax
        s.bc_lower = [pc.BC.periodic, pc.BC.periodic]
        s.bc_upper = [pc.BC.periodic, pc.BC.periodic]
        s.device_loop = device_loop
        claw = pc.Controller()
        claw.solution = pc.Solution(st, dom)
        claw.solver = s
        claw.output_format = None
        claw.keep_copy = True
        claw.tfinal = tfinal
        claw.num_output_times = 1
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        claw.run()
        torch.cuda.synchronize()
        return time.perf_counter() - t0, s.status['numsteps'], claw.frames[-1].state.q

    run(True), run(False)
    res = {mode: run(mode) for mode in (True, False)}
    el, n, qf = res[True]
    tot0, tot1 = q0.reshape(3, -1).sum(axis=1), qf.reshape(3, -1).sum(axis=1)
    scale = np.array([tot0[0], tot0[0], tot0[0]])
    ms = el / n * 1e3
    out = {"metric": "ms per accepted step (application loop)", "value": ms, "unit": "ms/step", "n_gpus": 1,
           "steps": n, "warmup": 1, "ms_per_step": ms, "higher_is_better": False, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "SharpClaw shallow water 1024x1024 periodic, WENO5 + SSP104, Controller.run "
                                  "of about 500 steps at CFL 0.5", "global_batch": 1, "parallelism": "single GPU"},
           "extra": {"gcell_updates_per_s": mx * my / (ms * 1e-3) / 1e9,
                     "host_loop_ms_per_step": res[False][0] / res[False][1] * 1e3,
                     "device_loop_speedup": res[False][0] / el,
                     "frames_identical": bool(np.array_equal(qf, res[False][2])),
                     "relative_change_of_totals_h_hu_hv": (np.abs(tot1 - tot0) / scale).tolist()}}
    print(json.dumps(out), flush=True)


def measure_adjoint(args):
    """Gradient of the receiver misfit of 2D vc acoustics at 512x512 (2000 steps
    at dt = 0.8 dx / max c, MC limiter, trans_inc, extrapolation boundaries;
    K in [1,4], rho in [1,2] smooth random fields; 64 receivers on y = 0.95;
    observed traces from K scaled by 1.05 in a centred disc) with respect to
    K, rho and the initial state, checkpointing every 45 steps.  Reports the
    gradient time (forward with checkpoints + reverse sweep) over the time of
    the plain forward run, and the adjoint step kernels against the forward
    step kernel."""
    import torch
    from pyclaw_amd.adjoint import MisfitProblem, receivers_on_row
    from pyclaw_amd.problems import adjoint_medium
    mx = my = 512
    N = args.adjoint_steps
    rng = np.random.default_rng(2025)
    q0, K, rho, K_true = adjoint_medium(rng, mx, my)
    dx = dy = 1.0 / mx
    dt = 0.8 * dx / float(np.sqrt(K / rho).max())
    rec = receivers_on_row(mx, my, 0.95, 64)
    P = MisfitProblem(mx, my, dx, dy, dt, N, rec, order=2, transverse=1, limiters=4, every=45)
    aux_true = P.aux_from(K_true, rho)
    obs, _, _ = P.forward(P.padded(q0), aux_true, keep_checkpoints=False)
    P.p_obs = obs
    aux = P.aux_from(K, rho)
    q0p = P.padded(q0)
    P.forward(q0p, aux, keep_checkpoints=False)        # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    P.forward(q0p, aux, keep_checkpoints=False)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    traces, cps, _ = P.forward(q0p, aux)
    torch.cuda.synchronize()
    t_fwd_cp = time.perf_counter() - t0
    t0 = time.perf_counter()
    lam, gaux = P.backward(traces, cps, aux)
    t_issue = time.perf_counter() - t0      # host time to issue the reverse sweep's launches
    torch.cuda.synchronize()
    t_bwd = time.perf_counter() - t0
    J = float(P.objective(traces))
    finite = bool(torch.isfinite(lam).all()) and bool(torch.isfinite(gaux).all())
    # every state kept in HBM (every=1: 2000 x 6.4 MB): no recomputation in the reverse sweep
    del cps
    P.every = 1
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    traces1, cps1, _ = P.forward(q0p, aux)
    torch.cuda.synchronize()
    t_fwd_all = time.perf_counter() - t0
    t0 = time.perf_counter()
    lam1, gaux1 = P.backward(traces1, cps1, aux)
    torch.cuda.synchronize()
    t_bwd_all = time.perf_counter() - t0
    same_all = bool(torch.equal(lam1, lam)) and bool(torch.equal(gaux1, gaux))
    del cps1
    P.every = 45
    # per-kernel averages on the launch stream: forward step vs adjoint step
    from pyclaw_amd.adjoint import _Launches
    st = torch.cuda.current_stream()
    go = _Launches(P, aux, st, P._cfl)
    lam2 = torch.zeros_like(lam)
    reps = 50
    e = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    a, b = q0p.clone(), torch.empty_like(q0p)
    ab = [a.data_ptr(), b.data_ptr()]
    e[0].record(st)
    for k in range(reps):
        go.step(ab[k % 2], ab[1 - k % 2])
    e[1].record(st)
    e[2].record(st)
    for k in range(reps):
        go.adjoint(ab[0], lam.data_ptr(), lam2.data_ptr(), gaux.data_ptr())
    e[3].record(st)
    torch.cuda.synchronize()
    us_fwd = e[0].elapsed_time(e[1]) * 1e3 / reps
    us_adj = e[2].elapsed_time(e[3]) * 1e3 / reps
    cells = mx * my
    adj_bytes = 120.0 * cells          # read q 24 + aux 16 + lam 24, write lam 24, gaux r/w 32
    ratio = (t_fwd_cp + t_bwd) / t_fwd
    out = {"metric": "adjoint/fwd time", "value": ratio, "unit": "x", "n_gpus": 1, "steps": N, "warmup": N,
           "ms_per_step": 1e3 * (t_fwd_cp + t_bwd) / N, "higher_is_better": False, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded smooth random media)",
           "config": {"workload": "vc acoustics 512x512, %d steps, misfit gradient w.r.t. K, rho, q0; MC, "
                                  "trans_inc, checkpoint every 45" % N, "global_batch": 1,
                      "parallelism": "single GPU"},
           "roofline": {"bound": "hbm", "achieved": adj_bytes / (us_adj * 1e-6) / 1e9, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": adj_bytes / (us_adj * 1e-6) / 1e9 / HBM_PEAK_GBS,
                        "traffic": None, "bytes_per_cell_update": 120, "kernel": "adjoint step (2 launches)"},
           "extra": {"forward_s": t_fwd, "forward_with_checkpoints_s": t_fwd_cp, "reverse_s": t_bwd,
                     "reverse_over_forward": t_bwd / t_fwd, "reverse_issue_s": t_issue,
                     "forward_step_us": us_fwd, "adjoint_step_us": us_adj,
                     "adjoint_over_forward_step": us_adj / us_fwd, "J": J, "finite": finite,
                     "checkpoints": (N + 44) // 45,
                     "all_states_in_hbm": {"forward_s": t_fwd_all, "reverse_s": t_bwd_all,
                                           "gradient_over_forward": (t_fwd_all + t_bwd_all) / t_fwd,
                                           "bitwise_equal_to_checkpointed": same_all}}}
    return out


def bench_adjoint(args):
    print(json.dumps(measure_adjoint(args)), flush=True)


def bench_copy(args):
    """Calibration for the HBM byte counters: a plain device-to-device copy of
    512 MiB, 5 times (FETCH_SIZE and WRITE_SIZE of that kernel should each
    read 524288 KiB per dispatch)."""
    import torch
    n = 512 * 1024 * 1024 // 8
    a = torch.ones(n, dtype=torch.float64, device='cuda')
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 5
    print(json.dumps({"metric": "copy GB/s", "value": 2 * n * 8 / (ms * 1e-3) / 1e9, "unit": "GB/s",
                      "n_gpus": 1, "steps": 5, "warmup": 0, "ms_per_step": ms, "higher_is_better": True,
                      "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
                      "config": {"workload": "device copy 512 MiB", "global_batch": 1,
                                 "parallelism": "single GPU"}}), flush=True)


OTHER = {'batched': bench_batched, 'sharpclaw': bench_sharpclaw, 'acoustics': bench_acoustics,
         'acoustics_app': bench_acoustics_app, 'sharpclaw_app': bench_sharpclaw_app, 'adjoint': bench_adjoint,
         'copy': bench_copy}


if __name__ == '__main__':
    main()
