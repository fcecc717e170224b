#!/usr/bin/env python
"""Benchmark: FP64 classic 2D Euler shock-bubble step at 4096x1024 on one GPU.

A "step" is one unsplit classic time step (Roe solver with entropy fix, MC
limiter on all five waves, transverse correction waves, inflow / wall /
extrapolation boundaries) over the whole grid: one launch of the fused HIP
kernel.  Inputs are the shock-bubble initial data, resident in HBM before the
timed region; dt is fixed at the value giving Courant number 0.8 on the
initial data (the kernel still computes and max-combines the Courant number
of every step; it is reported).

Prints one JSON line (see the keys below).  --gpus N>1 is not supported: this
build targets one GPU.
"""
import argparse
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


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=300)
    ap.add_argument('--warmup', type=int, default=10)
    ap.add_argument('--mx', type=int, default=4096)
    ap.add_argument('--my', type=int, default=1024)
    ap.add_argument('--precision', default='f64', choices=['f64', 'f32'])
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--cpu-sample-steps', type=int, default=3)
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


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU; multi-GPU runs are not part of this build "
              "(see DESIGN.md for how the grid would shard)", file=sys.stderr)
        sys.exit(2)
    import torch
    from pyclaw_amd import _lib
    from pyclaw_amd.problems import euler_max_speed, shock_state, shockbubble

    mx, my = args.mx, args.my
    q0, dx, dy = shockbubble(mx, my)
    smax = euler_max_speed(q0)
    dt = 0.8 * min(dx, dy) / smax
    dev = torch.device('cuda', 0)
    dtype = torch.float64 if args.precision == 'f64' else torch.float32
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
    prec = _lib.F64 if args.precision == 'f64' else _lib.F32
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    bufs = [qa, qb]

    def step(k):
        rc = L.claw2d_classic_step_ex(ctx.handle, grid, m, r, bc, prec, bufs[k % 2].data_ptr(),
                                      bufs[(k + 1) % 2].data_ptr(), None, dtd.data_ptr(), cfl.data_ptr(),
                                      _lib.STEP_ACCUMULATE_CFL, sh)
        _lib.check(rc, 'claw2d_classic_step_ex')

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
    bytes_per_cell = BYTES_PER_CELL if args.precision == 'f64' else BYTES_PER_CELL // 2
    achieved = bytes_per_cell * cells / (ms_per_step * 1e-3) / 1e9
    cfl_max = float(cfl.item())
    qfin = bufs[(args.warmup + args.steps) % 2]
    finite = bool(torch.isfinite(qfin).all().item())

    cpu = None
    if not args.no_cpu_baseline:
        v, el = cpu_baseline(q0, dx, dy, dt, mx, my, args.cpu_sample_steps)
        cpu = {"value": v, "unit": "Gcell-updates/s", "cores": 1, "kind": "port",
               "sample": "%d steps of the same 4096x1024 FP64 unsplit step (%.1f s), single-thread C "
                         "restatement of step2/flux2/limiter + Euler Roe solver (oracle/claw_oracle.c, "
                         "gcc -O2 -ffp-contract=off); the reference Euler solver cannot run this grid "
                         "(maxm2=1800)" % (args.cpu_sample_steps, el)}

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
        "dtype": args.precision,
        "data": "synthetic (shock-bubble initial data, generated in-process)",
        "config": {"workload": "classic 2D Euler shock-bubble %dx%d, unsplit, Roe+efix, MC on 5 waves, "
                               "transverse correction waves, fixed dt at CFL 0.8" % (mx, my),
                   "grid": [mx, my], "global_batch": 1, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "bytes_per_cell_update": bytes_per_cell,
                     "kernel": "classic_unsplit_kernel<Euler5<%s>>" % ('double' if args.precision == 'f64' else 'float')},
        "cpu_baseline": cpu,
        "extra": {"cfl_max_over_timed_steps": cfl_max, "finite": finite, "dt": dt, "wall_s": wall,
                  "hbm_pct": 100.0 * achieved / HBM_PEAK_GBS},
    }
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
