"""Batched runs: every instance of a batch (uneven sizes and coefficients, one
persistent work-queue launch per step, hipGraph replay) must equal the same
instance stepped alone through the plain single-grid entry point, bit for bit,
with identical dt sequences; and the device-side accept/reject rule must
reproduce the reference's evolve_to_time logic (solver.py:651-687)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _instances(n, seed, sizes=(64, 96, 130, 200, 61, 123)):
    from pyclaw_amd.problems import layered_vc_acoustics
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        mx = int(sizes[k % len(sizes)]); my = int(sizes[(k * 7 + 3) % len(sizes)])
        q, aux = layered_vc_acoustics(rng, mx, my)
        dx, dy = 1.0 / mx, 1.0 / my
        out.append(dict(q=q, aux=aux, dx=dx, dy=dy, dt_initial=0.9 * min(dx, dy) / aux[1].max()))
    return out


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_equals_one_by_one(use_graph):
    import pyclaw_amd as pc
    from pyclaw_amd.batch import BatchedClawSolver2D, run_one_by_one
    print("batch vs one-by-one, graph", use_graph, flush=True)
    inst = _instances(10, 1)
    b = BatchedClawSolver2D(pc.riemann.rp2_vc_acoustics, {}, inst)
    b.run(12, use_graph=use_graph)
    if use_graph:
        b.run(12, use_graph=True)       # replay of the cached graph
        nsteps = 24
    else:
        nsteps = 12
    ref = run_one_by_one(pc.riemann.rp2_vc_acoustics, {}, inst, nsteps)
    assert (b.naccepted.cpu().numpy() == nsteps).all()
    for k in range(len(inst)):
        assert np.array_equal(b.q(k), ref[k]), k


def test_batch_reject_then_accept_matches_reference_rule():
    """An instance started with a too-large dt: the first attempt is rejected
    (state kept, dt shrunk by cfl_desired/cfl), later attempts accepted."""
    import pyclaw_amd as pc
    from pyclaw_amd.batch import BatchedClawSolver2D
    inst = _instances(3, 2, sizes=(70, 90))
    inst[1]['dt_initial'] *= 3.0
    b = BatchedClawSolver2D(pc.riemann.rp2_vc_acoustics, {}, inst)
    nsteps = 6
    b.run(nsteps, use_graph=True)
    # oracle: the same rule on the CPU (solver.py:651-687, no end time)
    O.set_params(O.RP_VC_ACOUSTICS, {})
    for k, I in enumerate(inst):
        q = I['q'].copy(order='F'); dt = I['dt_initial']; t = 0.0; nacc = 0; nrej = 0
        for _ in range(nsteps):
            qbc = np.zeros((3, q.shape[1] + 4, q.shape[2] + 4), order='F'); qbc[:, 2:-2, 2:-2] = q
            O.fill_bc(qbc, 2, [1, 1], [1, 1], True)
            abc = np.zeros((2, q.shape[1] + 4, q.shape[2] + 4), order='F'); abc[:, 2:-2, 2:-2] = I['aux']
            O.fill_bc(abc, 2, [1, 1], [1, 1], False)
            qn, cfl = O.step2(O.RP_VC_ACOUSTICS, qbc, abc, I['dx'], I['dy'], dt, O.method_array(2, 1, False),
                              O.mthlim_array(4, 2), 2)
            if cfl <= 1.0:
                q = np.asfortranarray(qn[:, 2:-2, 2:-2]); t += dt; nacc += 1
            else:
                nrej += 1
            dt = min(1e99, dt * 0.9 / cfl)
        assert int(b.naccepted[k]) == nacc and int(b.nrejected[k]) == nrej
        assert float(b.t[k]) == t and float(b.dt[k]) == dt
        assert np.array_equal(b.q(k), q)
    assert int(b.nrejected[1]) == 1


@pytest.mark.parametrize("shape", ["tuned", (8, 1), (128, 2), (5, 4)])
def test_batch_shapes_equal_one_by_one(shape):
    """The tile shape (rows per band, wavefronts per block), set or picked by
    claw2d_batch_autotune, changes only scheduling: the batch still equals the
    one-by-one runs bit for bit, and tuning leaves the state untouched."""
    import pyclaw_amd as pc
    from pyclaw_amd.batch import BatchedClawSolver2D, run_one_by_one
    print("batch shape", shape, flush=True)
    inst = _instances(9, 3)
    b = BatchedClawSolver2D(pc.riemann.rp2_vc_acoustics, {}, inst)
    b.run(3, use_graph=True)
    q3 = [b.q(k) for k in range(len(inst))]
    if shape == "tuned":
        rows, wpb = b.autotune()
        assert rows in (8, 16, 32, 64, 128) and wpb in (1, 2, 4)
        for k in range(len(inst)):
            assert np.array_equal(b.q(k), q3[k]), k
    else:
        b.set_shape(*shape)
    b.run(9, use_graph=True)
    ref = run_one_by_one(pc.riemann.rp2_vc_acoustics, {}, inst, 12)
    assert (b.naccepted.cpu().numpy() == 12).all()
    for k in range(len(inst)):
        assert np.array_equal(b.q(k), ref[k]), k
