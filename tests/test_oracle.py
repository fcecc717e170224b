"""Pins the CPU oracle (oracle/claw_oracle.c), which every GPU parity test uses.

1. Bit-exact against the reference's own Fortran compiled from its sources
   (oracle/_ref, built by oracle/build_ref.sh where the reference is present):
   classic step2 / step2ds with every transverse mode and limiter, SharpClaw dq.
   The non-Euler Riemann solvers are external to the reference, so for those
   the reference sweeps run with this repository's solvers plugged in; the
   solvers themselves are pinned by 2. and by the end-to-end fixtures.
2. Closed forms and invariants: limiter formulas, wave sums, Roe property,
   conservation, fixed points, linearity, WENO5 polynomial exactness, SSP104
   order of accuracy.
"""
import numpy as np
import pytest

import oracle as O
from helpers import acoustic_state, euler_state, shallow_state, vc_aux

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="reference build (oracle/_ref) not present")


# ---------------------------------------------------------------------------
# 1. reference parity
# ---------------------------------------------------------------------------
@needs_ref
@pytest.mark.parametrize("trans", [0, 1, 2])
@pytest.mark.parametrize("lim", [0, 1, 2, 3, 4, 5, 7])
def test_step2_euler_matches_reference(trans, lim):
    rng = np.random.default_rng(100 + trans * 10 + lim)
    q = euler_state(rng, 37 + 4, 23 + 4)
    O.set_params(O.RP_EULER, dict(gamma=1.4, gamma1=0.4))
    O.ref_set_euler_params(1.4, 0.4)
    meth = O.method_array(2, trans, False)
    mth = O.mthlim_array(lim, 5)
    a, ca = O.step2(O.RP_EULER, np.asfortranarray(q), None, 0.1, 0.07, 0.01, meth, mth, 5)
    b, cb = O.ref_step2(O.RP_EULER, np.asfortranarray(q), None, 0.1, 0.07, 0.01, meth, mth, 5)
    assert np.array_equal(a, b)
    assert ca == cb


@needs_ref
def test_step2ds_euler_matches_reference():
    rng = np.random.default_rng(7)
    q = np.asfortranarray(euler_state(rng, 41, 27))
    O.set_params(O.RP_EULER, dict(gamma=1.4, gamma1=0.4))
    O.ref_set_euler_params(1.4, 0.4)
    meth = O.method_array(2, 1, True)
    mth = O.mthlim_array([4, 4, 4, 4, 2], 5)
    a = q.copy(order='F'); b = q.copy(order='F')
    c1 = O.step2ds(O.RP_EULER, q, a, None, 0.1, 0.07, 0.01, meth, mth, 5, 1)
    c2 = O.step2ds(O.RP_EULER, a, a, None, 0.1, 0.07, 0.01, meth, mth, 5, 2)
    _, d1 = O.ref_step2(O.RP_EULER, q, None, 0.1, 0.07, 0.01, meth, mth, 5, ids=1, qnew=b)
    _, d2 = O.ref_step2(O.RP_EULER, b, None, 0.1, 0.07, 0.01, meth, mth, 5, ids=2, qnew=b)
    assert np.array_equal(a, b) and c1 == d1 and c2 == d2


@needs_ref
@pytest.mark.parametrize("name", ["acoustics", "vc_acoustics", "shallow"])
@pytest.mark.parametrize("trans", [0, 1, 2])
def test_step2_sweeps_match_reference(name, trans):
    rng = np.random.default_rng(200 + trans)
    mx, my = 33, 29
    if name == "acoustics":
        rp, nw, pd, q, aux = O.RP_ACOUSTICS, 2, dict(rho=1., bulk=4., cc=2., zz=2.), acoustic_state(rng, mx, my), None
    elif name == "vc_acoustics":
        rp, nw, pd, q, aux = O.RP_VC_ACOUSTICS, 2, {}, acoustic_state(rng, mx, my), vc_aux(rng, mx, my)
    else:
        rp, nw, pd, q, aux = O.RP_SHALLOW, 3, dict(grav=1.), shallow_state(rng, mx, my), None
    O.set_params(rp, pd)
    meth = O.method_array(2, trans, False)
    mth = O.mthlim_array(4, nw)
    a, ca = O.step2(rp, np.asfortranarray(q), aux, 0.1, 0.07, 0.01, meth, mth, nw)
    b, cb = O.ref_step2(rp, np.asfortranarray(q), aux, 0.1, 0.07, 0.01, meth, mth, nw)
    assert np.array_equal(a, b) and ca == cb


@needs_ref
@pytest.mark.parametrize("name", ["euler", "shallow", "acoustics", "vc_acoustics"])
def test_sharpclaw_dq_matches_reference(name):
    rng = np.random.default_rng(300)
    mx, my = 31, 22
    aux = None
    if name == "euler":
        rp, nw, pd, q = O.RP_EULER, 5, dict(gamma=1.4, gamma1=0.4), euler_state(rng, mx + 6, my + 6)
        O.ref_set_euler_params(1.4, 0.4)
    elif name == "shallow":
        rp, nw, pd, q = O.RP_SHALLOW, 3, dict(grav=1.), shallow_state(rng, mx + 6, my + 6)
    elif name == "acoustics":
        rp, nw, pd, q = O.RP_ACOUSTICS, 2, dict(rho=1., bulk=4., cc=2., zz=2.), acoustic_state(rng, mx + 6, my + 6)
    else:
        rp, nw, pd, q = O.RP_VC_ACOUSTICS, 2, {}, acoustic_state(rng, mx + 6, my + 6)
        aux = np.asfortranarray(vc_aux(rng, mx + 6, my + 6))
    O.set_params(rp, pd)
    a, ca = O.sharp_dq(rp, np.asfortranarray(q), aux, 0.1, 0.07, 0.01, nw)
    b, cb = O.ref_sharp_dq(rp, np.asfortranarray(q), aux, 0.1, 0.07, 0.01, nw)
    assert np.array_equal(a, b) and ca == cb


# ---------------------------------------------------------------------------
# 2. closed forms and invariants
# ---------------------------------------------------------------------------
def test_philim_closed_forms():
    a = 2.0
    for r in [-3.0, -0.5, 0.0, 0.3, 0.5, 1.0, 1.7, 2.5, 4.0]:
        b = r * a
        assert O.philim(a, b, 1) == max(0., min(1., r))
        assert O.philim(a, b, 2) == max(0., min(1., 2 * r), min(2., r))
        assert O.philim(a, b, 3) == (r + abs(r)) / (1 + abs(r))
        assert O.philim(a, b, 4) == max(0., min((1 + r) / 2, 2., 2 * r))
        assert O.philim(a, b, 5) == r
        # ids outside 1..5 fall through to minmod (philim.f90:19-26)
        assert O.philim(a, b, 9) == O.philim(a, b, 1)
    # MC equals 2 for large r, 0 for r <= 0, 1 at r = 1; superbee 2 for r >= 2
    assert O.philim(1., 10., 4) == 2. and O.philim(1., -1., 4) == 0. and O.philim(1., 1., 4) == 1.
    assert O.philim(1., 3., 2) == 2.


def _flux_euler(q, ixy):
    rho, mu, mv, E = q[0], q[1], q[2], q[3]
    if ixy == 2:
        mu, mv = mv, mu
    u = mu / rho
    p = 0.4 * (E - 0.5 * (q[1] ** 2 + q[2] ** 2) / rho)
    f = np.zeros_like(q)
    f[0] = mu
    f[1 if ixy == 1 else 2] = mu * u + p
    f[2 if ixy == 1 else 1] = mv * u
    f[3] = u * (E + p)
    f[4] = q[4] * u
    return f


@pytest.mark.parametrize("ixy", [1, 2])
def test_euler_roe_property_and_wave_sum(ixy):
    rng = np.random.default_rng(11)
    lo = euler_state(rng, 50, 1)[:, :, 0]
    hi = euler_state(rng, 50, 1)[:, :, 0]
    O.set_params(O.RP_EULER, dict(gamma=1.4, gamma1=0.4, efix=0))
    w, s, am, ap = O.rpn2(O.RP_EULER, ixy, lo, hi, num_waves=5)
    dq = hi - lo
    assert np.allclose(w.sum(axis=1), dq, rtol=1e-12, atol=1e-12)
    # Roe property for the flow variables; the tracer wave carries the jump of
    # the tracer density at the Roe velocity
    sw = np.einsum('mwk,wk->mk', w, s)
    df = _flux_euler(hi, ixy) - _flux_euler(lo, ixy)
    assert np.allclose(sw[:4], df[:4], rtol=1e-11, atol=1e-11)
    assert np.allclose(am + ap, sw, rtol=1e-12, atol=1e-12)
    # with the entropy fix the total fluctuation is unchanged
    O.set_params(O.RP_EULER, dict(gamma=1.4, gamma1=0.4, efix=1))
    w2, s2, am2, ap2 = O.rpn2(O.RP_EULER, ixy, lo, hi, num_waves=5)
    assert np.allclose(am2 + ap2, sw, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("ixy", [1, 2])
def test_shallow_roe_property(ixy):
    rng = np.random.default_rng(12)
    lo = shallow_state(rng, 40, 1)[:, :, 0]
    hi = shallow_state(rng, 40, 1)[:, :, 0]
    O.set_params(O.RP_SHALLOW, dict(grav=1.))
    w, s, am, ap = O.rpn2(O.RP_SHALLOW, ixy, lo, hi, num_waves=3)
    assert np.allclose(w.sum(axis=1), hi - lo, atol=1e-13)

    def flux(q):
        h, hu, hv = q
        if ixy == 2:
            hu, hv = hv, hu
        f = np.zeros_like(q)
        f[0] = hu
        f[1 if ixy == 1 else 2] = hu * hu / h + 0.5 * h * h
        f[2 if ixy == 1 else 1] = hu * hv / h
        return f
    assert np.allclose(am + ap, flux(hi) - flux(lo), rtol=1e-11, atol=1e-12)


@pytest.mark.parametrize("ixy", [1, 2])
def test_acoustics_fluctuations(ixy):
    rng = np.random.default_rng(13)
    lo = rng.standard_normal((3, 30)); hi = rng.standard_normal((3, 30))
    rho, K = 1.7, 3.1
    c = np.sqrt(K / rho)
    O.set_params(O.RP_ACOUSTICS, dict(rho=rho, bulk=K, cc=c, zz=rho * c))
    w, s, am, ap = O.rpn2(O.RP_ACOUSTICS, ixy, lo, hi, num_waves=2)
    A = np.zeros((3, 3))
    n = 1 if ixy == 1 else 2
    A[0, n] = K
    A[n, 0] = 1. / rho
    # the two acoustic waves carry the p and normal-velocity jumps; the
    # zero-speed shear wave (transverse velocity) is not represented
    assert np.allclose(w.sum(axis=1)[[0, n]], (hi - lo)[[0, n]], atol=1e-13)
    assert np.allclose(am + ap, A @ (hi - lo), atol=1e-12)
    assert np.all(am[0] * 0 == 0) and np.allclose(s[0], -c) and np.allclose(s[1], c)


def test_vc_acoustics_wave_sum_and_speeds():
    rng = np.random.default_rng(14)
    lo = rng.standard_normal((3, 25)); hi = rng.standard_normal((3, 25))
    alo = vc_aux(rng, 25, 1)[:, :, 0]; ahi = vc_aux(rng, 25, 1)[:, :, 0]
    O.set_params(O.RP_VC_ACOUSTICS, {})
    w, s, am, ap = O.rpn2(O.RP_VC_ACOUSTICS, 1, lo, hi, alo, ahi, num_waves=2)
    assert np.allclose(w.sum(axis=1)[:2], (hi - lo)[:2], atol=1e-13)
    assert np.allclose(s[0], -alo[1]) and np.allclose(s[1], ahi[1])
    # each wave is an eigenvector of its own medium: p = -Z_lo u (left), p = Z_hi u (right)
    assert np.allclose(w[0, 0], -alo[0] * alo[1] * w[1, 0])
    assert np.allclose(w[0, 1], ahi[0] * ahi[1] * w[1, 1])


def _periodic_step(name, q, trans, split=False):
    ng = 2
    qbc = np.zeros((q.shape[0], q.shape[1] + 4, q.shape[2] + 4), order='F')
    qbc[:, ng:-ng, ng:-ng] = q
    O.fill_bc(qbc, ng, [2, 2], [2, 2], True)
    rp, nw, pd = {'euler': (O.RP_EULER, 5, dict(gamma=1.4, gamma1=0.4)),
                  'shallow': (O.RP_SHALLOW, 3, dict(grav=1.)),
                  'acoustics': (O.RP_ACOUSTICS, 2, dict(rho=1., bulk=4., cc=2., zz=2.))}[name]
    O.set_params(rp, pd)
    meth = O.method_array(2, trans, split)
    lim = O.mthlim_array(4, nw)
    if not split:
        qn, c = O.step2(rp, qbc, None, 0.1, 0.1, 0.02, meth, lim, nw)
    else:
        qn = qbc.copy(order='F')
        O.step2ds(rp, qbc, qn, None, 0.1, 0.1, 0.02, meth, lim, nw, 1)
        O.step2ds(rp, qn, qn, None, 0.1, 0.1, 0.02, meth, lim, nw, 2)
    return qn[:, ng:-ng, ng:-ng]


@pytest.mark.parametrize("name", ["euler", "shallow", "acoustics"])
@pytest.mark.parametrize("trans", [0, 1, 2])
def test_conservation_periodic(name, trans):
    rng = np.random.default_rng(15)
    q = {'euler': euler_state, 'shallow': shallow_state, 'acoustics': acoustic_state}[name](rng, 24, 19)
    qn = _periodic_step(name, q, trans)
    tot0 = q.sum(axis=(1, 2)); tot1 = qn.sum(axis=(1, 2))
    scale = np.abs(q).sum(axis=(1, 2))
    # the Euler solver's fifth field is a tracer advected in non-conservative
    # form (wave 5 = jump of the tracer at speed u), so only mass, momentum
    # and energy telescope
    n = 4 if name == 'euler' else q.shape[0]
    assert np.all(np.abs(tot1 - tot0)[:n] <= 1e-13 * scale[:n])


def test_constant_state_is_fixed_point():
    q = np.ones((5, 20, 15))
    q[1] = 0.3; q[2] = -0.2; q[3] = 2.5; q[4] = 0.1
    qn = _periodic_step('euler', q, 2)
    assert np.allclose(qn, q, rtol=0, atol=1e-15)


def test_acoustics_step_linear_without_limiter():
    rng = np.random.default_rng(16)
    q1 = acoustic_state(rng, 18, 14); q2 = acoustic_state(rng, 18, 14)
    ng = 2

    def step(q):
        qbc = np.zeros((3, 22, 18), order='F'); qbc[:, ng:-ng, ng:-ng] = q
        O.fill_bc(qbc, ng, [3, 1], [1, 3], True)
        O.set_params(O.RP_ACOUSTICS, dict(rho=1., bulk=4., cc=2., zz=2.))
        qn, _ = O.step2(O.RP_ACOUSTICS, qbc, None, 0.1, 0.1, 0.02, O.method_array(2, 2, False),
                        O.mthlim_array(0, 2), 2)
        return qn[:, ng:-ng, ng:-ng]
    lhs = step(2.0 * q1 - 3.0 * q2)
    rhs = 2.0 * step(q1) - 3.0 * step(q2)
    assert np.allclose(lhs, rhs, rtol=1e-12, atol=1e-12)


def test_weno5_polynomial_exactness_and_mirror():
    x = np.arange(16.0)
    # cell averages of P(x) = a + b x + c x^2 over unit cells centred at k
    a, b, c = 0.7, -1.3, 0.45
    avg = a + b * x + c * (x * x + 1. / 12.)
    ql, qr = O.weno5(avg[None, :])
    P = lambda xx: a + b * xx + c * xx * xx
    k = np.arange(2, 14)
    assert np.allclose(ql[0, k], P(k - 0.5), rtol=1e-12, atol=1e-12)
    assert np.allclose(qr[0, k], P(k + 0.5), rtol=1e-12, atol=1e-12)
    # mirror symmetry: reversing the data swaps left and right edge values
    rng = np.random.default_rng(17)
    v = rng.standard_normal(20)
    l1, r1 = O.weno5(v[None, :])
    l2, r2 = O.weno5(v[None, ::-1].copy())
    kk = np.arange(2, 18)
    assert np.allclose(l1[0, kk], r2[0, 19 - kk], rtol=1e-13, atol=1e-13)
    # a discontinuity: no new extrema beyond the data range (essentially non-oscillatory)
    step = np.where(np.arange(20) < 10, 1.0, 0.0)
    l3, r3 = O.weno5(step[None, :])
    assert l3[0, 2:18].max() <= 1.0 + 1e-6 and r3[0, 2:18].min() >= -1e-6


def _stability_function(integrator, z):
    """R(z) of the oracle's SharpClaw step on q' = z q (one step, dt = 1)."""
    from oracle import driver

    class P:
        cfl_max = 1e9
        time_integrator = integrator
    orig = driver.sharp_dq
    driver.sharp_dq = lambda P_, q, dt, t: (z * q, 0.0)
    try:
        q, _ = driver.sharp_step(P, np.ones((1, 1, 1)), 1.0, 0.0)
    finally:
        driver.sharp_dq = orig
    return float(q.ravel()[0])


def test_euler_and_ssp33_stability_polynomials():
    """sharpclaw/solver.py:199-218: forward Euler is 1 + z; Shu-Osher SSP33 is
    the cubic Taylor polynomial 1 + z + z^2/2 + z^3/6 (third order), with the
    strong-stability interval [-1, 0] of coefficient 1."""
    for z in [-1.0, -0.7, -0.3, 0.1, 0.4]:
        assert _stability_function('Euler', z) == 1.0 + z
        assert abs(_stability_function('SSP33', z) - (1 + z + z * z / 2 + z ** 3 / 6)) <= 1e-15
    e1 = abs(_stability_function('SSP33', 0.2) - np.exp(0.2))
    e2 = abs(_stability_function('SSP33', 0.1) - np.exp(0.1))
    assert 12 < e1 / e2 < 20      # ratio 2^4 for a third-order method
    assert abs(_stability_function('SSP33', -1.0)) <= 1.0


def test_unknown_integrator_raises():
    from oracle import driver

    class P:
        cfl_max = 1e9
        time_integrator = 'RK45'
    with pytest.raises(Exception, match='Unrecognized time integrator'):
        driver.sharp_step(P, np.ones((1, 1, 1)), 1.0, 0.0)


def test_ssp104_fourth_order():
    """The SSP104 recurrences of the reference (sharpclaw/solver.py:221-251),
    as restated in oracle.driver.sharp_step, applied to q' = lambda q, must
    agree with exp(z) to O(z^5)."""
    from oracle import driver

    def R(z):
        class P:  # minimal problem stub: dq = z*q
            cfl_max = 1e9
        import types
        orig = driver.sharp_dq
        driver.sharp_dq = lambda P_, q, dt, t: (z * q, 0.0)
        try:
            q, _ = driver.sharp_step(P, np.ones((1, 1, 1)), 1.0, 0.0)
        finally:
            driver.sharp_dq = orig
        return float(q.ravel()[0])
    e1 = abs(R(0.2) - np.exp(0.2))
    e2 = abs(R(0.1) - np.exp(0.1))
    assert e1 < 1e-5
    assert 20 < e1 / e2 < 45      # ratio 2^5 = 32 for a fourth-order method
    # strong stability region contains [-6, 0] (SSP coefficient 6 for 10 stages)
    assert abs(R(-6.0)) <= 1.0 + 1e-12


# ---------------------------------------------------------------------------
# 3. the multi-threaded sweeps equal the single-threaded ones bit for bit
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("name", ["acoustics", "vc_acoustics", "euler", "shallow"])
@pytest.mark.parametrize("threads", [2, 5, 8])
def test_threaded_oracle_is_bit_identical(name, threads):
    """Row (column) bands that recompute one slice on each side give every cell
    its contributions in the reference's order: step2, both step2ds calls (the
    second aliased as classic/solver.py:529-531 calls it) and SharpClaw dq."""
    rng = np.random.default_rng(400 + threads)
    mx, my = 37, 23
    aux = None
    if name == "euler":
        rp, nw, pd, make = O.RP_EULER, 5, dict(gamma=1.4, gamma1=0.4), euler_state
    elif name == "shallow":
        rp, nw, pd, make = O.RP_SHALLOW, 3, dict(grav=1.), shallow_state
    elif name == "acoustics":
        rp, nw, pd, make = O.RP_ACOUSTICS, 2, dict(rho=1., bulk=4., cc=2., zz=2.), acoustic_state
    else:
        rp, nw, pd, make = O.RP_VC_ACOUSTICS, 2, {}, acoustic_state
    O.set_params(rp, pd)
    q = np.asfortranarray(make(rng, mx + 4, my + 4))
    if name == "vc_acoustics":
        aux = np.asfortranarray(vc_aux(rng, mx + 4, my + 4))
    q3 = np.asfortranarray(make(rng, mx + 6, my + 6))
    aux3 = np.asfortranarray(vc_aux(rng, mx + 6, my + 6)) if aux is not None else None
    out = {}
    for nt in (1, threads):
        old = O.set_threads(nt)
        try:
            res = []
            for trans in (0, 1, 2):
                res.append(O.step2(rp, q, aux, 0.1, 0.07, 0.01, O.method_array(2, trans, False),
                                   O.mthlim_array(4, nw), nw))
            meth = O.method_array(2, 1, True)
            a = q.copy(order='F')
            c1 = O.step2ds(rp, q, a, aux, 0.1, 0.07, 0.01, meth, O.mthlim_array(4, nw), nw, 1)
            c2 = O.step2ds(rp, a, a, aux, 0.1, 0.07, 0.01, meth, O.mthlim_array(4, nw), nw, 2)
            res.append((a, (c1, c2)))
            res.append(O.sharp_dq(rp, q3, aux3, 0.1, 0.07, 0.01, nw))
            out[nt] = res
        finally:
            O.set_threads(old)
    for (a, ca), (b, cb) in zip(out[1], out[threads]):
        assert np.array_equal(a, b) and ca == cb
