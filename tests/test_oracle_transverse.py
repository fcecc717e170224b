"""Pins the oracle's transverse Riemann solvers and the shallow-water entropy fix.

The reference takes the acoustics, vc-acoustics and shallow-water solvers from
the external clawpack.riemann package (call sites
apps/acoustics_2d_homogeneous/acoustics.py:25-26,
apps/acoustics_2d_variable/acoustics.py:25-26, apps/shallow_2d/shallow2D.py:44-45),
which is absent here.  Their contract is fixed by how classic/flux2.f90:184-206
uses them: rpt2 splits a normal fluctuation asdq into the parts that move down
(bmasdq) and up (bpasdq) in the transverse direction, i.e.

    bmasdq = B^- asdq,  bpasdq = B^+ asdq,  B^+- = R diag(max/min(lambda, 0)) R^-1

with B the transverse flux Jacobian (for shallow water at the interface's Roe
state).  The checks below build B and its eigen-decomposition with numpy, from
the PDE alone, so a swapped bm/bp, a sign flip or a normal/transverse (mu/mv)
swap in the oracle fails them.

The shallow-water entropy fix (Harten-Hyman, structured like the reference's
Euler fix rpn2_euler_5wave.f:205-286) is checked by its defining properties:
supersonic data give a one-sided split, non-transonic data give the plain Roe
split, and a transonic wave is split into a left-going piece at the left
characteristic speed and a right-going piece at the right one, in the unique
proportion that conserves the wave's flux; the split is continuous across the
switches between these cases.
"""
import numpy as np
import pytest

import oracle as O


def _split(B, v):
    lam, R = np.linalg.eig(B)
    lam = lam.real
    R = R.real
    L = np.linalg.inv(R)
    Bm = R @ np.diag(np.minimum(lam, 0.0)) @ L
    Bp = R @ np.diag(np.maximum(lam, 0.0)) @ L
    return Bm @ v, Bp @ v, np.sort(lam)


# ---------------------------------------------------------------------------
# homogeneous acoustics rpt2
# ---------------------------------------------------------------------------
def _acoustics_B(ixy, rho, K):
    """Transverse flux Jacobian of p_t + K div(u) = 0, rho u_t + grad p = 0 in (p, u, v):
    for ixy=1 (x-normal sweep) the transverse direction is y, else x."""
    t = 2 if ixy == 1 else 1
    B = np.zeros((3, 3))
    B[0, t] = K
    B[t, 0] = 1.0 / rho
    return B


@pytest.mark.parametrize("ixy", [1, 2])
@pytest.mark.parametrize("imp", [1, 2])
def test_acoustics_rpt2_is_eigen_split_of_transverse_jacobian(ixy, imp):
    rng = np.random.default_rng(40 + ixy * 3 + imp)
    rho, K = 1.7, 3.1
    c = np.sqrt(K / rho)
    O.set_params(O.RP_ACOUSTICS, dict(rho=rho, bulk=K, cc=c, zz=rho * c))
    k = 24
    asdq = rng.standard_normal((3, k))
    lo = rng.standard_normal((3, k)); hi = rng.standard_normal((3, k))
    bm, bp = O.rpt2(O.RP_ACOUSTICS, ixy, lo, hi, asdq, imp=imp, num_waves=2)
    B = _acoustics_B(ixy, rho, K)
    for j in range(k):
        em, ep, lam = _split(B, asdq[:, j])
        assert np.allclose(lam, [-c, 0.0, c], atol=1e-14)
        assert np.allclose(bm[:, j], em, rtol=1e-13, atol=1e-13)
        assert np.allclose(bp[:, j], ep, rtol=1e-13, atol=1e-13)
        assert np.allclose(bm[:, j] + bp[:, j], B @ asdq[:, j], rtol=1e-13, atol=1e-13)
    # the normal velocity has no transverse flux: neither part carries it
    n = 1 if ixy == 1 else 2
    assert np.all(bm[n] == 0.0) and np.all(bp[n] == 0.0)


def test_acoustics_rpt2_equals_vc_rpt2_on_uniform_medium():
    """Homogeneous acoustics is vc-acoustics with one medium; vc-acoustics is
    pinned end to end by the reference's fort.q0020 / pressure_classic.txt runs
    (unsplit, transverse waves on).  The two must agree bit for bit."""
    rng = np.random.default_rng(47)
    rho, c = 1.3, 1.9
    O.set_params(O.RP_ACOUSTICS, dict(rho=rho, bulk=rho * c * c, cc=c, zz=rho * c))
    k = 16
    asdq = rng.standard_normal((3, k))
    q = rng.standard_normal((3, k))
    aux = np.empty((2, 2, k)); aux[0] = rho; aux[1] = c
    for ixy in (1, 2):
        for imp in (1, 2):
            a = O.rpt2(O.RP_ACOUSTICS, ixy, q, q, asdq, imp=imp, num_waves=2)
            O.set_params(O.RP_VC_ACOUSTICS, {})
            b = O.rpt2(O.RP_VC_ACOUSTICS, ixy, q, q, asdq, imp=imp, aux1=aux, aux2=aux, aux3=aux, num_waves=2)
            O.set_params(O.RP_ACOUSTICS, dict(rho=rho, bulk=rho * c * c, cc=c, zz=rho * c))
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


@pytest.mark.parametrize("ixy", [1, 2])
@pytest.mark.parametrize("imp", [1, 2])
def test_vc_acoustics_rpt2_interface_problems(ixy, imp):
    """Down-going part: the lower medium's -c eigenvector, with the weight that
    solves the jump problem between the lower row and the cell's own row; up-going
    part likewise with the upper row.  The cell is the interface's left cell for
    imp=1 (amdq) and right cell for imp=2 (apdq)."""
    rng = np.random.default_rng(50 + 2 * ixy + imp)
    k = 20
    asdq = rng.standard_normal((3, k))
    q = rng.standard_normal((3, k))
    aux1, aux2, aux3 = (np.stack([1 + 3 * rng.random((2, k)), 0.5 + 1.5 * rng.random((2, k))]) for _ in range(3))
    O.set_params(O.RP_VC_ACOUSTICS, {})
    bm, bp = O.rpt2(O.RP_VC_ACOUSTICS, ixy, q, q, asdq, imp=imp, aux1=aux1, aux2=aux2, aux3=aux3, num_waves=2)
    t = 2 if ixy == 1 else 1
    n = 3 - t
    side = 0 if imp == 1 else 1
    for j in range(k):
        Zm = aux1[0, side, j] * aux1[1, side, j]; cm = aux1[1, side, j]
        Z = aux2[0, side, j] * aux2[1, side, j]
        Zp = aux3[0, side, j] * aux3[1, side, j]; cp = aux3[1, side, j]
        p, vt = asdq[0, j], asdq[t, j]
        # asdq = a1 (-Zm, 1) + b (Z, 1) at the lower interface; asdq = b' (-Z, 1) + a2 (Zp, 1) at the upper
        a1 = np.linalg.solve([[-Zm, Z], [1.0, 1.0]], [p, vt])[0]
        a2 = np.linalg.solve([[-Z, Zp], [1.0, 1.0]], [p, vt])[1]
        em = np.zeros(3); em[0] = -cm * a1 * -Zm; em[t] = -cm * a1
        ep = np.zeros(3); ep[0] = cp * a2 * Zp; ep[t] = cp * a2
        assert np.allclose(bm[:, j], em, rtol=1e-12, atol=1e-13)
        assert np.allclose(bp[:, j], ep, rtol=1e-12, atol=1e-13)
    assert np.all(bm[n] == 0.0) and np.all(bp[n] == 0.0)


# ---------------------------------------------------------------------------
# shallow water
# ---------------------------------------------------------------------------
G = 1.0


def _roe_state(lo, hi):
    """Roe averages (h-bar arithmetic mean, sqrt(h)-weighted velocities), in x/y."""
    sl, sr = np.sqrt(lo[0]), np.sqrt(hi[0])
    hbar = 0.5 * (lo[0] + hi[0])
    ux = (lo[1] / sl + hi[1] / sr) / (sl + sr)
    uy = (lo[2] / sl + hi[2] / sr) / (sl + sr)
    return hbar, ux, uy


def _shallow_jacobian(direction, h, ux, uy):
    """Flux Jacobian in (h, hu, hv) of the shallow-water flux in direction 'x' or 'y'."""
    c2 = G * h
    if direction == 'x':
        return np.array([[0., 1., 0.], [c2 - ux * ux, 2 * ux, 0.], [-ux * uy, uy, ux]])
    return np.array([[0., 0., 1.], [-ux * uy, uy, ux], [c2 - uy * uy, 0., 2 * uy]])


def _shallow_states(rng, k):
    h = 0.5 + rng.random((2, k))
    u = 0.8 * rng.standard_normal((2, 2, k))
    lo = np.stack([h[0], h[0] * u[0, 0], h[0] * u[0, 1]])
    hi = np.stack([h[1], h[1] * u[1, 0], h[1] * u[1, 1]])
    return lo, hi


@pytest.mark.parametrize("ixy", [1, 2])
@pytest.mark.parametrize("imp", [1, 2])
def test_shallow_rpt2_is_eigen_split_of_roe_transverse_jacobian(ixy, imp):
    rng = np.random.default_rng(60 + 2 * ixy + imp)
    k = 30
    lo, hi = _shallow_states(rng, k)
    asdq = rng.standard_normal((3, k))
    O.set_params(O.RP_SHALLOW, dict(grav=G))
    bm, bp = O.rpt2(O.RP_SHALLOW, ixy, lo, hi, asdq, imp=imp, num_waves=3)
    for j in range(k):
        h, ux, uy = _roe_state(lo[:, j], hi[:, j])
        B = _shallow_jacobian('y' if ixy == 1 else 'x', h, ux, uy)
        em, ep, lam = _split(B, asdq[:, j])
        vt = uy if ixy == 1 else ux
        assert np.allclose(lam, np.sort([vt - np.sqrt(G * h), vt, vt + np.sqrt(G * h)]), rtol=1e-13, atol=1e-13)
        assert np.allclose(bm[:, j], em, rtol=1e-11, atol=1e-12)
        assert np.allclose(bp[:, j], ep, rtol=1e-11, atol=1e-12)
        assert np.allclose(bm[:, j] + bp[:, j], B @ asdq[:, j], rtol=1e-12, atol=1e-12)


def _shallow_flux(q, ixy):
    h, hu, hv = q
    un = (hu if ixy == 1 else hv) / h
    f = np.empty_like(q)
    f[0] = h * un
    f[1] = hu * un + (0.5 * G * h * h if ixy == 1 else 0.0)
    f[2] = hv * un + (0.5 * G * h * h if ixy == 2 else 0.0)
    return f


def _state(h, un, ut, ixy):
    """(h, hu, hv) with normal velocity un and transverse ut for the sweep direction."""
    return np.array([h, h * un, h * ut]) if ixy == 1 else np.array([h, h * ut, h * un])


def _roe_waves(lo, hi, ixy):
    """Waves and speeds of the Roe linearisation from numpy's eigenvectors of the
    normal Roe matrix (independent of the oracle's closed-form strengths)."""
    h, ux, uy = _roe_state(lo, hi)
    A = _shallow_jacobian('x' if ixy == 1 else 'y', h, ux, uy)
    lam, R = np.linalg.eig(A)
    order = np.argsort(lam.real)
    lam, R = lam.real[order], R.real[:, order]
    alpha = np.linalg.solve(R, hi - lo)
    return [alpha[p] * R[:, p] for p in range(3)], lam


def _efix(lo, hi, ixy):
    O.set_params(O.RP_SHALLOW, dict(grav=G))
    w, s, am, ap = O.rpn2(O.RP_SHALLOW, ixy, lo[:, None], hi[:, None], num_waves=3)
    return w[:, :, 0], s[:, 0], am[:, 0], ap[:, 0]


def _char_speed(q, ixy, family):
    un = (q[1] if ixy == 1 else q[2]) / q[0]
    return un - np.sqrt(G * q[0]) if family == 1 else un + np.sqrt(G * q[0])


@pytest.mark.parametrize("ixy", [1, 2])
def test_shallow_efix_supersonic_is_one_sided(ixy):
    rng = np.random.default_rng(70 + ixy)
    for _ in range(20):
        hl, hr = 0.5 + rng.random(2)
        # every characteristic speed of both states and of the Roe state positive
        lo = _state(hl, 2.5 + rng.random(), rng.standard_normal(), ixy)
        hi = _state(hr, 2.5 + rng.random(), rng.standard_normal(), ixy)
        w, s, am, ap = _efix(lo, hi, ixy)
        assert np.all(s > 0)
        assert np.all(am == 0.0)
        assert np.allclose(ap, _shallow_flux(hi, ixy) - _shallow_flux(lo, ixy), rtol=1e-12, atol=1e-12)
        # mirrored: all speeds negative
        lo = _state(hl, -2.5 - rng.random(), rng.standard_normal(), ixy)
        hi = _state(hr, -2.5 - rng.random(), rng.standard_normal(), ixy)
        w, s, am, ap = _efix(lo, hi, ixy)
        assert np.all(s < 0)
        assert np.all(ap == 0.0)
        assert np.allclose(am, _shallow_flux(hi, ixy) - _shallow_flux(lo, ixy), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("ixy", [1, 2])
def test_shallow_efix_subsonic_equals_roe_split(ixy):
    """No transonic rarefaction: the fix leaves the Roe split min(s,0)W / max(s,0)W."""
    rng = np.random.default_rng(80 + ixy)
    n = 0
    for _ in range(60):
        hl, hr = 0.8 + 0.4 * rng.random(2)
        lo = _state(hl, 0.3 * rng.standard_normal(), rng.standard_normal(), ixy)
        hi = _state(hr, 0.3 * rng.standard_normal(), rng.standard_normal(), ixy)
        w, s, am, ap = _efix(lo, hi, ixy)
        W, lam = _roe_waves(lo, hi, ixy)
        # skip data where a family's speed changes sign across its wave
        s0, s1 = _char_speed(lo, ixy, 1), _char_speed(lo + W[0], ixy, 1)
        s3l, s3r = _char_speed(hi - W[2], ixy, 3), _char_speed(hi, ixy, 3)
        if (s0 < 0) != (s1 < 0) or (s3l < 0) != (s3r < 0):
            continue
        n += 1
        assert np.allclose(s, lam, rtol=1e-13, atol=1e-13)
        roe_m = sum(min(lam[p], 0.0) * W[p] for p in range(3))
        roe_p = sum(max(lam[p], 0.0) * W[p] for p in range(3))
        assert np.allclose(am, roe_m, rtol=1e-12, atol=1e-13)
        assert np.allclose(ap, roe_p, rtol=1e-12, atol=1e-13)
    assert n >= 40


def _transonic_case(ixy, family, rng):
    """Left state at rest-ish and a fast shallow right state (1-rarefaction through
    sonic), or the mirror image for family 3."""
    while True:
        if family == 1:
            lo = _state(1.0 + 0.2 * rng.random(), 0.2 + 0.3 * rng.random(), rng.standard_normal(), ixy)
            hi = _state(0.3 + 0.2 * rng.random(), 1.6 + 0.5 * rng.random(), rng.standard_normal(), ixy)
        else:
            lo = _state(0.3 + 0.2 * rng.random(), -1.6 - 0.5 * rng.random(), rng.standard_normal(), ixy)
            hi = _state(1.0 + 0.2 * rng.random(), -0.2 - 0.3 * rng.random(), rng.standard_normal(), ixy)
        W, lam = _roe_waves(lo, hi, ixy)
        if family == 1:
            sl, sr = _char_speed(lo, ixy, 1), _char_speed(lo + W[0], ixy, 1)
        else:
            sl, sr = _char_speed(hi - W[2], ixy, 3), _char_speed(hi, ixy, 3)
        if sl < 0 < sr:
            return lo, hi, W, lam, sl, sr


@pytest.mark.parametrize("ixy", [1, 2])
@pytest.mark.parametrize("family", [1, 3])
def test_shallow_efix_transonic_harten_hyman_split(ixy, family):
    """A transonic wave p (speed s_hat, left/right characteristic speeds sl < 0 < sr)
    is split into beta*W moving at sl and (1-beta)*W moving at sr, with
    beta*sl + (1-beta)*sr = s_hat (the wave's flux s_hat*W is conserved)."""
    rng = np.random.default_rng(90 + 2 * ixy + family)
    for _ in range(15):
        lo, hi, W, lam, sl, sr = _transonic_case(ixy, family, rng)
        w, s, am, ap = _efix(lo, hi, ixy)
        p = 0 if family == 1 else 2
        others = [q for q in range(3) if q != p]
        rest = sum(min(lam[q], 0.0) * W[q] for q in others)
        part = am - rest                      # the transonic wave's left-going part
        k = int(np.argmax(np.abs(W[p])))
        sigma = part[k] / W[p][k]             # part = sigma * W_p
        assert np.allclose(part, sigma * W[p], rtol=1e-10, atol=1e-12)
        beta = sigma / sl
        assert -1e-12 <= beta <= 1 + 1e-12
        beta_hh = (sr - lam[p]) / (sr - sl)   # the unique flux-conserving weight
        assert abs(beta - beta_hh) <= 1e-10
        assert np.allclose(ap, sum(max(lam[q], 0.0) * W[q] for q in others) + (lam[p] - sigma) * W[p],
                           rtol=1e-10, atol=1e-12)
        assert np.allclose(am + ap, _shallow_flux(hi, ixy) - _shallow_flux(lo, ixy), rtol=1e-11, atol=1e-12)


@pytest.mark.parametrize("ixy", [1, 2])
def test_shallow_efix_continuous_across_switches(ixy):
    """amdq is continuous where the left state's 1-speed crosses zero (transonic ->
    supersonic right) and where the intermediate state's 1-speed crosses zero
    (transonic -> plain Roe split)."""
    hr, ur = 0.4, 1.9
    hl = 1.0
    cl = np.sqrt(G * hl)
    hi = _state(hr, ur, 0.3, ixy)
    prev = None
    # left 1-speed s0 = ul - cl swept through 0
    for d in np.linspace(-1e-3, 1e-3, 41):
        lo = _state(hl, cl + d, -0.2, ixy)
        am = _efix(lo, hi, ixy)[2]
        if prev is not None:
            assert np.abs(am - prev).max() <= 5e-4
        prev = am
    am_edge = _efix(_state(hl, cl - 1e-9, -0.2, ixy), hi, ixy)[2]
    assert np.abs(am_edge).max() <= 1e-8
    # intermediate 1-speed swept through 0: vary the right state's velocity
    lo = _state(1.0, 0.3, -0.2, ixy)
    found = False
    prev = None
    for ur2 in np.linspace(0.2, 2.0, 3601):
        hi = _state(0.6, ur2, 0.1, ixy)
        W, lam = _roe_waves(lo, hi, ixy)
        s1 = _char_speed(lo + W[0], ixy, 1)
        am = _efix(lo, hi, ixy)[2]
        if prev is not None:
            assert np.abs(am - prev[0]).max() <= 2e-3
            if (prev[1] < 0) != (s1 < 0):
                found = True
        prev = (am, s1)
    assert found
