"""Initial conditions and set-ups of the benchmarked 2D problems.

Each function returns plain numpy arrays in the reference's (m, i, j) layout,
so the same data feed the GPU path and the CPU oracle.

  acoustics_pulse   -- homogeneous acoustics ring (apps/acoustics_2d_homogeneous/acoustics.py:82-92)
  shockbubble       -- Euler shock/bubble (apps/euler_2d/shockbubble.py:17-87)
  shallow_hump      -- shallow-water Gaussian hump (apps/shallow_2d/shallow2D.py style)
"""
import math

import numpy as np

GAMMA = 1.4
GAMMA1 = GAMMA - 1.


def centers(lower, upper, n):
    d = (upper - lower) / float(n)
    return lower + (np.arange(n) + 0.5) * d, d


def acoustics_pulse(mx, my, lower=(-1., -1.), upper=(1., 1.), center=(0., 0.), width=0.2):
    """p = (|r-0.5|<=width)(1+cos(pi (r-0.5)/width)), u = v = 0 (acoustics.py:82-92),
    r measured from `center`."""
    x, _ = centers(lower[0], upper[0], mx)
    y, _ = centers(lower[1], upper[1], my)
    Y, X = np.meshgrid(y, x)
    r = np.sqrt((X - center[0]) ** 2 + (Y - center[1]) ** 2)
    q = np.zeros((3, mx, my), order='F')
    q[0] = (np.abs(r - 0.5) <= width) * (1. + np.cos(np.pi * (r - 0.5) / width))
    return q


def shock_state(pinf=5.):
    """Post-shock state from Rankine-Hugoniot at pressure ratio pinf (shockbubble.py:26-28)."""
    rinf = (GAMMA1 + pinf * (GAMMA + 1.)) / ((GAMMA + 1.) + GAMMA1 * pinf)
    vinf = 1. / np.sqrt(GAMMA) * (pinf - 1.) / np.sqrt(0.5 * ((GAMMA + 1.) / GAMMA) * pinf + 0.5 * GAMMA1 / GAMMA)
    einf = 0.5 * rinf * vinf ** 2 + pinf / GAMMA1
    return rinf, vinf, einf


def shockbubble(mx, my, lower=(0., 0.), upper=(2., 0.5), x0=0.5, y0=0., r0=0.2, rhoin=0.1, pinf=5.,
                xshock=0.2):
    """Shock/bubble initial data with the bubble-edge cells area-averaged by
    adaptive quadrature exactly as the reference does (shockbubble.py:17-57)."""
    from scipy import integrate
    x, dx = centers(lower[0], upper[0], mx)
    y, dy = centers(lower[1], upper[1], my)
    rhoout, pout, pin = 1., 1., 1.
    rinf, vinf, einf = shock_state(pinf)
    Y, X = np.meshgrid(y, x)
    r = np.sqrt((X - x0) ** 2 + (Y - y0) ** 2)
    q = np.zeros((5, mx, my), order='F')
    q[0] = rinf * (X < xshock) + rhoin * (r <= r0) + rhoout * (r > r0) * (X >= xshock)
    q[1] = rinf * vinf * (X < xshock)
    q[2] = 0.
    q[3] = einf * (X < xshock) + (pin * (r <= r0) + pout * (r > r0) * (X >= xshock)) / GAMMA1
    q[4] = 1. * (r <= r0)

    def ycirc(xx, ymin, ymax):
        if ((xx - x0) ** 2) < (r0 ** 2):
            return max(min(y0 + np.sqrt(r0 ** 2 - (xx - x0) ** 2), ymax) - ymin, 0.)
        return 0

    d2 = np.linalg.norm([dx, dy]) / 2.
    dx2, dy2 = dx / 2., dy / 2.
    for i, j in zip(*np.nonzero(np.abs(r - r0) < d2)):
        ydown = y[j] - dy2
        yup = y[j] + dy2
        infrac, _ = integrate.quad(ycirc, x[i] - dx2, x[i] + dx2, args=(ydown, yup), epsabs=1.e-8, epsrel=1.e-5)
        infrac = infrac / (dx * dy)
        q[0, i, j] = rhoin * infrac + rhoout * (1. - infrac)
        q[3, i, j] = (pin * infrac + pout * (1. - infrac)) / GAMMA1
        q[4, i, j] = 1. * infrac
    return q, dx, dy


def euler_max_speed(q):
    rho = q[0]
    u = q[1] / rho
    v = q[2] / rho
    p = GAMMA1 * (q[3] - 0.5 * rho * (u * u + v * v))
    c = np.sqrt(GAMMA * p / rho)
    return float(max((np.abs(u) + c).max(), (np.abs(v) + c).max()))


def shallow_hump(mx, my, x0, lower=(0., 0.), upper=(1., 1.), amp=0.5, width2=0.01):
    """h = 1 + amp*exp(-|x-x0|^2/width2), hu = hv = 0."""
    x, _ = centers(lower[0], upper[0], mx)
    y, _ = centers(lower[1], upper[1], my)
    Y, X = np.meshgrid(y, x)
    q = np.zeros((3, mx, my), order='F')
    q[0] = 1. + amp * np.exp(-((X - x0[0]) ** 2 + (Y - x0[1]) ** 2) / width2)
    return q


def sound_speed_acoustics(bulk, rho):
    return math.sqrt(bulk / rho)


def layered_vc_acoustics(rng, nx, ny, nlayers=None, amp=1.0, width=0.05):
    """One instance of the batched variable-coefficient acoustics workload on
    [0,1]^2: 2-6 horizontal layers (interfaces uniform in y), per-layer
    rho ~ U[1,4] and c ~ U[0.5,2]; pressure pulse amp*exp(-|x-x0|^2/width^2)
    at a uniform random x0; u = v = 0.  Returns q (3,nx,ny), aux (2,nx,ny)."""
    if nlayers is None:
        nlayers = int(rng.integers(2, 7))
    cuts = np.sort(rng.random(nlayers - 1))
    rho_l = 1.0 + 3.0 * rng.random(nlayers)
    c_l = 0.5 + 1.5 * rng.random(nlayers)
    x0 = rng.random(2)
    x, _ = centers(0.0, 1.0, nx)
    y, _ = centers(0.0, 1.0, ny)
    layer = np.searchsorted(cuts, y)
    aux = np.zeros((2, nx, ny), order='F')
    aux[0] = rho_l[layer][None, :]
    aux[1] = c_l[layer][None, :]
    Y, X = np.meshgrid(y, x)
    q = np.zeros((3, nx, ny), order='F')
    q[0] = amp * np.exp(-((X - x0[0]) ** 2 + (Y - x0[1]) ** 2) / width ** 2)
    return q, aux


def batched_sizes(rng, n, choices=tuple(range(64, 513, 32))):
    return [(int(rng.choice(choices)), int(rng.choice(choices))) for _ in range(n)]


def smooth_random_field(rng, nx, ny, lo, hi, nmodes=16, kmax=4):
    """Sum of nmodes random Fourier modes on [0,1]^2 (integer wave numbers
    0..kmax per axis, uniform phases and amplitudes), rescaled to [lo, hi]."""
    x, _ = centers(0.0, 1.0, nx)
    y, _ = centers(0.0, 1.0, ny)
    Y, X = np.meshgrid(y, x)
    f = np.zeros((nx, ny))
    for _ in range(nmodes):
        kx, ky = rng.integers(0, kmax + 1, size=2)
        f += rng.random() * np.cos(2.0 * np.pi * (kx * X + ky * Y) + 2.0 * np.pi * rng.random())
    f = (f - f.min()) / max(f.max() - f.min(), 1e-300)
    return lo + (hi - lo) * f


def adjoint_medium(rng, nx, ny, width=0.02, disc=0.15, scale=1.05):
    """The adjoint workload's data on [0,1]^2: K in [1,4] and rho in [1,2]
    (smooth random fields), a Gaussian pressure pulse exp(-r^2/width^2) at the
    centre, and the perturbed K (times `scale` inside a centred disc of radius
    `disc`) that produces the observed traces.  Returns q0, K, rho, K_true."""
    K = smooth_random_field(rng, nx, ny, 1.0, 4.0)
    rho = smooth_random_field(rng, nx, ny, 1.0, 2.0)
    x, _ = centers(0.0, 1.0, nx)
    y, _ = centers(0.0, 1.0, ny)
    Y, X = np.meshgrid(y, x)
    r2 = (X - 0.5) ** 2 + (Y - 0.5) ** 2
    q0 = np.zeros((3, nx, ny))
    q0[0] = np.exp(-r2 / width ** 2)
    K_true = K * np.where(r2 <= disc ** 2, scale, 1.0)
    return q0, K, rho, K_true
