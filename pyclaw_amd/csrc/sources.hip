// Source terms as device operations on the padded planar layout, gfx950.
//
// The Euler equations with cylindrical (radial) symmetry, the geometric
// source terms of the shock-bubble application
// (apps/euler_2d/shockbubble.py:89-124 step_Euler_radial, Clawpack style:
// one 2-stage Runge-Kutta step in place; :127-153 dq_Euler_radial, SharpClaw
// style: dt times the source).  aux component 0 holds the radius (the cell
// centre's y).  One thread per interior cell, x fastest (flat index over the
// mx x my interior), so each component plane is read and written in
// coalesced rows.  Every expression keeps the
// reference's operation order (numpy evaluates left to right), and the library
// is built without contraction, so FP64 results equal the reference's bit for
// bit; the factor (ndim - 1) = 1 of the reference is exact and left out.
#include <hip/hip_runtime.h>

#include "common.hpp"

namespace clawhip {

template <typename R>
__global__ __launch_bounds__(256) void euler_radial_step_kernel(GridDev g, R *q, const R *aux, R dt, R gamma1) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (long long)g.mx * g.my) return;
    const int i = (int)(k % g.mx), j = (int)(k / g.mx);
    const long long off = (long long)(j + g.ng) * g.ldq + (i + g.ng);
    R *q0 = q + off, *q1 = q0 + g.plane, *q2 = q1 + g.plane, *q3 = q2 + g.plane;
    const R rad = aux[off];
    const R a0 = *q0, a1 = *q1, a2 = *q2, a3 = *q3;
    const R dt2 = dt / R(2);
    // shockbubble.py:103-114: first stage to qstar
    R rho = a0, u = a1 / rho, v = a2 / rho;
    R press = gamma1 * (a3 - R(0.5) * rho * (u * u + v * v));
    const R c1 = dt2 / rad;
    const R s0 = a0 - c1 * a2;
    const R s1 = a1 - c1 * rho * u * v;
    const R s2 = a2 - c1 * rho * v * v;
    const R s3 = a3 - c1 * v * (a3 + press);
    // :116-124: second stage from qstar's fluxes
    rho = s0; u = s1 / rho; v = s2 / rho;
    press = gamma1 * (s3 - R(0.5) * rho * (u * u + v * v));
    const R c2 = dt / rad;
    *q0 = a0 - c2 * s2;
    *q1 = a1 - c2 * rho * u * v;
    *q2 = a2 - c2 * rho * v * v;
    *q3 = a3 - c2 * v * (s3 + press);
}

template <typename R>
__global__ __launch_bounds__(256) void euler_radial_dq_kernel(GridDev g, const R *q, const R *aux, R *dq, R dt,
                                                              R gamma1, int ne) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (long long)g.mx * g.my) return;
    const int i = (int)(k % g.mx), j = (int)(k / g.mx);
    const long long off = (long long)(j + g.ng) * g.ldq + (i + g.ng);
    const R rad = aux[off];
    const R a0 = q[off], a1 = q[g.plane + off], a2 = q[2 * g.plane + off], a3 = q[3 * g.plane + off];
    // shockbubble.py:141-151
    const R rho = a0, u = a1 / rho, v = a2 / rho;
    const R press = gamma1 * (a3 - R(0.5) * rho * (u * u + v * v));
    const R c = -dt / rad;
    dq[off] = c * a2;
    dq[g.plane + off] = c * rho * u * v;
    dq[2 * g.plane + off] = c * rho * v * v;
    dq[3 * g.plane + off] = c * v * (a3 + press);
    for (int m = 4; m < ne; ++m) dq[m * g.plane + off] = R(0);
}

static int radial_check(const ClawGrid *grid, int precision, int num_eqn, const void *q, const void *aux,
                        double dt) {
    if (!grid || !q || !aux) return CLAW_EINVAL;
    if (precision != CLAW_F64 && precision != CLAW_F32) return CLAW_EINVAL;
    if (num_eqn < 4 || grid->mx < 1 || grid->my < 1 || grid->ng < 0) return CLAW_EINVAL;
    if (grid->ldq < grid->mx + 2 * grid->ng || grid->plane < grid->ldq * (grid->my + 2 * grid->ng)) return CLAW_EINVAL;
    if (((long long)grid->mx * grid->my + 255) / 256 > 0x7fffffffLL) return CLAW_EINVAL;
    if (!(dt >= 0.0)) return CLAW_EINVAL;
    return CLAW_OK;
}

}  // namespace clawhip

using namespace clawhip;

extern "C" {

int claw2d_euler_radial_step(const ClawGrid *grid, int precision, int num_eqn, void *q, const void *aux, double dt,
                             double gamma1, void *stream) {
    const int rc = radial_check(grid, precision, num_eqn, q, aux, dt);
    if (rc != CLAW_OK) return rc;
    const GridDev g{grid->mx, grid->my, grid->ng, grid->ldq, grid->plane};
    const dim3 blocks((unsigned)(((long long)grid->mx * grid->my + 255) / 256));
    hipStream_t st = (hipStream_t)stream;
    if (precision == CLAW_F64)
        hipLaunchKernelGGL(euler_radial_step_kernel<double>, blocks, dim3(256), 0, st, g, (double *)q,
                           (const double *)aux, dt, gamma1);
    else
        hipLaunchKernelGGL(euler_radial_step_kernel<float>, blocks, dim3(256), 0, st, g, (float *)q,
                           (const float *)aux, (float)dt, (float)gamma1);
    return hipGetLastError() == hipSuccess ? CLAW_OK : CLAW_EDEVICE;
}

int claw2d_euler_radial_dq(const ClawGrid *grid, int precision, int num_eqn, const void *q, const void *aux,
                           void *dq, double dt, double gamma1, void *stream) {
    const int rc = radial_check(grid, precision, num_eqn, q, aux, dt);
    if (rc != CLAW_OK) return rc;
    if (!dq || dq == q) return CLAW_EINVAL;
    const GridDev g{grid->mx, grid->my, grid->ng, grid->ldq, grid->plane};
    const dim3 blocks((unsigned)(((long long)grid->mx * grid->my + 255) / 256));
    hipStream_t st = (hipStream_t)stream;
    if (precision == CLAW_F64)
        hipLaunchKernelGGL(euler_radial_dq_kernel<double>, blocks, dim3(256), 0, st, g, (const double *)q,
                           (const double *)aux, (double *)dq, dt, gamma1, num_eqn);
    else
        hipLaunchKernelGGL(euler_radial_dq_kernel<float>, blocks, dim3(256), 0, st, g, (const float *)q,
                           (const float *)aux, (float *)dq, (float)dt, (float)gamma1, num_eqn);
    return hipGetLastError() == hipSuccess ? CLAW_OK : CLAW_EDEVICE;
}

}  // extern "C"
