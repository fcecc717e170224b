/*
 * claw_oracle.c -- CPU restatement of the reference's 2D time-step kernels.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle for the HIP
 * kernels under pyclaw_amd/csrc.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product path never calls it.
 *
 * Every routine restates a Fortran routine of the reference with the same
 * loop order and the same floating-point operation order (built with
 * -O2 -ffp-contract=off so that no multiply-add is fused), so that results
 * can be compared bit for bit with the reference compiled from its own
 * sources (oracle/build_ref.sh) on the same inputs.
 *
 * Array layout follows the reference (Fortran column-major, interleaved
 * components): q(meqn, 1-ng:mx+ng, 1-ng:my+ng) is stored as
 *   q[((j+ng-1)*(mx+2ng) + (i+ng-1))*meqn + m]      (i, j 1-based, m 0-based)
 *
 * The Riemann solvers use the Fortran calling convention (every argument by
 * reference) so that they can also be handed to the reference's step2/flux2
 * as the rpn2/rpt2 procedure arguments.  Their scalar parameters live in
 * orc_cparam, the counterpart of the reference's /cparam/ common block.
 *
 * Parity status of each function (how it is pinned) is listed in the
 * repository's design notes; in short:
 *   philim/limiter/flux2/step2/step2ds, Euler rpn2/rpt2, SharpClaw
 *   flux2/flux1/weno5: bit-exact against the reference compiled by amdflang.
 *   Acoustics, vc-acoustics and shallow-water Riemann solvers: the reference
 *   takes them from the external clawpack.riemann package, absent here; they
 *   are written from the linear/Roe eigen-decompositions and pinned by
 *   closed-form invariants and the reference's end-to-end fixtures.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* Riemann-solver parameters (reference: common /cparam/)             */
/* ------------------------------------------------------------------ */
typedef struct {
    double rho, bulk, cc, zz;   /* acoustics            */
    double gamma, gamma1;       /* Euler                */
    double grav;                /* shallow water        */
    int efix;                   /* Euler: entropy fix   */
} orc_cparam_t;

orc_cparam_t orc_cparam = {1.0, 1.0, 1.0, 1.0, 1.4, 0.4, 1.0, 1};

void orc_set_cparam(double rho, double bulk, double cc, double zz,
                    double gamma, double gamma1, double grav, int efix)
{
    orc_cparam.rho = rho; orc_cparam.bulk = bulk; orc_cparam.cc = cc;
    orc_cparam.zz = zz; orc_cparam.gamma = gamma; orc_cparam.gamma1 = gamma1;
    orc_cparam.grav = grav; orc_cparam.efix = efix;
}

typedef void (*rpn2_fn)(const int *ixy, const int *maxm, const int *meqn,
                        const int *mwaves, const int *mbc, const int *mx,
                        const double *ql, const double *qr, const double *auxl,
                        const double *auxr, double *wave, double *s,
                        double *amdq, double *apdq, const int *num_aux);
typedef void (*rpt2_fn)(const int *ixy, const int *maxm, const int *meqn,
                        const int *mwaves, const int *mbc, const int *mx,
                        const double *ql, const double *qr, const double *aux1,
                        const double *aux2, const double *aux3, const int *imp,
                        const double *asdq, double *bmasdq, double *bpasdq,
                        const int *num_aux);

/* slice accessors: arrays dimensioned (n, 1-mbc:maxm+mbc), Fortran index i */
#define V(a, n, m, i)  (a)[((i) + mbc - 1) * (n) + (m)]
#define W(a, meqn, mwaves, m, mw, i) (a)[(((i) + mbc - 1) * (mwaves) + (mw)) * (meqn) + (m)]

static inline double dmax1(double a, double b) { return a > b ? a : b; }
static inline double dmin1(double a, double b) { return a < b ? a : b; }

/* ------------------------------------------------------------------ */
/* Acoustics (constant coefficients).  External to the reference      */
/* (clawpack.riemann rp2_acoustics); restated from the eigenvectors   */
/* r1=(-Z,1,0) s=-c, r2=(Z,1,0) s=+c of the normal flux Jacobian.      */
/* ------------------------------------------------------------------ */
void orc_rpn2_acoustics_(const int *ixy, const int *maxm, const int *meqn,
                         const int *mwaves, const int *pmbc, const int *pmx,
                         const double *ql, const double *qr, const double *auxl,
                         const double *auxr, double *wave, double *s,
                         double *amdq, double *apdq, const int *num_aux)
{
    (void)maxm; (void)auxl; (void)auxr; (void)num_aux;
    const int mbc = *pmbc, mx = *pmx, ne = *meqn, nw = *mwaves;
    const int mu = (*ixy == 1) ? 1 : 2, mv = (*ixy == 1) ? 2 : 1;
    const double zz = orc_cparam.zz, cc = orc_cparam.cc;
    for (int i = 2 - mbc; i <= mx + mbc; ++i) {
        double d0 = V(ql, ne, 0, i) - V(qr, ne, 0, i - 1);
        double d1 = V(ql, ne, mu, i) - V(qr, ne, mu, i - 1);
        double a1 = (-d0 + zz * d1) / (2.0 * zz);
        double a2 = (d0 + zz * d1) / (2.0 * zz);
        W(wave, ne, nw, 0, 0, i) = -a1 * zz;
        W(wave, ne, nw, mu, 0, i) = a1;
        W(wave, ne, nw, mv, 0, i) = 0.0;
        V(s, nw, 0, i) = -cc;
        W(wave, ne, nw, 0, 1, i) = a2 * zz;
        W(wave, ne, nw, mu, 1, i) = a2;
        W(wave, ne, nw, mv, 1, i) = 0.0;
        V(s, nw, 1, i) = cc;
    }
    for (int m = 0; m < ne; ++m)
        for (int i = 2 - mbc; i <= mx + mbc; ++i) {
            V(amdq, ne, m, i) = V(s, nw, 0, i) * W(wave, ne, nw, m, 0, i);
            V(apdq, ne, m, i) = V(s, nw, 1, i) * W(wave, ne, nw, m, 1, i);
        }
}

void orc_rpt2_acoustics_(const int *ixy, const int *maxm, const int *meqn,
                         const int *mwaves, const int *pmbc, const int *pmx,
                         const double *ql, const double *qr, const double *aux1,
                         const double *aux2, const double *aux3, const int *imp,
                         const double *asdq, double *bmasdq, double *bpasdq,
                         const int *num_aux)
{
    (void)maxm; (void)mwaves; (void)ql; (void)qr; (void)aux1; (void)aux2;
    (void)aux3; (void)imp; (void)num_aux;
    const int mbc = *pmbc, mx = *pmx, ne = *meqn;
    const int mu = (*ixy == 1) ? 1 : 2, mv = (*ixy == 1) ? 2 : 1;
    const double zz = orc_cparam.zz, cc = orc_cparam.cc;
    for (int i = 2 - mbc; i <= mx + mbc; ++i) {
        double a1 = (-V(asdq, ne, 0, i) + zz * V(asdq, ne, mv, i)) / (2.0 * zz);
        double a2 = (V(asdq, ne, 0, i) + zz * V(asdq, ne, mv, i)) / (2.0 * zz);
        V(bmasdq, ne, 0, i) = cc * a1 * zz;
        V(bmasdq, ne, mu, i) = 0.0;
        V(bmasdq, ne, mv, i) = -cc * a1;
        V(bpasdq, ne, 0, i) = cc * a2 * zz;
        V(bpasdq, ne, mu, i) = 0.0;
        V(bpasdq, ne, mv, i) = cc * a2;
    }
}

/* ------------------------------------------------------------------ */
/* Variable-coefficient acoustics, aux = (rho, c) per cell.  External  */
/* to the reference (clawpack.riemann rp2_vc_acoustics); restated from */
/* the interface-impedance decomposition (left-going wave in the left  */
/* medium, right-going wave in the right medium).                      */
/* ------------------------------------------------------------------ */
void orc_rpn2_vc_acoustics_(const int *ixy, const int *maxm, const int *meqn,
                            const int *mwaves, const int *pmbc, const int *pmx,
                            const double *ql, const double *qr, const double *auxl,
                            const double *auxr, double *wave, double *s,
                            double *amdq, double *apdq, const int *num_aux)
{
    (void)maxm;
    const int mbc = *pmbc, mx = *pmx, ne = *meqn, nw = *mwaves, na = *num_aux;
    const int mu = (*ixy == 1) ? 1 : 2, mv = (*ixy == 1) ? 2 : 1;
    for (int i = 2 - mbc; i <= mx + mbc; ++i) {
        double d0 = V(ql, ne, 0, i) - V(qr, ne, 0, i - 1);
        double d1 = V(ql, ne, mu, i) - V(qr, ne, mu, i - 1);
        double zi = V(auxl, na, 0, i) * V(auxl, na, 1, i);
        double zim = V(auxr, na, 0, i - 1) * V(auxr, na, 1, i - 1);
        double a1 = (-d0 + zi * d1) / (zim + zi);
        double a2 = (d0 + zim * d1) / (zim + zi);
        W(wave, ne, nw, 0, 0, i) = -a1 * zim;
        W(wave, ne, nw, mu, 0, i) = a1;
        W(wave, ne, nw, mv, 0, i) = 0.0;
        V(s, nw, 0, i) = -V(auxr, na, 1, i - 1);
        W(wave, ne, nw, 0, 1, i) = a2 * zi;
        W(wave, ne, nw, mu, 1, i) = a2;
        W(wave, ne, nw, mv, 1, i) = 0.0;
        V(s, nw, 1, i) = V(auxl, na, 1, i);
    }
    for (int m = 0; m < ne; ++m)
        for (int i = 2 - mbc; i <= mx + mbc; ++i) {
            V(amdq, ne, m, i) = V(s, nw, 0, i) * W(wave, ne, nw, m, 0, i);
            V(apdq, ne, m, i) = V(s, nw, 1, i) * W(wave, ne, nw, m, 1, i);
        }
}

void orc_rpt2_vc_acoustics_(const int *ixy, const int *maxm, const int *meqn,
                            const int *mwaves, const int *pmbc, const int *pmx,
                            const double *ql, const double *qr, const double *aux1,
                            const double *aux2, const double *aux3, const int *imp,
                            const double *asdq, double *bmasdq, double *bpasdq,
                            const int *num_aux)
{
    (void)maxm; (void)mwaves; (void)ql; (void)qr;
    const int mbc = *pmbc, mx = *pmx, ne = *meqn, na = *num_aux;
    const int mu = (*ixy == 1) ? 1 : 2, mv = (*ixy == 1) ? 2 : 1;
    for (int i = 2 - mbc; i <= mx + mbc; ++i) {
        int i1 = (*imp == 1) ? i - 1 : i;
        double cm = V(aux1, na, 1, i1);
        double cp = V(aux3, na, 1, i1);
        double zm = V(aux1, na, 0, i1) * V(aux1, na, 1, i1);
        double zz = V(aux2, na, 0, i1) * V(aux2, na, 1, i1);
        double zp = V(aux3, na, 0, i1) * V(aux3, na, 1, i1);
        double a1 = (-V(asdq, ne, 0, i) + V(asdq, ne, mv, i) * zz) / (zm + zz);
        double a2 = (V(asdq, ne, 0, i) + V(asdq, ne, mv, i) * zz) / (zz + zp);
        V(bmasdq, ne, 0, i) = cm * a1 * zm;
        V(bmasdq, ne, mu, i) = 0.0;
        V(bmasdq, ne, mv, i) = -cm * a1;
        V(bpasdq, ne, 0, i) = cp * a2 * zp;
        V(bpasdq, ne, mu, i) = 0.0;
        V(bpasdq, ne, mv, i) = cp * a2;
    }
}

/* ------------------------------------------------------------------ */
/* Euler 5-wave Roe solver with entropy fix.                          */
/* Follows development/rp_approaches/rpn2_euler_5wave.f:5-302 and     */
/* rpt2_euler_5wave.f (transverse) -- same expressions, same order.   */
/* The Roe averages are kept per slice (reference: common /comroe/).  */
/* ------------------------------------------------------------------ */
/* per thread: the multi-threaded sweeps run one slice per thread at a time */
static __thread double *roe_u2v2, *roe_u, *roe_v, *roe_enth, *roe_a, *roe_g1a2, *roe_euv;
static __thread int roe_cap;
static __thread int roe_off;

static void roe_reserve(int n, int mbc)
{
    if (n > roe_cap) {
        free(roe_u2v2); free(roe_u); free(roe_v); free(roe_enth);
        free(roe_a); free(roe_g1a2); free(roe_euv);
        roe_u2v2 = malloc(sizeof(double) * n); roe_u = malloc(sizeof(double) * n);
        roe_v = malloc(sizeof(double) * n); roe_enth = malloc(sizeof(double) * n);
        roe_a = malloc(sizeof(double) * n); roe_g1a2 = malloc(sizeof(double) * n);
        roe_euv = malloc(sizeof(double) * n);
        roe_cap = n;
    }
    roe_off = mbc;
}
#define RI(i) ((i) + roe_off)

void orc_rpn2_euler5_(const int *ixy, const int *maxm, const int *meqn,
                      const int *mwaves, const int *pmbc, const int *pmx,
                      const double *ql, const double *qr, const double *auxl,
                      const double *auxr, double *wave, double *s,
                      double *amdq, double *apdq, const int *num_aux)
{
    (void)auxl; (void)auxr; (void)num_aux;
    const int mbc = *pmbc, mx = *pmx, ne = *meqn, nw = *mwaves;
    const double gamma = orc_cparam.gamma, gamma1 = orc_cparam.gamma1;
    const int mu = (*ixy == 1) ? 1 : 2, mv = (*ixy == 1) ? 2 : 1;
    roe_reserve(*maxm + 2 * mbc + 2, mbc);

    for (int i = 2 - mbc; i <= mx + mbc; ++i) {
        double rhsqrtl = sqrt(V(qr, ne, 0, i - 1));
        double rhsqrtr = sqrt(V(ql, ne, 0, i));
        double pl = gamma1 * (V(qr, ne, 3, i - 1) - 0.5 * (V(qr, ne, 1, i - 1) * V(qr, ne, 1, i - 1) +
                    V(qr, ne, 2, i - 1) * V(qr, ne, 2, i - 1)) / V(qr, ne, 0, i - 1));
        double pr = gamma1 * (V(ql, ne, 3, i) - 0.5 * (V(ql, ne, 1, i) * V(ql, ne, 1, i) +
                    V(ql, ne, 2, i) * V(ql, ne, 2, i)) / V(ql, ne, 0, i));
        double rhsq2 = rhsqrtl + rhsqrtr;
        double u = (V(qr, ne, mu, i - 1) / rhsqrtl + V(ql, ne, mu, i) / rhsqrtr) / rhsq2;
        double v = (V(qr, ne, mv, i - 1) / rhsqrtl + V(ql, ne, mv, i) / rhsqrtr) / rhsq2;
        double enth = ((V(qr, ne, 3, i - 1) + pl) / rhsqrtl + (V(ql, ne, 3, i) + pr) / rhsqrtr) / rhsq2;
        double u2v2 = u * u + v * v;
        double a2 = gamma1 * (enth - 0.5 * u2v2);
        double a = sqrt(a2);
        roe_u[RI(i)] = u; roe_v[RI(i)] = v; roe_enth[RI(i)] = enth;
        roe_u2v2[RI(i)] = u2v2; roe_a[RI(i)] = a;
        roe_g1a2[RI(i)] = gamma1 / a2;
        roe_euv[RI(i)] = enth - u2v2;
    }

    for (int i = 2 - mbc; i <= mx + mbc; ++i) {
        double u = roe_u[RI(i)], v = roe_v[RI(i)], a = roe_a[RI(i)], enth = roe_enth[RI(i)];
        double d0 = V(ql, ne, 0, i) - V(qr, ne, 0, i - 1);
        double d1 = V(ql, ne, mu, i) - V(qr, ne, mu, i - 1);
        double d2 = V(ql, ne, mv, i) - V(qr, ne, mv, i - 1);
        double d3 = V(ql, ne, 3, i) - V(qr, ne, 3, i - 1);
        double a3 = roe_g1a2[RI(i)] * (roe_euv[RI(i)] * d0 + u * d1 + v * d2 - d3);
        double a2 = d2 - v * d0;
        double a4 = (d1 + (a - u) * d0 - a * a3) / (2.0 * a);
        double a1 = d0 - a3 - a4;
        W(wave, ne, nw, 0, 0, i) = a1;
        W(wave, ne, nw, mu, 0, i) = a1 * (u - a);
        W(wave, ne, nw, mv, 0, i) = a1 * v;
        W(wave, ne, nw, 3, 0, i) = a1 * (enth - u * a);
        W(wave, ne, nw, 4, 0, i) = 0.0;
        V(s, nw, 0, i) = u - a;

        W(wave, ne, nw, 0, 1, i) = 0.0;
        W(wave, ne, nw, mu, 1, i) = 0.0;
        W(wave, ne, nw, mv, 1, i) = a2;
        W(wave, ne, nw, 3, 1, i) = a2 * v;
        W(wave, ne, nw, 4, 1, i) = 0.0;
        V(s, nw, 1, i) = u;

        W(wave, ne, nw, 0, 2, i) = a3;
        W(wave, ne, nw, mu, 2, i) = a3 * u;
        W(wave, ne, nw, mv, 2, i) = a3 * v;
        W(wave, ne, nw, 3, 2, i) = a3 * 0.5 * roe_u2v2[RI(i)];
        W(wave, ne, nw, 4, 2, i) = 0.0;
        V(s, nw, 2, i) = u;

        W(wave, ne, nw, 0, 3, i) = a4;
        W(wave, ne, nw, mu, 3, i) = a4 * (u + a);
        W(wave, ne, nw, mv, 3, i) = a4 * v;
        W(wave, ne, nw, 3, 3, i) = a4 * (enth + u * a);
        W(wave, ne, nw, 4, 3, i) = 0.0;
        V(s, nw, 3, i) = u + a;

        W(wave, ne, nw, 0, 4, i) = 0.0;
        W(wave, ne, nw, mu, 4, i) = 0.0;
        W(wave, ne, nw, mv, 4, i) = 0.0;
        W(wave, ne, nw, 3, 4, i) = 0.0;
        W(wave, ne, nw, 4, 4, i) = V(ql, ne, 4, i) - V(qr, ne, 4, i - 1);
        V(s, nw, 4, i) = u;
    }

    if (!orc_cparam.efix) {
        for (int m = 0; m < ne; ++m)
            for (int i = 2 - mbc; i <= mx + mbc; ++i) {
                V(amdq, ne, m, i) = 0.0;
                V(apdq, ne, m, i) = 0.0;
                for (int mw = 0; mw < nw; ++mw) {
                    if (V(s, nw, mw, i) < 0.0)
                        V(amdq, ne, m, i) = V(amdq, ne, m, i) + V(s, nw, mw, i) * W(wave, ne, nw, m, mw, i);
                    else
                        V(apdq, ne, m, i) = V(apdq, ne, m, i) + V(s, nw, mw, i) * W(wave, ne, nw, m, mw, i);
                }
            }
        return;
    }

    for (int i = 2 - mbc; i <= mx + mbc; ++i) {
        double sfract;
        /* 1-wave */
        double rhoim1 = V(qr, ne, 0, i - 1);
        double pim1 = gamma1 * (V(qr, ne, 3, i - 1) - 0.5 * (V(qr, ne, mu, i - 1) * V(qr, ne, mu, i - 1) +
                      V(qr, ne, mv, i - 1) * V(qr, ne, mv, i - 1)) / rhoim1);
        double cim1 = sqrt(gamma * pim1 / rhoim1);
        double s0 = V(qr, ne, mu, i - 1) / rhoim1 - cim1;
        if (s0 >= 0.0 && V(s, nw, 0, i) > 0.0) {
            for (int m = 0; m < ne; ++m) V(amdq, ne, m, i) = 0.0;
            continue;
        }
        double rho1 = V(qr, ne, 0, i - 1) + W(wave, ne, nw, 0, 0, i);
        double rhou1 = V(qr, ne, mu, i - 1) + W(wave, ne, nw, mu, 0, i);
        double rhov1 = V(qr, ne, mv, i - 1) + W(wave, ne, nw, mv, 0, i);
        double en1 = V(qr, ne, 3, i - 1) + W(wave, ne, nw, 3, 0, i);
        double p1 = gamma1 * (en1 - 0.5 * (rhou1 * rhou1 + rhov1 * rhov1) / rho1);
        double c1 = sqrt(gamma * p1 / rho1);
        double s1 = rhou1 / rho1 - c1;
        if (s0 < 0.0 && s1 > 0.0)
            sfract = s0 * (s1 - V(s, nw, 0, i)) / (s1 - s0);
        else if (V(s, nw, 0, i) < 0.0)
            sfract = V(s, nw, 0, i);
        else
            sfract = 0.0;
        for (int m = 0; m < ne; ++m) V(amdq, ne, m, i) = sfract * W(wave, ne, nw, m, 0, i);
        /* contact */
        if (V(s, nw, 1, i) >= 0.0) continue;
        for (int m = 0; m < ne; ++m) {
            V(amdq, ne, m, i) = V(amdq, ne, m, i) + V(s, nw, 1, i) * W(wave, ne, nw, m, 1, i);
            V(amdq, ne, m, i) = V(amdq, ne, m, i) + V(s, nw, 2, i) * W(wave, ne, nw, m, 2, i);
            V(amdq, ne, m, i) = V(amdq, ne, m, i) + V(s, nw, 4, i) * W(wave, ne, nw, m, 4, i);
        }
        /* 4-wave */
        double rhoi = V(ql, ne, 0, i);
        double pi = gamma1 * (V(ql, ne, 3, i) - 0.5 * (V(ql, ne, mu, i) * V(ql, ne, mu, i) +
                    V(ql, ne, mv, i) * V(ql, ne, mv, i)) / rhoi);
        double ci = sqrt(gamma * pi / rhoi);
        double s3 = V(ql, ne, mu, i) / rhoi + ci;
        double rho2 = V(ql, ne, 0, i) - W(wave, ne, nw, 0, 3, i);
        double rhou2 = V(ql, ne, mu, i) - W(wave, ne, nw, mu, 3, i);
        double rhov2 = V(ql, ne, mv, i) - W(wave, ne, nw, mv, 3, i);
        double en2 = V(ql, ne, 3, i) - W(wave, ne, nw, 3, 3, i);
        double p2 = gamma1 * (en2 - 0.5 * (rhou2 * rhou2 + rhov2 * rhov2) / rho2);
        double c2 = sqrt(gamma * p2 / rho2);
        double s2 = rhou2 / rho2 + c2;
        if (s2 < 0.0 && s3 > 0.0)
            sfract = s2 * (s3 - V(s, nw, 3, i)) / (s3 - s2);
        else if (V(s, nw, 3, i) < 0.0)
            sfract = V(s, nw, 3, i);
        else
            continue;
        for (int m = 0; m < ne; ++m)
            V(amdq, ne, m, i) = V(amdq, ne, m, i) + sfract * W(wave, ne, nw, m, 3, i);
    }
    for (int m = 0; m < ne; ++m)
        for (int i = 2 - mbc; i <= mx + mbc; ++i) {
            double df = 0.0;
            for (int mw = 0; mw < nw; ++mw) df = df + V(s, nw, mw, i) * W(wave, ne, nw, m, mw, i);
            V(apdq, ne, m, i) = df - V(amdq, ne, m, i);
        }
}

void orc_rpt2_euler5_(const int *ixy, const int *maxm, const int *meqn,
                      const int *mwaves, const int *pmbc, const int *pmx,
                      const double *ql, const double *qr, const double *aux1,
                      const double *aux2, const double *aux3, const int *imp,
                      const double *asdq, double *bmasdq, double *bpasdq,
                      const int *num_aux)
{
    (void)maxm; (void)mwaves; (void)ql; (void)qr; (void)aux1; (void)aux2;
    (void)aux3; (void)imp; (void)num_aux;
    const int mbc = *pmbc, mx = *pmx, ne = *meqn;
    const int mu = (*ixy == 1) ? 1 : 2, mv = (*ixy == 1) ? 2 : 1;
    double waveb[4][5], sb[4];
    for (int i = 2 - mbc; i <= mx + mbc; ++i) {
        double u = roe_u[RI(i)], v = roe_v[RI(i)], a = roe_a[RI(i)], enth = roe_enth[RI(i)];
        double a3 = roe_g1a2[RI(i)] * (roe_euv[RI(i)] * V(asdq, ne, 0, i) + u * V(asdq, ne, mu, i) +
                    v * V(asdq, ne, mv, i) - V(asdq, ne, 3, i));
        double a2 = V(asdq, ne, mu, i) - u * V(asdq, ne, 0, i);
        double a4 = (V(asdq, ne, mv, i) + (a - v) * V(asdq, ne, 0, i) - a * a3) / (2.0 * a);
        double a1 = V(asdq, ne, 0, i) - a3 - a4;

        waveb[0][0] = a1; waveb[0][mu] = a1 * u; waveb[0][mv] = a1 * (v - a);
        waveb[0][3] = a1 * (enth - v * a); waveb[0][4] = 0.0; sb[0] = v - a;

        waveb[1][0] = a3; waveb[1][mu] = a3 * u + a2; waveb[1][mv] = a3 * v;
        waveb[1][3] = a3 * 0.5 * roe_u2v2[RI(i)] + a2 * u; waveb[1][4] = 0.0; sb[1] = v;

        waveb[2][0] = a4; waveb[2][mu] = a4 * u; waveb[2][mv] = a4 * (v + a);
        waveb[2][3] = a4 * (enth + v * a); waveb[2][4] = 0.0; sb[2] = v + a;

        waveb[3][0] = 0.0; waveb[3][mu] = 0.0; waveb[3][mv] = 0.0; waveb[3][3] = 0.0;
        waveb[3][4] = V(asdq, ne, 4, i); sb[3] = v;

        for (int m = 0; m < ne; ++m) {
            V(bmasdq, ne, m, i) = 0.0;
            V(bpasdq, ne, m, i) = 0.0;
            for (int mw = 0; mw < 4; ++mw) {
                V(bmasdq, ne, m, i) = V(bmasdq, ne, m, i) + dmin1(sb[mw], 0.0) * waveb[mw][m];
                V(bpasdq, ne, m, i) = V(bpasdq, ne, m, i) + dmax1(sb[mw], 0.0) * waveb[mw][m];
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* Shallow water, Roe solver with entropy fix.  External to the        */
/* reference (clawpack.riemann rp2_shallow_roe_with_efix).  Restated   */
/* from the Roe linearisation (h-bar arithmetic mean, sqrt(h)-weighted */
/* velocities) with the transonic-rarefaction fix structured like the  */
/* reference's Euler solver.  Parity unpinned against the reference.   */
/* ------------------------------------------------------------------ */
void orc_rpn2_shallow_roe_(const int *ixy, const int *maxm, const int *meqn,
                           const int *mwaves, const int *pmbc, const int *pmx,
                           const double *ql, const double *qr, const double *auxl,
                           const double *auxr, double *wave, double *s,
                           double *amdq, double *apdq, const int *num_aux)
{
    (void)maxm; (void)auxl; (void)auxr; (void)num_aux;
    const int mbc = *pmbc, mx = *pmx, ne = *meqn, nw = *mwaves;
    const int mu = (*ixy == 1) ? 1 : 2, mv = (*ixy == 1) ? 2 : 1;
    const double grav = orc_cparam.grav;
    for (int i = 2 - mbc; i <= mx + mbc; ++i) {
        double hl = V(qr, ne, 0, i - 1), hr = V(ql, ne, 0, i);
        double hbar = (hl + hr) * 0.5;
        double hsqrtl = sqrt(hl), hsqrtr = sqrt(hr);
        double hsq2 = hsqrtl + hsqrtr;
        double u = (V(qr, ne, mu, i - 1) / hsqrtl + V(ql, ne, mu, i) / hsqrtr) / hsq2;
        double v = (V(qr, ne, mv, i - 1) / hsqrtl + V(ql, ne, mv, i) / hsqrtr) / hsq2;
        double a = sqrt(grav * hbar);
        double d0 = hr - hl;
        double d1 = V(ql, ne, mu, i) - V(qr, ne, mu, i - 1);
        double d2 = V(ql, ne, mv, i) - V(qr, ne, mv, i - 1);
        double a1 = ((u + a) * d0 - d1) / (2.0 * a);
        double a2 = -v * d0 + d2;
        double a3 = (-(u - a) * d0 + d1) / (2.0 * a);
        W(wave, ne, nw, 0, 0, i) = a1;
        W(wave, ne, nw, mu, 0, i) = a1 * (u - a);
        W(wave, ne, nw, mv, 0, i) = a1 * v;
        V(s, nw, 0, i) = u - a;
        W(wave, ne, nw, 0, 1, i) = 0.0;
        W(wave, ne, nw, mu, 1, i) = 0.0;
        W(wave, ne, nw, mv, 1, i) = a2;
        V(s, nw, 1, i) = u;
        W(wave, ne, nw, 0, 2, i) = a3;
        W(wave, ne, nw, mu, 2, i) = a3 * (u + a);
        W(wave, ne, nw, mv, 2, i) = a3 * v;
        V(s, nw, 2, i) = u + a;
    }
    for (int i = 2 - mbc; i <= mx + mbc; ++i) {
        double sfract;
        double hl = V(qr, ne, 0, i - 1);
        double s0 = V(qr, ne, mu, i - 1) / hl - sqrt(grav * hl);
        if (s0 >= 0.0 && V(s, nw, 0, i) > 0.0) {
            for (int m = 0; m < ne; ++m) V(amdq, ne, m, i) = 0.0;
            goto done;
        }
        {
            double h1 = hl + W(wave, ne, nw, 0, 0, i);
            double hu1 = V(qr, ne, mu, i - 1) + W(wave, ne, nw, mu, 0, i);
            double s1 = hu1 / h1 - sqrt(grav * h1);
            if (s0 < 0.0 && s1 > 0.0)
                sfract = s0 * (s1 - V(s, nw, 0, i)) / (s1 - s0);
            else if (V(s, nw, 0, i) < 0.0)
                sfract = V(s, nw, 0, i);
            else
                sfract = 0.0;
            for (int m = 0; m < ne; ++m) V(amdq, ne, m, i) = sfract * W(wave, ne, nw, m, 0, i);
        }
        if (V(s, nw, 1, i) >= 0.0) goto done;
        for (int m = 0; m < ne; ++m)
            V(amdq, ne, m, i) = V(amdq, ne, m, i) + V(s, nw, 1, i) * W(wave, ne, nw, m, 1, i);
        {
            double hr = V(ql, ne, 0, i);
            double s03 = V(ql, ne, mu, i) / hr + sqrt(grav * hr);
            double h3 = hr - W(wave, ne, nw, 0, 2, i);
            double hu3 = V(ql, ne, mu, i) - W(wave, ne, nw, mu, 2, i);
            double s3 = hu3 / h3 + sqrt(grav * h3);
            if (s3 < 0.0 && s03 > 0.0)
                sfract = s3 * (s03 - V(s, nw, 2, i)) / (s03 - s3);
            else if (V(s, nw, 2, i) < 0.0)
                sfract = V(s, nw, 2, i);
            else
                goto done;
            for (int m = 0; m < ne; ++m)
                V(amdq, ne, m, i) = V(amdq, ne, m, i) + sfract * W(wave, ne, nw, m, 2, i);
        }
    done:;
    }
    for (int m = 0; m < ne; ++m)
        for (int i = 2 - mbc; i <= mx + mbc; ++i) {
            double df = 0.0;
            for (int mw = 0; mw < nw; ++mw) df = df + V(s, nw, mw, i) * W(wave, ne, nw, m, mw, i);
            V(apdq, ne, m, i) = df - V(amdq, ne, m, i);
        }
}

void orc_rpt2_shallow_roe_(const int *ixy, const int *maxm, const int *meqn,
                           const int *mwaves, const int *pmbc, const int *pmx,
                           const double *ql, const double *qr, const double *aux1,
                           const double *aux2, const double *aux3, const int *imp,
                           const double *asdq, double *bmasdq, double *bpasdq,
                           const int *num_aux)
{
    (void)maxm; (void)mwaves; (void)aux1; (void)aux2; (void)aux3; (void)imp; (void)num_aux;
    const int mbc = *pmbc, mx = *pmx, ne = *meqn;
    const int mu = (*ixy == 1) ? 1 : 2, mv = (*ixy == 1) ? 2 : 1;
    const double grav = orc_cparam.grav;
    for (int i = 2 - mbc; i <= mx + mbc; ++i) {
        double hl = V(qr, ne, 0, i - 1), hr = V(ql, ne, 0, i);
        double hbar = (hl + hr) * 0.5;
        double hsqrtl = sqrt(hl), hsqrtr = sqrt(hr);
        double hsq2 = hsqrtl + hsqrtr;
        double u = (V(qr, ne, mu, i - 1) / hsqrtl + V(ql, ne, mu, i) / hsqrtr) / hsq2;
        double v = (V(qr, ne, mv, i - 1) / hsqrtl + V(ql, ne, mv, i) / hsqrtr) / hsq2;
        double a = sqrt(grav * hbar);
        double q0 = V(asdq, ne, 0, i), qu = V(asdq, ne, mu, i), qv = V(asdq, ne, mv, i);
        double a1 = (0.5 / a) * ((v + a) * q0 - qv);
        double a2 = qu - u * q0;
        double a3 = (0.5 / a) * (-(v - a) * q0 + qv);
        double wb[3][3], sb[3];
        wb[0][0] = a1; wb[0][mu] = a1 * u; wb[0][mv] = a1 * (v - a); sb[0] = v - a;
        wb[1][0] = 0.0; wb[1][mu] = a2; wb[1][mv] = 0.0; sb[1] = v;
        wb[2][0] = a3; wb[2][mu] = a3 * u; wb[2][mv] = a3 * (v + a); sb[2] = v + a;
        for (int m = 0; m < ne; ++m) {
            V(bmasdq, ne, m, i) = 0.0;
            V(bpasdq, ne, m, i) = 0.0;
            for (int mw = 0; mw < 3; ++mw) {
                V(bmasdq, ne, m, i) = V(bmasdq, ne, m, i) + dmin1(sb[mw], 0.0) * wb[mw][m];
                V(bpasdq, ne, m, i) = V(bpasdq, ne, m, i) + dmax1(sb[mw], 0.0) * wb[mw][m];
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* philim / limiter: classic/philim.f90:4-57, classic/limiter.f90:4-61 */
/* ------------------------------------------------------------------ */
double orc_philim(double a, double b, int meth)
{
    double r = b / a;
    switch (meth) {
    case 2: /* superbee */
        return dmax1(dmax1(0.0, dmin1(1.0, 2.0 * r)), dmin1(2.0, r));
    case 3: /* van Leer */
        return (r + fabs(r)) / (1.0 + fabs(r));
    case 4: { /* MC */
        double c = (1.0 + r) / 2.0;
        return dmax1(0.0, dmin1(dmin1(c, 2.0), 2.0 * r));
    }
    case 5: /* Beam-Warming */
        return r;
    default: /* minmod; the computed GOTO falls through here for ids outside 1..5 */
        return dmax1(0.0, dmin1(1.0, r));
    }
}

static void limiter(int maxm, int meqn, int mwaves, int mbc, int mx, double *wave,
                    const double *s, const int *mthlim)
{
    (void)maxm;
    for (int mw = 0; mw < mwaves; ++mw) {
        if (mthlim[mw] == 0) continue;
        double dotr = 0.0;
        for (int i = 0; i <= mx + 1; ++i) {
            double wnorm2 = 0.0;
            double dotl = dotr;
            dotr = 0.0;
            for (int m = 0; m < meqn; ++m) {
                double w = W(wave, meqn, mwaves, m, mw, i);
                wnorm2 = wnorm2 + w * w;
                dotr = dotr + w * W(wave, meqn, mwaves, m, mw, i + 1);
            }
            if (i == 0) continue;
            if (wnorm2 == 0.0) continue;
            double wlimitr;
            if (V(s, mwaves, mw, i) > 0.0)
                wlimitr = orc_philim(wnorm2, dotl, mthlim[mw]);
            else
                wlimitr = orc_philim(wnorm2, dotr, mthlim[mw]);
            for (int m = 0; m < meqn; ++m)
                W(wave, meqn, mwaves, m, mw, i) = wlimitr * W(wave, meqn, mwaves, m, mw, i);
        }
    }
}

/* ------------------------------------------------------------------ */
/* flux2: classic/flux2.f90:5-210                                     */
/* method[] is the reference's method(1:7) (0-based here).           */
/* ------------------------------------------------------------------ */
typedef struct {
    double *wave, *s, *amdq, *apdq, *cqxx, *bmasdq, *bpasdq;
} flux_work_t;

static void flux2(int ixy, int maxm, int meqn, int mwaves, int num_aux, int mbc, int mx,
                  const double *q1d, const double *dtdx1d, const double *aux1,
                  const double *aux2, const double *aux3, const int *method,
                  const int *mthlim, double *qadd, double *fadd, double *gadd,
                  double *cfl1d, flux_work_t *wk, rpn2_fn rpn2, rpt2_fn rpt2, int use_fwave)
{
    double *wave = wk->wave, *s = wk->s, *amdq = wk->amdq, *apdq = wk->apdq;
    double *cqxx = wk->cqxx, *bmasdq = wk->bmasdq, *bpasdq = wk->bpasdq;
    const int ne = meqn;
    /* gadd(meqn, 2, 1-mbc:maxm+mbc) */
#define G(m, k, i) gadd[(((i) + mbc - 1) * 2 + (k)) * ne + (m)]
#define DT(i) dtdx1d[(i) + mbc - 1]
    int limit = 0;
    for (int mw = 0; mw < mwaves; ++mw) if (mthlim[mw] > 0) limit = 1;

    for (int i = 1 - mbc; i <= mx + mbc; ++i)
        for (int m = 0; m < ne; ++m) {
            V(qadd, ne, m, i) = 0.0; V(fadd, ne, m, i) = 0.0;
            G(m, 0, i) = 0.0; G(m, 1, i) = 0.0;
        }

    int ixy_ = ixy, maxm_ = maxm, meqn_ = meqn, mwaves_ = mwaves, mbc_ = mbc, mx_ = mx, na_ = num_aux;
    rpn2(&ixy_, &maxm_, &meqn_, &mwaves_, &mbc_, &mx_, q1d, q1d, aux2, aux2, wave, s, amdq, apdq, &na_);

    /* forall: every right-hand side of a statement is evaluated before any assignment */
    for (int i = 1; i <= mx + 1; ++i)
        for (int m = 0; m < ne; ++m)
            V(qadd, ne, m, i) = V(qadd, ne, m, i) - DT(i) * V(apdq, ne, m, i);
    for (int i = 1; i <= mx + 1; ++i)
        for (int m = 0; m < ne; ++m)
            V(qadd, ne, m, i - 1) = V(qadd, ne, m, i - 1) - DT(i - 1) * V(amdq, ne, m, i);

    double c = 0.0;
    for (int mw = 0; mw < mwaves; ++mw)
        for (int i = 1; i <= mx + 1; ++i)
            c = dmax1(dmax1(c, DT(i) * V(s, mwaves, mw, i)), -DT(i - 1) * V(s, mwaves, mw, i));
    *cfl1d = c;

    if (method[1] != 1) {
        if (limit) limiter(maxm, ne, mwaves, mbc, mx, wave, s, mthlim);
        for (int i = 2 - mbc; i <= mx + mbc; ++i) {
            double dtdxave = 0.5 * (DT(i - 1) + DT(i));
            for (int m = 0; m < ne; ++m) {
                double cq = 0.0;
                for (int mw = 0; mw < mwaves; ++mw) {
                    double sv = V(s, mwaves, mw, i);
                    double sg = use_fwave ? (sv >= 0.0 ? 1.0 : -1.0) : fabs(sv);
                    cq = cq + sg * (1.0 - fabs(sv) * dtdxave) * W(wave, ne, mwaves, m, mw, i);
                }
                V(cqxx, ne, m, i) = cq;
                V(fadd, ne, m, i) = V(fadd, ne, m, i) + 0.5 * cq;
            }
        }
    }

    if (method[2] <= 0) return;

    if (method[1] > 1 && method[2] == 2) {
        for (int i = 1; i <= mx + 1; ++i)
            for (int m = 0; m < ne; ++m) {
                V(amdq, ne, m, i) = V(amdq, ne, m, i) + V(cqxx, ne, m, i);
                V(apdq, ne, m, i) = V(apdq, ne, m, i) - V(cqxx, ne, m, i);
            }
    }

    int one = 1, two = 2;
    rpt2(&ixy_, &maxm_, &meqn_, &mwaves_, &mbc_, &mx_, q1d, q1d, aux1, aux2, aux3, &one,
         amdq, bmasdq, bpasdq, &na_);
    for (int i = 1; i <= mx + 1; ++i)
        for (int m = 0; m < ne; ++m) {
            G(m, 0, i - 1) = G(m, 0, i - 1) - 0.5 * DT(i - 1) * V(bmasdq, ne, m, i);
            G(m, 1, i - 1) = G(m, 1, i - 1) - 0.5 * DT(i - 1) * V(bpasdq, ne, m, i);
        }
    rpt2(&ixy_, &maxm_, &meqn_, &mwaves_, &mbc_, &mx_, q1d, q1d, aux1, aux2, aux3, &two,
         apdq, bmasdq, bpasdq, &na_);
    for (int i = 1; i <= mx + 1; ++i)
        for (int m = 0; m < ne; ++m) {
            G(m, 0, i) = G(m, 0, i) - 0.5 * DT(i) * V(bmasdq, ne, m, i);
            G(m, 1, i) = G(m, 1, i) - 0.5 * DT(i) * V(bpasdq, ne, m, i);
        }
#undef G
#undef DT
}

/* ------------------------------------------------------------------ */
/* step2 (unsplit) and step2ds (split): classic/step2.f90:2-251,      */
/* classic/step2ds.f90:2-258.  qold/qnew/aux are full padded arrays.  */
/* ------------------------------------------------------------------ */
#define QF(a, ne, i, j) ((a) + ((size_t)((j) + mbc - 1) * (size_t)(mx + 2 * mbc) + (size_t)((i) + mbc - 1)) * (ne))

static void alloc_work(flux_work_t *wk, int maxm, int mbc, int meqn, int mwaves)
{
    size_t n = (size_t)(maxm + 2 * mbc);
    wk->wave = calloc(n * meqn * mwaves, sizeof(double));
    wk->s = calloc(n * mwaves, sizeof(double));
    wk->amdq = calloc(n * meqn, sizeof(double));
    wk->apdq = calloc(n * meqn, sizeof(double));
    wk->cqxx = calloc(n * meqn, sizeof(double));
    wk->bmasdq = calloc(n * meqn, sizeof(double));
    wk->bpasdq = calloc(n * meqn, sizeof(double));
}
static void free_work(flux_work_t *wk)
{
    free(wk->wave); free(wk->s); free(wk->amdq); free(wk->apdq);
    free(wk->cqxx); free(wk->bmasdq); free(wk->bpasdq);
}

/* Shared sweep driver.  mode 0: unsplit step2; 1: step2ds ids=1; 2: step2ds ids=2.
 * dir 1 runs the x sweeps (slices are rows j, step2.f90:92-165 / step2ds.f90:96-167) over
 * slices s0..s1; dir 2 the y sweeps (slices are columns i, step2.f90:173-248 /
 * step2ds.f90:177-252).  Only updates of rows (dir 1) or columns (dir 2) w0..w1 are
 * applied, so that disjoint write ranges can run on different threads: a range that
 * also runs the one slice on each side of it receives, for each of its cells, the
 * contributions of the slices j-1, j, j+1 in the reference's order (step2ds slices
 * update only their own row or column). */
typedef struct {
    int mode, maxm, meqn, mwaves, num_aux, mbc, mx, my;
    const double *qold; double *qnew; const double *aux;
    double dx, dy, dt;
    const int *method, *mthlim;
    int use_fwave;
    rpn2_fn rpn2; rpt2_fn rpt2;
} sweep_args_t;

static double sweep_dir(const sweep_args_t *A, int dir, int s0, int s1, int w0, int w1)
{
    const int mode = A->mode, maxm = A->maxm, ne = A->meqn, mwaves = A->mwaves, num_aux = A->num_aux;
    const int mbc = A->mbc, mx = A->mx, my = A->my;
    const double *qold = A->qold, *aux = A->aux;
    double *qnew = A->qnew;
    const int na = num_aux > 0 ? num_aux : 1;
    size_t n = (size_t)(maxm + 2 * mbc);
    double *q1d = calloc(n * ne, sizeof(double));
    double *qadd = calloc(n * ne, sizeof(double));
    double *fadd = calloc(n * ne, sizeof(double));
    double *gadd = calloc(n * ne * 2, sizeof(double));
    double *dtdx1d = calloc(n, sizeof(double));
    double *dtdy1d = calloc(n, sizeof(double));
    double *aux1 = calloc(n * na, sizeof(double));
    double *aux2 = calloc(n * na, sizeof(double));
    double *aux3 = calloc(n * na, sizeof(double));
    flux_work_t wk;
    alloc_work(&wk, maxm, mbc, ne, mwaves);
    double cfl = 0.0, cfl1d;
    const double dtdx = A->dt / A->dx, dtdy = A->dt / A->dy;
    for (int i = 1 - mbc; i <= maxm + mbc; ++i) {
        dtdx1d[i + mbc - 1] = dtdx;
        dtdy1d[i + mbc - 1] = dtdy;
    }
#define G(m, k, i) gadd[(((i) + mbc - 1) * 2 + (k)) * ne + (m)]
#define AUXF(ma, i, j) (QF(aux, na, i, j))[ma]
#define WR(r) ((r) >= w0 && (r) <= w1)

    if (dir == 1) {
        for (int j = s0; j <= s1; ++j) {
            for (int i = 1 - mbc; i <= mx + mbc; ++i)
                for (int m = 0; m < ne; ++m) V(q1d, ne, m, i) = QF(qold, ne, i, j)[m];
            if (num_aux > 0) {
                for (int i = 1 - mbc; i <= mx + mbc; ++i)
                    for (int ma = 0; ma < num_aux; ++ma) {
                        V(aux2, na, ma, i) = AUXF(ma, i, j);
                        if (mode == 0 || j != 1 - mbc) V(aux1, na, ma, i) = AUXF(ma, i, j - 1);
                        if (mode == 0 || j != my + mbc) V(aux3, na, ma, i) = AUXF(ma, i, j + 1);
                    }
            }
            flux2(1, maxm, ne, mwaves, na, mbc, mx, q1d, dtdx1d, aux1, aux2, aux3, A->method, A->mthlim,
                  qadd, fadd, gadd, &cfl1d, &wk, A->rpn2, A->rpt2, A->use_fwave);
            cfl = dmax1(cfl, cfl1d);
            if (mode == 0) {
                for (int m = 0; m < ne; ++m)
                    for (int i = 1; i <= mx; ++i) {
                        if (WR(j))
                            QF(qnew, ne, i, j)[m] = QF(qnew, ne, i, j)[m] + V(qadd, ne, m, i)
                                - dtdx * (V(fadd, ne, m, i + 1) - V(fadd, ne, m, i))
                                - dtdy * (G(m, 1, i) - G(m, 0, i));
                        if (WR(j - 1))
                            QF(qnew, ne, i, j - 1)[m] = QF(qnew, ne, i, j - 1)[m] - dtdy * G(m, 0, i);
                        if (WR(j + 1))
                            QF(qnew, ne, i, j + 1)[m] = QF(qnew, ne, i, j + 1)[m] + dtdy * G(m, 1, i);
                    }
            } else if (WR(j)) {
                for (int i = 1; i <= mx; ++i)
                    for (int m = 0; m < ne; ++m)
                        QF(qnew, ne, i, j)[m] = QF(qnew, ne, i, j)[m] + V(qadd, ne, m, i)
                            - dtdx * (V(fadd, ne, m, i + 1) - V(fadd, ne, m, i));
            }
        }
    } else {
        for (int i = s0; i <= s1; ++i) {
            for (int j = 1 - mbc; j <= my + mbc; ++j)
                for (int m = 0; m < ne; ++m) V(q1d, ne, m, j) = QF(qold, ne, i, j)[m];
            if (num_aux > 0) {
                for (int j = 1 - mbc; j <= my + mbc; ++j)
                    for (int ma = 0; ma < num_aux; ++ma) {
                        V(aux2, na, ma, j) = AUXF(ma, i, j);
                        if (mode == 0 || i != 1 - mbc) V(aux1, na, ma, j) = AUXF(ma, i - 1, j);
                        if (mode == 0 || i != mx + mbc) V(aux3, na, ma, j) = AUXF(ma, i + 1, j);
                    }
            }
            flux2(2, maxm, ne, mwaves, na, mbc, my, q1d, dtdy1d, aux1, aux2, aux3, A->method, A->mthlim,
                  qadd, fadd, gadd, &cfl1d, &wk, A->rpn2, A->rpt2, A->use_fwave);
            cfl = dmax1(cfl, cfl1d);
            if (mode == 0) {
                for (int m = 0; m < ne; ++m)
                    for (int j = 1; j <= my; ++j) {
                        if (WR(i))
                            QF(qnew, ne, i, j)[m] = QF(qnew, ne, i, j)[m] + (V(qadd, ne, m, j)
                                - dtdy * (V(fadd, ne, m, j + 1) - V(fadd, ne, m, j))
                                - dtdx * (G(m, 1, j) - G(m, 0, j)));
                        if (WR(i - 1))
                            QF(qnew, ne, i - 1, j)[m] = QF(qnew, ne, i - 1, j)[m] - dtdx * G(m, 0, j);
                        if (WR(i + 1))
                            QF(qnew, ne, i + 1, j)[m] = QF(qnew, ne, i + 1, j)[m] + dtdx * G(m, 1, j);
                    }
            } else if (WR(i)) {
                for (int j = 1; j <= my; ++j)
                    for (int m = 0; m < ne; ++m)
                        QF(qnew, ne, i, j)[m] = QF(qnew, ne, i, j)[m] + V(qadd, ne, m, j)
                            - dtdy * (V(fadd, ne, m, j + 1) - V(fadd, ne, m, j));
            }
        }
    }
#undef G
#undef AUXF
#undef WR
    free_work(&wk);
    free(q1d); free(qadd); free(fadd); free(gadd); free(dtdx1d); free(dtdy1d);
    free(aux1); free(aux2); free(aux3);
    return cfl;
}

/* Slices of each sweep: step2 runs rows/columns 0..n+1 (step2.f90:92,173); step2ds all
 * of them, ghosts included (step2ds.f90:96,177).  Updated rows/columns lie in 1-mbc..n+mbc. */
static void sweep_slices(int mode, int mbc, int n, int *s0, int *s1)
{
    *s0 = (mode == 0) ? 0 : 1 - mbc;
    *s1 = (mode == 0) ? n + 1 : n + mbc;
}

static double classic_sweeps(const sweep_args_t *A, int nthreads)
{
    double cfl = 0.0;
    for (int dir = 1; dir <= 2; ++dir) {
        if ((A->mode == 1 && dir == 2) || (A->mode == 2 && dir == 1)) continue;
        const int n = dir == 1 ? A->my : A->mx;
        int s0, s1;
        sweep_slices(A->mode, A->mbc, n, &s0, &s1);
        const int w0 = 1 - A->mbc, w1 = n + A->mbc;
        if (nthreads <= 1) {
            cfl = dmax1(cfl, sweep_dir(A, dir, s0, s1, w0, w1));
            continue;
        }
        /* disjoint write ranges of about equal size; each also runs the slice on either side */
        const int nw = w1 - w0 + 1;
        const int nb = nthreads < nw ? nthreads : nw;
        double c = 0.0;
#pragma omp parallel for num_threads(nb) schedule(static, 1) reduction(max : c)
        for (int b = 0; b < nb; ++b) {
            const int a0 = w0 + (int)((long long)nw * b / nb), a1 = w0 + (int)((long long)nw * (b + 1) / nb) - 1;
            /* step2 slices update the rows/columns on either side of them, step2ds
             * slices only their own (and in the aliased ids=2 call a neighbour's
             * column may already be updated, so it must not be re-read) */
            const int h = A->mode == 0 ? 1 : 0;
            int t0 = a0 - h < s0 ? s0 : a0 - h, t1 = a1 + h > s1 ? s1 : a1 + h;
            if (t0 <= t1) {
                const double cb = sweep_dir(A, dir, t0, t1, a0, a1);
                c = c > cb ? c : cb;
            }
        }
        cfl = dmax1(cfl, c);
    }
    return cfl;
}

/* rp ids for the exported entry points */
static void pick_rp(int rp, rpn2_fn *n, rpt2_fn *t)
{
    switch (rp) {
    case 0: *n = orc_rpn2_acoustics_; *t = orc_rpt2_acoustics_; break;
    case 1: *n = orc_rpn2_vc_acoustics_; *t = orc_rpt2_vc_acoustics_; break;
    case 2: *n = orc_rpn2_euler5_; *t = orc_rpt2_euler5_; break;
    default: *n = orc_rpn2_shallow_roe_; *t = orc_rpt2_shallow_roe_; break;
    }
}

/* One unsplit step: qnew (which must equal qold on entry) is updated in place. */
double orc_step2_mt(int rp, int maxm, int meqn, int mwaves, int num_aux, int mbc, int mx, int my,
                    const double *qold, double *qnew, const double *aux, double dx, double dy,
                    double dt, const int *method, const int *mthlim, int use_fwave, int nthreads)
{
    sweep_args_t A = {0, maxm, meqn, mwaves, num_aux, mbc, mx, my, qold, qnew, aux, dx, dy, dt,
                      method, mthlim, use_fwave, 0, 0};
    pick_rp(rp, &A.rpn2, &A.rpt2);
    return classic_sweeps(&A, nthreads);
}

double orc_step2(int rp, int maxm, int meqn, int mwaves, int num_aux, int mbc, int mx, int my,
                 const double *qold, double *qnew, const double *aux, double dx, double dy,
                 double dt, const int *method, const int *mthlim, int use_fwave)
{
    return orc_step2_mt(rp, maxm, meqn, mwaves, num_aux, mbc, mx, my, qold, qnew, aux, dx, dy, dt,
                        method, mthlim, use_fwave, 1);
}

/* One call of step2ds (ids = 1 or 2).  With ids = 2 qold may alias qnew: each
 * column is read before any write to it, and only by its own slice. */
double orc_step2ds_mt(int rp, int maxm, int meqn, int mwaves, int num_aux, int mbc, int mx, int my,
                      const double *qold, double *qnew, const double *aux, double dx, double dy,
                      double dt, const int *method, const int *mthlim, int ids, int use_fwave, int nthreads)
{
    sweep_args_t A = {ids, maxm, meqn, mwaves, num_aux, mbc, mx, my, qold, qnew, aux, dx, dy, dt,
                      method, mthlim, use_fwave, 0, 0};
    pick_rp(rp, &A.rpn2, &A.rpt2);
    return classic_sweeps(&A, nthreads);
}

double orc_step2ds(int rp, int maxm, int meqn, int mwaves, int num_aux, int mbc, int mx, int my,
                   const double *qold, double *qnew, const double *aux, double dx, double dy,
                   double dt, const int *method, const int *mthlim, int ids, int use_fwave)
{
    return orc_step2ds_mt(rp, maxm, meqn, mwaves, num_aux, mbc, mx, my, qold, qnew, aux, dx, dy, dt,
                          method, mthlim, ids, use_fwave, 1);
}

/* Direct access to one Riemann solver over a slice (for unit tests). */
void orc_rpn2(int rp, int ixy, int maxm, int meqn, int mwaves, int mbc, int mx, const double *ql,
              const double *qr, const double *auxl, const double *auxr, double *wave, double *s,
              double *amdq, double *apdq, int num_aux)
{
    rpn2_fn n; rpt2_fn t; pick_rp(rp, &n, &t);
    n(&ixy, &maxm, &meqn, &mwaves, &mbc, &mx, ql, qr, auxl, auxr, wave, s, amdq, apdq, &num_aux);
}

void orc_rpt2(int rp, int ixy, int maxm, int meqn, int mwaves, int mbc, int mx, const double *ql,
              const double *qr, const double *aux1, const double *aux2, const double *aux3,
              int imp, const double *asdq, double *bmasdq, double *bpasdq, int num_aux)
{
    rpn2_fn n; rpt2_fn t; pick_rp(rp, &n, &t);
    t(&ixy, &maxm, &meqn, &mwaves, &mbc, &mx, ql, qr, aux1, aux2, aux3, &imp, asdq, bmasdq,
      bpasdq, &num_aux);
}

/* Address of a Riemann solver, so a harness can hand it to other code. */
void *orc_rp_pointer(int rp, int which)
{
    rpn2_fn n; rpt2_fn t; pick_rp(rp, &n, &t);
    return which == 0 ? (void *)n : (void *)t;
}

/* ------------------------------------------------------------------ */
/* Boundary fill of a padded array: solver.py:299-434 (q) and         */
/* :438-576 (aux).  Order: x lower, x upper, y lower, y upper; each   */
/* x fill covers all rows, each y fill all columns, so corners get    */
/* the y rule applied to x-filled ghosts.                             */
/* bc codes: 1 extrap, 2 periodic, 3 wall; any other code leaves the  */
/* ghost strip untouched (custom fills are done by the caller).       */
/* negate: 1 for q (wall negates the normal component), 0 for aux.    */
/* ------------------------------------------------------------------ */
void orc_fill_bc(double *a, int ne, int mx, int my, int mbc, const int *bc_lower,
                 const int *bc_upper, int negate)
{
    const int nxp = mx + 2 * mbc, nyp = my + 2 * mbc;
#define P(ii, jj) (a + ((size_t)(jj) * nxp + (ii)) * ne)   /* 0-based padded */
    for (int idim = 0; idim < 2; ++idim) {
        int n = idim == 0 ? mx : my;
        int nt = idim == 0 ? nyp : nxp;   /* transverse extent (full, padded) */
        for (int side = 0; side < 2; ++side) {
            int bc = side == 0 ? bc_lower[idim] : bc_upper[idim];
            for (int k = 0; k < mbc; ++k) {
                int g, src;   /* padded indices along idim */
                if (side == 0) {
                    g = k;
                    if (bc == 1) src = mbc;
                    else if (bc == 2) src = n + k;               /* qbc[-2ng+k] */
                    else if (bc == 3) src = 2 * mbc - 1 - k;
                    else continue;
                } else {
                    g = n + 2 * mbc - 1 - k;                    /* qbc[-k-1] */
                    if (bc == 1) src = n + mbc - 1;              /* qbc[-ng-1] */
                    else if (bc == 2) src = mbc + (mbc - 1 - k); /* qbc[ng:2ng] */
                    else if (bc == 3) src = n + k;               /* qbc[-2ng+k] */
                    else continue;
                }
                for (int t = 0; t < nt; ++t) {
                    double *dst = idim == 0 ? P(g, t) : P(t, g);
                    const double *sp = idim == 0 ? P(src, t) : P(t, src);
                    for (int m = 0; m < ne; ++m) dst[m] = sp[m];
                    if (bc == 3 && negate) dst[idim + 1] = -sp[idim + 1];
                }
            }
        }
    }
#undef P
}

/* ------------------------------------------------------------------ */
/* SharpClaw: weno.f90:5-83 (weno5), flux1_2.f90:2-196 (flux1),       */
/* sharpclaw/flux2.f90:2-101 (flux2).  lim_type=2, char_decomp=0,     */
/* tfluct_solver off, index_capa 0.                                   */
/* ------------------------------------------------------------------ */
/* q, ql, qr: (meqn, 1:n) with n = slice length incl. ghosts, 1-based k */
static void weno5_slice(const double *q, double *ql, double *qr, int meqn, int kmin, int kmax)
{
    const double eps = (double)1.0e-36f;
#define Q(m, k) q[((k) - 1) * meqn + (m)]
    for (int k = kmin; k <= kmax; ++k)
        for (int m = 0; m < meqn; ++m) {
            double qm2 = Q(m, k - 2), qm1 = Q(m, k - 1), q0 = Q(m, k), qp1 = Q(m, k + 1), qp2 = Q(m, k + 2);
            double sigma0 = 3.3333333333333333333333333333333333 * (q0 * q0)
                - 10.333333333333333333333333333333333 * q0 * qp1
                + 3.6666666666666666666666666666666667 * q0 * qp2
                + 8.3333333333333333333333333333333333 * (qp1 * qp1)
                - 6.3333333333333333333333333333333333 * qp1 * qp2
                + 1.3333333333333333333333333333333333 * (qp2 * qp2);
            double sigma1 = 4.3333333333333333333333333333333333 * (q0 * q0)
                - 4.3333333333333333333333333333333333 * q0 * qp1
                - 4.3333333333333333333333333333333333 * q0 * qm1
                + 1.3333333333333333333333333333333333 * (qp1 * qp1)
                + 1.6666666666666666666666666666666667 * qp1 * qm1
                + 1.3333333333333333333333333333333333 * (qm1 * qm1);
            double sigma2 = 3.3333333333333333333333333333333333 * (q0 * q0)
                - 10.333333333333333333333333333333333 * q0 * qm1
                + 3.6666666666666666666666666666666667 * q0 * qm2
                + 8.3333333333333333333333333333333333 * (qm1 * qm1)
                - 6.3333333333333333333333333333333333 * qm1 * qm2
                + 1.3333333333333333333333333333333333 * (qm2 * qm2);
            double t0 = eps + sigma0, t1 = eps + sigma1, t2 = eps + sigma2;
            double omega0 = 0.1 / (t0 * t0);
            double omega1 = 0.6 / (t1 * t1);
            double omega2 = 0.3 / (t2 * t2);
            double omega3 = 0.3 / (t0 * t0);
            double omega4 = 0.6 / (t1 * t1);
            double omega5 = 0.1 / (t2 * t2);
            double fr0 = 1.8333333333333333333333333333333333 * q0
                - 1.1666666666666666666666666666666667 * qp1
                + 0.33333333333333333333333333333333333 * qp2;
            double fr1 = 0.83333333333333333333333333333333333 * q0
                - 0.16666666666666666666666666666666667 * qp1
                + 0.33333333333333333333333333333333333 * qm1;
            double fr2 = 0.33333333333333333333333333333333333 * q0
                + 0.83333333333333333333333333333333333 * qm1
                - 0.16666666666666666666666666666666667 * qm2;
            double fr3 = 0.33333333333333333333333333333333333 * q0
                + 0.83333333333333333333333333333333333 * qp1
                - 0.16666666666666666666666666666666667 * qp2;
            double fr4 = 0.83333333333333333333333333333333333 * q0
                + 0.33333333333333333333333333333333333 * qp1
                - 0.16666666666666666666666666666666667 * qm1;
            double fr5 = 1.8333333333333333333333333333333333 * q0
                - 1.1666666666666666666666666666666667 * qm1
                + 0.33333333333333333333333333333333333 * qm2;
            ql[(k - 1) * meqn + m] = (fr0 * omega0 + fr1 * omega1 + fr2 * omega2) / (omega0 + omega1 + omega2);
            qr[(k - 1) * meqn + m] = (fr3 * omega3 + fr4 * omega4 + fr5 * omega5) / (omega3 + omega4 + omega5);
        }
#undef Q
}

/* flux1 over one slice.  q1d, dq1d, aux: (ne|na, 1-mbc:mx+mbc) (Fortran i). */
static double sharp_flux1(const double *q1d, double *dq1d, const double *aux, double dt, double dx,
                          int ixy, int num_aux, int meqn, int mwaves, int mx, int mbc, int maxnx,
                          rpn2_fn rpn2)
{
    const int ne = meqn, na = num_aux > 0 ? num_aux : 1;
    const int n = mx + 2 * mbc;
    double *ql = calloc((size_t)n * ne, sizeof(double)), *qr = calloc((size_t)n * ne, sizeof(double));
    double *wave = calloc((size_t)n * ne * mwaves, sizeof(double)), *s = calloc((size_t)n * mwaves, sizeof(double));
    double *amdq = calloc((size_t)n * ne, sizeof(double)), *apdq = calloc((size_t)n * ne, sizeof(double));
    double *amdq2 = calloc((size_t)n * ne, sizeof(double)), *apdq2 = calloc((size_t)n * ne, sizeof(double));
    double *auxl = calloc((size_t)n * na, sizeof(double)), *auxr = calloc((size_t)n * na, sizeof(double));
    const double dtdx = dt / dx;

    for (int i = 1 - mbc; i <= mx + mbc; ++i)
        for (int m = 0; m < ne; ++m) V(dq1d, ne, m, i) = 0.0;

    /* weno5 indexes the slice 1..n; only entries consumed by interfaces 1..mx+1 matter */
    weno5_slice(q1d, ql, qr, ne, 3 > mbc - 1 ? 3 : mbc - 1, n - 2);

    int ixy_ = ixy, maxm_ = maxnx, ne_ = ne, nw_ = mwaves, mbc_ = mbc, mx_ = mx, na_ = na;
    /* rpn2 runs over interfaces 2-mbc..mx+mbc; entries 1 and n of ql/qr are zero here
       (the reference leaves them unset); they only reach interfaces that are not used. */
    rpn2(&ixy_, &maxm_, &ne_, &nw_, &mbc_, &mx_, ql, qr, aux, aux, wave, s, amdq, apdq, &na_);

    double cfl = 0.0;
    for (int mw = 0; mw < mwaves; ++mw)
        for (int i = 1; i <= mx + 1; ++i)
            cfl = dmax1(dmax1(cfl, dtdx * V(s, mwaves, mw, i)), -dtdx * V(s, mwaves, mw, i));

    for (int i = 1 - mbc + 1; i <= mx + mbc; ++i)
        for (int m = 0; m < ne; ++m) {
            V(qr, ne, m, i - 1) = V(ql, ne, m, i);
            V(ql, ne, m, i) = V(qr, ne, m, i);
        }
    if (num_aux > 0)
        for (int i = 1 - mbc + 1; i <= mx + mbc; ++i)
            for (int m = 0; m < num_aux; ++m) {
                V(auxr, na, m, i - 1) = V(aux, na, m, i);
                V(auxl, na, m, i) = V(aux, na, m, i);
            }
    rpn2(&ixy_, &maxm_, &ne_, &nw_, &mbc_, &mx_, ql, qr, auxl, auxr, wave, s, amdq2, apdq2, &na_);

    for (int i = 1; i <= mx; ++i)
        for (int m = 0; m < ne; ++m)
            V(dq1d, ne, m, i) = V(dq1d, ne, m, i) - dtdx * (V(amdq, ne, m, i + 1) + V(apdq, ne, m, i)
                                + V(amdq2, ne, m, i) + V(apdq2, ne, m, i));
    free(ql); free(qr); free(wave); free(s); free(amdq); free(apdq); free(amdq2); free(apdq2);
    free(auxl); free(auxr);
    return cfl;
}

/* SharpClaw dq for the whole patch (sharpclaw/flux2.f90:2-101).  q, aux padded (mbc
   ghosts, already filled); dq padded, overwritten (zeroed first).  Returns the CFL
   number.  Each x slice (row j) adds only to its own row of dq and each y slice to
   its own column, so the slices of one direction run on nthreads threads; the y
   slices start after every x slice has added its part, as in the reference. */
double orc_sharp_dq_mt(int rp, int meqn, int mwaves, int num_aux, int mbc, int mx, int my,
                       const double *q, const double *aux, double *dq, double dx, double dy, double dt,
                       int nthreads)
{
    rpn2_fn rpn2; rpt2_fn rpt2u; pick_rp(rp, &rpn2, &rpt2u); (void)rpt2u;
    const int ne = meqn, na = num_aux > 0 ? num_aux : 1;
    const int maxnx = mx > my ? mx : my;
    size_t tot = (size_t)(mx + 2 * mbc) * (my + 2 * mbc) * ne;
    memset(dq, 0, tot * sizeof(double));
    size_t nmax = (size_t)(maxnx + 2 * mbc);
    double cfl = 0.0;
    if (nthreads < 1) nthreads = 1;
    for (int dir = 1; dir <= 2; ++dir) {
        const int n0 = dir == 1 ? my : mx;
        double c = 0.0;
#pragma omp parallel num_threads(nthreads) reduction(max : c)
        {
            double *q1d = calloc(nmax * ne, sizeof(double)), *dq1d = calloc(nmax * ne, sizeof(double));
            double *a1d = calloc(nmax * na, sizeof(double));
#pragma omp for schedule(static)
            for (int k = 0; k <= n0 + 1; ++k) {
                double cs;
                if (dir == 1) {
                    const int j = k;
                    for (int i = 1 - mbc; i <= mx + mbc; ++i) {
                        for (int m = 0; m < ne; ++m) V(q1d, ne, m, i) = QF(q, ne, i, j)[m];
                        if (num_aux > 0) for (int m = 0; m < num_aux; ++m) V(a1d, na, m, i) = QF(aux, na, i, j)[m];
                    }
                    cs = sharp_flux1(q1d, dq1d, a1d, dt, dx, 1, num_aux, ne, mwaves, mx, mbc, maxnx, rpn2);
                    for (int i = 1; i <= mx; ++i)
                        for (int m = 0; m < ne; ++m) QF(dq, ne, i, j)[m] = QF(dq, ne, i, j)[m] + V(dq1d, ne, m, i);
                } else {
                    const int i = k;
                    for (int j = 1 - mbc; j <= my + mbc; ++j) {
                        for (int m = 0; m < ne; ++m) V(q1d, ne, m, j) = QF(q, ne, i, j)[m];
                        if (num_aux > 0) for (int m = 0; m < num_aux; ++m) V(a1d, na, m, j) = QF(aux, na, i, j)[m];
                    }
                    cs = sharp_flux1(q1d, dq1d, a1d, dt, dy, 2, num_aux, ne, mwaves, my, mbc, maxnx, rpn2);
                    for (int j = 1; j <= my; ++j)
                        for (int m = 0; m < ne; ++m) QF(dq, ne, i, j)[m] = QF(dq, ne, i, j)[m] + V(dq1d, ne, m, j);
                }
                c = c > cs ? c : cs;
            }
            free(q1d); free(dq1d); free(a1d);
        }
        cfl = dmax1(cfl, c);
    }
    return cfl;
}

double orc_sharp_dq(int rp, int meqn, int mwaves, int num_aux, int mbc, int mx, int my,
                    const double *q, const double *aux, double *dq, double dx, double dy, double dt)
{
    return orc_sharp_dq_mt(rp, meqn, mwaves, num_aux, mbc, mx, my, q, aux, dq, dx, dy, dt, 1);
}

/* WENO5 reconstruction over a slice q(meqn, 1:n) for k = 3..n-2 (unit tests). */
void orc_weno5(const double *q, double *ql, double *qr, int meqn, int n)
{
    weno5_slice(q, ql, qr, meqn, 3, n - 2);
}
