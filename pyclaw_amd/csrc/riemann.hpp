// Riemann solvers for the fused 2D wave-propagation kernels (gfx950).
//
// Each solver is a struct with compile-time sizes and two device functions:
//   rpn<DIR>(...)  normal solve at one interface:  waves, speeds, A^-dq, A^+dq
//   rpt<DIR>(...)  transverse split of one fluctuation into B^-A*dq, B^+A*dq
// DIR = 0 for an x-interface (normal momentum is component 1), DIR = 1 for a
// y-interface (normal momentum is component 2), as mu/mv in the reference.
// "lo" is the cell below/left of the interface (the reference's qr(i-1)),
// "hi" the cell above/right (ql(i)).
//
// Operation order in every expression follows the reference Fortran (or, for
// the solvers the reference takes from clawpack.riemann, this repository's
// restatement of them), so that an FP64 build compiled without contraction
// gives the same bits as the CPU reference.
//   Euler 5-wave Roe + entropy fix: development/rp_approaches/rpn2_euler_5wave.f:87-298,
//                                   rpt2_euler_5wave.f:349-395
//   acoustics / vc-acoustics / shallow-water Roe: external to the reference;
//   call sites apps/acoustics_2d_homogeneous/acoustics.py:25-26,
//   apps/acoustics_2d_variable/acoustics.py:25-26, apps/shallow_2d/shallow2D.py:44-45.
//
// WNZ(m, mw) marks wave components that are not identically zero; kernels use
// it to skip zero terms in cross-lane exchanges and dot products (adding an
// exact zero does not change a floating-point sum).
#pragma once
#include <hip/hip_runtime.h>

#include "common.hpp"

namespace clawhip {

enum RiemannKind { RP_ACOUSTICS = 0, RP_VC_ACOUSTICS = 1, RP_EULER = 2, RP_SHALLOW = 3 };

template <typename R> __device__ __forceinline__ R rmax(R a, R b) { return a > b ? a : b; }
template <typename R> __device__ __forceinline__ R rmin(R a, R b) { return a < b ? a : b; }

// Division policy of the Euler solver.  By default every quotient is the
// reference's (x / y, correctly rounded, in the reference's order).  A build
// that defines CLAW_SHARED_RECIPROCALS (the fused-multiply-add FP64 build)
// divides once where the reference divides several values by the same
// denominator and multiplies by that reciprocal: a few ulp from the
// reference's quotients, about a third fewer FP64 division sequences.
#ifdef CLAW_SHARED_RECIPROCALS
constexpr bool SHARED_RCP = true;
#else
constexpr bool SHARED_RCP = false;
#endif

// 1/x in the shared-reciprocal build: the hardware reciprocal refined by two
// Newton steps (within about one ulp of 1/x for normal x; used for
// positive Roe-solver quantities only).  Other builds never call it.
template <typename R> __device__ __forceinline__ R rcp(R x) { return R(1) / x; }
template <> __device__ __forceinline__ double rcp<double>(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, e, r);
}

// Square root in the shared-reciprocal build: Goldschmidt's iteration from the
// hardware reciprocal square root y ~ 1/sqrt(x), carrying g -> sqrt(x) and
// h -> 1/(2 sqrt(x)): one coupled step r = 1/2 - g h, g += g r, h += h r, then
// two residual corrections g += (x - g^2) h.  Within about one ulp for normal
// positive x (the quantities whose roots the solvers take: densities, depths,
// squared sound speeds); 0, +inf and NaN pass through as sqrt returns them.
// Without the library's rescaling of inputs below 2^-767 it takes 13
// instructions instead of 18.  Other builds use the correctly rounded sqrt.
__device__ __forceinline__ double sqrt_goldschmidt(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    const double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    return (x == 0.0 || x == __builtin_inf() || x != x) ? x : g;
}

template <typename R> __device__ __forceinline__ R rsqrt_exact(R x);
template <> __device__ __forceinline__ double rsqrt_exact<double>(double x) {
    if constexpr (SHARED_RCP) return sqrt_goldschmidt(x);
    return __builtin_sqrt(x);
}
template <> __device__ __forceinline__ float rsqrt_exact<float>(float x) { return __builtin_sqrtf(x); }

// Two quotients with one denominator, q1 = n1/d and q2 = n2/d.  The
// shared-reciprocal build multiplies both numerators by rcp(d).  Otherwise
// each is the IEEE quotient n/d on its own.  (Sharing the denominator's part
// of the division sequence between the two is exact only when both numerators
// leave the denominator scaled alike, and the per-lane test for the other
// case splits the code into small blocks: measured slower than two whole
// divisions.)
template <typename R>
__device__ __forceinline__ void div2(R n1, R n2, R d, R &q1, R &q2) {
    if constexpr (SHARED_RCP) {
        const R r = rcp<R>(d);
        q1 = n1 * r;
        q2 = n2 * r;
    } else {
        q1 = n1 / d;
        q2 = n2 / d;
    }
}
// Three quotients with one denominator, as div2: q_k = n_k/d.
template <typename R>
__device__ __forceinline__ void div3(R n1, R n2, R n3, R d, R &q1, R &q2, R &q3) {
    if constexpr (SHARED_RCP) {
        const R r = rcp<R>(d);
        q1 = n1 * r;
        q2 = n2 * r;
        q3 = n3 * r;
    } else {
        q1 = n1 / d;
        q2 = n2 / d;
        q3 = n3 / d;
    }
}
// ---------------------------------------------------------------------------
// Constant-coefficient acoustics: q = (p, u, v), parameters rho, bulk, c, Z.
// ---------------------------------------------------------------------------
template <typename R> struct Acoustics {
    static constexpr int NE = 3, NW = 2, NA = 0;
    static constexpr bool HAS_EFIX = false;   // no entropy fix
    struct Params { R rho, bulk, cc, zz; };
    struct Cache {};
    static constexpr bool SPEEDS_FROM_CACHE = false;
    static constexpr bool HAS_STATE = false;   // no per-state quantities shared between solves
    static constexpr bool CELL_STATE = false;  // no per-cell quantities shared between solves
    __host__ __device__ static constexpr bool SDIST(int) { return true; }   // every wave speed may differ
    // limiters run in every lane (zero waves keep phi = 1): the waves of
    // these solvers are rarely zero across a whole wavefront
    static constexpr bool SKIP_ZERO_WAVES = false;
    struct CellState {};
    __device__ static __forceinline__ void speeds(const Cache &, R (&)[NW]) {}
    template <int DIR> static constexpr int MU = DIR == 0 ? 1 : 2;
    template <int DIR> static constexpr int MV = DIR == 0 ? 2 : 1;
    template <int DIR> __host__ __device__ static constexpr bool WNZ(int m, int mw) {
        return m != (DIR == 0 ? 2 : 1);
    }
    template <int DIR>
    __device__ static __forceinline__ void rpn(const Params &P, const R (&lo)[NE], const R (&hi)[NE],
                                               const R *alo, const R *ahi, R (&wave)[NW][NE], R (&s)[NW],
                                               R (&amdq)[NE], R (&apdq)[NE], Cache &) {
        constexpr int mu = MU<DIR>, mv = MV<DIR>;
        R d0 = hi[0] - lo[0];
        R d1 = hi[mu] - lo[mu];
        R a1 = (-d0 + P.zz * d1) / (R(2) * P.zz);
        R a2 = (d0 + P.zz * d1) / (R(2) * P.zz);
        wave[0][0] = -a1 * P.zz; wave[0][mu] = a1; wave[0][mv] = R(0); s[0] = -P.cc;
        wave[1][0] = a2 * P.zz; wave[1][mu] = a2; wave[1][mv] = R(0); s[1] = P.cc;
#pragma unroll
        for (int m = 0; m < NE; ++m) { amdq[m] = s[0] * wave[0][m]; apdq[m] = s[1] * wave[1][m]; }
    }
    template <int DIR>
    __device__ static __forceinline__ void rpt(const Params &P, int imp, const R (&asdq)[NE], const R *a1_,
                                               const R *a2_, const R *a3_, R (&bm)[NE], R (&bp)[NE],
                                               const Cache &) {
        constexpr int mu = MU<DIR>, mv = MV<DIR>;
        R a1 = (-asdq[0] + P.zz * asdq[mv]) / (R(2) * P.zz);
        R a2 = (asdq[0] + P.zz * asdq[mv]) / (R(2) * P.zz);
        bm[0] = P.cc * a1 * P.zz; bm[mu] = R(0); bm[mv] = -P.cc * a1;
        bp[0] = P.cc * a2 * P.zz; bp[mu] = R(0); bp[mv] = P.cc * a2;
    }
};

// ---------------------------------------------------------------------------
// Variable-coefficient acoustics: aux = (rho, c) per cell.
// ---------------------------------------------------------------------------
template <typename R> struct VCAcoustics {
    static constexpr int NE = 3, NW = 2, NA = 2;
    static constexpr bool HAS_EFIX = false;   // no entropy fix
    struct Params { R unused; };
    struct Cache {};
    static constexpr bool SPEEDS_FROM_CACHE = false;
    static constexpr bool HAS_STATE = false;   // no per-state quantities shared between solves
    static constexpr bool CELL_STATE = false;  // no per-cell quantities shared between solves
    __host__ __device__ static constexpr bool SDIST(int) { return true; }   // every wave speed may differ
    // limiters run in every lane (zero waves keep phi = 1): the waves of
    // these solvers are rarely zero across a whole wavefront
    static constexpr bool SKIP_ZERO_WAVES = false;
    struct CellState {};
    __device__ static __forceinline__ void speeds(const Cache &, R (&)[NW]) {}
    template <int DIR> static constexpr int MU = DIR == 0 ? 1 : 2;
    template <int DIR> static constexpr int MV = DIR == 0 ? 2 : 1;
    template <int DIR> __host__ __device__ static constexpr bool WNZ(int m, int mw) {
        return m != (DIR == 0 ? 2 : 1);
    }
    template <int DIR>
    __device__ static __forceinline__ void rpn(const Params &, const R (&lo)[NE], const R (&hi)[NE],
                                               const R *alo, const R *ahi, R (&wave)[NW][NE], R (&s)[NW],
                                               R (&amdq)[NE], R (&apdq)[NE], Cache &) {
        constexpr int mu = MU<DIR>, mv = MV<DIR>;
        R d0 = hi[0] - lo[0];
        R d1 = hi[mu] - lo[mu];
        R zi = ahi[0] * ahi[1];
        R zim = alo[0] * alo[1];
        R a1 = (-d0 + zi * d1) / (zim + zi);
        R a2 = (d0 + zim * d1) / (zim + zi);
        wave[0][0] = -a1 * zim; wave[0][mu] = a1; wave[0][mv] = R(0); s[0] = -alo[1];
        wave[1][0] = a2 * zi; wave[1][mu] = a2; wave[1][mv] = R(0); s[1] = ahi[1];
#pragma unroll
        for (int m = 0; m < NE; ++m) { amdq[m] = s[0] * wave[0][m]; apdq[m] = s[1] * wave[1][m]; }
    }
    // am/a0/ap: aux of the cell that receives the fluctuation (a0) and of its
    // transverse neighbours below (am) and above (ap).
    template <int DIR>
    __device__ static __forceinline__ void rpt(const Params &, int imp, const R (&asdq)[NE], const R *am,
                                               const R *a0, const R *ap, R (&bm)[NE], R (&bp)[NE],
                                               const Cache &) {
        constexpr int mu = MU<DIR>, mv = MV<DIR>;
        R cm = am[1], cp = ap[1];
        R zm = am[0] * am[1];
        R zz = a0[0] * a0[1];
        R zp = ap[0] * ap[1];
        R a1 = (-asdq[0] + asdq[mv] * zz) / (zm + zz);
        R a2 = (asdq[0] + asdq[mv] * zz) / (zz + zp);
        bm[0] = cm * a1 * zm; bm[mu] = R(0); bm[mv] = -cm * a1;
        bp[0] = cp * a2 * zp; bp[mu] = R(0); bp[mv] = cp * a2;
    }
};

// ---------------------------------------------------------------------------
// Euler equations with a tracer: q = (rho, rho u, rho v, E, rho*phi... tracer)
// 5 waves: u-a, shear, entropy, u+a, tracer.
// ---------------------------------------------------------------------------
template <typename R> struct Euler5 {
    static constexpr int NE = 5, NW = 5, NA = 0;
    static constexpr bool HAS_EFIX = true;    // Params::efix switches the entropy fix
    struct Params { R gamma, gamma1; int efix; };
    // Roe averages kept for the transverse solves of the same interface; the
    // other quantities rpt needs are recomputed with rpn's expressions
    // (g1a2 = gamma1 / a^2 and, in the shared-reciprocal build, i2a = 1/(2a):
    // rpn's values, so a transverse solve of a carried interface does not
    // divide again)
    struct Cache { R u, v, enth, a, g1a2, i2a; };
    // the wave speeds of an interface follow from its cache (rpn's expressions)
    static constexpr bool SPEEDS_FROM_CACHE = true;
    static constexpr bool HAS_STATE = false;   // no per-state quantities shared between solves
    __device__ static __forceinline__ void speeds(const Cache &C, R (&s)[NW]) {
        s[0] = C.u - C.a; s[1] = C.u; s[2] = C.u; s[3] = C.u + C.a; s[4] = C.u;
    }
    template <int DIR> static constexpr int MU = DIR == 0 ? 1 : 2;
    template <int DIR> static constexpr int MV = DIR == 0 ? 2 : 1;
    template <int DIR> __host__ __device__ static constexpr bool WNZ(int m, int mw) {
        // wave 0,2,3: components 0..3; wave 1 (shear): mv and 3; wave 4 (tracer): 4
        return (mw == 0 || mw == 2 || mw == 3) ? (m < 4)
             : (mw == 1) ? (m == (DIR == 0 ? 2 : 1) || m == 3)
             : (m == 4);
    }
    // The quantities of one cell that every Riemann problem it is part of
    // computes, as the left or the right state, in x and in y alike (the
    // reference's expressions, rpn2_euler_5wave.f:87-104 and the entropy
    // fix's :205-286): rs = sqrt(rho); ms = (rho u, rho v)/sqrt(rho) and
    // hs = (E + p)/sqrt(rho), the terms of the Roe averages; un = (rho u,
    // rho v)/rho and c = sqrt(gamma p/rho), the entropy fix's characteristic
    // speeds.  A kernel that meets a cell in four problems computes them once;
    // each is the same operation on the same operands either way (the
    // pressure's sum of squared momenta is commutative in the two directions).
    struct CellState { R rs, ms[2], hs, un[2], c; };
    static constexpr bool CELL_STATE = true;
    // speeds u - a, u, u, u + a, u: waves 2 and 4 repeat wave 1's speed
    __host__ __device__ static constexpr bool SDIST(int mw) { return mw != 2 && mw != 4; }
    // a wave family whose waves are zero in every lane of a wavefront skips
    // its limiter (flows with quiescent regions, such as a shock entering
    // gas at rest, have many)
    static constexpr bool SKIP_ZERO_WAVES = true;
    __device__ static __forceinline__ CellState cell(const Params &P, const R (&q)[NE]) {
        CellState S;
        const R gamma1 = P.gamma1;
        S.rs = rsqrt_exact<R>(q[0]);
        R p;
        if constexpr (SHARED_RCP) {
            const R ir = rcp<R>(q[0]), irs = rcp<R>(S.rs);
            p = gamma1 * (q[3] - R(0.5) * (q[1] * q[1] + q[2] * q[2]) * ir);
            S.ms[0] = q[1] * irs; S.ms[1] = q[2] * irs;
            S.hs = (q[3] + p) * irs;
            S.un[0] = q[1] * ir; S.un[1] = q[2] * ir;
            S.c = P.efix ? rsqrt_exact<R>(P.gamma * p * ir) : R(0);
        } else {
            R ke;
            div3<R>(R(0.5) * (q[1] * q[1] + q[2] * q[2]), q[1], q[2], q[0], ke, S.un[0], S.un[1]);
            p = gamma1 * (q[3] - ke);
            div3<R>(q[1], q[2], q[3] + p, S.rs, S.ms[0], S.ms[1], S.hs);
            S.c = P.efix ? rsqrt_exact<R>(P.gamma * p / q[0]) : R(0);
        }
        return S;
    }
    // a cell record moved one lane up (lane L receives lane L-1's)
    __device__ static __forceinline__ CellState cell_up(const CellState &S) {
        return CellState{lane_up(S.rs), {lane_up(S.ms[0]), lane_up(S.ms[1])}, lane_up(S.hs),
                         {lane_up(S.un[0]), lane_up(S.un[1])}, lane_up(S.c)};
    }
    template <int DIR>
    __device__ static __forceinline__ void rpn(const Params &P, const R (&lo)[NE], const R (&hi)[NE],
                                               const R *al, const R *ar, R (&wave)[NW][NE], R (&s)[NW],
                                               R (&amdq)[NE], R (&apdq)[NE], Cache &C) {
        rpn_c<DIR>(P, lo, hi, cell(P, lo), cell(P, hi), wave, s, amdq, apdq, C);
    }
    template <int DIR>
    __device__ static __forceinline__ void rpn_c(const Params &P, const R (&lo)[NE], const R (&hi)[NE],
                                                 const CellState &Sl, const CellState &Sr, R (&wave)[NW][NE],
                                                 R (&s)[NW], R (&amdq)[NE], R (&apdq)[NE], Cache &C) {
        constexpr int mu = MU<DIR>, mv = MV<DIR>;
        const R gamma1 = P.gamma1;
        // u = (lo[mu]/sqrt(rho_l) + hi[mu]/sqrt(rho_r)) / (sqrt(rho_l) + sqrt(rho_r)), v and H alike
        R u, v, enth;
        div3<R>(Sl.ms[mu - 1] + Sr.ms[mu - 1], Sl.ms[mv - 1] + Sr.ms[mv - 1], Sl.hs + Sr.hs, Sl.rs + Sr.rs,
                u, v, enth);
        R u2v2 = u * u + v * v;
        R a2 = gamma1 * (enth - R(0.5) * u2v2);
        R a = rsqrt_exact<R>(a2);
        R g1a2 = SHARED_RCP ? gamma1 * rcp<R>(a2) : gamma1 / a2;
        R euv = enth - u2v2;
        const R i2a = SHARED_RCP ? rcp<R>(R(2) * a) : R(0);
        C.u = u; C.v = v; C.enth = enth; C.a = a; C.g1a2 = g1a2; C.i2a = i2a;

        R d0 = hi[0] - lo[0];
        R d1 = hi[mu] - lo[mu];
        R d2 = hi[mv] - lo[mv];
        R d3 = hi[3] - lo[3];
        R a3 = g1a2 * (euv * d0 + u * d1 + v * d2 - d3);
        R a2w = d2 - v * d0;
        R a4 = SHARED_RCP ? (d1 + (a - u) * d0 - a * a3) * i2a : (d1 + (a - u) * d0 - a * a3) / (R(2) * a);
        R a1 = d0 - a3 - a4;
        wave[0][0] = a1; wave[0][mu] = a1 * (u - a); wave[0][mv] = a1 * v;
        wave[0][3] = a1 * (enth - u * a); wave[0][4] = R(0); s[0] = u - a;
        wave[1][0] = R(0); wave[1][mu] = R(0); wave[1][mv] = a2w; wave[1][3] = a2w * v;
        wave[1][4] = R(0); s[1] = u;
        wave[2][0] = a3; wave[2][mu] = a3 * u; wave[2][mv] = a3 * v;
        wave[2][3] = a3 * R(0.5) * u2v2; wave[2][4] = R(0); s[2] = u;
        wave[3][0] = a4; wave[3][mu] = a4 * (u + a); wave[3][mv] = a4 * v;
        wave[3][3] = a4 * (enth + u * a); wave[3][4] = R(0); s[3] = u + a;
        wave[4][0] = R(0); wave[4][mu] = R(0); wave[4][mv] = R(0); wave[4][3] = R(0);
        wave[4][4] = hi[4] - lo[4]; s[4] = u;

        // Sums that start at +0 never become -0, so a term s * w with a wave
        // entry w that is structurally zero adds +-0 and changes nothing (for
        // finite data): such terms are skipped (WNZ).
        if (!P.efix) {
            // s < 0 goes to amdq, s >= 0 to apdq; min(s, 0) and max(s, 0) put
            // an exact no-op term (+-0) in the other sum
            R sm[NW], sp[NW];
#pragma unroll
            for (int mw = 0; mw < NW; ++mw) { sm[mw] = fmin_ieee(s[mw], R(0)); sp[mw] = fmax_ieee(s[mw], R(0)); }
#pragma unroll
            for (int m = 0; m < NE; ++m) {
                R am = R(0), ap = R(0);
#pragma unroll
                for (int mw = 0; mw < NW; ++mw) {
                    if (!WNZ<DIR>(m, mw)) continue;
                    am = am + sm[mw] * wave[mw][m];
                    ap = ap + sp[mw] * wave[mw][m];
                }
                amdq[m] = am; apdq[m] = ap;
            }
            return;
        }
        const R gamma = P.gamma;
        // rpn2_euler_5wave.f:205-286: amdq with the transonic-rarefaction fix
        do {
            R sfract;
            // the left state's u - c (its pressure is the pl of the Roe averages)
            const R s0 = Sl.un[mu - 1] - Sl.c;
            if (s0 >= R(0) && s[0] > R(0)) {
#pragma unroll
                for (int m = 0; m < NE; ++m) amdq[m] = R(0);
                break;
            }
            R rho1 = lo[0] + wave[0][0];
            R rhou1 = lo[mu] + wave[0][mu];
            R rhov1 = lo[mv] + wave[0][mv];
            R en1 = lo[3] + wave[0][3];
            R s1;
            if constexpr (SHARED_RCP) {
                const R i1 = rcp<R>(rho1);
                R p1 = gamma1 * (en1 - R(0.5) * (rhou1 * rhou1 + rhov1 * rhov1) * i1);
                s1 = rhou1 * i1 - rsqrt_exact<R>(gamma * p1 * i1);
            } else {
                R p1 = gamma1 * (en1 - R(0.5) * (rhou1 * rhou1 + rhov1 * rhov1) / rho1);
                R c1 = rsqrt_exact<R>(gamma * p1 / rho1);
                s1 = rhou1 / rho1 - c1;
            }
            if (s0 < R(0) && s1 > R(0)) sfract = s0 * (s1 - s[0]) / (s1 - s0);
            else if (s[0] < R(0)) sfract = s[0];
            else sfract = R(0);
#pragma unroll
            for (int m = 0; m < NE; ++m) amdq[m] = sfract * wave[0][m];
            if (s[1] >= R(0)) break;
            // here s[1] = s[2] = s[4] = u < 0: a structurally zero entry adds -0,
            // which changes no value (not even -0 or +0)
#pragma unroll
            for (int m = 0; m < NE; ++m) {
                if (WNZ<DIR>(m, 1)) amdq[m] = amdq[m] + s[1] * wave[1][m];
                if (WNZ<DIR>(m, 2)) amdq[m] = amdq[m] + s[2] * wave[2][m];
                if (WNZ<DIR>(m, 4)) amdq[m] = amdq[m] + s[4] * wave[4][m];
            }
            R s3, s2;
            R rho2 = hi[0] - wave[3][0];
            R rhou2 = hi[mu] - wave[3][mu];
            R rhov2 = hi[mv] - wave[3][mv];
            R en2 = hi[3] - wave[3][3];
            s3 = Sr.un[mu - 1] + Sr.c;   // the right state's u + c
            if constexpr (SHARED_RCP) {
                const R i2 = rcp<R>(rho2);
                R p2 = gamma1 * (en2 - R(0.5) * (rhou2 * rhou2 + rhov2 * rhov2) * i2);
                s2 = rhou2 * i2 + rsqrt_exact<R>(gamma * p2 * i2);
            } else {
                R p2 = gamma1 * (en2 - R(0.5) * (rhou2 * rhou2 + rhov2 * rhov2) / rho2);
                R c2 = rsqrt_exact<R>(gamma * p2 / rho2);
                s2 = rhou2 / rho2 + c2;
            }
            if (s2 < R(0) && s3 > R(0)) sfract = s2 * (s3 - s[3]) / (s3 - s2);
            else if (s[3] < R(0)) sfract = s[3];
            else break;
#pragma unroll
            for (int m = 0; m < NE; ++m) amdq[m] = amdq[m] + sfract * wave[3][m];
        } while (0);
#pragma unroll
        for (int m = 0; m < NE; ++m) {
            R df = R(0);
#pragma unroll
            for (int mw = 0; mw < NW; ++mw)
                if (WNZ<DIR>(m, mw)) df = df + s[mw] * wave[mw][m];
            apdq[m] = df - amdq[m];
        }
    }
    template <int DIR>
    __device__ static __forceinline__ void rpt(const Params &P, int, const R (&asdq)[NE], const R *, const R *,
                                               const R *, R (&bm)[NE], R (&bp)[NE], const Cache &C) {
        constexpr int mu = MU<DIR>, mv = MV<DIR>;
        const R u = C.u, v = C.v, a = C.a, enth = C.enth;
        // the same expressions as in rpn (rpn2_euler_5wave.f:100-104), so the
        // values are identical; gamma1 / a^2 is rpn's own
        const R u2v2 = u * u + v * v;
        const R g1a2 = C.g1a2;
        const R euv = enth - u2v2;
        R a3 = g1a2 * (euv * asdq[0] + u * asdq[mu] + v * asdq[mv] - asdq[3]);
        R a2 = asdq[mu] - u * asdq[0];
        R a4;
        if constexpr (SHARED_RCP)   // 1/(2a) is rpn's, shared by the interface's two transverse solves
            a4 = (asdq[mv] + (a - v) * asdq[0] - a * a3) * C.i2a;
        else
            a4 = (asdq[mv] + (a - v) * asdq[0] - a * a3) / (R(2) * a);
        R a1 = asdq[0] - a3 - a4;
        R wb[4][NE], sb[4];
        wb[0][0] = a1; wb[0][mu] = a1 * u; wb[0][mv] = a1 * (v - a); wb[0][3] = a1 * (enth - v * a);
        wb[0][4] = R(0); sb[0] = v - a;
        wb[1][0] = a3; wb[1][mu] = a3 * u + a2; wb[1][mv] = a3 * v; wb[1][3] = a3 * R(0.5) * u2v2 + a2 * u;
        wb[1][4] = R(0); sb[1] = v;
        wb[2][0] = a4; wb[2][mu] = a4 * u; wb[2][mv] = a4 * (v + a); wb[2][3] = a4 * (enth + v * a);
        wb[2][4] = R(0); sb[2] = v + a;
        wb[3][0] = R(0); wb[3][mu] = R(0); wb[3][mv] = R(0); wb[3][3] = R(0); wb[3][4] = asdq[4];
        sb[3] = v;
        // waves 0-2 have components 0-3, wave 3 (tracer) component 4 only; the
        // sums start at +0, so the structurally zero terms are skipped, and
        // the sign of a zero min(s, 0) or max(s, 0) changes nothing
#pragma unroll
        for (int m = 0; m < NE; ++m) {
            R b_m = R(0), b_p = R(0);
#pragma unroll
            for (int mw = 0; mw < 4; ++mw) {
                if ((mw == 3) != (m == 4)) continue;
                b_m = b_m + fmin_ieee(sb[mw], R(0)) * wb[mw][m];
                b_p = b_p + fmax_ieee(sb[mw], R(0)) * wb[mw][m];
            }
            bm[m] = b_m; bp[m] = b_p;
        }
    }
};

// ---------------------------------------------------------------------------
// Shallow water: q = (h, hu, hv), Roe solver with entropy fix, gravity grav.
// ---------------------------------------------------------------------------
template <typename R> struct ShallowRoe {
    static constexpr int NE = 3, NW = 3, NA = 0;
    static constexpr bool HAS_EFIX = false;   // the entropy fix is always on
    static constexpr bool CELL_STATE = false;  // the per-state quantities below depend on the direction
    __host__ __device__ static constexpr bool SDIST(int) { return true; }
    static constexpr bool SKIP_ZERO_WAVES = false;
    struct CellState {};
    struct Params { R grav; };
    struct Cache { R u, v, a; };
    static constexpr bool SPEEDS_FROM_CACHE = false;
    __device__ static __forceinline__ void speeds(const Cache &, R (&)[NW]) {}
    template <int DIR> static constexpr int MU = DIR == 0 ? 1 : 2;
    template <int DIR> static constexpr int MV = DIR == 0 ? 2 : 1;
    template <int DIR> __host__ __device__ static constexpr bool WNZ(int m, int mw) {
        return mw != 1 || m == (DIR == 0 ? 2 : 1);
    }
    // The quantities of one state that each Riemann problem it is part of (as
    // the left or the right state) computes: sqrt(h); hu/sqrt(h) and
    // hv/sqrt(h), the terms of the Roe velocities; hu/h and sqrt(g h), the
    // entropy fix's characteristic speed.  A caller that meets a state in two
    // problems (SharpClaw's interface and in-cell solves) computes them once;
    // each is the same operation on the same operands either way.
    struct State { R hs, umh, vmh, un, sgh; };
    static constexpr bool HAS_STATE = true;
    // a state record moved one lane up (lane L receives lane L-1's)
    __device__ static __forceinline__ State state_up(const State &S) {
        return State{lane_up(S.hs), lane_up(S.umh), lane_up(S.vmh), lane_up(S.un), lane_up(S.sgh)};
    }
    template <int DIR>
    __device__ static __forceinline__ State state(const Params &P, const R (&q)[NE]) {
        constexpr int mu = MU<DIR>, mv = MV<DIR>;
        State S;
        S.hs = rsqrt_exact<R>(q[0]);
        div2<R>(q[mu], q[mv], S.hs, S.umh, S.vmh);
        S.un = SHARED_RCP ? q[mu] * rcp<R>(q[0]) : q[mu] / q[0];
        // with g = 1, g h is h and sqrt(g h) is sqrt(h) (also for g > 0 where
        // g h equals h: h = 0 or inf); the test is on the uniform g, so it
        // does not split the code per lane
        S.sgh = P.grav == R(1) ? S.hs : rsqrt_exact<R>(P.grav * q[0]);
        return S;
    }
    template <int DIR>
    __device__ static __forceinline__ void rpn(const Params &P, const R (&lo)[NE], const R (&hi)[NE],
                                               const R *, const R *, R (&wave)[NW][NE], R (&s)[NW],
                                               R (&amdq)[NE], R (&apdq)[NE], Cache &C) {
        rpn_s<DIR>(P, lo, hi, state<DIR>(P, lo), state<DIR>(P, hi), wave, s, amdq, apdq, C);
    }
    template <int DIR>
    __device__ static __forceinline__ void rpn_s(const Params &P, const R (&lo)[NE], const R (&hi)[NE],
                                                 const State &Sl, const State &Sr, R (&wave)[NW][NE], R (&s)[NW],
                                                 R (&amdq)[NE], R (&apdq)[NE], Cache &C) {
        constexpr int mu = MU<DIR>, mv = MV<DIR>;
        const R grav = P.grav;
        R hl = lo[0], hr = hi[0];
        R hbar = (hl + hr) * R(0.5);
        R hsqrtl = Sl.hs, hsqrtr = Sr.hs;
        R hsq2 = hsqrtl + hsqrtr;
        // u = (hu_l/sqrt(h_l) + hu_r/sqrt(h_r)) / (sqrt(h_l) + sqrt(h_r)), v alike
        R u, v;
        div2<R>(Sl.umh + Sr.umh, Sl.vmh + Sr.vmh, hsq2, u, v);
        R a = rsqrt_exact<R>(grav * hbar);
        C.u = u; C.v = v; C.a = a;
        R d0 = hr - hl;
        R d1 = hi[mu] - lo[mu];
        R d2 = hi[mv] - lo[mv];
        R a1, a3;
        div2<R>((u + a) * d0 - d1, -(u - a) * d0 + d1, R(2) * a, a1, a3);
        R a2 = -v * d0 + d2;
        wave[0][0] = a1; wave[0][mu] = a1 * (u - a); wave[0][mv] = a1 * v; s[0] = u - a;
        wave[1][0] = R(0); wave[1][mu] = R(0); wave[1][mv] = a2; s[1] = u;
        wave[2][0] = a3; wave[2][mu] = a3 * (u + a); wave[2][mv] = a3 * v; s[2] = u + a;
        // depths are positive: the shared-reciprocal build multiplies by 1/h
        do {
            R sfract;
            R s0 = Sl.un - Sl.sgh;
            if (s0 >= R(0) && s[0] > R(0)) {
#pragma unroll
                for (int m = 0; m < NE; ++m) amdq[m] = R(0);
                break;
            }
            R h1 = hl + wave[0][0];
            R hu1 = lo[mu] + wave[0][mu];
            R s1 = (SHARED_RCP ? hu1 * rcp<R>(h1) : hu1 / h1) - rsqrt_exact<R>(grav * h1);
            if (s0 < R(0) && s1 > R(0)) sfract = s0 * (s1 - s[0]) / (s1 - s0);
            else if (s[0] < R(0)) sfract = s[0];
            else sfract = R(0);
#pragma unroll
            for (int m = 0; m < NE; ++m) amdq[m] = sfract * wave[0][m];
            if (s[1] >= R(0)) break;
            // s[1] = u < 0: a structurally zero entry adds -0, which changes no value
#pragma unroll
            for (int m = 0; m < NE; ++m)
                if (WNZ<DIR>(m, 1)) amdq[m] = amdq[m] + s[1] * wave[1][m];
            R s03 = Sr.un + Sr.sgh;
            R h3 = hr - wave[2][0];
            R hu3 = hi[mu] - wave[2][mu];
            R s3 = (SHARED_RCP ? hu3 * rcp<R>(h3) : hu3 / h3) + rsqrt_exact<R>(grav * h3);
            if (s3 < R(0) && s03 > R(0)) sfract = s3 * (s03 - s[2]) / (s03 - s3);
            else if (s[2] < R(0)) sfract = s[2];
            else break;
#pragma unroll
            for (int m = 0; m < NE; ++m) amdq[m] = amdq[m] + sfract * wave[2][m];
        } while (0);
#pragma unroll
        for (int m = 0; m < NE; ++m) {
            R df = R(0);
#pragma unroll
            for (int mw = 0; mw < NW; ++mw)
                if (WNZ<DIR>(m, mw)) df = df + s[mw] * wave[mw][m];
            apdq[m] = df - amdq[m];
        }
    }
    template <int DIR>
    __device__ static __forceinline__ void rpt(const Params &, int, const R (&asdq)[NE], const R *, const R *,
                                               const R *, R (&bm)[NE], R (&bp)[NE], const Cache &C) {
        constexpr int mu = MU<DIR>, mv = MV<DIR>;
        const R u = C.u, v = C.v, a = C.a;
        R q0 = asdq[0], qu = asdq[mu], qv = asdq[mv];
        const R h2a = SHARED_RCP ? R(0.5) * rcp<R>(a) : R(0.5) / a;
        R a1 = h2a * ((v + a) * q0 - qv);
        R a2 = qu - u * q0;
        R a3 = h2a * (-(v - a) * q0 + qv);
        R wb[3][NE], sb[3];
        wb[0][0] = a1; wb[0][mu] = a1 * u; wb[0][mv] = a1 * (v - a); sb[0] = v - a;
        wb[1][0] = R(0); wb[1][mu] = a2; wb[1][mv] = R(0); sb[1] = v;
        wb[2][0] = a3; wb[2][mu] = a3 * u; wb[2][mv] = a3 * (v + a); sb[2] = v + a;
        // wave 1 has component mu only; the sums start at +0, so the
        // structurally zero terms are skipped, and the sign of a zero
        // min(s, 0) or max(s, 0) changes nothing
#pragma unroll
        for (int m = 0; m < NE; ++m) {
            R b_m = R(0), b_p = R(0);
#pragma unroll
            for (int mw = 0; mw < 3; ++mw) {
                if (mw == 1 && m != mu) continue;
                b_m = b_m + fmin_ieee(sb[mw], R(0)) * wb[mw][m];
                b_p = b_p + fmax_ieee(sb[mw], R(0)) * wb[mw][m];
            }
            bm[m] = b_m; bp[m] = b_p;
        }
    }
};

}  // namespace clawhip
