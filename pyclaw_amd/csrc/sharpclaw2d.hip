// SharpClaw 2D: fused WENO5 reconstruction + Riemann solves + Runge-Kutta
// stage update, for gfx950.
//
// Replaces the reference's dq evaluation (src/pyclaw/sharpclaw/flux2.f90:2-101
// -> flux1_2.f90:2-196 -> reconstruct.f90:85-116 -> weno.f90:5-83) and the
// numpy stage combinations of SSP104 / SSP33 / Euler
// (src/pyclaw/sharpclaw/solver.py:185-255).  One launch evaluates
//   dq(m,i,j) = (0 + dq1d_x(m,i)) + dq1d_y(m,j)
//   dq1d(m,i) = 0 - dtdx*(((amdq(i+1) + apdq(i)) + amdq2(i)) + apdq2(i))
// (flux1_2.f90:190-193, flux2.f90:59-61, 91-93) and combines it with the
// stage registers in the kernel epilogue, so the only HBM traffic is the
// stage's inputs and outputs.
//
// Work decomposition: one wavefront per strip of 64 columns (58 produced:
// WENO needs two cells on each side, the interface fluctuations one more),
// marching up a band of rows with a five-row register window for the
// y-direction reconstruction.  x-neighbours come from other lanes
// (ds_bpermute), y-neighbours from the lane's own registers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <new>
#include <type_traits>

#include "common.hpp"
#include "ctx.hpp"
#include "riemann.hpp"

// The kernels and their launcher are compiled twice: here without
// floating-point contraction (the reference's arithmetic bit for bit), in
// namespace clawhip::exact, and from sharpclaw2d_fma.hip with contraction and
// shared reciprocals, in clawhip::fused, for the CLAW_F64_FMA precision.  The
// C entry points are compiled here only.
#ifndef CLAW_SHARP_VARIANT
#define CLAW_SHARP_VARIANT exact
#define CLAW_SHARP_MAIN 1
#define CLAW_SHARP_F64_KEY CLAW_F64
#endif

namespace clawhip {
namespace CLAW_SHARP_VARIANT {

constexpr int SX_USEFUL = WAVE - 6;
constexpr int SNG = 3;   // SharpClaw ghost width ((weno_order+1)/2, sharpclaw/solver.py:159)

// Stage epilogues (sharpclaw/solver.py:201-251); q0 is the step's initial state.
enum StageMode {
    SM_DQ = 0,          // out = dq
    SM_EULER = 1,       // out = in + dq
    SM_SSP104_A = 2,    // out = in + dq/6.
    SM_SSP104_MID = 3,  // t = in + dq/6.; s2 = q0/25. + (9./25)*t; out = 15.*s2 - 5.*t
    SM_SSP104_END = 4,  // out = (s2 + 0.6*in) + 0.1*dq
    SM_SSP33_2 = 5,     // out = 0.75*q0 + 0.25*(in + dq)
    SM_SSP33_3 = 6      // out = 1./3.*q0 + 2./3.*(in + dq)
};

// The solver's per-state quantities (a placeholder type for solvers without them).
template <class RP, bool HAS> struct StateOf { using type = char; };
template <class RP> struct StateOf<RP, true> { using type = typename RP::State; };

// weno.f90:32-79 for one component: left (ql) and right (qr) edge values of
// the cell whose average is q0, from the averages qm2..qp2.
template <typename R>
__device__ __forceinline__ void weno5(R qm2, R qm1, R q0, R qp1, R qp2, R &ql, R &qr) {
    const R eps = R((double)1.0e-36f);
    // The smoothness indicators are expanded quadratic forms whose O(q^2)
    // terms cancel to O(dq^2) on smooth data; they are never contracted, so
    // every build computes them with the reference's roundings.
    R sigma0, sigma1, sigma2;
    {
#pragma clang fp contract(off)
        sigma0 = R(3.3333333333333333333333333333333333) * (q0 * q0)
            - R(10.333333333333333333333333333333333) * q0 * qp1
            + R(3.6666666666666666666666666666666667) * q0 * qp2
            + R(8.3333333333333333333333333333333333) * (qp1 * qp1)
            - R(6.3333333333333333333333333333333333) * qp1 * qp2
            + R(1.3333333333333333333333333333333333) * (qp2 * qp2);
        sigma1 = R(4.3333333333333333333333333333333333) * (q0 * q0)
            - R(4.3333333333333333333333333333333333) * q0 * qp1
            - R(4.3333333333333333333333333333333333) * q0 * qm1
            + R(1.3333333333333333333333333333333333) * (qp1 * qp1)
            + R(1.6666666666666666666666666666666667) * qp1 * qm1
            + R(1.3333333333333333333333333333333333) * (qm1 * qm1);
        sigma2 = R(3.3333333333333333333333333333333333) * (q0 * q0)
            - R(10.333333333333333333333333333333333) * q0 * qm1
            + R(3.6666666666666666666666666666666667) * q0 * qm2
            + R(8.3333333333333333333333333333333333) * (qm1 * qm1)
            - R(6.3333333333333333333333333333333333) * qm1 * qm2
            + R(1.3333333333333333333333333333333333) * (qm2 * qm2);
    }
    R t0 = eps + sigma0, t1 = eps + sigma1, t2 = eps + sigma2;
    R t00 = t0 * t0, t11 = t1 * t1, t22 = t2 * t2;
    R omega0, omega1, omega2, omega3, omega4, omega5;
    if constexpr (SHARED_RCP) {
        // one reciprocal per smoothness indicator, by the hardware reciprocal
        // and two Newton steps: (eps + sigma)^2 >= 1e-72 is positive and normal
        const R i0 = rcp<R>(t00), i1 = rcp<R>(t11), i2 = rcp<R>(t22);
        omega0 = R(0.1) * i0; omega1 = R(0.6) * i1; omega2 = R(0.3) * i2;
        omega3 = R(0.3) * i0; omega4 = R(0.6) * i1; omega5 = R(0.1) * i2;
    } else {
        // omega0 = 0.1/t00, omega3 = 0.3/t00 and omega2 = 0.3/t22, omega5 =
        // 0.1/t22 share their denominators' part of the division sequence;
        // omega1 and omega4 are the same quotient 0.6/t11
        div2<R>(R(0.1), R(0.3), t00, omega0, omega3);
        omega1 = R(0.6) / t11; omega4 = omega1;
        div2<R>(R(0.3), R(0.1), t22, omega2, omega5);
    }
    R fr0 = R(1.8333333333333333333333333333333333) * q0 - R(1.1666666666666666666666666666666667) * qp1
        + R(0.33333333333333333333333333333333333) * qp2;
    R fr1 = R(0.83333333333333333333333333333333333) * q0 - R(0.16666666666666666666666666666666667) * qp1
        + R(0.33333333333333333333333333333333333) * qm1;
    R fr2 = R(0.33333333333333333333333333333333333) * q0 + R(0.83333333333333333333333333333333333) * qm1
        - R(0.16666666666666666666666666666666667) * qm2;
    R fr3 = R(0.33333333333333333333333333333333333) * q0 + R(0.83333333333333333333333333333333333) * qp1
        - R(0.16666666666666666666666666666666667) * qp2;
    R fr4 = R(0.83333333333333333333333333333333333) * q0 + R(0.33333333333333333333333333333333333) * qp1
        - R(0.16666666666666666666666666666666667) * qm1;
    R fr5 = R(1.8333333333333333333333333333333333) * q0 - R(1.1666666666666666666666666666666667) * qm1
        + R(0.33333333333333333333333333333333333) * qm2;
    if constexpr (SHARED_RCP) {   // the weight sums are positive and normal as well
        ql = (fr0 * omega0 + fr1 * omega1 + fr2 * omega2) * rcp<R>(omega0 + omega1 + omega2);
        qr = (fr3 * omega3 + fr4 * omega4 + fr5 * omega5) * rcp<R>(omega3 + omega4 + omega5);
    } else {
        ql = (fr0 * omega0 + fr1 * omega1 + fr2 * omega2) / (omega0 + omega1 + omega2);
        qr = (fr3 * omega3 + fr4 * omega4 + fr5 * omega5) / (omega3 + omega4 + omega5);
    }
}

template <class RP, typename R>
struct SharpParams {
    const R *qin;     // state dq is evaluated on
    const R *q0;      // step's initial state (SSP104_MID, SSP33)
    R *s2;            // SSP104 second register (written by MID, read by END)
    R *out;           // stage output
    const R *aux;
    GridDev g;
    BCDev bq, ba;
    typename RP::Params P;
    const double *dt;
    double *cfl;      // this stage's Courant number (max-combined)
    unsigned long long *cgrp;   // a context's Courant group words (cfl_group_max), or null
    double dx, dy;
    int mode;
    int H, nsx, nsy, nsyb, wpb;
    int fill;         // > 0: the paired fill decomposition with this share (fill_piece, common.hpp)
};

// Rows [R0, R1) of column strip sx; cfl is max-combined with the Courant
// numbers of the interfaces this band owns.
template <class RP, typename R>
__device__ __forceinline__ void sharp_band(const SharpParams<RP, R> &p, const int sx, const int R0, const int R1,
                                           R &cfl) {
    constexpr int NE = RP::NE, NW = RP::NW;
    constexpr bool HAS_AUX = RP::NA > 0;
    constexpr int NA = HAS_AUX ? RP::NA : 1;
    using Cache = typename RP::Cache;
    // per-state quantities shared by the two Riemann problems a reconstructed
    // state takes part in (riemann.hpp); solvers without them recompute
    constexpr bool ST = RP::HAS_STATE;
    using State = typename std::conditional<ST, typename StateOf<RP, ST>::type, char>::type;

    const int lane = threadIdx.x & (WAVE - 1);
    const GridDev g = p.g;
    const int c0 = sx * SX_USEFUL;
    const int i = c0 - 3 + lane;
    const bool produce = lane >= 3 && lane < 3 + SX_USEFUL && i < g.mx;

    const double dtd = *p.dt;   // uniform constants in scalar registers (whole wavefront active)
    const R dtdx = uniform(R(dtd / p.dx)), dtdy = uniform(R(dtd / p.dy));
    const auto &P = p.P;
    const int mode = p.mode;

    // five-row window of q (rows w-2..w+2) and aux (rows w-1, w)
    R win[5][NE];
    R aw[5][NA];
    const int wlast = R1 + (R1 == g.my ? 1 : 0);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        load_q<NE>(p.qin, g, p.bq, i, R0 - 3 + k, win[k + 1]);
        if constexpr (HAS_AUX) load_aux<NA>(p.aux, g, p.ba, i, R0 - 3 + k, aw[k + 1]);
    }
    R qrY_prev[NE], apY_prev[NE], am2Y_prev[NE], ap2Y_prev[NE];
#pragma unroll
    for (int m = 0; m < NE; ++m) { qrY_prev[m] = apY_prev[m] = am2Y_prev[m] = ap2Y_prev[m] = R(0); }
    State srY_prev{};
#pragma unroll
    for (int k = 0; k < 5; ++k)
#pragma unroll
        for (int a = 0; a < NA; ++a) if (!HAS_AUX) aw[k][a] = R(1);

    for (int w = R0 - 1; w <= wlast; ++w) {
        // shift the window and load row w+2
#pragma unroll
        for (int k = 0; k < 4; ++k) {
#pragma unroll
            for (int m = 0; m < NE; ++m) win[k][m] = win[k + 1][m];
#pragma unroll
            for (int a = 0; a < NA; ++a) aw[k][a] = aw[k + 1][a];
        }
        load_q<NE>(p.qin, g, p.bq, i, w + 2, win[4]);
        if constexpr (HAS_AUX) load_aux<NA>(p.aux, g, p.ba, i, w + 2, aw[4]);
        // win[k] = row w-2+k

        // ---- y reconstruction at (i, w) and the y-interface (i, w-1/2) --------
        R qlY[NE], qrY[NE];
#pragma unroll
        for (int m = 0; m < NE; ++m) weno5<R>(win[0][m], win[1][m], win[2][m], win[3][m], win[4][m], qlY[m], qrY[m]);
        R amY[NE], apY[NE], am2Y[NE], ap2Y[NE];
        State slY{}, srY{};
        if constexpr (ST) { srY = RP::template state<1>(P, qrY); }
        if (w >= R0) {
            R wave[NW][NE], s[NW];
            Cache c;
            if constexpr (ST) {
                slY = RP::template state<1>(P, qlY);
                RP::template rpn_s<1>(P, qrY_prev, qlY, srY_prev, slY, wave, s, amY, apY, c);
            } else {
                RP::template rpn<1>(P, qrY_prev, qlY, aw[1], aw[2], wave, s, amY, apY, c);
            }
            if (w <= g.my && i >= -1 && i <= g.mx) {
#pragma unroll
                for (int mw = 0; mw < NW; ++mw) cfl = fmax_ieee(cfl, dtdy * rabs(s[mw]));
            }
            if constexpr (ST) {
                RP::template rpn_s<1>(P, qlY, qrY, slY, srY, wave, s, am2Y, ap2Y, c);
            } else {
                RP::template rpn<1>(P, qlY, qrY, aw[2], aw[2], wave, s, am2Y, ap2Y, c);
            }
        }

        // ---- output row o = w-1: x part and combination --------------------------
        const int o = w - 1;
        const bool out_row = (o >= R0 && o < R1);
        const bool cfl_row = (o == -1 && R0 == 0) || (o == g.my && R1 == g.my);
        if (out_row || cfl_row) {
            // row o is win[1]; WENO in x at (i, o)
            R xm2[NE], xm1[NE], xp1[NE], xp2[NE], qlX[NE], qrX[NE], qrXl[NE];
#pragma unroll
            for (int m = 0; m < NE; ++m) {
                xm2[m] = lane_up(lane_up(win[1][m])); xm1[m] = lane_up(win[1][m]);
                xp1[m] = lane_dn(win[1][m]); xp2[m] = lane_dn(lane_dn(win[1][m]));
                weno5<R>(xm2[m], xm1[m], win[1][m], xp1[m], xp2[m], qlX[m], qrX[m]);
            }
#pragma unroll
            for (int m = 0; m < NE; ++m) qrXl[m] = lane_up(qrX[m]);
            R alx[NA];
            if constexpr (HAS_AUX) {
#pragma unroll
                for (int a = 0; a < NA; ++a) alx[a] = lane_up(aw[1][a]);
            }
            R amX[NE], apX[NE], am2X[NE], ap2X[NE];
            {
                R wave[NW][NE], s[NW];
                Cache c;
                if constexpr (ST) {
                    // the left neighbour's right-edge state and its quantities come
                    // from the lane to the left
                    const State srX = RP::template state<0>(P, qrX), slX = RP::template state<0>(P, qlX);
                    const State srXl = RP::state_up(srX);
                    RP::template rpn_s<0>(P, qrXl, qlX, srXl, slX, wave, s, amX, apX, c);
                    if (i >= 0 && i <= g.mx && lane >= 3 && lane <= WAVE - 3) {
#pragma unroll
                        for (int mw = 0; mw < NW; ++mw) cfl = fmax_ieee(cfl, dtdx * rabs(s[mw]));
                    }
                    RP::template rpn_s<0>(P, qlX, qrX, slX, srX, wave, s, am2X, ap2X, c);
                } else {
                    RP::template rpn<0>(P, qrXl, qlX, alx, aw[1], wave, s, amX, apX, c);
                    if (i >= 0 && i <= g.mx && lane >= 3 && lane <= WAVE - 3) {
#pragma unroll
                        for (int mw = 0; mw < NW; ++mw) cfl = fmax_ieee(cfl, dtdx * rabs(s[mw]));
                    }
                    RP::template rpn<0>(P, qlX, qrX, aw[1], aw[1], wave, s, am2X, ap2X, c);
                }
            }
            if (out_row) {
                const long long off = (long long)(o + g.ng) * g.ldq + (i + g.ng);
#pragma unroll
                for (int m = 0; m < NE; ++m) {
                    R amXr = lane_dn(amX[m]);
                    R dqx = R(0) - dtdx * (((amXr + apX[m]) + am2X[m]) + ap2X[m]);
                    R dqy = R(0) - dtdy * (((amY[m] + apY_prev[m]) + am2Y_prev[m]) + ap2Y_prev[m]);
                    R dq = (R(0) + dqx) + dqy;
                    if (produce) {
                        const long long at = m * g.plane + off;
                        R v;
                        switch (mode) {
                        case SM_DQ: v = dq; break;
                        case SM_EULER: v = p.qin[at] + dq; break;
                        case SM_SSP104_A: v = p.qin[at] + dq / R(6); break;
                        case SM_SSP104_MID: {
                            R t = p.qin[at] + dq / R(6);
                            R s2v = p.q0[at] / R(25) + R(9. / 25) * t;
                            p.s2[at] = s2v;
                            v = R(15) * s2v - R(5) * t;
                            break;
                        }
                        case SM_SSP104_END: v = (p.s2[at] + R(0.6) * p.qin[at]) + R(0.1) * dq; break;
                        case SM_SSP33_2: v = R(0.75) * p.q0[at] + R(0.25) * (p.qin[at] + dq); break;
                        default: v = R(1. / 3.) * p.q0[at] + R(2. / 3.) * (p.qin[at] + dq); break;
                        }
                        p.out[at] = v;
                    }
                }
            }
        }
        // ---- carry the y-interface data to the next row ---------------------------
#pragma unroll
        for (int m = 0; m < NE; ++m) {
            qrY_prev[m] = qrY[m];
            if (w >= R0) { apY_prev[m] = apY[m]; am2Y_prev[m] = am2Y[m]; ap2Y_prev[m] = ap2Y[m]; }
        }
        srY_prev = srY;
    }
}

// One wavefront's work: a band of p.H rows of one strip, or (p.fill > 0)
// its piece of the paired fill decomposition (fill_piece, common.hpp; share
// p.fill), one band either way.
template <class RP, typename R>
__device__ __forceinline__ void sharp_stage_body(const SharpParams<RP, R> &p) {
    const int lane = threadIdx.x & (WAVE - 1);
    // Courant number: the reference's max(cfl, dtdx * s, -dtdx * s) is
    // dtdx * |s| exactly (dtdx > 0), one IEEE max per wave
    R cfl = R(0);
    int sx, R0, R1;
    bool active = true;
    if (p.fill > 0) {
        // the younger pieces first in each strip (shallow water 1024x1024:
        // 0.973 Gcell/s against 0.948 older first)
        fill_piece<false, false>(p.nsx, p.g.my, p.fill & 255, p.fill >> 8, sx, R0, R1);
    } else {
        const int vb = xcd_block();
        sx = vb / p.nsyb;
        const int sy = (vb % p.nsyb) * p.wpb + (threadIdx.x >> 6);
        R0 = sy * p.H;
        R1 = min(R0 + p.H, p.g.my);
        // wavefronts past the grid's bands still count themselves in below
        active = sx < p.nsx && sy < p.nsy;
    }
    if (active) sharp_band<RP, R>(p, sx, R0, R1, cfl);
    const double c = wave_max((double)cfl);
    if (p.cgrp) cfl_group_max(p.cfl, p.cgrp, c);
    else if (lane == 0) atomic_max_nonneg(p.cfl, c);
}

template <class RP, typename R>
__global__ __launch_bounds__(256) void sharp_stage_kernel(const SharpParams<RP, R> p) {
    sharp_stage_body<RP, R>(p);
}

// The same kernel compiled for three wavefronts per SIMD (168 VGPRs: the FP64
// acoustics instantiations fit, shallow water and Euler spill); the SharpClaw
// autotuner times both builds.  The default build's occupancy follows its
// register use (FP64 shallow water: 222 VGPRs, two wavefronts per SIMD).
template <class RP, typename R>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3)))
void sharp_stage_kernel_w3(const SharpParams<RP, R> p) {
    sharp_stage_body<RP, R>(p);
}

__global__ void zero_slot_kernel(double *p) { *p = 0.0; }
__global__ void zero_slots_kernel(double *p, int n) {
    const int k = threadIdx.x;
    if (k < n) p[k] = 0.0;
}

template <class RP, typename R>
static typename RP::Params make_sharp_params(const ClawRiemann *rp);
template <> Acoustics<double>::Params make_sharp_params<Acoustics<double>, double>(const ClawRiemann *rp) {
    return {rp->p[0], rp->p[1], rp->p[2], rp->p[3]};
}
template <> Acoustics<float>::Params make_sharp_params<Acoustics<float>, float>(const ClawRiemann *rp) {
    return {(float)rp->p[0], (float)rp->p[1], (float)rp->p[2], (float)rp->p[3]};
}
template <> VCAcoustics<double>::Params make_sharp_params<VCAcoustics<double>, double>(const ClawRiemann *) { return {0.0}; }
template <> VCAcoustics<float>::Params make_sharp_params<VCAcoustics<float>, float>(const ClawRiemann *) { return {0.0f}; }
template <> Euler5<double>::Params make_sharp_params<Euler5<double>, double>(const ClawRiemann *rp) {
    return {rp->p[0], rp->p[1], rp->efix};
}
template <> Euler5<float>::Params make_sharp_params<Euler5<float>, float>(const ClawRiemann *rp) {
    return {(float)rp->p[0], (float)rp->p[1], rp->efix};
}
template <> ShallowRoe<double>::Params make_sharp_params<ShallowRoe<double>, double>(const ClawRiemann *rp) { return {rp->p[0]}; }
template <> ShallowRoe<float>::Params make_sharp_params<ShallowRoe<float>, float>(const ClawRiemann *rp) { return {(float)rp->p[0]}; }

// The paired fill decomposition's shape (fill_piece, common.hpp): the older
// wavefront's share of a SIMD pair, 2 units of rows for the younger one's 1,
// the younger pieces 15/16 of a unit.  The younger piece's row count is
// rounded down per strip, and strips of 170 and 171 units at 1024x1024 give
// 6- and 5-row pieces at 16/16 but 5 rows everywhere at 15/16 (pieces of one
// size balance the SIMDs better).  Shallow water 1024x1024, Gcell-updates/s,
// with the mirrored pairs and the Courant group slots: 13/16 (4 rows):
// 0.992-0.995; 14/16 and 15/16 (5 rows): 1.003-1.008; 16/16 (5 and 6 rows):
// 0.975; 17/16 and 18/16 (6 rows): 0.992-0.994.  Before those: share 2,
// 12/16: 0.938; 14/16: 0.972; 16/16: 0.970; 18/16: 0.968; 20/16: 0.964;
// share 3: 0.938; share 1 (equal pieces): 0.933; equal shares running over
// strip ends: 0.91.
constexpr int SHARP_FILL_A = 2, SHARP_FILL_C = 15;

// Default launch shape.  Band height 0 is the paired fill decomposition: one
// wavefront per slot of the build (CUs x the wavefronts per CU that the
// kernel's register and LDS use allow), the two wavefronts of a SIMD
// splitting their rows so that both finish together.
static int sharp_band_height(int rows, int strips_x, int &w, int &wpb) {
    const long long work = (long long)rows * strips_x;
    if (work >= 7LL * 2048) {
        // shares of at least ~7 rows at two wavefronts per SIMD (1024x1024
        // shallow water: 9 rows each, 0.87 Gcell/s against 0.79 for 7-row
        // bands of the three-per-SIMD build in four-wavefront blocks)
        w = 1;
        wpb = 1;
        return 0;
    }
    long long target = 2048;
    int h = (int)std::max<long long>(16, (work + target - 1) / target);
    return std::min(h, std::max(rows, 1));
}

static TuneKey sharp_key(const ClawGrid *grid, const ClawRiemann *rp, int precision) {
    return TuneKey{rp->kind, SHARP_KEY, 0, precision, grid->mx, grid->my};
}

// band height, build (1: default, 3: three waves per SIMD) and wavefronts per
// block for this call: the tuner's candidate, the tuned value, or the default
static int sharp_band(Ctx *ctx, const ClawGrid *grid, const ClawRiemann *rp, int precision, int strips, int &w,
                      int &wpb) {
    w = 1;
    wpb = 4;
    if (ctx && ctx->force_band != 0) {   // a tuner candidate; -1 is the fill decomposition
        w = ctx->force_w;
        if (ctx->force_wpb > 0) wpb = ctx->force_wpb;
        return ctx->force_band > 0 ? std::min(ctx->force_band, grid->my) : 0;
    }
    if (ctx) {
        const TuneKey key = sharp_key(grid, rp, precision);
        auto iw = ctx->wpb.find(key);
        if (iw != ctx->wpb.end()) wpb = iw->second;
        auto it = ctx->band.find(key);
        if (it != ctx->band.end()) { w = it->second.second; return std::min(it->second.first, grid->my); }
    }
    int dw = wpb;
    const int h = sharp_band_height(grid->my, strips, w, dw);
    if (!(ctx && ctx->wpb.count(sharp_key(grid, rp, precision)))) wpb = dw;
    return h;
}

// Wavefront slots of a build on the current (or the context's) device: CUs x
// the blocks per CU that the runtime's occupancy query gives for the kernel
// (its registers and LDS) x wavefronts per block.  The query's answer is kept
// per instantiation, build and block size.
template <class RP, typename R>
static int sharp_slots(Ctx *ctx, int w, int wpb) {
    static std::atomic<int> per_cu[2][5];
    const int b = w == 3 ? 1 : 0;
    int nb = per_cu[b][wpb].load(std::memory_order_relaxed);
    if (nb <= 0) {
        const void *fn = w == 3 ? reinterpret_cast<const void *>(&sharp_stage_kernel_w3<RP, R>)
                                : reinterpret_cast<const void *>(&sharp_stage_kernel<RP, R>);
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, WAVE * wpb, 0) != hipSuccess || nb <= 0)
            nb = std::max(1, 4 * (w == 3 ? 3 : 1) / wpb);
        per_cu[b][wpb].store(nb, std::memory_order_relaxed);
    }
    int dev = 0, cus = 0;
    if (ctx) dev = ctx->device;
    else (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
    return cus * nb * wpb;
}

template <class RP, typename R>
static int launch_sharp(Ctx *ctx, const ClawGrid *grid, const ClawRiemann *rp, const ClawBC *bc, int mode,
                        const void *qin, const void *q0, void *s2, void *out, const void *aux,
                        const double *dt, double *cfl, bool zero_cfl, hipStream_t st) {
    int wpb = 4;
    SharpParams<RP, R> p;
    p.qin = (const R *)qin; p.q0 = (const R *)q0; p.s2 = (R *)s2; p.out = (R *)out; p.aux = (const R *)aux;
    p.g = GridDev{grid->mx, grid->my, grid->ng, grid->ldq, grid->plane};
    p.bq = BCDev{{bc->lower[0], bc->lower[1]}, {bc->upper[0], bc->upper[1]}};
    p.ba = BCDev{{bc->aux_lower[0], bc->aux_lower[1]}, {bc->aux_upper[0], bc->aux_upper[1]}};
    p.P = make_sharp_params<RP, R>(rp);
    p.dt = dt; p.cfl = cfl; p.dx = grid->dx; p.dy = grid->dy; p.mode = mode;
    // without a context the stage's wavefronts max-combine into *cfl directly
    p.cgrp = ctx ? ctx->cfl_groups : nullptr;
    p.nsx = (grid->mx + SX_USEFUL - 1) / SX_USEFUL;
    int w = 1;
    p.H = sharp_band(ctx, grid, rp, std::is_same<R, float>::value ? CLAW_F32 : CLAW_SHARP_F64_KEY, p.nsx, w, wpb);
    p.fill = 0;
    int fill_blocks = 0;
    if (p.H == 0) {
        // the paired fill decomposition: one-wavefront blocks, one per slot of
        // the build; the older wavefront of each SIMD pair takes
        // SHARP_FILL_A units of rows for the younger one's one (equal
        // shares in the three-per-SIMD build, where there are no pairs);
        // where it cannot tile the grid, bands for about 2048 wavefronts
        wpb = 1;
        fill_blocks = sharp_slots<RP, R>(ctx, w, 1) & ~15;
        const int A = w == 3 ? 1 : SHARP_FILL_A, C = w == 3 ? 16 : SHARP_FILL_C;
        if (fill_usable(fill_blocks, p.nsx, grid->my, A, C)) {
            p.fill = fill_param(A, C);
            p.H = grid->my;
        } else {
            const long long work = (long long)grid->my * p.nsx;
            p.H = (int)std::min<long long>(grid->my, std::max<long long>(4, (work + 2047) / 2048));
        }
    }
    p.nsy = (grid->my + p.H - 1) / p.H;
    p.wpb = wpb;
    p.nsyb = (p.nsy + wpb - 1) / wpb;
    const int nblocks = p.fill > 0 ? fill_blocks : p.nsx * p.nsyb;
    if (zero_cfl) hipLaunchKernelGGL(zero_slot_kernel, dim3(1), dim3(1), 0, st, cfl);
    if (w == 3)
        hipLaunchKernelGGL((sharp_stage_kernel_w3<RP, R>), dim3(nblocks), dim3(WAVE * wpb), 0, st, p);
    else
        hipLaunchKernelGGL((sharp_stage_kernel<RP, R>), dim3(nblocks), dim3(WAVE * wpb), 0, st, p);
    return hipGetLastError() == hipSuccess ? CLAW_OK : CLAW_EDEVICE;
}

static int sharp_validate(const ClawGrid *grid, const ClawRiemann *rp, const ClawBC *bc, int precision,
                          const void *aux) {
    if (!grid || !rp || !bc) return CLAW_EINVAL;
    if (grid->mx < 3 || grid->my < 3 || grid->ng < SNG) return CLAW_EINVAL;
    if (grid->ldq < grid->mx + 2 * grid->ng || grid->plane < grid->ldq * (grid->my + 2 * grid->ng)) return CLAW_EINVAL;
    for (int s = 0; s < 2; ++s) {
        if (bc->lower[s] < 0 || bc->lower[s] > 3 || bc->upper[s] < 0 || bc->upper[s] > 3) return CLAW_EINVAL;
        if ((bc->lower[s] == CLAW_BC_PERIODIC) != (bc->upper[s] == CLAW_BC_PERIODIC)) return CLAW_EINVAL;
    }
    if (precision != CLAW_F64 && precision != CLAW_F32 && precision != CLAW_F64_FMA) return CLAW_EINVAL;
    if (rp->kind == CLAW_RP_VC_ACOUSTICS && !aux) return CLAW_EINVAL;
    if (rp->kind < 0 || rp->kind > 3) return CLAW_EINVAL;
    return CLAW_OK;
}

static int sharp_dispatch(Ctx *ctx, const ClawGrid *grid, const ClawRiemann *rp, const ClawBC *bc, int precision,
                          int mode, const void *qin, const void *q0, void *s2, void *out, const void *aux,
                          const double *dt, double *cfl, bool zero_cfl, hipStream_t st) {
    const bool f32 = precision == CLAW_F32;
#define SDISPATCH(RPT)                                                                                        \
    return f32 ? launch_sharp<RPT<float>, float>(ctx, grid, rp, bc, mode, qin, q0, s2, out, aux, dt, cfl, zero_cfl, st) \
               : launch_sharp<RPT<double>, double>(ctx, grid, rp, bc, mode, qin, q0, s2, out, aux, dt, cfl, zero_cfl, st)
    switch (rp->kind) {
    case CLAW_RP_ACOUSTICS: SDISPATCH(Acoustics);
    case CLAW_RP_VC_ACOUSTICS: SDISPATCH(VCAcoustics);
    case CLAW_RP_EULER: SDISPATCH(Euler5);
    default: SDISPATCH(ShallowRoe);
    }
#undef SDISPATCH
}

#ifndef CLAW_SHARP_MAIN
// FP64 launches of this build (arguments validated by the C entry points).
int sharp_dispatch_f64(Ctx *ctx, const ClawGrid *grid, const ClawRiemann *rp, const ClawBC *bc, int mode,
                       const void *qin, const void *q0, void *s2, void *out, const void *aux, const double *dt,
                       double *cfl, bool zero_cfl, hipStream_t st) {
    return sharp_dispatch(ctx, grid, rp, bc, CLAW_F64, mode, qin, q0, s2, out, aux, dt, cfl, zero_cfl, st);
}
#endif
}  // namespace CLAW_SHARP_VARIANT
}  // namespace clawhip

#ifdef CLAW_SHARP_MAIN
namespace clawhip {
namespace fused {
int sharp_dispatch_f64(Ctx *ctx, const ClawGrid *grid, const ClawRiemann *rp, const ClawBC *bc, int mode,
                       const void *qin, const void *q0, void *s2, void *out, const void *aux, const double *dt,
                       double *cfl, bool zero_cfl, hipStream_t st);
}  // namespace fused
}  // namespace clawhip

using namespace clawhip;
using namespace clawhip::exact;

// the build for the precision: CLAW_F64_FMA runs the fused kernels
static int sharp_route(Ctx *ctx, const ClawGrid *grid, const ClawRiemann *rp, const ClawBC *bc, int precision,
                       int mode, const void *qin, const void *q0, void *s2, void *out, const void *aux,
                       const double *dt, double *cfl, bool zero_cfl, hipStream_t st) {
    if (precision == CLAW_F64_FMA)
        return fused::sharp_dispatch_f64(ctx, grid, rp, bc, mode, qin, q0, s2, out, aux, dt, cfl, zero_cfl, st);
    return sharp_dispatch(ctx, grid, rp, bc, precision, mode, qin, q0, s2, out, aux, dt, cfl, zero_cfl, st);
}

extern "C" {

int claw2d_sharpclaw_dq(void *ctx, const ClawGrid *grid, const ClawRiemann *rp, const ClawBC *bc, int precision,
                        const void *q, void *dq, const void *aux, const double *dt_dev, double *cfl_dev,
                        void *stream) {
    int rc = sharp_validate(grid, rp, bc, precision, aux);
    if (rc) return rc;
    if (!q || !dq || q == dq || !dt_dev || !cfl_dev) return CLAW_EINVAL;
    return sharp_route((Ctx *)ctx, grid, rp, bc, precision, SM_DQ, q, nullptr, nullptr, dq, aux, dt_dev, cfl_dev, true,
                          (hipStream_t)stream);
}

int claw2d_sharpclaw_stage(void *ctx, const ClawGrid *grid, const ClawRiemann *rp, const ClawBC *bc,
                           int precision, int mode, const void *qin, const void *q0, void *s2, void *out,
                           const void *aux, const double *dt_dev, double *cfl_dev, void *stream) {
    int rc = sharp_validate(grid, rp, bc, precision, aux);
    if (rc) return rc;
    if (mode < SM_DQ || mode > SM_SSP33_3) return CLAW_EINVAL;
    if (!qin || !out || qin == out || !dt_dev || !cfl_dev) return CLAW_EINVAL;
    if ((mode == SM_SSP104_MID || mode == SM_SSP33_2 || mode == SM_SSP33_3) && !q0) return CLAW_EINVAL;
    if ((mode == SM_SSP104_MID || mode == SM_SSP104_END) && !s2) return CLAW_EINVAL;
    if (mode == SM_SSP104_MID && (s2 == qin || s2 == out)) return CLAW_EINVAL;
    return sharp_route((Ctx *)ctx, grid, rp, bc, precision, mode, qin, q0, s2, out, aux, dt_dev, cfl_dev, true,
                          (hipStream_t)stream);
}

int claw2d_sharpclaw_ssp104_step(void *ctx, const ClawGrid *grid, const ClawRiemann *rp, const ClawBC *bc,
                                 int precision, const void *q, void *qnew, void *s1a, void *s1b, void *s2,
                                 const void *aux, const double *dt_dev, double *cfl10_dev, void *stream) {
    int rc = sharp_validate(grid, rp, bc, precision, aux);
    if (rc) return rc;
    if (!q || !qnew || !s1a || !s1b || !s2 || !dt_dev || !cfl10_dev) return CLAW_EINVAL;
    const void *bufs[5] = {q, qnew, s1a, s1b, s2};
    for (int a = 0; a < 5; ++a)
        for (int b = a + 1; b < 5; ++b)
            if (bufs[a] == bufs[b]) return CLAW_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    // the ten stages' Courant slots are cleared by one launch
    hipLaunchKernelGGL(zero_slots_kernel, dim3(1), dim3(WAVE), 0, st, cfl10_dev, 10);
    // stage k reads `in` and writes `out`; s1 ping-pongs between s1a and s1b
    void *s1[2] = {s1a, s1b};
    int k = 0, cur = 0;
#define STAGE(MODE, IN, OUT)                                                                                   \
    do {                                                                                                       \
        rc = sharp_route((Ctx *)ctx, grid, rp, bc, precision, MODE, IN, q, s2, OUT, aux, dt_dev, cfl10_dev + k, false, st); \
        if (rc) return rc;                                                                                     \
        ++k;                                                                                                   \
    } while (0)
    STAGE(SM_SSP104_A, q, s1[cur]);
    for (int r = 0; r < 3; ++r) { STAGE(SM_SSP104_A, s1[cur], s1[cur ^ 1]); cur ^= 1; }
    STAGE(SM_SSP104_MID, s1[cur], s1[cur ^ 1]); cur ^= 1;
    for (int r = 0; r < 4; ++r) { STAGE(SM_SSP104_A, s1[cur], s1[cur ^ 1]); cur ^= 1; }
    STAGE(SM_SSP104_END, s1[cur], qnew);
#undef STAGE
    return CLAW_OK;
}

int claw2d_sharpclaw_set_launch(void *ctx, const ClawGrid *grid, const ClawRiemann *rp, int precision, int band,
                                int waves_per_simd) {
    return claw2d_sharpclaw_set_launch_ex(ctx, grid, rp, precision, band, waves_per_simd, 4);
}

int claw2d_sharpclaw_set_launch_ex(void *ctx, const ClawGrid *grid, const ClawRiemann *rp, int precision, int band,
                                   int waves_per_simd, int waves_per_block) {
    if (!ctx || !grid || !rp || band < 0 || (waves_per_simd != 1 && waves_per_simd != 3)) return CLAW_EINVAL;
    if (waves_per_block != 1 && waves_per_block != 2 && waves_per_block != 4) return CLAW_EINVAL;
    Ctx *c = (Ctx *)ctx;
    const TuneKey key = sharp_key(grid, rp, precision);
    c->band[key] = std::make_pair(band, waves_per_simd);
    c->wpb[key] = waves_per_block;
    return CLAW_OK;
}

int claw2d_sharpclaw_get_launch(void *ctx, const ClawGrid *grid, const ClawRiemann *rp, int precision, int *band,
                                int *waves_per_simd, int *waves_per_block) {
    if (!ctx || !grid || !rp || !band || !waves_per_simd || !waves_per_block) return CLAW_EINVAL;
    *band = sharp_band((Ctx *)ctx, grid, rp, precision, (grid->mx + SX_USEFUL - 1) / SX_USEFUL, *waves_per_simd,
                       *waves_per_block);
    return CLAW_OK;
}

int claw2d_sharpclaw_autotune(void *ctx, const ClawGrid *grid, const ClawRiemann *rp, const ClawBC *bc,
                              int precision, const void *q, void *dq, const void *aux, const double *dt_dev,
                              double *cfl_dev, void *stream, int *band_out) {
    if (!ctx) return CLAW_EINVAL;
    int rc = sharp_validate(grid, rp, bc, precision, aux);
    if (rc) return rc;
    if (!q || !dq || q == dq || !dt_dev || !cfl_dev) return CLAW_EINVAL;
    Ctx *c = (Ctx *)ctx;
    hipStream_t st = (hipStream_t)stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        if (e0) (void)hipEventDestroy(e0);
        return CLAW_EDEVICE;
    }
    // row-band heights (ending with the balanced heights for one and two
    // wavefronts per SIMD, for the default build, and for three, for the
    // three-per-SIMD build, then the fill decomposition, -1) x (default or
    // three-per-SIMD build) x (1 or 4 wavefronts per block)
    const int strips = (grid->mx + SX_USEFUL - 1) / SX_USEFUL;
    constexpr int NC = 13;
    const int cand[NC] = {4, 6, 8, 12, 16, 24, 32, 48, 64, balanced_band(c->device, grid->my, strips, 1),
                          balanced_band(c->device, grid->my, strips, 2),
                          balanced_band(c->device, grid->my, strips, 3), -1};
    int best = 0, best_w = 1, best_wpb = 4;
    bool have = false;
    float best_ms = 0.f;
    for (int k = 0; k < 4 * NC; ++k) {
        const int h = cand[k % NC], w = (k / NC) % 2 == 0 ? 1 : 3, wpb = k < 2 * NC ? 4 : 1;
        if (h > grid->my && h != cand[0]) continue;
        if ((k % NC) >= 9 && (k % NC) <= 10 && w != 1) continue;
        if ((k % NC) == 11 && w != 3) continue;
        c->force_band = h;
        c->force_w = w;
        c->force_wpb = wpb;
        rc = sharp_route(c, grid, rp, bc, precision, SM_DQ, q, nullptr, nullptr, dq, aux, dt_dev, cfl_dev, false, st);
        if (rc) break;
        const int reps = 5;
        (void)hipEventRecord(e0, st);
        for (int r = 0; r < reps && !rc; ++r)
            rc = sharp_route(c, grid, rp, bc, precision, SM_DQ, q, nullptr, nullptr, dq, aux, dt_dev, cfl_dev,
                                false, st);
        (void)hipEventRecord(e1, st);
        if (rc || hipEventSynchronize(e1) != hipSuccess) { rc = rc ? rc : CLAW_EDEVICE; break; }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (!have || ms < best_ms) { best = h > 0 ? h : 0; best_w = w; best_wpb = wpb; best_ms = ms; have = true; }
    }
    c->force_band = 0;
    c->force_w = 0;
    c->force_wpb = 0;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    c->band[sharp_key(grid, rp, precision)] = std::make_pair(best, best_w);
    c->wpb[sharp_key(grid, rp, precision)] = best_wpb;
    if (band_out) *band_out = best_w == 3 ? -best : best;
    return CLAW_OK;
}

}  // extern "C"
namespace clawhip {
namespace exact {   // defined with the classic loop (classic2d.hip)
int loop_gauges(const ClawLoopGauges *G, int precision, const ClawLoopState *cur, const ClawLoopState *next,
                const void *q, hipStream_t st);
}
}
extern "C" {

int claw2d_sharpclaw_loop(void *ctx, const ClawGrid *grid, const ClawRiemann *rp, const ClawBC *bc,
                          int precision, void *qa, void *qb, void *s1a, void *s1b, void *s2, const void *aux,
                          const ClawLoopRule *rule, ClawLoopState *states, double *cfl, long long nbytes, int k0,
                          int niter, const ClawLoopGauges *gauges, void *stream) {
    if (!rule || rule->ncfl != 10 || !states || !cfl || !qa || !qb || niter < 0 || k0 < 0) return CLAW_EINVAL;
    if (gauges && gauges->n > 0 && precision == CLAW_F32) return CLAW_EINVAL;
    void *q[2] = {qa, qb};
    for (int i = 0; i < niter; ++i) {
        const int a = (k0 + i) & 1;
        double *ca = cfl + 10 * a, *cb = cfl + 10 * (1 - a);
        int rc = claw2d_sharpclaw_ssp104_step(ctx, grid, rp, bc, precision, q[a], q[1 - a], s1a, s1b, s2, aux,
                                              &states[a].dt, ca, stream);
        if (rc == CLAW_OK)
            rc = claw2d_loop_advance(rule, &states[a], &states[1 - a], ca, cb, q[a], q[1 - a], nbytes, stream);
        if (rc == CLAW_OK) rc = loop_gauges(gauges, precision, &states[a], &states[1 - a], q[1 - a], (hipStream_t)stream);
        if (rc != CLAW_OK) return rc;
    }
    return CLAW_OK;
}

}  // extern "C"
#endif  // CLAW_SHARP_MAIN
