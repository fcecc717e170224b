// Discrete adjoint of the unsplit classic step for variable-coefficient
// acoustics (q = (p, u, v), aux = (rho, c) per cell), gfx950.
//
// Forward map (the reference's step2, classic/step2.f90:92-248, with flux2,
// classic/flux2.f90:100-206, the limiter, classic/limiter.f90:31-61 and
// philim.f90:18-55, and the vc-acoustics normal and transverse solvers; dt is
// fixed): q_{n+1} = q_n + sum over interfaces of their contributions.  For an
// x-interface (I, J) between cells lo = (I-1, J) and hi = (I, J), with
// dtdn = dt/dx, gam = dt/dy * dt/(2 dx):
//     lo: -dtdn A-dq - dtdn f          hi: -dtdn A+dq + dtdn f      (f = cq/2)
//     column of lo (I-1) at rows J-1, J, J+1:  +gam bmA, gam (bpA - bmA), -gam bpA
//     column of hi (I)   likewise with bmP, bpP
// where A-dq = s1 W1, A+dq = s2 W2 (raw waves), cq = sum_p |s_p|(1 - |s_p| dtdn) W~_p
// (limited waves), (bmA, bpA) = rpt(A-dq [+cq]) with the aux of column I-1 and
// (bmP, bpP) = rpt(A+dq [-cq]) with the aux of column I.  y-interfaces are the
// same with x and y exchanged.  Ghost values are functions of interior values
// through the boundary codes (copies, negated normal velocity at walls).
//
// Reverse sweep, written by hand: given lam = dJ/dq_{n+1} (zero outside the
// interior, since ghost cells of q_{n+1} are discarded), each interface turns
// lam into the adjoints of its outputs, then goes back through the transverse
// solver (transposed, plus its derivative with respect to the three aux cells),
// the correction flux, the limiter (phi and its derivative with respect to the
// ratio theta = W.W_upwind / W.W, which also reaches the upwind neighbour's raw
// wave), the fluctuations and the normal solver (transposed, plus its aux
// derivative).  The results are gathered per cell instead of scattered:
// a thread owns one position, recomputes every interface whose inputs include
// that position and sums their contributions in a fixed order, so the sweep is
// deterministic and needs no atomics.  A block stages its tile of positions
// and the cells within 3 of it (q with the boundary map applied, aux, lam) in
// LDS first, so the recomputation reads LDS, not global memory.  Tiles cover
// the ghost frame too, and each ghost's contribution is added to the ring cell
// it was copied from: through LDS inside the tile when no side is periodic
// (the tiling puts every ghost in the tile of its source cell), otherwise
// through a workspace and a second, small launch.
//
// Requirements: c > 0 in every cell (so wave 1 moves left and is limited with
// the interface to its right, wave 2 with the interface to its left); boundary
// codes extrap, periodic or wall (stored ghost cells are treated as constants).
#include <hip/hip_runtime.h>

#include "common.hpp"

namespace clawhip {

namespace adj {

constexpr int ANG = 2;   // ghost width of the classic step

struct Params {
    const long long *rec_off = nullptr;   // receivers (distinct interior cells, component-0 offsets)
    const double *rec_tr = nullptr, *rec_obs = nullptr;   // their traces and observed values at this step
    int nrec = 0;
    const double *q;      // q_n, padded planar layout
    const double *aux;    // (rho, c), padded planar layout
    const double *lam;    // lam_{n+1}, interior cells read only
    double *lam_out;      // lam_n
    double *gaux;         // accumulated dJ/d(rho, c)
    double *frame;        // workspace: 5 planes, ghost positions only
    GridDev g;
    BCDev bq, ba;
    double dtdx, dtdy;
    int order, trans;
    int mth[2];
};

struct CellQ { double p, un, ut; };          // pressure, normal and transverse velocity
struct CellA { double rho, c; };

// (n, t) -> (i, j): DIR 0 normal = x, DIR 1 normal = y
template <int DIR> __host__ __device__ __forceinline__ int ci(int n, int t) { return DIR == 0 ? n : t; }
template <int DIR> __host__ __device__ __forceinline__ int cj(int n, int t) { return DIR == 0 ? t : n; }

// Cell data come through an accessor (the parameters plus loadq/loada/loadl):
// GlobalAcc reads the padded arrays through the boundary map, TileAcc reads a
// block's tile staged in LDS.  Both return the same values.
struct GlobalAcc : Params {
    __host__ __device__ __forceinline__ void loadq(int i, int j, double &p, double &u, double &v) const {
        SrcCell s = bc_source(i, j, g, bq);
        p = q[s.off];
        u = q[g.plane + s.off];
        v = q[2 * g.plane + s.off];
        if (s.fx) u = -u;
        if (s.fy) v = -v;
    }
    __host__ __device__ __forceinline__ void loada(int i, int j, double &rho, double &c) const {
        SrcCell s = bc_source(i, j, g, ba);
        rho = aux[s.off];
        c = aux[g.plane + s.off];
    }
    template <int DIR>
    __host__ __device__ __forceinline__ void face(int, int, const CellQ &ql, const CellQ &qh, const CellA &alo,
                                                  const CellA &ahi, double &a1, double &a2, double &iD) const;
    // 1/(z(n, t-1) + z(n, t)), (n, t) in the normal/transverse coordinates of DIR
    template <int DIR>
    __host__ __device__ __forceinline__ double tiD(int n, int t) const {
        double r0, c0, r1, c1;
        loada(DIR == 0 ? n : t - 1, DIR == 0 ? t - 1 : n, r0, c0);
        loada(DIR == 0 ? n : t, DIR == 0 ? t : n, r1, c1);
        return 1.0 / (r0 * c0 + r1 * c1);
    }
    // lam_{n+1}: zero outside the interior (ghost cells of q_{n+1} are discarded)
    __host__ __device__ __forceinline__ void loadl(int i, int j, double &p, double &u, double &v) const {
        if (i < 0 || j < 0 || i >= g.mx || j >= g.my) { p = u = v = 0.0; return; }
        const long long off = (long long)(j + g.ng) * g.ldq + (i + g.ng);
        p = lam[off];
        u = lam[g.plane + off];
        v = lam[2 * g.plane + off];
    }
};

// Tile of positions handled by one block, and the cells its interfaces read
// (3 around every position).
constexpr int TW = 64, TH = 4, HALO = 3, BOXW = TW + 2 * HALO, BOXH = TH + 2 * HALO, BOX = BOXW * BOXH;

struct TileAcc : Params {
    const double *s;   // 8 planes of BOX values: q (3, boundary map applied), aux (2), lam (3)
    const double *f;   // 6 planes of BOX values: (a1, a2, 1/D) of the x-face and of the y-face whose
                       // upper/right cell is the box cell (the box's first column/row has none)
    int bx, by;        // cell coordinates of the box's first element
    template <int DIR>
    __host__ __device__ __forceinline__ void face(int N, int T, const CellQ &, const CellQ &, const CellA &,
                                                  const CellA &, double &a1, double &a2, double &iD) const {
        const int k = at(DIR == 0 ? N : T, DIR == 0 ? T : N) + DIR * 3 * BOX;
        a1 = f[k]; a2 = f[BOX + k]; iD = f[2 * BOX + k];
    }
    // the transverse face's 1/D: a y-face for DIR 0, an x-face for DIR 1
    template <int DIR>
    __host__ __device__ __forceinline__ double tiD(int n, int t) const {
        return DIR == 0 ? f[5 * BOX + at(n, t)] : f[2 * BOX + at(t, n)];
    }
    __host__ __device__ __forceinline__ int at(int i, int j) const { return (j - by) * BOXW + (i - bx); }
    __host__ __device__ __forceinline__ void loadq(int i, int j, double &p, double &u, double &v) const {
        const int k = at(i, j);
        p = s[k]; u = s[BOX + k]; v = s[2 * BOX + k];
    }
    __host__ __device__ __forceinline__ void loada(int i, int j, double &rho, double &c) const {
        const int k = at(i, j);
        rho = s[3 * BOX + k]; c = s[4 * BOX + k];
    }
    __host__ __device__ __forceinline__ void loadl(int i, int j, double &p, double &u, double &v) const {
        const int k = at(i, j);
        p = s[5 * BOX + k]; u = s[6 * BOX + k]; v = s[7 * BOX + k];
    }
};

template <int DIR, class PT>
__host__ __device__ __forceinline__ CellQ getq(const PT &P, int n, int t) {
    double p, u, v;
    P.loadq(ci<DIR>(n, t), cj<DIR>(n, t), p, u, v);
    CellQ r;
    r.p = p;
    r.un = DIR == 0 ? u : v;
    r.ut = DIR == 0 ? v : u;
    return r;
}

template <int DIR, class PT>
__host__ __device__ __forceinline__ CellA geta(const PT &P, int n, int t) {
    CellA r;
    P.loada(ci<DIR>(n, t), cj<DIR>(n, t), r.rho, r.c);
    return r;
}

template <int DIR, class PT>
__host__ __device__ __forceinline__ CellQ getl(const PT &P, int n, int t) {
    double p, u, v;
    P.loadl(ci<DIR>(n, t), cj<DIR>(n, t), p, u, v);
    CellQ r;
    r.p = p;
    r.un = DIR == 0 ? u : v;
    r.ut = DIR == 0 ? v : u;
    return r;
}

// raw waves of interface N (between cells N-1 and N) in transverse line T
struct Raw {
    double w1p, w1u, w2p, w2u;   // W1 = (w1p, w1u, 0), W2 = (w2p, w2u, 0)
    double s1, s2;               // -c(lo), c(hi)
    double a1, a2, d1, iD;       // wave strengths, normal-velocity jump, 1/(zlo + zhi)
    double zlo, zhi;
    CellA alo, ahi;
};

// the normal solve's wave strengths (as the forward computes them) and 1/D
template <int DIR>
__host__ __device__ __forceinline__ void strengths(const CellQ &ql, const CellQ &qh, const CellA &alo,
                                                   const CellA &ahi, double &a1, double &a2, double &iD) {
    const double d0 = qh.p - ql.p;
    const double d1 = qh.un - ql.un;
    const double zhi = ahi.rho * ahi.c, zlo = alo.rho * alo.c;
    const double D = zlo + zhi;
    a1 = (-d0 + zhi * d1) / D;
    a2 = (d0 + zlo * d1) / D;
    iD = 1.0 / D;
}

template <int DIR>
__host__ __device__ __forceinline__ void GlobalAcc::face(int, int, const CellQ &ql, const CellQ &qh, const CellA &alo,
                                                         const CellA &ahi, double &a1, double &a2, double &iD) const {
    strengths<DIR>(ql, qh, alo, ahi, a1, a2, iD);
}

template <int DIR, class PT>
__host__ __device__ __forceinline__ Raw raw(const PT &P, int N, int T) {
    const CellQ ql = getq<DIR>(P, N - 1, T), qh = getq<DIR>(P, N, T);
    Raw r;
    r.alo = geta<DIR>(P, N - 1, T);
    r.ahi = geta<DIR>(P, N, T);
    r.zhi = r.ahi.rho * r.ahi.c;
    r.zlo = r.alo.rho * r.alo.c;
    P.template face<DIR>(N, T, ql, qh, r.alo, r.ahi, r.a1, r.a2, r.iD);
    r.d1 = qh.un - ql.un;
    r.w1p = -r.a1 * r.zlo; r.w1u = r.a1;
    r.w2p = r.a2 * r.zhi;  r.w2u = r.a2;
    r.s1 = -r.alo.c; r.s2 = r.ahi.c;
    return r;
}

// phi(theta) and dphi/dtheta with the forward's branch choices (philim.f90:18-55)
__host__ __device__ __forceinline__ void philim_d(double r, int meth, double &phi, double &dphi) {
    switch (meth) {
    case 2: {   // superbee: max(max(0, min(1, 2r)), min(2, r))
        const double A = 1.0 < 2.0 * r ? 1.0 : 2.0 * r, dA = 1.0 < 2.0 * r ? 0.0 : 2.0;
        const double B = 2.0 < r ? 2.0 : r, dB = 2.0 < r ? 0.0 : 1.0;
        const double t = 0.0 > A ? 0.0 : A, dt = 0.0 > A ? 0.0 : dA;
        phi = t > B ? t : B; dphi = t > B ? dt : dB;
        return;
    }
    case 3: {   // van Leer: (r + |r|) / (1 + |r|)
        const double ar = __builtin_fabs(r), sg = r > 0.0 ? 1.0 : (r < 0.0 ? -1.0 : 0.0);
        phi = (r + ar) / (1.0 + ar);
        dphi = ((1.0 + sg) * (1.0 + ar) - (r + ar) * sg) / ((1.0 + ar) * (1.0 + ar));
        return;
    }
    case 4: {   // MC: max(0, min(min((1+r)/2, 2), 2r))
        const double c = (1.0 + r) / 2.0;
        const double m1 = c < 2.0 ? c : 2.0, dm1 = c < 2.0 ? 0.5 : 0.0;
        const double m2 = m1 < 2.0 * r ? m1 : 2.0 * r, dm2 = m1 < 2.0 * r ? dm1 : 2.0;
        phi = 0.0 > m2 ? 0.0 : m2; dphi = 0.0 > m2 ? 0.0 : dm2;
        return;
    }
    case 5: phi = r; dphi = 1.0; return;
    default: {  // minmod (ids outside 1..5 as well): max(0, min(1, r))
        const double m = 1.0 < r ? 1.0 : r, dm = 1.0 < r ? 0.0 : 1.0;
        phi = 0.0 > m ? 0.0 : m; dphi = 0.0 > m ? 0.0 : dm;
        return;
    }
    }
}

// limiter state of one wave (wp, wu) with its upwind neighbour (np, nu).
// The derivative terms are kept in scaled form, W = s u with s = max|W_k|,
// so that 1/|W|^2 is never formed: tiny wave tails (|W|^2 subnormal) would
// otherwise make it overflow.  With e_s = (g . u) phi' / |u|^2:
//   dtheta/dW contribution  e_s (Wn/s - 2 theta u),   push to Wn  e_s u.
struct Lim { double phi, dphi, th; bool on, d; double up, uu, nps, nus, un2; };

__host__ __device__ __forceinline__ Lim limit(double wp, double wu, double np, double nu, int meth) {
    Lim L;
    const double wn2 = wp * wp + wu * wu;
    L.on = (meth != 0) && (wn2 != 0.0);
    L.d = false;
    L.phi = 1.0; L.dphi = 0.0; L.th = 0.0;
    L.up = L.uu = L.nps = L.nus = 0.0; L.un2 = 1.0;
    if (L.on) {
        const double dot = wp * np + wu * nu;
        L.th = dot / wn2;
        philim_d(L.th, meth, L.phi, L.dphi);
        L.d = L.dphi != 0.0;
        if (L.d) {
            const double aw = __builtin_fabs(wp), bw = __builtin_fabs(wu);
            const double is = 1.0 / (aw > bw ? aw : bw);
            L.up = wp * is; L.uu = wu * is;
            L.nps = np * is; L.nus = nu * is;
            L.un2 = L.up * L.up + L.uu * L.uu;
        }
    }
    return L;
}

// transposed vc-acoustics transverse solve.  Forward (asdq t = (t0, ., tt)):
//   a1 = (-t0 + tt zz)/(zm + zz),  a2 = (t0 + tt zz)/(zz + zp)
//   bm = (cm a1 zm, 0, -cm a1),    bp = (cp a2 zp, 0, cp a2)
// In: adjoints of bm, bp (p and transverse components).  Out: adjoint of t
// (p and transverse components) and of (rho, c) of the three cells.
struct RptRev { double gt0, gtt; double gm_rho, gm_c, g0_rho, g0_c, gp_rho, gp_c; };

__host__ __device__ __forceinline__ RptRev rpt_rev(double t0, double tt, CellA am, CellA a0, CellA ap, double iD1,
                                          double iD2, double gbm0, double gbmt, double gbp0, double gbpt) {
    // iD1 = 1/(zm + zz), iD2 = 1/(zz + zp): the neighbouring faces' values
    const double zm = am.rho * am.c, zz = a0.rho * a0.c, zp = ap.rho * ap.c;
    const double a1 = (-t0 + tt * zz) * iD1, a2 = (t0 + tt * zz) * iD2;
    const double cm = am.c, cp = ap.c;
    const double ga1 = cm * zm * gbm0 - cm * gbmt;
    const double gcm = a1 * zm * gbm0 - a1 * gbmt;
    double gzm = cm * a1 * gbm0;
    const double ga2 = cp * zp * gbp0 + cp * gbpt;
    const double gcp = a2 * zp * gbp0 + a2 * gbpt;
    double gzp = cp * a2 * gbp0;
    const double i1 = ga1 * iD1, i2 = ga2 * iD2;
    RptRev r;
    r.gt0 = -i1 + i2;
    r.gtt = zz * (i1 + i2);
    const double gzz = i1 * (tt - a1) + i2 * (tt - a2);
    gzm -= i1 * a1;
    gzp -= i2 * a2;
    r.gm_rho = gzm * am.c; r.gm_c = gcm + gzm * am.rho;
    r.g0_rho = gzz * a0.c; r.g0_c = gzz * a0.rho;
    r.gp_rho = gzp * ap.c; r.gp_c = gcp + gzp * ap.rho;
    return r;
}

// Everything one interface sends backwards, given its raw waves, its two
// limiter states and lam around it.
struct Own {
    double gw1p, gw1u, gw2p, gw2u;     // own part of the adjoint of the raw waves
    double push1p, push1u;             // to the raw wave 1 of interface N+1
    double push2p, push2u;             // to the raw wave 2 of interface N-1
    double gclo, gchi;                 // adjoints of c(lo), c(hi) through the speeds
    RptRev rA, rP;                     // transverse solves (column N-1, column N)
};

template <int DIR, class PT>
__host__ __device__ __forceinline__ Own own(const PT &P, int N, int T, const Raw &F, const Lim &L1, const Lim &L2) {
    const double dtdn = DIR == 0 ? P.dtdx : P.dtdy;
    const double dtdt = DIR == 0 ? P.dtdy : P.dtdx;
    const CellQ lc = getl<DIR>(P, N - 1, T), hc = getl<DIR>(P, N, T);
    // adjoints of A-dq, A+dq (p, normal, transverse) and of cq
    double gAm0 = -dtdn * lc.p, gAmu = -dtdn * lc.un, gAmt = -dtdn * lc.ut;
    double gAp0 = -dtdn * hc.p, gApu = -dtdn * hc.un, gApt = -dtdn * hc.ut;
    double gq0 = 0.0, gqu = 0.0;
    if (P.order == 2) { gq0 = 0.5 * dtdn * (hc.p - lc.p); gqu = 0.5 * dtdn * (hc.un - lc.un); }
    // limited waves and cq (needed by trans_cor's transverse inputs)
    const double k1 = -F.s1 * (1.0 - (-F.s1) * dtdn), k2 = F.s2 * (1.0 - F.s2 * dtdn);
    const double l1p = L1.phi * F.w1p, l1u = L1.phi * F.w1u, l2p = L2.phi * F.w2p, l2u = L2.phi * F.w2u;
    Own o;
    o.rA = RptRev{0, 0, 0, 0, 0, 0, 0, 0};
    o.rP = RptRev{0, 0, 0, 0, 0, 0, 0, 0};
    if (P.trans > 0) {
        const double gam = dtdt * (0.5 * dtdn);
        const CellQ lm = getl<DIR>(P, N - 1, T - 1), lp = getl<DIR>(P, N - 1, T + 1);
        const CellQ hm = getl<DIR>(P, N, T - 1), hp = getl<DIR>(P, N, T + 1);
        double tA0 = F.s1 * F.w1p, tP0 = F.s2 * F.w2p;
        if (P.order == 2 && P.trans == 2) {
            const double cq0 = k1 * l1p + k2 * l2p;
            tA0 = tA0 + cq0; tP0 = tP0 - cq0;
        }
        o.rA = rpt_rev(tA0, 0.0, geta<DIR>(P, N - 1, T - 1), F.alo, geta<DIR>(P, N - 1, T + 1),
                       P.template tiD<DIR>(N - 1, T), P.template tiD<DIR>(N - 1, T + 1),
                       gam * (lm.p - lc.p), gam * (lm.ut - lc.ut), gam * (lc.p - lp.p), gam * (lc.ut - lp.ut));
        o.rP = rpt_rev(tP0, 0.0, geta<DIR>(P, N, T - 1), F.ahi, geta<DIR>(P, N, T + 1),
                       P.template tiD<DIR>(N, T), P.template tiD<DIR>(N, T + 1),
                       gam * (hm.p - hc.p), gam * (hm.ut - hc.ut), gam * (hc.p - hp.p), gam * (hc.ut - hp.ut));
        gAm0 += o.rA.gt0; gAmt += o.rA.gtt;
        gAp0 += o.rP.gt0; gApt += o.rP.gtt;
        if (P.order == 2 && P.trans == 2) { gq0 += o.rA.gt0 - o.rP.gt0; }
    }
    (void)gAmt; (void)gApt;   // the waves have no transverse component
    // fluctuations: A-dq = s1 W1, A+dq = s2 W2
    o.gw1p = F.s1 * gAm0; o.gw1u = F.s1 * gAmu;
    o.gw2p = F.s2 * gAp0; o.gw2u = F.s2 * gApu;
    o.gclo = -(gAm0 * F.w1p + gAmu * F.w1u);
    o.gchi = gAp0 * F.w2p + gApu * F.w2u;
    o.push1p = o.push1u = o.push2p = o.push2u = 0.0;
    if (P.order == 2) {
        // cq = k1 W~1 + k2 W~2, k = |s| (1 - |s| dtdn)
        const double g1p = k1 * gq0, g1u = k1 * gqu, g2p = k2 * gq0, g2u = k2 * gqu;
        o.gclo += (1.0 - 2.0 * (-F.s1) * dtdn) * (gq0 * l1p + gqu * l1u);
        o.gchi += (1.0 - 2.0 * F.s2 * dtdn) * (gq0 * l2p + gqu * l2u);
        // W~ = phi(theta) W, theta = W.Wn / W.W
        o.gw1p += L1.phi * g1p; o.gw1u += L1.phi * g1u;
        o.gw2p += L2.phi * g2p; o.gw2u += L2.phi * g2u;
        if (L1.d) {
            const double e = (g1p * L1.up + g1u * L1.uu) * L1.dphi / L1.un2;
            o.gw1p += e * (L1.nps - 2.0 * L1.th * L1.up); o.gw1u += e * (L1.nus - 2.0 * L1.th * L1.uu);
            o.push1p = e * L1.up; o.push1u = e * L1.uu;
        }
        if (L2.d) {
            const double e = (g2p * L2.up + g2u * L2.uu) * L2.dphi / L2.un2;
            o.gw2p += e * (L2.nps - 2.0 * L2.th * L2.up); o.gw2u += e * (L2.nus - 2.0 * L2.th * L2.uu);
            o.push2p = e * L2.up; o.push2u = e * L2.uu;
        }
    }
    return o;
}

// transposed normal solve: adjoints of (d0, d1) and of zlo, zhi from the
// adjoint of the raw waves
struct RpnRev { double gd0, gd1, gzlo, gzhi; };

__host__ __device__ __forceinline__ RpnRev rpn_rev(const Raw &F, double g1p, double g1u, double g2p, double g2u) {
    const double ga1 = -F.zlo * g1p + g1u;
    const double ga2 = F.zhi * g2p + g2u;
    const double i1 = ga1 * F.iD, i2 = ga2 * F.iD;
    RpnRev r;
    r.gd0 = -i1 + i2;
    r.gd1 = F.zhi * i1 + F.zlo * i2;
    r.gzhi = F.a2 * g2p + i1 * (F.d1 - F.a1) - i2 * F.a2;
    r.gzlo = -F.a1 * g1p - i1 * F.a1 + i2 * (F.d1 - F.a2);
    return r;
}

// The limiter push of one wave of interface N to its upwind neighbour, as
// the scalar e_s with push = e_s * u, W = s u (wave 0: wave 1, else wave 2):
// e_s = (adjoint of the limited wave . u) phi'(theta) / u.u.
template <int DIR, class PT>
__host__ __device__ __forceinline__ double cq_adjoint(const PT &P, int N, int T, const Raw &F, const Lim &L1,
                                             const Lim &L2, int wave) {
    const double dtdn = DIR == 0 ? P.dtdx : P.dtdy;
    const CellQ lc = getl<DIR>(P, N - 1, T), hc = getl<DIR>(P, N, T);
    double gq0 = 0.5 * dtdn * (hc.p - lc.p);
    const double gqu = 0.5 * dtdn * (hc.un - lc.un);
    const double k1 = -F.s1 * (1.0 - (-F.s1) * dtdn), k2 = F.s2 * (1.0 - F.s2 * dtdn);
    if (P.trans == 2) {
        // trans_cor: the transverse solves take A-dq + cq and A+dq - cq
        const double dtdt = DIR == 0 ? P.dtdy : P.dtdx;
        const double gam = dtdt * (0.5 * dtdn);
        const CellQ lm = getl<DIR>(P, N - 1, T - 1), lp = getl<DIR>(P, N - 1, T + 1);
        const CellQ hm = getl<DIR>(P, N, T - 1), hp = getl<DIR>(P, N, T + 1);
        const double cq0 = k1 * (L1.phi * F.w1p) + k2 * (L2.phi * F.w2p);
        const RptRev rA = rpt_rev(F.s1 * F.w1p + cq0, 0.0, geta<DIR>(P, N - 1, T - 1), F.alo,
                                  geta<DIR>(P, N - 1, T + 1), P.template tiD<DIR>(N - 1, T),
                                  P.template tiD<DIR>(N - 1, T + 1), gam * (lm.p - lc.p), gam * (lm.ut - lc.ut),
                                  gam * (lc.p - lp.p), gam * (lc.ut - lp.ut));
        const RptRev rP = rpt_rev(F.s2 * F.w2p - cq0, 0.0, geta<DIR>(P, N, T - 1), F.ahi, geta<DIR>(P, N, T + 1),
                                  P.template tiD<DIR>(N, T), P.template tiD<DIR>(N, T + 1),
                                  gam * (hm.p - hc.p), gam * (hm.ut - hc.ut), gam * (hc.p - hp.p),
                                  gam * (hc.ut - hp.ut));
        gq0 += rA.gt0 - rP.gt0;
    }
    const Lim &L = wave == 0 ? L1 : L2;
    const double k = wave == 0 ? k1 : k2;
    return (k * gq0 * L.up + k * gqu * L.uu) * L.dphi / L.un2;
}

// Transverse-solve aux partial at position (gn, gt) from interfaces of the
// lines gt-1 and gt+1: rptP of interface (gn, T) and rptA of (gn+1, T) use
// the column gn, with (gn, gt) as the cell above (T = gt-1) or below (T = gt+1).
template <int DIR, class PT>
__host__ __device__ __forceinline__ void rpt_side(const PT &P, int gn, int gt, double &grho, double &gc) {
    const double dtdn = DIR == 0 ? P.dtdx : P.dtdy;
    const double dtdt = DIR == 0 ? P.dtdy : P.dtdx;
    const double gam = dtdt * (0.5 * dtdn);
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        const int T = side == 0 ? gt - 1 : gt + 1;
        const CellQ lm = getl<DIR>(P, gn, T - 1), lc = getl<DIR>(P, gn, T), lp = getl<DIR>(P, gn, T + 1);
        const double gbm0 = gam * (lm.p - lc.p), gbmt = gam * (lm.ut - lc.ut);
        const double gbp0 = gam * (lc.p - lp.p), gbpt = gam * (lc.ut - lp.ut);
        const CellA am = geta<DIR>(P, gn, T - 1), a0 = geta<DIR>(P, gn, T), ap = geta<DIR>(P, gn, T + 1);
        const double iD1 = P.template tiD<DIR>(gn, T), iD2 = P.template tiD<DIR>(gn, T + 1);
        // interface (gn, T): A+dq enters column gn; interface (gn+1, T): A-dq enters column gn
#pragma unroll
        for (int which = 0; which < 2; ++which) {
            const int N = which == 0 ? gn : gn + 1;
            const Raw F = raw<DIR>(P, N, T);
            double t0 = which == 0 ? F.s2 * F.w2p : F.s1 * F.w1p;
            if (P.order == 2 && P.trans == 2) {
                const Raw Fm = raw<DIR>(P, N - 1, T), Fp = raw<DIR>(P, N + 1, T);
                const Lim L1 = limit(F.w1p, F.w1u, Fp.w1p, Fp.w1u, P.mth[0]);
                const Lim L2 = limit(F.w2p, F.w2u, Fm.w2p, Fm.w2u, P.mth[1]);
                const double k1 = -F.s1 * (1.0 - (-F.s1) * dtdn), k2 = F.s2 * (1.0 - F.s2 * dtdn);
                const double cq0 = k1 * (L1.phi * F.w1p) + k2 * (L2.phi * F.w2p);
                t0 = which == 0 ? t0 - cq0 : t0 + cq0;
            }
            const RptRev r = rpt_rev(t0, 0.0, am, a0, ap, iD1, iD2, gbm0, gbmt, gbp0, gbpt);
            if (side == 0) { grho += r.gp_rho; gc += r.gp_c; }
            else { grho += r.gm_rho; gc += r.gm_c; }
        }
    }
}

// Raw adjoint contributions at position (i, j) (interior or ghost) from every
// interface normal to DIR whose inputs include that position.
template <int DIR, class PT>
__host__ __device__ __forceinline__ void position_dir(const PT &P, int i, int j, double (&gq)[3], double &grho,
                                             double &gc) {
    const int gn = DIR == 0 ? i : j, gt = DIR == 0 ? j : i;
    // interfaces gn-1 .. gn+2 along the line gt
    const Raw R0 = raw<DIR>(P, gn - 1, gt), R1 = raw<DIR>(P, gn, gt);
    const Raw R2 = raw<DIR>(P, gn + 1, gt), R3 = raw<DIR>(P, gn + 2, gt);
    // wave 1 is limited with the interface to its right, wave 2 with the one to its left
    const Lim L1_1 = limit(R1.w1p, R1.w1u, R2.w1p, R2.w1u, P.mth[0]);
    const Lim L2_1 = limit(R1.w2p, R1.w2u, R0.w2p, R0.w2u, P.mth[1]);
    const Lim L1_2 = limit(R2.w1p, R2.w1u, R3.w1p, R3.w1u, P.mth[0]);
    const Lim L2_2 = limit(R2.w2p, R2.w2u, R1.w2p, R1.w2u, P.mth[1]);
    const Own O1 = own<DIR>(P, gn, gt, R1, L1_1, L2_1);
    const Own O2 = own<DIR>(P, gn + 1, gt, R2, L1_2, L2_2);
    double pin1p = 0.0, pin1u = 0.0, pin2p = 0.0, pin2u = 0.0;
    if (P.order == 2) {
        // wave 1 of interface gn-1 limited with gn; wave 2 of gn+2 with gn+1
        const Lim L1_0 = limit(R0.w1p, R0.w1u, R1.w1p, R1.w1u, P.mth[0]);
        const Lim L2_3 = limit(R3.w2p, R3.w2u, R2.w2p, R2.w2u, P.mth[1]);
        // the other wave of these two interfaces matters only through trans_cor's cq
        Lim L2_0 = limit(0.0, 0.0, 0.0, 0.0, 0), L1_3 = L2_0;
        if (P.trans == 2) {
            const Raw Rm = raw<DIR>(P, gn - 2, gt), Rp = raw<DIR>(P, gn + 3, gt);
            L2_0 = limit(R0.w2p, R0.w2u, Rm.w2p, Rm.w2u, P.mth[1]);
            L1_3 = limit(R3.w1p, R3.w1u, Rp.w1p, Rp.w1u, P.mth[0]);
        }
        if (L1_0.d) {
            const double g = cq_adjoint<DIR>(P, gn - 1, gt, R0, L1_0, L2_0, 0);
            pin1p = g * L1_0.up; pin1u = g * L1_0.uu;
        }
        if (L2_3.d) {
            const double g = cq_adjoint<DIR>(P, gn + 2, gt, R3, L1_3, L2_3, 1);
            pin2p = g * L2_3.up; pin2u = g * L2_3.uu;
        }
    }
    // interface gn: the position is its upper/right cell
    {
        const RpnRev r = rpn_rev(R1, O1.gw1p + pin1p, O1.gw1u + pin1u, O1.gw2p + O2.push2p, O1.gw2u + O2.push2u);
        gq[0] += r.gd0;
        gq[DIR == 0 ? 1 : 2] += r.gd1;
        grho += r.gzhi * R1.ahi.c;
        gc += r.gzhi * R1.ahi.rho + O1.gchi + O1.rP.g0_c;
        grho += O1.rP.g0_rho;
    }
    // interface gn+1: the position is its lower/left cell
    {
        const RpnRev r = rpn_rev(R2, O2.gw1p + O1.push1p, O2.gw1u + O1.push1u, O2.gw2p + pin2p, O2.gw2u + pin2u);
        gq[0] -= r.gd0;
        gq[DIR == 0 ? 1 : 2] -= r.gd1;
        grho += r.gzlo * R2.alo.c;
        gc += r.gzlo * R2.alo.rho + O2.gclo + O2.rA.g0_c;
        grho += O2.rA.g0_rho;
    }
    if (P.trans > 0) rpt_side<DIR>(P, gn, gt, grho, gc);
}

template <class PT>
__host__ __device__ __forceinline__ void position_raw(const PT &P, int i, int j, double (&gq)[3], double &grho,
                                             double &gc) {
    gq[0] = gq[1] = gq[2] = 0.0;
    grho = 0.0; gc = 0.0;
    position_dir<0>(P, i, j, gq, grho, gc);
    position_dir<1>(P, i, j, gq, grho, gc);
}

// Writes the raw contributions at position (i, j) of the padded region
// [-ANG, mx+ANG) x [-ANG, my+ANG): interior cells get lam_n = lam_{n+1} + raw
// and their aux gradient; ghost positions go to the frame workspace, folded
// back into their source cells by the fold launch (fold_field).
template <class PT>
__host__ __device__ __forceinline__ void position_write(const PT &A, int i, int j) {
    double gq[3], grho, gc;
    position_raw(A, i, j, gq, grho, gc);
    const long long off = (long long)(j + A.g.ng) * A.g.ldq + (i + A.g.ng);
    if (i >= 0 && j >= 0 && i < A.g.mx && j < A.g.my) {
        for (int m = 0; m < 3; ++m) A.lam_out[m * A.g.plane + off] = A.lam[m * A.g.plane + off] + gq[m];
        A.gaux[off] += grho;
        A.gaux[A.g.plane + off] += gc;
    } else {
        A.frame[off] = gq[0];
        A.frame[A.g.plane + off] = gq[1];
        A.frame[2 * A.g.plane + off] = gq[2];
        A.frame[3 * A.g.plane + off] = grho;
        A.frame[4 * A.g.plane + off] = gc;
    }
}

// cells within ANG of the boundary: the only ones ghost values are copied from
__host__ __device__ __forceinline__ long long ring_count(const Params &P) {
    const int mx = P.g.mx, my = P.g.my;
    const int rb = my < ANG ? my : ANG, rt = (my - rb) < ANG ? (my - rb) : ANG;
    const int cl = mx < ANG ? mx : ANG, cr = (mx - cl) < ANG ? (mx - cl) : ANG;
    return (long long)(rb + rt) * mx + (long long)(my - rb - rt) * (cl + cr);
}

__host__ __device__ __forceinline__ void ring_cell(const Params &P, long long k, int &i, int &j) {
    const int mx = P.g.mx, my = P.g.my;
    const int rb = my < ANG ? my : ANG, rt = (my - rb) < ANG ? (my - rb) : ANG;
    const int cl = mx < ANG ? mx : ANG, cr = (mx - cl) < ANG ? (mx - cl) : ANG;
    if (k < (long long)rb * mx) {
        j = (int)(k / mx); i = (int)(k % mx);
    } else if (k < (long long)(rb + rt) * mx) {
        const long long kk = k - (long long)rb * mx;
        j = my - rt + (int)(kk / mx); i = (int)(kk % mx);
    } else {
        const long long kk = k - (long long)(rb + rt) * mx;
        const int c = (int)(kk % (cl + cr));
        j = rb + (int)(kk / (cl + cr));
        i = c < cl ? c : mx - cr + (c - cl);
    }
}

// Ghost positions that the boundary map copies from ring cell (i, j), for
// field m: the candidates are (gi[a], gj[b]) with gi in {i, -2, -1, mx, mx+1}
// and gj in {j, -2, -1, my, my+1}, (a, b) != (0, 0); the map is separable (x
// index from the x codes, y index from the y codes; bc_source) and a wall
// negates the normal momentum it reflects.  A candidate contributes when
// hx[a] && hy[b], negated when nx[a] != ny[b].
struct FoldMap {
    int gi[5], gj[5];
    bool hx[5], hy[5], nx[5], ny[5];
    __host__ __device__ __forceinline__ FoldMap(const Params &P, int i, int j, int m) {
        const int mx = P.g.mx, my = P.g.my;
        const BCDev &bc = m < 3 ? P.bq : P.ba;
        gi[0] = i; gi[1] = -2; gi[2] = -1; gi[3] = mx; gi[4] = mx + 1;
        gj[0] = j; gj[1] = -2; gj[2] = -1; gj[3] = my; gj[4] = my + 1;
#pragma unroll
        for (int a = 0; a < 5; ++a) {
            bool f = false;
            const int cx = a <= 2 ? bc.lo[0] : bc.hi[0], cy = a <= 2 ? bc.lo[1] : bc.hi[1];
            hx[a] = a == 0 || bc_map1(gi[a], mx, cx, f) == i;
            nx[a] = a != 0 && m == 1 && cx == CLAW_BC_WALL;
            hy[a] = a == 0 || bc_map1(gj[a], my, cy, f) == j;
            ny[a] = a != 0 && m == 2 && cy == CLAW_BC_WALL;
        }
    }
};

// Sum, in a fixed order (b outer, a inner), of field m's frame contributions
// over the ghost positions that the boundary map copies from ring cell
// (i, j).  Every candidate is a ghost position inside the frame, so all 24
// loads are issued unconditionally and the match selects them: no load waits
// on another.
__host__ __device__ __forceinline__ double fold_field(const Params &P, int i, int j, int m) {
    const FoldMap F(P, i, j, m);
    const double *src = P.frame + (long long)m * P.g.plane;
    double acc = 0.0;
#pragma unroll
    for (int b = 0; b < 5; ++b) {
#pragma unroll
        for (int a = 0; a < 5; ++a) {
            if (a == 0 && b == 0) continue;
            const double v = src[(long long)(F.gj[b] + P.g.ng) * P.g.ldq + (F.gi[a] + P.g.ng)];
            const bool neg = F.nx[a] != F.ny[b];
            acc += (F.hx[a] && F.hy[b]) ? (neg ? -v : v) : 0.0;
        }
    }
    return acc;
}

// Tiles over the positions [-ANG, n+ANG) of one axis: width w from -ANG,
// except that the last tile starts at n-ANG when it would otherwise start to
// the right of it, so that it holds the high ghost positions together with
// the ring cells they are copied from (the first tile holds the low ones).
// Tile t owns the positions from its origin up to the next tile's origin.
__host__ __device__ __forceinline__ int tile_origin(int t, int nt, int n, int w) {
    const int o = w * t - ANG;
    return (t == nt - 1 && o > n - ANG) ? n - ANG : o;
}

__host__ __device__ __forceinline__ int tile_end(int t, int nt, int n, int w) {
    return t == nt - 1 ? n + ANG : tile_origin(t + 1, nt, n, w);
}

// number of tiles along x and y
__host__ __device__ __forceinline__ int tiles_x(const Params &P) { return (P.g.mx + 2 * ANG + TW - 1) / TW; }
__host__ __device__ __forceinline__ int tiles_y(const Params &P) { return (P.g.my + 2 * ANG + TH - 1) / TH; }

// Whether every ghost position lies in the same tile as the ring cells it is
// copied from: no periodic side (a periodic ghost copies from the far side)
// and, when an axis has more than one tile, a last tile that starts at or
// beyond ANG (the first tile then owns the low ring cells [0, ANG)).
__host__ __device__ __forceinline__ bool fold_in_tile(const Params &P, int ntx, int nty) {
    for (int s = 0; s < 2; ++s)
        if (P.bq.lo[s] == CLAW_BC_PERIODIC || P.bq.hi[s] == CLAW_BC_PERIODIC || P.ba.lo[s] == CLAW_BC_PERIODIC ||
            P.ba.hi[s] == CLAW_BC_PERIODIC)
            return false;
    if (ntx > 1 && tile_origin(ntx - 1, ntx, P.g.mx, TW) < ANG) return false;
    if (nty > 1 && tile_origin(nty - 1, nty, P.g.my, TH) < ANG) return false;
    return true;
}

// Field m of the ghost contributions folded into ring cell (i, j) from xg, the
// tile's ghost positions (5 planes of TW x TH, origin (tx0, ty0)), in the order
// of fold_field.
__host__ __device__ __forceinline__ double fold_tile(const Params &P, const double *xg, int tx0, int ty0, int i, int j,
                                                     int m) {
    const FoldMap F(P, i, j, m);
    double acc = 0.0;
#pragma unroll
    for (int b = 0; b < 5; ++b) {
#pragma unroll
        for (int a = 0; a < 5; ++a) {
            if (a == 0 && b == 0) continue;
            if (F.hx[a] && F.hy[b]) {
                const double v = xg[m * TW * TH + (F.gj[b] - ty0) * TW + (F.gi[a] - tx0)];
                acc += F.nx[a] != F.ny[b] ? -v : v;
            }
        }
    }
    return acc;
}

// Receiver r whose cell the tile owns goes to slot[(rj - ty0) * TW + (ri - tx0)]
// (receivers are distinct cells: no slot is written twice).
__host__ __device__ __forceinline__ void receiver_slots(const Params &P, int tx0, int ty0, int txe, int tye, int *slot,
                                                        int t0, int nt) {
    for (int r = t0; r < P.nrec; r += nt) {
        const long long ro = P.rec_off[r];
        const int rj = (int)(ro / P.g.ldq) - P.g.ng, ri = (int)(ro % P.g.ldq) - P.g.ng;
        if (ri >= tx0 && ri < txe && rj >= ty0 && rj < tye) slot[(rj - ty0) * TW + (ri - tx0)] = r;
    }
}

// Interior cell (i, j) with its raw contributions: lam_n = lam_{n+1} + raw
// (lam holds the cell's lam_{n+1}), the aux gradient (ga0, ga1: its values
// before this step) accumulates, a ring cell then adds its ghosts'
// contributions and a receiver cell its residual (trace minus observed) -- the
// same operations in the same order as the position writes followed by the
// fold launch.
__host__ __device__ __forceinline__ void position_finish(const Params &P, const double *xg, int tx0, int ty0, int i,
                                                         int j, const double (&lam)[3], double ga0, double ga1,
                                                         const double (&gq)[3], double grho, double gc, int rec) {
    const long long off = (long long)(j + P.g.ng) * P.g.ldq + (i + P.g.ng);
    double v[5];
    for (int m = 0; m < 3; ++m) v[m] = lam[m] + gq[m];
    v[3] = ga0 + grho;
    v[4] = ga1 + gc;
    if (i < ANG || j < ANG || i >= P.g.mx - ANG || j >= P.g.my - ANG)
        for (int m = 0; m < 5; ++m) v[m] = v[m] + fold_tile(P, xg, tx0, ty0, i, j, m);
    if (rec >= 0) v[0] += P.rec_tr[rec] - P.rec_obs[rec];
    for (int m = 0; m < 3; ++m) P.lam_out[m * P.g.plane + off] = v[m];
    P.gaux[off] = v[3];
    P.gaux[P.g.plane + off] = v[4];
}

// Wave strengths and 1/D of every face inside the staged box (faces need both
// cells in the box, so the box's first column has no x-face, its first row no
// y-face), computed once per tile from the staged cells.
template <int NT>
__host__ __device__ __forceinline__ void tile_faces(const TileAcc &A, int t0) {
    // thread t0 of NT handles box cells t0, t0 + NT, ...; all its faces are
    // computed before any is stored
    constexpr int NPER = (BOX + NT - 1) / NT;
    double *f = const_cast<double *>(A.f);
    double v[NPER][6];
#pragma unroll
    for (int r = 0; r < NPER; ++r) {
        const int k = t0 + r * NT;
        if (k >= BOX) break;
        const int bi = k % BOXW, bj = k / BOXW;
        const int i = A.bx + bi, j = A.by + bj;
        if (bi > 0)
            strengths<0>(getq<0>(A, i - 1, j), getq<0>(A, i, j), geta<0>(A, i - 1, j), geta<0>(A, i, j), v[r][0],
                         v[r][1], v[r][2]);
        if (bj > 0)
            strengths<1>(getq<1>(A, j - 1, i), getq<1>(A, j, i), geta<1>(A, j - 1, i), geta<1>(A, j, i), v[r][3],
                         v[r][4], v[r][5]);
    }
#pragma unroll
    for (int r = 0; r < NPER; ++r) {
        const int k = t0 + r * NT;
        if (k >= BOX) break;
        if (k % BOXW > 0) { f[k] = v[r][0]; f[BOX + k] = v[r][1]; f[2 * BOX + k] = v[r][2]; }
        if (k / BOXW > 0) { f[3 * BOX + k] = v[r][3]; f[4 * BOX + k] = v[r][4]; f[5 * BOX + k] = v[r][5]; }
    }
}

// Stages q (boundary map applied), aux and lam for the box whose first cell is
// (bx, by): thread t0 of NT handles box cells t0, t0 + NT, ...; all its loads
// are issued before any LDS store, so their latencies overlap.
template <int NT>
__host__ __device__ __forceinline__ void stage_box(const Params &P, double *tile, int bx, int by, int t0) {
    constexpr int NPER = (BOX + NT - 1) / NT;
    double v[NPER][8];
#pragma unroll
    for (int r = 0; r < NPER; ++r) {
        const int k = t0 + r * NT;
        if (k >= BOX) break;
        const int i = bx + k % BOXW, j = by + k / BOXW;
        const SrcCell sq = bc_source(i, j, P.g, P.bq);
        const double u = P.q[P.g.plane + sq.off], w = P.q[2 * P.g.plane + sq.off];
        v[r][0] = P.q[sq.off];
        v[r][1] = sq.fx ? -u : u;
        v[r][2] = sq.fy ? -w : w;
        const SrcCell sa = bc_source(i, j, P.g, P.ba);
        v[r][3] = P.aux[sa.off];
        v[r][4] = P.aux[P.g.plane + sa.off];
        const bool in = i >= 0 && j >= 0 && i < P.g.mx && j < P.g.my;
        const long long off = in ? (long long)(j + P.g.ng) * P.g.ldq + (i + P.g.ng) : 0;
        v[r][5] = in ? P.lam[off] : 0.0;
        v[r][6] = in ? P.lam[P.g.plane + off] : 0.0;
        v[r][7] = in ? P.lam[2 * P.g.plane + off] : 0.0;
    }
#pragma unroll
    for (int r = 0; r < NPER; ++r) {
        const int k = t0 + r * NT;
        if (k >= BOX) break;
#pragma unroll
        for (int m = 0; m < 8; ++m) tile[m * BOX + k] = v[r][m];
    }
}

// One block per TW x TH tile of the padded region (tile_origin), a 1D grid
// of tiles_x * tiles_y blocks taken in row-major tile order after the XCD
// remap, so each XCD's L2 holds a band of tile rows and the halo rows that
// neighbouring tiles share: stage q (boundary map applied), aux and lam for
// the tile plus HALO cells around it in LDS, then one thread per owned
// position.  FOLD (fold_in_tile holds): the
// ghost positions' contributions go through LDS to the ring cells of the same
// tile and the receivers' residuals are added here, so no fold launch follows;
// otherwise ghost positions write the frame workspace for the fold launch.
static_assert(5 * TW * TH <= 5 * BOX, "the ghosts' planes must not reach the staged lam planes");

template <bool FOLD>
__global__ __launch_bounds__(256) void adjoint_tile_kernel(const Params P) {
    __shared__ double tile[8 * BOX];
    __shared__ double faces[6 * BOX];
    __shared__ int rslot[FOLD ? TW * TH : 1];
    TileAcc A;
    static_cast<Params &>(A) = P;
    const int ntx = tiles_x(P), nty = tiles_y(P);
    const int vb = xcd_block(), bxi = vb % ntx, byi = vb / ntx;
    const int tx0 = tile_origin(bxi, ntx, P.g.mx, TW), ty0 = tile_origin(byi, nty, P.g.my, TH);
    const int txe = tile_end(bxi, ntx, P.g.mx, TW), tye = tile_end(byi, nty, P.g.my, TH);
    A.s = tile;
    A.f = faces;
    A.bx = tx0 - HALO;
    A.by = ty0 - HALO;
    if (FOLD) rslot[threadIdx.x] = -1;
    stage_box<256>(P, tile, A.bx, A.by, threadIdx.x);
    __syncthreads();
    if (FOLD) receiver_slots(P, tx0, ty0, txe, tye, rslot, threadIdx.x, 256);
    tile_faces<256>(A, threadIdx.x);
    __syncthreads();
    const int i = tx0 + (threadIdx.x & (TW - 1)), j = ty0 + (threadIdx.x / TW);
    const bool own = i < txe && j < tye;
    if (!FOLD) {
        if (own) position_write(A, i, j);
        return;
    }
    const bool in = i >= 0 && j >= 0 && i < P.g.mx && j < P.g.my;
    // the aux gradient's old values load while the position is computed
    double ga0 = 0.0, ga1 = 0.0;
    if (own && in) {
        const long long off = (long long)(j + P.g.ng) * P.g.ldq + (i + P.g.ng);
        ga0 = P.gaux[off];
        ga1 = P.gaux[P.g.plane + off];
    }
    double gq[3] = {0.0, 0.0, 0.0}, grho = 0.0, gc = 0.0;
    if (own) position_raw(A, i, j, gq, grho, gc);
    // Only tiles in the first or last row or column hold ghost positions and
    // ring cells (the first tile of an axis starts at -ANG, the last at or
    // before n-ANG); the test is uniform over the block.
    double *xg = tile;   // planes 0-1 of the staged box: q, read for the last time by position_raw
    if (bxi == 0 || bxi == ntx - 1 || byi == 0 || byi == nty - 1) {
        __syncthreads();
        if (own && !in) {
            xg[threadIdx.x] = gq[0];
            xg[TW * TH + threadIdx.x] = gq[1];
            xg[2 * TW * TH + threadIdx.x] = gq[2];
            xg[3 * TW * TH + threadIdx.x] = grho;
            xg[4 * TW * TH + threadIdx.x] = gc;
        }
        __syncthreads();
    }
    if (own && in) {
        const int k = A.at(i, j);
        const double lam[3] = {tile[5 * BOX + k], tile[6 * BOX + k], tile[7 * BOX + k]};
        position_finish(P, xg, tx0, ty0, i, j, lam, ga0, ga1, gq, grho, gc, rslot[threadIdx.x]);
    }
}

// Thread t < 5 * nring folds field m = t / nring of ring cell t % nring (lam
// for m < 3, the aux gradient for m = 3, 4); the misfit's residuals (trace
// minus observed) then enter lam at the receiver cells: a ring cell's pressure
// thread adds its own after its fold, every other receiver has a thread of its
// own (t >= 5 * nring).  Receivers are distinct cells, so no element is written
// by two threads.  The ring threads find their receiver with a wave-wide scan
// of the receiver offsets (64 per load, broadcast lane by lane).
__global__ __launch_bounds__(128) void adjoint_fold_kernel(const Params P) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long nring = ring_count(P);
    const int lane = threadIdx.x & (WAVE - 1);
    const long long wave_t0 = t - lane;
    const bool fold = t < 5 * nring;
    const int m = fold ? (int)(t / nring) : 0;
    int i = 0, j = 0;
    if (fold) ring_cell(P, t - m * nring, i, j);
    const long long off = (long long)(j + P.g.ng) * P.g.ldq + (i + P.g.ng);
    int hit = -1;
    if (P.nrec > 0 && wave_t0 < nring) {   // wave-uniform: some lane folds a pressure
        for (int r0 = 0; r0 < P.nrec; r0 += WAVE) {
            const int r = r0 + lane;
            const long long ro = r < P.nrec ? P.rec_off[r] : -1;
            const int lo = (int)(unsigned)(ro & 0xffffffffLL), hi = (int)(ro >> 32);
            const int cnt = P.nrec - r0 < WAVE ? P.nrec - r0 : WAVE;
            for (int s = 0; s < cnt; ++s) {
                const long long o = ((long long)__builtin_amdgcn_readlane(hi, s) << 32) |
                                    (long long)(unsigned)__builtin_amdgcn_readlane(lo, s);
                if (o == off) hit = r0 + s;
            }
        }
    }
    if (fold) {
        const double acc = fold_field(P, i, j, m);
        double *dst = m < 3 ? P.lam_out + (long long)m * P.g.plane + off : P.gaux + (long long)(m - 3) * P.g.plane + off;
        double v = *dst + acc;
        if (m == 0 && hit >= 0) v += P.rec_tr[hit] - P.rec_obs[hit];
        *dst = v;
        return;
    }
    const long long r = t - 5 * nring;
    if (r >= P.nrec) return;
    const long long ro = P.rec_off[r];
    const int rj = (int)(ro / P.g.ldq) - P.g.ng, ri = (int)(ro % P.g.ldq) - P.g.ng;
    if (ri < ANG || rj < ANG || ri >= P.g.mx - ANG || rj >= P.g.my - ANG) return;   // a ring cell: done above
    P.lam_out[ro] += P.rec_tr[r] - P.rec_obs[r];
}

}  // namespace adj

// Receiver traces and their adjoint sources: out[k] = a[off[k]]; a[off[k]] += v[k] - w[k].
__global__ void gather_kernel(const double *a, const long long *off, int n, double *out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[k] = a[off[k]];
}

__global__ void misfit_source_kernel(double *a, const long long *off, int n, const double *v, const double *w) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) a[off[k]] += v[k] - w[k];
}

}  // namespace clawhip

using namespace clawhip;

extern "C" long long claw2d_vc_adjoint_workspace_bytes(const ClawGrid *grid);

static int vc_adjoint_launch(const ClawGrid *grid, const ClawMethod *meth, const ClawBC *bc, const double *q,
                             const double *aux, const double *lam_in, double *lam_out, double *gaux, double dt,
                             void *workspace, long long workspace_bytes, const long long *rec_off, int nrec,
                             const double *rec_tr, const double *rec_obs, void *stream) {
    if (!grid || !meth || !bc || !q || !aux || !lam_in || !lam_out || !gaux || !workspace) return CLAW_EINVAL;
    if (nrec < 0 || (nrec > 0 && (!rec_off || !rec_tr || !rec_obs))) return CLAW_EINVAL;
    if (lam_in == lam_out) return CLAW_EINVAL;
    if (grid->mx < 2 || grid->my < 2 || grid->ng < adj::ANG) return CLAW_EINVAL;
    if (grid->ldq < grid->mx + 2 * grid->ng || grid->plane < grid->ldq * (grid->my + 2 * grid->ng)) return CLAW_EINVAL;
    if (workspace_bytes < claw2d_vc_adjoint_workspace_bytes(grid)) return CLAW_EINVAL;
    if (meth->split || meth->fwave || (meth->order != 1 && meth->order != 2)) return CLAW_EINVAL;
    if (meth->transverse < 0 || meth->transverse > 2 || meth->num_waves != 2) return CLAW_EINVAL;
    for (int s = 0; s < 2; ++s) {
        const int codes[4] = {bc->lower[s], bc->upper[s], bc->aux_lower[s], bc->aux_upper[s]};
        for (int c : codes)
            if (c < CLAW_BC_EXTRAP || c > CLAW_BC_WALL) return CLAW_EINVAL;   // stored ghosts are not differentiated
        if ((bc->lower[s] == CLAW_BC_PERIODIC) != (bc->upper[s] == CLAW_BC_PERIODIC)) return CLAW_EINVAL;
        if ((bc->aux_lower[s] == CLAW_BC_PERIODIC) != (bc->aux_upper[s] == CLAW_BC_PERIODIC)) return CLAW_EINVAL;
    }
    adj::Params P;
    P.q = q; P.aux = aux; P.lam = lam_in; P.lam_out = lam_out; P.gaux = gaux; P.frame = (double *)workspace;
    P.g = GridDev{grid->mx, grid->my, grid->ng, grid->ldq, grid->plane};
    P.bq = BCDev{{bc->lower[0], bc->lower[1]}, {bc->upper[0], bc->upper[1]}};
    P.ba = BCDev{{bc->aux_lower[0], bc->aux_lower[1]}, {bc->aux_upper[0], bc->aux_upper[1]}};
    P.dtdx = dt / grid->dx; P.dtdy = dt / grid->dy;
    P.order = meth->order; P.trans = meth->transverse;
    P.mth[0] = meth->mthlim[0]; P.mth[1] = meth->mthlim[1];
    P.rec_off = rec_off; P.rec_tr = rec_tr; P.rec_obs = rec_obs; P.nrec = nrec;
    hipStream_t st = (hipStream_t)stream;
    const int ntx = adj::tiles_x(P), nty = adj::tiles_y(P);
    const dim3 tiles((unsigned)(ntx * nty));
    if (adj::fold_in_tile(P, ntx, nty)) {
        hipLaunchKernelGGL(adj::adjoint_tile_kernel<true>, tiles, dim3(256), 0, st, P);
    } else {
        hipLaunchKernelGGL(adj::adjoint_tile_kernel<false>, tiles, dim3(256), 0, st, P);
        const long long nfold = 5 * adj::ring_count(P) + nrec;
        hipLaunchKernelGGL(adj::adjoint_fold_kernel, dim3((unsigned)((nfold + 127) / 128)), dim3(128), 0, st, P);
    }
    return hipGetLastError() == hipSuccess ? CLAW_OK : CLAW_EDEVICE;
}

extern "C" {

long long claw2d_vc_adjoint_workspace_bytes(const ClawGrid *grid) {
    if (!grid) return -1;
    return 5LL * grid->plane * (long long)sizeof(double);
}


int claw2d_vc_adjoint_step(const ClawGrid *grid, const ClawMethod *meth, const ClawBC *bc, const double *q,
                           const double *aux, const double *lam_in, double *lam_out, double *gaux, double dt,
                           void *workspace, long long workspace_bytes, void *stream) {
    return vc_adjoint_launch(grid, meth, bc, q, aux, lam_in, lam_out, gaux, dt, workspace, workspace_bytes, nullptr, 0,
                             nullptr, nullptr, stream);
}

int claw2d_vc_adjoint_step_misfit(const ClawGrid *grid, const ClawMethod *meth, const ClawBC *bc, const double *q,
                                  const double *aux, const double *lam_in, double *lam_out, double *gaux, double dt,
                                  void *workspace, long long workspace_bytes, const long long *offsets, int nrec,
                                  const double *traces, const double *observed, void *stream) {
    return vc_adjoint_launch(grid, meth, bc, q, aux, lam_in, lam_out, gaux, dt, workspace, workspace_bytes, offsets,
                             nrec, traces, observed, stream);
}

int claw2d_gather(const double *a, const long long *offsets, int n, double *out, void *stream) {
    if (!a || !offsets || !out || n < 0) return CLAW_EINVAL;
    if (n == 0) return CLAW_OK;
    hipLaunchKernelGGL(gather_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, a, offsets, n, out);
    return hipGetLastError() == hipSuccess ? CLAW_OK : CLAW_EDEVICE;
}

int claw2d_misfit_source(double *a, const long long *offsets, int n, const double *traces, const double *observed,
                         void *stream) {
    if (!a || !offsets || !traces || !observed || n < 0) return CLAW_EINVAL;
    if (n == 0) return CLAW_OK;
    hipLaunchKernelGGL(misfit_source_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, a, offsets,
                       n, traces, observed);
    return hipGetLastError() == hipSuccess ? CLAW_OK : CLAW_EDEVICE;
}

}  // extern "C"
