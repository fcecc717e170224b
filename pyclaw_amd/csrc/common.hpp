// Shared device helpers: lane exchange, boundary synthesis, the wave limiter.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/clawhip.h"

namespace clawhip {

constexpr int WAVE = 64;

// ---- lane exchange (ds_bpermute; addresses are byte offsets lane*4) -------
__device__ __forceinline__ float lane_get(float v, int addr) {
    return __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(v)));
}
__device__ __forceinline__ double lane_get(double v, int addr) {
    long long b = __double_as_longlong(v);
    int lo = __builtin_amdgcn_ds_bpermute(addr, (int)(b & 0xffffffffll));
    int hi = __builtin_amdgcn_ds_bpermute(addr, (int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// ---- neighbour exchange by whole-wavefront DPP rotates (VALU, no LDS) -----
// lane_up: lane L receives lane L-1's value (lane 0 receives lane 63's);
// lane_dn: lane L receives lane L+1's value (lane 63 receives lane 0's).
// The same values as lane_get with addresses ((lane -/+ 1) & 63) * 4.
// A whole-wavefront rotate gives every lane a source, so no lane keeps the
// destination's old value: the move takes none (no zeroing of the destination).
constexpr int DPP_WAVE_ROL1 = 0x134, DPP_WAVE_ROR1 = 0x13C;
template <int CTRL> __device__ __forceinline__ int dpp_move(int v) {
    return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, false);
}
template <int CTRL> __device__ __forceinline__ float dpp_move(float v) {
    return __int_as_float(dpp_move<CTRL>(__float_as_int(v)));
}
template <int CTRL> __device__ __forceinline__ double dpp_move(double v) {
    long long b = __double_as_longlong(v);
    int lo = dpp_move<CTRL>((int)(b & 0xffffffffll));
    int hi = dpp_move<CTRL>((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
// A value that is the same in every lane, moved to scalar registers.  FP64
// arithmetic on uniform values runs in the vector unit, and the compiler then
// keeps the results in vector registers; a loop-invariant uniform constant
// costs two VGPRs per double unless it is moved.  Call with every lane of the
// wavefront active (the first active lane's value is taken).
__device__ __forceinline__ double uniform(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffll));
    const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ float uniform(float v) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

template <typename T> __device__ __forceinline__ T lane_up(T v) { return dpp_move<DPP_WAVE_ROR1>(v); }
template <typename T> __device__ __forceinline__ T lane_dn(T v) { return dpp_move<DPP_WAVE_ROL1>(v); }

// Virtual block index of a 1D grid: consecutive workgroups are dispatched
// round-robin over the 8 XCDs; remap so each XCD (own L2) gets a contiguous
// range of blocks.
__device__ __forceinline__ int xcd_block() {
    const int nb = gridDim.x, b = blockIdx.x;
    const int per = nb / 8, rem = nb % 8, xcd = b % 8, loc = b / 8;
    return xcd < rem ? xcd * (per + 1) + loc : rem * (per + 1) + (xcd - rem) * per + loc;
}

// The paired fill decomposition of a marching kernel's (strip, row) space:
// one-wavefront blocks, one per wavefront slot of a build that holds two
// wavefronts per SIMD.  Blocks are dispatched to the XCDs in turn (block b
// to XCD b % 8) and each XCD fills one slot of every SIMD before the
// second, so blocks loc and loc + half of an XCD (loc = b / 8, half =
// gridDim.x / 16) share a SIMD, and the first of them is the older
// wavefront there, which the SIMD's issue arbitration favours (oldest
// first): with equal shares the older one finishes after about two thirds
// of the younger one's time, and the younger one then runs alone.  So a
// SIMD's work is split about A : 1.  With P = gridDim.x / 2 pairs and S
// strips of my rows, strip s holds U_s = floor(P (A+1) (s+1) / S) -
// floor(P (A+1) s / S) units, n_s = floor(P (s+1) / S) - floor(P s / S)
// older pieces (A units each) and m_s = U_s - A n_s younger ones (one unit
// each).  The younger pieces are y = floor(C my / (16 U_s)) rows each (C:
// sixteenths of a unit; rounded down to even with EVEN), all alike; the
// older ones share the rest of the strip, their boundaries rounded down (to
// even with EVEN), and take the remainder.  OLDER_FIRST puts the older
// pieces at the start of the strip, else the younger ones.  Each kind is
// kept together: neighbouring pieces of one kind go to blocks dispatched one
// after the other, which share an XCD's L2 for the rows between them
// (alternating the kinds ran 7% slower).
// EVEN puts every band start on an even row, as with bands of an even
// height.  Every piece lies in one strip, so a wavefront marches one band;
// the pieces tile every strip exactly whatever the placement, and only the
// speed depends on the dispatch order.  Usable where fill_usable holds (so
// no piece is empty).
template <bool EVEN, bool OLDER_FIRST>
__device__ __forceinline__ void fill_piece(const int S, const int my, const int A, const int C, int &sx, int &R0,
                                           int &R1) {
    const int b = blockIdx.x, half = gridDim.x >> 4, loc = b >> 3;
    const bool young = loc >= half;
    // pair index of the SIMD; the younger wavefront counts the pairs from the
    // other end, so the two wavefronts of a SIMD work in mirrored parts of the
    // grid and a costly region is shared by the SIMDs instead of loading some
    // of them twice (Euler shock-bubble after 300 steps: +2.2%)
    const int idx0 = (b & 7) * half + (young ? loc - half : loc);
    const int idx = young ? (gridDim.x >> 1) - 1 - idx0 : idx0;
    const int P = gridDim.x >> 1, UT = P * (A + 1);
    int s;
    if (!young) {
        s = ((idx + 1) * S - 1) / P;           // floor(P s / S) <= idx < floor(P (s+1) / S)
    } else {
        // younger pieces before strip s: floor(UT s / S) - A floor(P s / S),
        // non-decreasing in s, within a strip of s * P / S
        s = min(idx * S / P, S - 1);
        while (s > 0 && UT * s / S - A * (P * s / S) > idx) --s;
        while (s < S - 1 && UT * (s + 1) / S - A * (P * (s + 1) / S) <= idx) ++s;
    }
    constexpr int MASK = EVEN ? ~1 : ~0;
    const int O0 = P * s / S, U0 = UT * s / S, Us = UT * (s + 1) / S - U0;
    const int no = P * (s + 1) / S - O0, ny = Us - A * no;
    const int y = (C * my / (16 * Us)) & MASK;
    sx = s;
    if (OLDER_FIRST) {
        // older pieces in [0, Lo), younger ones after (the last piece takes an
        // odd last row)
        const int Lo = ny > 0 ? (my - ny * y) & MASK : my;
        if (young) {
            const int l = idx - (U0 - A * O0);
            R0 = Lo + l * y;
            R1 = l + 1 == ny ? my : R0 + y;
        } else {
            const int l = idx - O0;
            R0 = (l * Lo / no) & MASK;
            R1 = l + 1 == no ? Lo : ((l + 1) * Lo / no) & MASK;
        }
    } else {
        const int Ly = ny * y, Lo = my - Ly;
        if (young) {
            R0 = (idx - (U0 - A * O0)) * y;
            R1 = R0 + y;
        } else {
            const int l = idx - O0;
            R0 = Ly + ((l * Lo / no) & MASK);
            R1 = l + 1 == no ? my : Ly + (((l + 1) * Lo / no) & MASK);
        }
    }
}

// Whether the paired fill decomposition can tile `rows` rows of `strips`
// strips with `blocks` one-wavefront blocks, share A and younger-piece
// scale C (sixteenths of a unit): every strip gets at least one older piece,
// no strip has fewer units than its older pieces take, younger pieces are at
// least two rows and leave the older ones at least two rows each, and the
// kernel's integer arithmetic stays within 32 bits.
inline bool fill_usable(int blocks, int strips, int rows, int A, int C) {
    if (blocks < 16 || blocks % 16 != 0 || strips < 1 || rows < 1 || A < 1 || A > 255 || C < 1 || C > 64)
        return false;
    const long long P = blocks / 2, S = strips, UT = P * (A + 1);
    if (UT * S >= (1LL << 31) || P * (S + 1) >= (1LL << 31) || 64LL * rows >= (1LL << 31)) return false;
    for (long long s = 0; s < S; ++s) {
        const long long no = P * (s + 1) / S - P * s / S, us = UT * (s + 1) / S - UT * s / S, ny = us - A * no;
        if (no < 1 || ny < 0 || (long long)rows * no >= (1LL << 31)) return false;
        const long long y = (C * rows / (16 * us)) & ~1LL;
        if (y < 2 || rows - ny * y < 2 * no + 1) return false;
    }
    return true;
}

// The fill parameter of a kernel launch: share A and younger-piece scale C.
inline int fill_param(int A, int C) { return A | (C << 8); }

inline int device_cus(int device) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus <= 0) cus = 256;
    return cus;
}

// ---- grid description on the device ---------------------------------------
struct GridDev {
    int mx, my, ng;
    long long ldq, plane;
};

struct BCDev {
    int lo[2], hi[2];
};

__host__ __device__ __forceinline__ int iclamp(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Map a 1-D ghost index to its source (solver.py:386-434 for one side).
// Returns the source index in [0, n-1]; sets flip for wall.
// (Written as selects: with uniform arguments it is straight-line scalar code.)
__host__ __device__ __forceinline__ int bc_map1(int k, int n, int code, bool &flip) {
    const bool lo = k < 0;
    const int per = lo ? k + n : k - n;
    const int wal = lo ? -1 - k : 2 * n - 1 - k;
    const int ext = lo ? 0 : n - 1;
    const int s = code == CLAW_BC_PERIODIC ? per : (code == CLAW_BC_WALL ? wal : ext);
    flip = (code == CLAW_BC_WALL);
    return iclamp(s, 0, n - 1);
}

// Element offset of the value that the reference's ghost fill (x sides first,
// then y sides over the full padded width) leaves at cell (i, j), plus which
// momentum components a wall reflection negates.  Out-of-range lanes are
// clamped into the array (their results are never used).
struct SrcCell { long long off; bool fx, fy; };

__host__ __device__ __forceinline__ SrcCell bc_source(int i, int j, const GridDev &g, const BCDev &bc) {
    SrcCell r;
    r.fx = false; r.fy = false;
    int ii = i, jj = j;
    bool stored_all = false;
    if (j < 0 || j >= g.my) {
        int code = j < 0 ? bc.lo[1] : bc.hi[1];
        if (code == CLAW_BC_STORED) {
            stored_all = true;
            jj = iclamp(j, -g.ng, g.my + g.ng - 1);
            ii = iclamp(i, -g.ng, g.mx + g.ng - 1);
        } else {
            jj = bc_map1(j, g.my, code, r.fy);
        }
    }
    if (!stored_all && (i < 0 || i >= g.mx)) {
        int code = i < 0 ? bc.lo[0] : bc.hi[0];
        if (code == CLAW_BC_STORED) ii = iclamp(i, -g.ng, g.mx + g.ng - 1);
        else ii = bc_map1(i, g.mx, code, r.fx);
    }
    r.off = (long long)(jj + g.ng) * g.ldq + (ii + g.ng);
    return r;
}

// The same source cell as bc_source, split into a column part (per lane; a
// lane that keeps its column computes it once) and a row part (the same for
// every lane of a row: scalar work).  A y ghost row with stored values takes
// the clamped column (x rule skipped); any other row takes the column's x
// rule.  The clamped column is kept as an offset from the x rule's column
// (dst), so choosing between the two is an add of a selected value, which
// compiles to registers (a choice between two fields of one record can be
// compiled as a load from a chosen address, which puts the record in
// scratch memory).
struct ColSrc { int ii, dst; bool fx; };
struct RowSrc { int jj; bool fy, st; };

__host__ __device__ __forceinline__ ColSrc col_source(int i, const GridDev &g, const BCDev &bc) {
    ColSrc c;
    c.fx = false;
    c.ii = i;
    const int ii_st = iclamp(i, -g.ng, g.mx + g.ng - 1);
    if (i < 0 || i >= g.mx) {
        const int code = i < 0 ? bc.lo[0] : bc.hi[0];
        if (code == CLAW_BC_STORED) c.ii = ii_st;
        else c.ii = bc_map1(i, g.mx, code, c.fx);
    }
    c.dst = ii_st - c.ii;
    return c;
}

__host__ __device__ __forceinline__ RowSrc row_source(int j, const GridDev &g, const BCDev &bc) {
    RowSrc r;
    const bool out = j < 0 || j >= g.my;
    const int code = j < 0 ? bc.lo[1] : bc.hi[1];
    bool wall;
    const int jm = bc_map1(j, g.my, code, wall);
    r.st = out && code == CLAW_BC_STORED;
    r.jj = !out ? j : (r.st ? iclamp(j, -g.ng, g.my + g.ng - 1) : jm);
    r.fy = out && !r.st && wall;
    return r;
}

// Element e of the plane that starts at base.  O32 (every plane smaller than
// 4 GiB): the byte offset is formed in 32 bits, so the access is a uniform
// 64-bit base (the plane) plus one 32-bit per-lane offset, with no per-lane
// 64-bit address.
template <bool O32, typename R>
__device__ __forceinline__ const R &plane_at(const R *base, long long e) {
    if constexpr (O32)
        return *reinterpret_cast<const R *>(reinterpret_cast<const char *>(base) +
                                            (unsigned)e * (unsigned)sizeof(R));
    else
        return base[e];
}
template <bool O32, typename R>
__device__ __forceinline__ R &plane_at(R *base, long long e) {
    if constexpr (O32)
        return *reinterpret_cast<R *>(reinterpret_cast<char *>(base) + (unsigned)e * (unsigned)sizeof(R));
    else
        return base[e];
}

template <int NE, typename R, bool O32 = false>
__device__ __forceinline__ void load_q(const R *__restrict__ q, const GridDev &g, const BCDev &bc,
                                       int i, int j, R (&out)[NE]) {
    SrcCell s = bc_source(i, j, g, bc);
#pragma unroll
    for (int m = 0; m < NE; ++m) {
        R v = plane_at<O32>(q + m * g.plane, s.off);
        if ((m == 1 && s.fx) || (m == 2 && s.fy)) v = -v;
        out[m] = v;
    }
}

// load_q / load_aux for a cell given as column and row parts (see col_source)
template <int NE, typename R, bool O32 = false>
__device__ __forceinline__ void load_q_cr(const R *__restrict__ q, const GridDev &g, const ColSrc &c,
                                          const RowSrc &r, R (&out)[NE]) {
    const long long off = (long long)(r.jj + g.ng) * g.ldq + (c.ii + (r.st ? c.dst : 0) + g.ng);
    const bool fx = !r.st && c.fx;
#pragma unroll
    for (int m = 0; m < NE; ++m) {
        R v = plane_at<O32>(q + m * g.plane, off);
        if ((m == 1 && fx) || (m == 2 && r.fy)) v = -v;
        out[m] = v;
    }
}

template <int NA, typename R, bool O32 = false>
__device__ __forceinline__ void load_aux_cr(const R *__restrict__ a, const GridDev &g, const ColSrc &c,
                                            const RowSrc &r, R (&out)[NA]) {
    const long long off = (long long)(r.jj + g.ng) * g.ldq + (c.ii + (r.st ? c.dst : 0) + g.ng);
#pragma unroll
    for (int m = 0; m < NA; ++m) out[m] = plane_at<O32>(a + m * g.plane, off);
}

template <int NA, typename R, bool O32 = false>
__device__ __forceinline__ void load_aux(const R *__restrict__ a, const GridDev &g, const BCDev &bc,
                                         int i, int j, R (&out)[NA]) {
    SrcCell s = bc_source(i, j, g, bc);
#pragma unroll
    for (int m = 0; m < NA; ++m) out[m] = plane_at<O32>(a + m * g.plane, s.off);
}

__device__ __forceinline__ double rabs(double x) { return __builtin_fabs(x); }
__device__ __forceinline__ float rabs(float x) { return __builtin_fabsf(x); }

// IEEE maximum/minimum (one v_max/v_min instruction; a NaN operand yields the
// other).  Where a select a > b ? a : b differs from it only in the sign of a
// zero result, and that zero is only ever added to a sum that starts at +0
// (so the sign changes nothing), these are exact replacements.
__device__ __forceinline__ double fmax_ieee(double a, double b) { return __builtin_fmax(a, b); }
__device__ __forceinline__ float fmax_ieee(float a, float b) { return __builtin_fmaxf(a, b); }
__device__ __forceinline__ double fmin_ieee(double a, double b) { return __builtin_fmin(a, b); }
__device__ __forceinline__ float fmin_ieee(float a, float b) { return __builtin_fminf(a, b); }

// ---- wave limiter (classic/philim.f90:4-57) ---------------------------------
template <typename R> __device__ __forceinline__ R rmax2(R a, R b) { return a > b ? a : b; }
template <typename R> __device__ __forceinline__ R rmin2(R a, R b) { return a < b ? a : b; }
// The limiter's max/min: selects (the reference's results for every input,
// NaN included) by default; the fused-multiply-add build, which does not
// promise bitwise equality, uses IEEE max/min (one instruction each; the same
// value for finite ratios up to the sign of a zero phi).
#ifdef CLAW_SHARED_RECIPROCALS
template <typename R> __device__ __forceinline__ R lmax(R a, R b) { return fmax_ieee(a, b); }
template <typename R> __device__ __forceinline__ R lmin(R a, R b) { return fmin_ieee(a, b); }
#else
template <typename R> __device__ __forceinline__ R lmax(R a, R b) { return rmax2<R>(a, b); }
template <typename R> __device__ __forceinline__ R lmin(R a, R b) { return rmin2<R>(a, b); }
#endif

// The limiter's ratio r = b/a (a = |W|^2 > 0).  The fused-multiply-add build
// multiplies b by the hardware reciprocal of a refined by two Newton steps
// (within about one ulp of b/a) instead of running the IEEE division
// sequence; a below 2^-1000 is raised to 2^-1000 first, where the
// reciprocal would overflow.  That changes r only for a wave under 2^-500 in
// size, and only towards zero with its sign kept, so phi changes by at most
// 2 and the limited wave by at most 2^-499.
#ifdef CLAW_SHARED_RECIPROCALS
template <typename R> __device__ __forceinline__ R lim_ratio(R b, R a) { return b / a; }
template <> __device__ __forceinline__ double lim_ratio<double>(double b, double a) {
    const double x = __builtin_fmax(a, 0x1p-1000);
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    return b * r;
}
#else
template <typename R> __device__ __forceinline__ R lim_ratio(R b, R a) { return b / a; }
#endif

template <typename R>
__device__ __forceinline__ R philim(R a, R b, int meth) {
    R r = lim_ratio<R>(b, a);
    switch (meth) {
    case 2: return lmax<R>(lmax<R>(R(0), lmin<R>(R(1), R(2) * r)), lmin<R>(R(2), r));
    case 3: { R ar = rabs(r); return (r + ar) / (R(1) + ar); }
    case 4: { R c = (R(1) + r) / R(2); return lmax<R>(R(0), lmin<R>(lmin<R>(c, R(2)), R(2) * r)); }
    case 5: return r;
    default: return lmax<R>(R(0), lmin<R>(R(1), r));
    }
}


// Deterministic non-negative max: the bit pattern of a non-negative IEEE value
// orders like the value, so an integer atomic max is exact and order-free.
__device__ __forceinline__ void atomic_max_nonneg(double *p, double v) {
    if (!(v > 0.0)) return;
    atomicMax(reinterpret_cast<unsigned long long *>(p), (unsigned long long)__double_as_longlong(v));
}

// The same, issued by every lane of a wavefront that holds the same v (a
// wave_max result) without a lane-dependent branch.  Code that runs inside a
// work-queue loop must not guard its per-wave atomics with `lane == 0`: two
// such guards in a row (this one at the end of an item, the queue fetch at the
// start of the next) let the compiler thread the branch so that lanes other
// than 0 skip the fetch and loop on a stale item that only lane 0 could end.
__device__ __forceinline__ void atomic_max_nonneg_all_lanes(double *p, double v) {
    const unsigned long long bits = v > 0.0 ? (unsigned long long)__double_as_longlong(v) : 0ull;
    atomicMax(reinterpret_cast<unsigned long long *>(p), bits);
}

// The Courant number of a whole launch, combined through CFL_GROUPS group
// slots.  Device-scope atomics on one address are performed one after the
// other (about 13 ns each on this part): one per wavefront on the output
// address held up the end of a 927-wavefront launch by 12 us.  Here the
// wavefronts of block b form group b % CFL_GROUPS; each max-combines its
// value into its group's slot, and a wavefront whose value raised the slot
// also max-combines it into *out (a group's maximum raises its slot, or an
// equal value did and was passed on, so *out gets the launch's maximum;
// later wavefronts rarely raise a slot).  Each then counts itself in, and
// the group's last wavefront leaves slot and count at zero for the next
// launch.  The count is issued after the max has returned (its increment is
// computed from the max's result), so every max of a group is performed
// before the last count and so before the reset, without a release fence
// (which would write back the L2).  grp: 2 * CFL_GROUPS * CFL_GROUP_STRIDE
// zeroed words of a context; launches that use the same words must not
// overlap.  v is the wavefront's value (the same in every lane); every
// wavefront of the launch must call this exactly once.
constexpr int CFL_GROUPS = 32, CFL_GROUP_STRIDE = 32;   // 256 bytes between words
constexpr size_t CFL_GROUP_BYTES = 2 * CFL_GROUPS * CFL_GROUP_STRIDE * sizeof(unsigned long long);
__device__ __forceinline__ void cfl_group_max(double *out, unsigned long long *grp, double v) {
    if ((threadIdx.x & (WAVE - 1)) != 0) return;
    const unsigned nb = gridDim.x, g = blockIdx.x % CFL_GROUPS;
    const unsigned long long n =
        (unsigned long long)(blockDim.x / WAVE) * (nb / CFL_GROUPS + (g < nb % CFL_GROUPS ? 1u : 0u));
    unsigned long long *slot = grp + g * CFL_GROUP_STRIDE, *cnt = grp + (CFL_GROUPS + g) * CFL_GROUP_STRIDE;
    const unsigned long long bits = v > 0.0 ? (unsigned long long)__double_as_longlong(v) : 0ull;
    const unsigned long long old = __hip_atomic_fetch_max(slot, bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (bits > old) atomicMax(reinterpret_cast<unsigned long long *>(out), bits);
    // slots hold bit patterns of non-negative values (< 2^63): the increment is 1
    const unsigned long long t = __hip_atomic_fetch_add(cnt, 1ull + (old >> 63), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (t + 1 == n) {
        (void)__hip_atomic_exchange(slot, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        (void)__hip_atomic_exchange(cnt, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        double o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    return v;
}

}  // namespace clawhip
