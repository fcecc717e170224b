// Library context shared by the kernel files: scratch buffers and the launch
// shapes picked by the autotuners, per problem shape.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <utility>

namespace clawhip {

// Launch shapes picked by the autotuners, per problem shape (SharpClaw keys
// use order = SHARP_KEY).
struct TuneKey {
    int kind, order, trans, precision, mx, my;
    bool operator<(const TuneKey &o) const {
        const int a[6] = {kind, order, trans, precision, mx, my}, b[6] = {o.kind, o.order, o.trans, o.precision, o.mx, o.my};
        return std::lexicographical_compare(a, a + 6, b, b + 6);
    }
};

struct Ctx {
    int device;
    void *scratch = nullptr;   // split intermediate state
    size_t scratch_bytes = 0;
    std::map<TuneKey, std::pair<int, int>> band;   // tuned (band height, waves per SIMD)
    std::map<TuneKey, int> wpb;                     // tuned wavefronts per block (classic step)
    int force_band = 0, force_w = 0, force_wpb = 0; // set while a tuner times a candidate
    bool wide_offsets = false;   // unsplit step: 64-bit per-lane offsets even for planes under 4 GiB
    unsigned long long *cfl_groups = nullptr;   // Courant group slots and counts (cfl_group_max), zeroed
};

constexpr int SHARP_KEY = 100;

// Band height that gives about `rounds` bands per wavefront slot of the device
// (CUs x 4 SIMDs x waves per SIMD): whole rounds of wavefronts, none left
// over for a partial one.
inline int balanced_band(int device, int rows, int strips_x, int waves_per_simd, int rounds = 1) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus <= 0)
        cus = 256;
    const long long slots = (long long)cus * 4 * waves_per_simd * rounds;
    const long long bands = std::max<long long>(1, slots / std::max(strips_x, 1));
    return (int)std::max<long long>(4, (rows + bands - 1) / bands);
}

}  // namespace clawhip
