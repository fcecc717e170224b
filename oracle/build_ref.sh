#!/bin/bash
# Compiles the reference's own 2D Fortran kernels, where they lie under
# /root/reference, into oracle/_ref/ (test infrastructure only).  oracle/_ref/ is
# git-ignored; like the other built libraries it is part of the tree copied for
# GPU runs, where nothing loads it: the tests that use it are CPU tests.
#   libref_classic.so   : classic/{step2,step2ds,flux2,limiter,philim}.f90 +
#                         development/rp_approaches/{rpn2,rpt2}_euler_5wave.f
#   libref_sharpclaw.so : sharpclaw/{ClawParams,weno,reconstruct,tfluct,evec,
#                         workspace2,flux2,flux1_2}.f90 (the reference's sharpclaw2
#                         source list) + ref_sharp_setup.f90 (module-state harness)
# The file lists follow src/pyclaw/classic/setup.py and src/pyclaw/sharpclaw/setup.py.
# -ffp-contract=off and no -march: plain x86-64 arithmetic, no fused multiply-add.
set -euo pipefail
REF=${REF:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
FC=${FC:-/opt/rocm/llvm/bin/amdflang}
if [ ! -d "$REF/src/pyclaw/classic" ]; then
    echo "reference sources not found at $REF; skipping oracle/_ref build" >&2
    exit 0
fi
mkdir -p "$OUT"
FFLAGS="-O2 -fPIC -ffp-contract=off"
C="$REF/src/pyclaw/classic"
D="$REF/development/rp_approaches"
S="$REF/src/pyclaw/sharpclaw"
TMP=$(mktemp -d)
trap 'rm -rf "$TMP"' EXIT
cd "$TMP"
$FC $FFLAGS -shared -o "$OUT/libref_classic.so" \
    "$C/limiter.f90" "$C/philim.f90" "$C/flux2.f90" "$C/step2ds.f90" "$C/step2.f90" \
    "$D/rpn2_euler_5wave.f" "$D/rpt2_euler_5wave.f"
$FC $FFLAGS -shared -o "$OUT/libref_sharpclaw.so" \
    "$S/ClawParams.f90" "$S/weno.f90" "$S/reconstruct.f90" "$S/tfluct.f90" "$S/evec.f90" \
    "$S/workspace2.f90" "$S/flux2.f90" "$S/flux1_2.f90" "$HERE/ref_sharp_setup.f90"
echo "built $OUT/libref_classic.so $OUT/libref_sharpclaw.so"
