// FP64 classic kernels with floating-point contraction: a*b + c is computed
// as one fused multiply-add with a single rounding wherever the compiler can
// fuse it, and the Euler solver divides once per shared denominator
// (CLAW_SHARED_RECIPROCALS in riemann.hpp).  Same kernels, launch shapes and
// operation order otherwise as classic2d.hip (which this file compiles a
// second time, in namespace clawhip::fused); the results differ from the
// reference's CPU arithmetic in the last bits only.
// Selected by precision CLAW_F64_FMA in claw2d_classic_step_ex.
#pragma clang fp contract(fast)
#define CLAW_SHARED_RECIPROCALS 1
#define CLAW_CLASSIC_VARIANT fused
#define CLAW_CLASSIC_F64_KEY CLAW_F64_FMA
#include "classic2d.hip"
