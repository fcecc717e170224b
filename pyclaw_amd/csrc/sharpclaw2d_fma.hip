// FP64 SharpClaw stage kernels with floating-point contraction (a*b + c as
// one fused multiply-add where the compiler can fuse it) and one division
// per shared denominator (the WENO5 weights' smoothness indicators, the Roe
// solvers' 1/rho-type quantities; CLAW_SHARED_RECIPROCALS).  Same kernels and
// launch shapes as sharpclaw2d.hip, which this file compiles a second time,
// in namespace clawhip::fused; selected by precision CLAW_F64_FMA.
#pragma clang fp contract(fast)
#define CLAW_SHARED_RECIPROCALS 1
#define CLAW_SHARP_VARIANT fused
#define CLAW_SHARP_F64_KEY CLAW_F64_FMA
#include "sharpclaw2d.hip"
