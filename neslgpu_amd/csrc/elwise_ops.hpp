// elwise_ops.hpp -- the elementwise operators of the CVL as device functors
// (ref nesl/cvl/serial/elwise.c:128-298, defins.h:116-129), shared by the
// elementwise entry points of elwise.hip and by generated fused-operation
// libraries, so that a fused expression computes exactly what the sequence
// of CVL calls it replaces computes.
//
// Arithmetic follows C on the host: int overflow wraps (computed on
// unsigned), int / and % by zero give 0 (the serial CVL would trap),
// floating point is compiled without contraction (-ffp-contract=off) so
// + - * / and sqrt round exactly as the serial CVL; libm transcendentals may
// differ from the host's in the last ulp.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

namespace cvlhip {

#define CVLHIP_F2(NAME, TI, TO, EXPR)                                                      \
    struct NAME {                                                                          \
        using In = TI;                                                                     \
        using Out = TO;                                                                    \
        __device__ __forceinline__ TO operator()(TI a, TI b) const { return (TO)(EXPR); }  \
    };
#define CVLHIP_F1(NAME, TI, TO, EXPR)                                                      \
    struct NAME {                                                                          \
        using In = TI;                                                                     \
        using Out = TO;                                                                    \
        __device__ __forceinline__ TO operator()(TI a) const { return (TO)(EXPR); }        \
    };

#define UADD(a, b) (int)((unsigned)(a) + (unsigned)(b))
#define USUB(a, b) (int)((unsigned)(a) - (unsigned)(b))
#define UMUL(a, b) (int)((unsigned)(a) * (unsigned)(b))

CVLHIP_F2(MaxZf, int, int, a > b ? a : b)
CVLHIP_F2(MinZf, int, int, a < b ? a : b)
CVLHIP_F2(AddZf, int, int, UADD(a, b))
CVLHIP_F2(SubZf, int, int, USUB(a, b))
CVLHIP_F2(MulZf, int, int, UMUL(a, b))
CVLHIP_F2(DivZf, int, int, b != 0 ? a / b : 0)
CVLHIP_F2(GrtZf, int, int, a > b)
CVLHIP_F2(LesZf, int, int, a < b)
CVLHIP_F2(GeqZf, int, int, a >= b)
CVLHIP_F2(LeqZf, int, int, a <= b)
CVLHIP_F2(EqlZf, int, int, a == b)
CVLHIP_F2(NeqZf, int, int, a != b)
CVLHIP_F2(LshZf, int, int, (int)((unsigned)a << (b & 31)))
CVLHIP_F2(RshZf, int, int, a >> (b & 31))
CVLHIP_F2(ModZf, int, int, b != 0 ? a % b : 0)
CVLHIP_F2(AndZf, int, int, a & b)
CVLHIP_F2(IorZf, int, int, a | b)
CVLHIP_F2(XorZf, int, int, a ^ b)
CVLHIP_F2(MaxDf, double, double, a > b ? a : b)
CVLHIP_F2(MinDf, double, double, a < b ? a : b)
CVLHIP_F2(AddDf, double, double, a + b)
CVLHIP_F2(SubDf, double, double, a - b)
CVLHIP_F2(MulDf, double, double, a * b)
CVLHIP_F2(DivDf, double, double, a / b)
CVLHIP_F2(GrtDf, double, int, a > b)
CVLHIP_F2(LesDf, double, int, a < b)
CVLHIP_F2(GeqDf, double, int, a >= b)
CVLHIP_F2(LeqDf, double, int, a <= b)
CVLHIP_F2(EqlDf, double, int, a == b)
CVLHIP_F2(NeqDf, double, int, a != b)
CVLHIP_F2(EqlBf, int, int, a == b)
CVLHIP_F2(NeqBf, int, int, a != b)
CVLHIP_F2(AndBf, int, int, (a && b))
CVLHIP_F2(IorBf, int, int, (a || b))
CVLHIP_F2(XorBf, int, int, a ^ b)

CVLHIP_F1(NotBf, int, int, !a)
CVLHIP_F1(NotZf, int, int, ~a)
CVLHIP_F1(IntDf, double, int, (int)a)
CVLHIP_F1(IntBf, int, int, a)
CVLHIP_F1(DblZf, int, double, (double)a)
CVLHIP_F1(BooZf, int, int, a != 0)
CVLHIP_F1(FlrDf, double, int, (int)floor(a))
CVLHIP_F1(CeiDf, double, int, (int)ceil(a))
CVLHIP_F1(TrnDf, double, int, (int)a)
CVLHIP_F1(RouDf, double, int, (int)(a + 0.5))
CVLHIP_F1(ExpDf, double, double, exp(a))
CVLHIP_F1(LogDf, double, double, log(a))
CVLHIP_F1(SqtDf, double, double, sqrt(a))
CVLHIP_F1(SinDf, double, double, sin(a))
CVLHIP_F1(CosDf, double, double, cos(a))
CVLHIP_F1(TanDf, double, double, tan(a))
CVLHIP_F1(AsnDf, double, double, asin(a))
CVLHIP_F1(AcsDf, double, double, acos(a))
CVLHIP_F1(AtnDf, double, double, atan(a))
CVLHIP_F1(SnhDf, double, double, sinh(a))
CVLHIP_F1(CshDf, double, double, cosh(a))
CVLHIP_F1(TnhDf, double, double, tanh(a))
CVLHIP_F1(CpyZf, int, int, a)
CVLHIP_F1(CpyDf, double, double, a)

}  // namespace cvlhip
