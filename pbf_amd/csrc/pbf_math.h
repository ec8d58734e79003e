// Correctly rounded FP32 primitives for the exact solver passes (gfx950).
//
// The library is built with -ffp-contract=off; every function here returns
// the IEEE round-to-nearest-even result of the operation it names, for the
// operand ranges stated, using fewer instructions than the compiler's general
// sequences (which also handle denormals, infinities and NaN). Operands outside
// the stated ranges take the compiler's general sequence. The self-test
// kernels in pbf_selftest.hip check the fast forms against the general ones,
// exhaustively over every float of the ranges where that is cheap.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pbf {

// Bounds of the fast ranges: denominators in [2^-60, 2^30], nonzero numerators
// in [2^-96, 2^60], so quotients and Markstein remainders stay normal.
constexpr float kDenLo = 0x1p-60f;
constexpr float kDenHi = 0x1p30f;
constexpr float kNumLo = 0x1p-96f;
constexpr float kNumHi = 0x1p60f;
constexpr float kSqrtLo = 0x1p-96f;  // below this the general sqrt scales its operand

// RN(sqrt(x)) for x in [2^-96, +inf) finite: the hardware square root (within
// one ulp) corrected by the signs of the residuals of its two neighbours,
// which is the compiler's sequence without the rescaling of tiny operands.
__device__ __forceinline__ float sqrt_rn_fast(float x) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const float sm = __int_as_float(__float_as_int(s) - 1);
    const float sp = __int_as_float(__float_as_int(s) + 1);
    const float rm = fmaf(-sm, s, x);
    const float rp = fmaf(-sp, s, x);
    float r = rm <= 0.0f ? sm : s;
    r = rp > 0.0f ? sp : r;
    return r;
}

// True in every lane when any lane of the wave has p: a wave-uniform branch
// condition, so that rare fallbacks cost no per-lane mask bookkeeping.
__device__ __forceinline__ bool wave_any(bool p) { return __ballot(p) != 0ull; }

// Lane mask of a single comparison: the mask the comparison itself produces.
// Combining such masks with scalar operations, rather than taking the ballot
// of a combined predicate, saves the two vector instructions the compiler
// otherwise spends turning the predicate into a register and back.
__device__ __forceinline__ uint64_t lanes(bool p) { return __builtin_amdgcn_ballot_w64(p); }
__device__ __forceinline__ uint64_t active_lanes() { return __builtin_amdgcn_read_exec(); }

// RN(sqrt(x)) of a squared distance x (not +inf: already compared with a
// finite bound). Operands below the fast range (and NaN) take the general
// sequence, in the lanes that have them, under a wave-uniform branch; the
// empty volatile asm keeps the compiler from speculating that sequence into
// every call (it would otherwise if-convert the rare branch and run both
// square roots for every operand). TAME: every operand is 0 or at least 2^-96
// (the squares of coordinate differences that are 0 or at least 2^-48), and
// the fast form alone is exact for both (at 0 both neighbours' residuals fail
// their tests and the hardware root of 0 is returned; the self-test checks
// it), so the test is skipped.
__device__ __forceinline__ float sqrt_rn_dist(float x, bool tame) {
    float r = sqrt_rn_fast(x);
    if (!tame) {
        if (lanes(x >= kSqrtLo) != active_lanes()) {
            asm volatile("" ::: "memory");
            if (!(x >= kSqrtLo)) r = sqrtf(x);
        }
    }
    return r;
}

// RN(1 / b) for |b| in [2^-60, 2^60): the hardware reciprocal refined by one
// Newton step in FMA form. The self-test checks it against the IEEE division
// for every float of that range.
__device__ __forceinline__ float rcp_rn_fast(float b) {
    const float y0 = __builtin_amdgcn_rcpf(b);
    const float e = fmaf(-b, y0, 1.0f);
    return fmaf(e, y0, y0);
}

// Markstein's correction: for y = RN(1/b) and q = RN(a y) within one ulp of
// a/b, the remainder r = a - b q is exact (one FMA) and RN(q + r y) = RN(a/b)
// when nothing underflows (Markstein 1990; Muller et al., Handbook of
// Floating-Point Arithmetic, Thm. 4.10). Valid for |a| in [2^-96, 2^60] and b
// in [2^-60, 2^30]. A zero a gives +0 from the two FMAs whatever its sign;
// SIGNED_ZERO returns a instead (the signed zero a / b gives, b > 0).
template <bool SIGNED_ZERO = true>
__device__ __forceinline__ float div_markstein(float a, float b, float y) {
    const float q = a * y;
    const float r = fmaf(-q, b, a);
    const float q1 = fmaf(r, y, q);
    if constexpr (SIGNED_ZERO) return a == 0.0f ? a : q1;
    return q1;
}

// True when a is zero or |a| lies in the fast numerator range.
__device__ __forceinline__ bool fast_numerator(float a) {
    const float aa = fabsf(a);
    return (aa == 0.0f) | ((aa >= kNumLo) & (aa <= kNumHi));
}

// True when a is zero or |a| >= kNumLo (callers that know |a| <= kNumHi).
__device__ __forceinline__ bool fast_numerator_lo(float a) {
    const float aa = fabsf(a);
    return (aa == 0.0f) | (aa >= kNumLo);
}

// RN(a_k / b), k = 0..2, for b > 0: the fast path for every lane, replaced by
// the IEEE division in the lanes whose operands leave the fast ranges.
// BOUNDED: the caller guarantees b <= kDenHi and |a_k| <= kNumHi, so only the
// lower bounds are tested. SIGNED_ZERO false: a zero quotient may come out as
// +0 where a / b is -0 (every other quotient is a / b bit for bit), for
// callers whose results do not depend on the sign of a zero quotient.
template <bool BOUNDED = false, bool SIGNED_ZERO = true>
__device__ __forceinline__ void div3_rn(float a0, float a1, float a2, float b, float &q0, float &q1, float &q2) {
    const float y = rcp_rn_fast(b);
    q0 = div_markstein<SIGNED_ZERO>(a0, b, y);
    q1 = div_markstein<SIGNED_ZERO>(a1, b, y);
    q2 = div_markstein<SIGNED_ZERO>(a2, b, y);
    const bool fast = BOUNDED ? (b >= kDenLo) & fast_numerator_lo(a0) & fast_numerator_lo(a1) & fast_numerator_lo(a2)
                              : (b >= kDenLo) & (b <= kDenHi) & fast_numerator(a0) & fast_numerator(a1) &
                                    fast_numerator(a2);
    // the same test as wave_any(!fast), from the lane masks of the single
    // comparisons (zero or >= kNumLo, per numerator)
    uint64_t ok = lanes(b >= kDenLo);
    if (BOUNDED) {
        ok &= (lanes(fabsf(a0) == 0.0f) | lanes(fabsf(a0) >= kNumLo)) &
              (lanes(fabsf(a1) == 0.0f) | lanes(fabsf(a1) >= kNumLo)) &
              (lanes(fabsf(a2) == 0.0f) | lanes(fabsf(a2) >= kNumLo));
    } else {
        ok &= lanes(b <= kDenHi);
        ok &= (lanes(fabsf(a0) == 0.0f) | (lanes(fabsf(a0) >= kNumLo) & lanes(fabsf(a0) <= kNumHi))) &
              (lanes(fabsf(a1) == 0.0f) | (lanes(fabsf(a1) >= kNumLo) & lanes(fabsf(a1) <= kNumHi))) &
              (lanes(fabsf(a2) == 0.0f) | (lanes(fabsf(a2) >= kNumLo) & lanes(fabsf(a2) <= kNumHi)));
    }
    if (ok != active_lanes()) {
        asm volatile("" ::: "memory");
        if (!fast) {
            q0 = a0 / b;
            q1 = a1 / b;
            q2 = a2 / b;
        }
    }
}

}  // namespace pbf
