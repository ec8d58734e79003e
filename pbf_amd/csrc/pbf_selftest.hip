// Self-tests of the fast correctly rounded primitives of pbf_math.h against
// the compiler's general IEEE sequences, on the device that runs them:
//   reciprocal  every float b with |b| in [2^-60, 2^60)           (exhaustive)
//   square root every float x in [2^-96, 2^8)                      (exhaustive)
//   division    pseudo-random a, b of the fast ranges, including the
//               solver's operands a = s dx, b = l in (0, 1]         (sampled;
//               the bounded form on the classes inside its bounds, and its
//               form without the sign of a zero quotient up to that sign)
#include "pbf_internal.h"
#include "pbf_math.h"

namespace pbf {
namespace {

constexpr int ST_BLOCK = 256;

__device__ __forceinline__ void count_mismatch(bool bad, unsigned long long *counter) {
    const unsigned long long m = __ballot(bad);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(counter, (unsigned long long)__popcll(m));
}

// Index k -> float with biased exponent e0 + k / 2^23 and mantissa k % 2^23.
__device__ __forceinline__ float float_at(uint64_t k, uint32_t e0, bool neg) {
    const uint32_t e = e0 + (uint32_t)(k >> 23);
    const uint32_t bits = (e << 23) | (uint32_t)(k & 0x7FFFFFu) | (neg ? 0x80000000u : 0u);
    return __uint_as_float(bits);
}

__global__ __launch_bounds__(ST_BLOCK) void k_test_rcp(uint64_t count, uint32_t e0, unsigned long long *bad) {
    const uint64_t stride = (uint64_t)gridDim.x * ST_BLOCK;
    for (uint64_t base = (uint64_t)blockIdx.x * ST_BLOCK; base < count; base += stride) {
        const uint64_t k = base + threadIdx.x;
        bool mis = false;
        if (k < count) {
            const float b = float_at(k >> 1, e0, k & 1);
            mis = __float_as_uint(rcp_rn_fast(b)) != __float_as_uint(1.0f / b);
        }
        count_mismatch(mis, bad);
    }
}

__global__ __launch_bounds__(ST_BLOCK) void k_test_sqrt(uint64_t count, uint32_t e0, unsigned long long *bad) {
    const uint64_t stride = (uint64_t)gridDim.x * ST_BLOCK;
    for (uint64_t base = (uint64_t)blockIdx.x * ST_BLOCK; base < count; base += stride) {
        const uint64_t k = base + threadIdx.x;
        bool mis = false;
        if (k < count) {
            const float x = float_at(k, e0, false);
            mis = __float_as_uint(sqrt_rn_fast(x)) != __float_as_uint(sqrtf(x));
            // and the zero squared distance the tame-position form feeds it
            if (k == 0) mis = mis | (__float_as_uint(sqrt_rn_fast(0.0f)) != 0u);
        }
        count_mismatch(mis, bad);
    }
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(ST_BLOCK) void k_test_div(uint64_t count, uint64_t seed, unsigned long long *bad) {
    const uint64_t stride = (uint64_t)gridDim.x * ST_BLOCK;
    for (uint64_t base = (uint64_t)blockIdx.x * ST_BLOCK; base < count; base += stride) {
        const uint64_t k = base + threadIdx.x;
        bool mis = false;
        if (k < count) {
            const uint64_t r1 = mix64(seed + 2 * k + 1), r2 = mix64(seed + 2 * k + 2);
            float a, b;
            switch (k & 3) {
                case 0: {  // solver operands: b = l in (2^-12, 1], a = s * d, |s| <= 14.33, |d| <= 1
                    b = __uint_as_float(0x39800000u + (uint32_t)(r1 % (0x3F800001u - 0x39800000u)));
                    const float t = 1.0f - b;
                    const float s = (-3.0f * 4.774648292756860f * t) * t;
                    const float d = (float)((int32_t)(r2 >> 32)) * 0x1p-31f;
                    a = s * d;
                    break;
                }
                case 1: {  // arbitrary bit patterns of the fast ranges
                    b = __uint_as_float(0x21800000u + (uint32_t)(r1 % (0x4E800001u - 0x21800000u)));
                    a = __uint_as_float(0x0F800000u + (uint32_t)(r2 % (0x5D800001u - 0x0F800000u)) |
                                        ((uint32_t)(r2 >> 63) << 31));
                    break;
                }
                case 2: {  // numerator and denominator with close mantissas (quotients near 1)
                    b = __uint_as_float(0x3F800000u | (uint32_t)(r1 & 0x7FFFFFu));
                    a = __uint_as_float((0x3F800000u | (uint32_t)(r1 & 0x7FFFFFu)) + (uint32_t)(r2 % 64) - 32u);
                    break;
                }
                default: {  // signed zero numerators
                    b = __uint_as_float(0x3A800000u + (uint32_t)(r1 % (0x3F800001u - 0x3A800000u)));
                    a = (r2 & 1) ? -0.0f : 0.0f;
                    break;
                }
            }
            float q0, q1, q2;
            div3_rn(a, a, a, b, q0, q1, q2);
            mis = __float_as_uint(q0) != __float_as_uint(a / b);
            if ((k & 3) != 1) {  // classes within b <= 2^30, |a| <= 2^60: the bounded form too
                div3_rn<true>(a, a, a, b, q0, q1, q2);
                mis |= __float_as_uint(q0) != __float_as_uint(a / b);
                // and the form that may drop the sign of a zero quotient (the
                // solver's gradient): equal bits, or two zeros
                div3_rn<true, false>(a, a, a, b, q0, q1, q2);
                const float ref = a / b;
                mis |= (__float_as_uint(q0) != __float_as_uint(ref)) & !((q0 == 0.0f) & (ref == 0.0f));
            }
        }
        count_mismatch(mis, bad);
    }
}

// The neighbour passes' workgroup renumbering as the kernels compute it:
// out[2 b] = xcd_block<NB>(b, nb), out[2 b + 1] = the XCD (XCC_ID hardware
// register, read-only) that ran workgroup b.
template <int NB>
__global__ __launch_bounds__(64) void k_xcd_map(uint32_t nb, uint32_t *out) {
    const uint32_t b = blockIdx.x;
    if (threadIdx.x != 0 || b >= nb) return;
    uint32_t xcc;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID, 0, 4)" : "=s"(xcc));
    out[2 * (size_t)b] = xcd_block<NB>(b, nb);
    out[2 * (size_t)b + 1] = xcc;
}

}  // namespace
}  // namespace pbf

extern "C" int pbf_selftest_xcd_map(int32_t width, uint32_t nblocks, uint32_t *out) {
    using namespace pbf;
    if (!out || nblocks == 0) {
        set_error("null output or no workgroups");
        return PBF_ERR_INVALID;
    }
    uint32_t *d = nullptr;
    PBF_HIP(hipMalloc(&d, 2 * (size_t)nblocks * sizeof(uint32_t)));
    hipError_t e = hipSuccess;
    switch (width) {
        case 64: hipLaunchKernelGGL(k_xcd_map<64>, dim3(nblocks), dim3(64), 0, nullptr, nblocks, d); break;
        case 128: hipLaunchKernelGGL(k_xcd_map<128>, dim3(nblocks), dim3(64), 0, nullptr, nblocks, d); break;
        case 256: hipLaunchKernelGGL(k_xcd_map<256>, dim3(nblocks), dim3(64), 0, nullptr, nblocks, d); break;
        case 512: hipLaunchKernelGGL(k_xcd_map<512>, dim3(nblocks), dim3(64), 0, nullptr, nblocks, d); break;
        default:
            (void)hipFree(d);
            set_error("width must be 64, 128, 256 or 512");
            return PBF_ERR_INVALID;
    }
    e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpy(out, d, 2 * (size_t)nblocks * sizeof(uint32_t), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return hip_fail(e, "pbf_selftest_xcd_map");
    return PBF_OK;
}

extern "C" int pbf_selftest_math(uint64_t div_samples, uint64_t seed, uint64_t out[6]) {
    using namespace pbf;
    if (!out) {
        set_error("null output");
        return PBF_ERR_INVALID;
    }
    unsigned long long *bad = nullptr;
    PBF_HIP(hipMalloc(&bad, 3 * sizeof(unsigned long long)));
    hipError_t e = hipMemset(bad, 0, 3 * sizeof(unsigned long long));
    const dim3 grid(8192), block(ST_BLOCK);
    // |b| in [2^-60, 2^60): biased exponents 67 .. 186, both signs
    const uint64_t n_rcp = 2ull * 120ull << 23;
    // x in [2^-96, 2^8): biased exponents 31 .. 134
    const uint64_t n_sqrt = 104ull << 23;
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_test_rcp, grid, block, 0, nullptr, n_rcp, 67u, bad);
        hipLaunchKernelGGL(k_test_sqrt, grid, block, 0, nullptr, n_sqrt, 31u, bad + 1);
        hipLaunchKernelGGL(k_test_div, grid, block, 0, nullptr, div_samples, seed, bad + 2);
        e = hipGetLastError();
    }
    unsigned long long host[3] = {0, 0, 0};
    if (e == hipSuccess) e = hipMemcpy(host, bad, sizeof host, hipMemcpyDeviceToHost);
    (void)hipFree(bad);
    if (e != hipSuccess) return hip_fail(e, "pbf_selftest_math");
    out[0] = host[0];
    out[1] = host[1];
    out[2] = host[2];
    out[3] = n_rcp;
    out[4] = n_sqrt;
    out[5] = div_samples;
    return PBF_OK;
}
