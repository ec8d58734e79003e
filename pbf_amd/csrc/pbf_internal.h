// Internal declarations shared by the HIP translation units of libpbf_amd.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "pbf.h"

namespace pbf {

// Constants of one solver configuration, resolved on the host once per
// parameter change. FP32 values are computed with the same float operations
// the reference's shaders use (left-associated products of h).
struct SolverConsts {
    int32_t n;
    int32_t gx, gy, gz;
    float h, hh;          // h, h*h
    float h9;             // h*h*h*h*h*h*h*h*h (Wpoly6 denominator)
    float h2cut;          // sq above this has fl(sqrt(sq)) > h for sure
    float h2loose;        // h2cut (1 + 2^-20): bound for the fused-multiply-add distance estimate
    int32_t h_is_one;     // h == 1: the h-products are identities
    float inv_rho0, eps, gravity, dt, k, scale, xsph_c, vort_eps;
    float lo[3], hi[3], gdir[3];
    int32_t extforce;
    uint32_t key_mask;
    uint32_t epoch;       // tag of this step's end-table writes
};

SolverConsts make_consts(uint32_t n, const int32_t grid[3], const pbf_params &p, int key_bits);

// ------------------------------------------------------------ launchers
void launch_predict(const SolverConsts &c, const float4 *pos, const float4 *vel, float4 *rec, uint32_t *keys,
                    hipStream_t st);
void launch_gather_records(int32_t n, const uint32_t *perm, const float4 *rec_in, float4 *rec_out,
                           hipStream_t st);
// out[g] = src[id of rec[g]]
void launch_gather_by_id(int32_t n, const float4 *rec, const float4 *src, float4 *out, hipStream_t st);
void launch_keys_of_records(const SolverConsts &c, const float4 *rec, uint32_t *keys, hipStream_t st);
void launch_fill_i32(int32_t *p, int32_t value, size_t count, hipStream_t st);
void launch_findcells(const SolverConsts &c, const float4 *sorted, int32_t *start, unsigned long long *end64,
                      hipStream_t st);
void launch_neighbourcells(const SolverConsts &c, const float4 *sorted, const int32_t *start,
                           const unsigned long long *end64, int32_t *nbr9, hipStream_t st);
// Per-particle list of the neighbours within h at one set of positions:
// written by the lambda pass (vorticity pass A), replayed by the position
// update (vorticity pass B). entries [cap][n] of 16-bit (row, offset) codes.
// All pointers null: the replaying pass scans its candidates.
struct PairList {
    uint16_t *entries = nullptr;
    uint16_t *count = nullptr;
    int32_t cap = 0;
};
void launch_lambda(const SolverConsts &c, const float4 *rec, const int32_t *nbr9, float *lambda, const PairList &pl,
                   hipStream_t st);
void launch_updatepos(const SolverConsts &c, const float4 *rec_in, const float *lambda, const int32_t *nbr9,
                      const PairList &pl, float4 *rec_out, hipStream_t st);
void launch_update(const SolverConsts &c, const float4 *rec, float4 *pos, float4 *vel, float4 *vel_sorted,
                   hipStream_t st);
void launch_vorticity_a(const SolverConsts &c, const float4 *rec, const int32_t *nbr9, const float4 *vel_sorted,
                        const float4 *vel_id, float4 *vxsph, float4 *omega, float *vort, const PairList &pl,
                        hipStream_t st);
void launch_vorticity_b(const SolverConsts &c, const float4 *rec, const int32_t *nbr9, const PairList &pl,
                        const float4 *vxsph, const float4 *omega, const float *vort, float4 *vel_id, hipStream_t st);
void launch_nbr9_to_12(int32_t n, const int32_t *nbr9, int32_t *nbr12, hipStream_t st);
void launch_nbr12_to_9(int32_t n, const int32_t *nbr12, int32_t *nbr9, hipStream_t st);
void launch_end64_to_i32(size_t cells, const unsigned long long *end64, int32_t *end, hipStream_t st);
void launch_i32_to_end64(size_t cells, const int32_t *end, unsigned long long *end64, hipStream_t st);

// ------------------------------------------------------------ radix sort
// Stable LSD sort of (key, index) pairs by the low key_bits bits, 8 bits per
// pass. Produces perm[g] = input index at sorted slot g.
struct RadixScratch {
    int32_t capacity = 0;
    uint32_t *keys_alt = nullptr;
    uint32_t *vals = nullptr;
    uint32_t *vals_alt = nullptr;
    uint32_t *counts = nullptr;  // 256 * tiles, digit-major
    uint32_t *totals = nullptr;  // 256 digit totals
    int32_t tiles = 0;
};
bool radix_alloc(RadixScratch &s, int32_t n);
void radix_free(RadixScratch &s);
// keys is clobbered. The permutation is written to perm, which must not alias
// the scratch buffers.
void radix_sort_perm(RadixScratch &s, uint32_t *keys, int32_t n, int key_bits, uint32_t *perm, hipStream_t st);

// ------------------------------------------------------------ errors
void set_error(const std::string &msg);
int hip_fail(hipError_t e, const char *what);

}  // namespace pbf

#define PBF_HIP(call)                                            \
    do {                                                         \
        hipError_t e_ = (call);                                  \
        if (e_ != hipSuccess) return pbf::hip_fail(e_, #call);   \
    } while (0)
