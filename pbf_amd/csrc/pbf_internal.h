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
    // Quantized positions of the candidate filter: a coordinate in units of
    // 1/qscale (a power of two with h qscale in [2048, 4096)), truncated, low
    // 16 bits; a pair with fl(sqrt(sq)) <= h has a quantized squared
    // distance of at most qcut.
    float qscale;
    int32_t qcut;
};

SolverConsts make_consts(uint32_t n, const int32_t grid[3], const pbf_params &p, int key_bits);

// One scene's slice of the particle arrays and its own constants. Sorted and
// id-order arrays hold each scene's particles in the same contiguous range
// [off, off + n); records carry scene-local ids, the neighbour table and the
// cell tables scene-local particle indices, so every array restricted to a
// scene equals the plain solver's array for that scene.
struct Seg {
    int32_t off, n;
    int32_t gx, gy, gz;
    int32_t cell_base;          // first entry of the scene's cells in the cell tables
    uint32_t key_mask, key_hi;  // key = key_hi | (cell hash & key_mask)
    float lo[3], hi[3], gdir[3];
    float xsph_c;
};

Seg seg_of(const SolverConsts &c);

// Where a kernel finds the scene of global slot i: one scene for the whole
// launch (kernel argument), or a table indexed by a per-slot scene number.
struct DomOne {
    Seg s;
    __device__ __forceinline__ Seg at(int32_t) const { return s; }
};
struct DomMany {
    const Seg *segs;
    const uint16_t *scene_of;
    __device__ __forceinline__ Seg at(int32_t i) const { return segs[scene_of[i]]; }
};

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2.
// Renumbering them so that every XCD runs chunks of kXcdChunkParticles
// consecutive sorted particles keeps a chunk's neighbour rows in one L2, while
// consecutive chunks go to different XCDs, so a dense region of the sorted
// array (whose particles have many more neighbours than the rest) is shared by
// all XCDs instead of landing on one. Bijective on [0, gridDim.x): full groups
// of 8 chunks are permuted, the remainder keeps its numbers.
constexpr uint32_t kXcdChunkParticles = 2048;

// b: the workgroup's number among nb (counted from a multiple of 8).
template <int NB>
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb) {
    constexpr uint32_t chunk = kXcdChunkParticles / NB;
    static_assert(chunk >= 1 && chunk * NB == kXcdChunkParticles, "chunk of whole workgroups");
    const uint32_t group = 8u * chunk;
    if (b >= nb / group * group) return b;
    const uint32_t x = b % 8u, k = b / 8u;
    return (k / chunk) * group + x * chunk + (k % chunk);
}

// Quantized position of a record for the candidate filter: each coordinate
// times qscale, clamped into the int32 range (NaN to 0), truncated, low 16
// bits; x and y in the first word, z in the low half of the second. The scale
// is a power of two, so the product is exact and each quantized coordinate is
// within one unit of the scaled one; clamping only brings values closer.
__device__ __forceinline__ uint32_t quantize1(float x, float qscale) {
    const float v = fminf(fmaxf(x * qscale, -2147483648.0f), 2147483520.0f);
    return (uint32_t)(int32_t)v & 0xFFFFu;
}
__device__ __forceinline__ uint2 quantize_pos(const float4 &r, float qscale) {
    return make_uint2(quantize1(r.x, qscale) | (quantize1(r.y, qscale) << 16), quantize1(r.z, qscale));
}

// A record with a coordinate in (0, 2^-25) in magnitude. When no record of a
// set of positions has one (every coordinate is 0 or at least 2^-25, so every
// coordinate difference is 0 or at least 2^-48 in magnitude), then for h = 1
// every pair within h (0 < r < 1, so h - r >= 2^-24 and the spiky factor is 0
// or at least 2^-45 in magnitude) has gradient numerators that are 0 or at
// least 2^-93 and a distance of at least 2^-48: inside the fast ranges of the
// division, whose range tests the pair passes can then skip. The writers of
// each set of positions record in a flag word whether any record is tiny.
__device__ __forceinline__ bool tiny_coord(float x) {
    const float a = fabsf(x);
    return (a > 0.0f) & (a < 0x1p-25f);
}
__device__ __forceinline__ void note_tiny(uint32_t *flag, const float4 &r) {
    if (flag && (tiny_coord(r.x) | tiny_coord(r.y) | tiny_coord(r.z))) *flag = 1u;
}

// ------------------------------------------------------------ launchers
// highlight (or null): the SPH handle's highlight words by id, of which the
// predict pass keeps bit 0 (clearhighlight.glsl) and the neighbour-cell pass
// then marks the candidates of the flagged particles (highlight.glsl).
// tiny (or null): ntiny flag words cleared for the step (positions_tame).
void launch_predict(const SolverConsts &c, const float4 *pos, const float4 *vel, float4 *rec, uint32_t *keys,
                    hipStream_t st, uint32_t *highlight = nullptr, uint32_t *tiny = nullptr, int32_t ntiny = 0);
void launch_gather_records(int32_t n, const uint32_t *perm, const float4 *rec_in, float4 *rec_out,
                           hipStream_t st);
// out[g] = src[id of rec[g]]
void launch_gather_by_id(int32_t n, const float4 *rec, const float4 *src, float4 *out, hipStream_t st);
void launch_keys_of_records(const SolverConsts &c, const float4 *rec, uint32_t *keys, hipStream_t st);
void launch_fill_i32(int32_t *p, int32_t value, size_t count, hipStream_t st);
// FindNeighbourCells' clear of the start table (src/NeighbourCellFinder.cpp:
// 115-126), which also advances the device-side tag of the step's end writes.
struct DenseList;
void launch_begin_cells(int32_t *start, size_t cells, uint32_t *epoch, hipStream_t st,
                        const DenseList *dl = nullptr);
void launch_findcells(const SolverConsts &c, const float4 *sorted, int32_t *start, unsigned long long *end64,
                      const uint32_t *epoch, hipStream_t st);
// dl: also list the step's dense particles and count the dense levels.
void launch_neighbourcells(const SolverConsts &c, const float4 *sorted, const int32_t *start,
                           const unsigned long long *end64, int32_t *nbr9, hipStream_t st,
                           const DenseList *dl = nullptr, uint32_t *highlight = nullptr);
// Per-particle list of the neighbours within h at one set of positions:
// written by the lambda pass (vorticity pass A), replayed by the position
// update (vorticity pass B). entries [cap][n] of 16-bit (row, offset) codes.
// All pointers null: the replaying pass scans its candidates.
// Capacity of the per-particle neighbour list that the lambda pass records for
// the position update (particles with more neighbours within h scan instead).
constexpr int32_t kListCap = 128;
// Rows allocated for the lists: a replay requests the rows of its next batch
// (up to 16 past the list's end) before it knows which of them are entries.
constexpr int32_t kListRows = kListCap + 16;
// List count of a particle with more than kListCap neighbours within h (no list).
constexpr uint16_t kListNone = 0xFFFF;

// Zero entries kept after the n particles of the one-scene solver's sorted
// records, lambda and |omega| arrays: a neighbour row starts below n and holds
// fewer than 128 entries, scanned in batches of at most 130, so no index reaches n + kPad.
constexpr int32_t kPad = 256;
// Solver iterations a parameter block may ask for.
constexpr int32_t kMaxIterations = 1000;

struct PairList {
    uint16_t *entries = nullptr;
    uint16_t *count = nullptr;
    int32_t cap = 0;
};
// The particles of a step with the most neighbour candidates, handed from
// the one-thread-per-particle path to the dense-particle workgroups of the same
// launches. Each step, the neighbour-cell pass keeps every particle's
// candidate count in cand and counts the particles above each of kDenseLevels
// candidate levels (levels[level * kDenseCopies + copy], cleared by
// launch_begin_cells); launch_dense_select then sets threshold to the lowest
// level with at most kDenseBudget particles above it (the highest level if
// none). Which particles are dense changes no result: both paths compute
// every particle bit for bit alike.
constexpr int kDenseLevels = 8;
constexpr int kDenseCopies = 256;
#ifndef PBF_DENSE_BUDGET
#define PBF_DENSE_BUDGET 32768
#endif
constexpr uint32_t kDenseBudget = PBF_DENSE_BUDGET;
struct DenseList {
    uint16_t *cand = nullptr;        // n entries
    uint32_t *levels = nullptr;      // kDenseLevels * kDenseCopies words
    int32_t *threshold = nullptr;    // 2 words: the threshold, the number of particles above it
};
void launch_dense_select(const DenseList &dl, hipStream_t st);
// padded: the arrays the neighbour loops gather from (rec, q and nrec) hold
// kPad zero entries after the n particles, so a batch of loads that runs past
// the end of a row stays in bounds.
// dl: the step's dense list, whose particles the launch's first workgroups
// handle, kDenseGroup lanes per particle. Null: every particle takes the
// one-thread-per-particle path.
// Values of a step that the adjoint's reverse pass reads, written as extra
// outputs of the solver passes when given (all null in the plain step).
struct SolverTape {
    float *sden = nullptr;     // lambda: sum |g_j|^2 + |sum g_j|^2 + eps, sorted slot
    float4 *G = nullptr;       // lambda: sum g_j (g_j = gradW / rho0)
    uint8_t *clamp = nullptr;  // position update: bit a set where axis a was not clamped, bit 3 + a at the upper wall
    float4 *eta = nullptr;     // vorticity B: eta before normalisation
};
// q: the quantized positions of rec, slot for slot (launch_quantize, or the
// sort's last pass and the position update's q_out), read by the candidate
// scans of lambda and vorticity A; padded like rec.
// tiny (or null): the flag word of rec's positions; when it is zero and h = 1,
// the pair passes skip the division's range tests (every operand is inside
// the fast ranges). Writers of positions set the flag of their output set.
// nrec: neighbour records for the replaying pass, nrec[i] = (position of
// rec[i], the pass's per-neighbour operand): lambda_i from launch_lambda,
// |omega_i| from launch_vorticity_a. The replaying pass reads its neighbours
// from it, one 16-byte load per pair. Padded like rec (the padding zero).
void launch_lambda(const SolverConsts &c, const float4 *rec, const uint2 *q, const int32_t *nbr9, float *lambda,
                   float4 *nrec, const PairList &pl, hipStream_t st, bool padded = false,
                   const DenseList *dl = nullptr, const SolverTape *tape = nullptr, const uint32_t *tiny = nullptr);
// The step's closing update (update.glsl:16-28: velocities from the
// positions before and after the step, both stored by particle id; vel_sorted,
// if set, also keeps the velocities in sorted order). Given to the last
// position update of a step, which performs it for each particle as soon as
// that particle's final position is known, in place of launch_update.
struct StepUpdate {
    float4 *pos = nullptr;
    float4 *vel = nullptr;
    float4 *vel_sorted = nullptr;
};
// q_out: the quantized positions of rec_out, written beside it (null: none).
void launch_updatepos(const SolverConsts &c, const float4 *rec_in, const float *lambda, const float4 *nrec,
                      const int32_t *nbr9, const PairList &pl, float4 *rec_out, uint2 *q_out, hipStream_t st,
                      bool padded = false,
                      const DenseList *dl = nullptr, const StepUpdate *up = nullptr,
                      const SolverTape *tape = nullptr, const uint32_t *tiny_in = nullptr,
                      uint32_t *tiny_out = nullptr);
void launch_update(const SolverConsts &c, const float4 *rec, float4 *pos, float4 *vel, float4 *vel_sorted,
                   hipStream_t st);
// vort (or null): |omega| by sorted slot (the reference's vorticities[]);
// the second pass reads it from nrec.
void launch_vorticity_a(const SolverConsts &c, const float4 *rec, const uint2 *q, const int32_t *nbr9,
                        const float4 *vel_sorted, const float4 *vel_id, float4 *vxsph, float4 *omega, float *vort,
                        float4 *nrec, const PairList &pl, hipStream_t st, bool padded = false,
                        const DenseList *dl = nullptr, const uint32_t *tiny = nullptr);
// (x, y, z of rec[i], v[i]) to out[i], i < count: the neighbour records of a
// replaying pass from a caller's records and per-particle operand.
void launch_pack_records(int32_t count, const float4 *rec, const float *v, float4 *out, hipStream_t st);
// q[i] = the quantized position of rec[i], i < count.
void launch_quantize(const SolverConsts &c, const float4 *rec, uint2 *q, int32_t count, hipStream_t st);
void launch_vorticity_b(const SolverConsts &c, const float4 *rec, const int32_t *nbr9, const PairList &pl,
                        const float4 *vxsph, const float4 *omega, const float4 *nrec, float4 *vel_id, hipStream_t st,
                        bool padded = false, const DenseList *dl = nullptr, const SolverTape *tape = nullptr,
                        const uint32_t *tiny = nullptr);
// Launch shape of the batched form: workgroup size of the neighbour passes
// (64, 128, 256, 512), workgroup size (64 .. 1024, a power of two) and particles
// per thread (1 .. 4) of the streaming passes, the scanning passes' LDS list
// depth, and persistent neighbour launches. The one-scene form uses the
// default.
struct LaunchShape {
    int32_t nb = 64;
    int32_t sb = 256;
    int32_t ppt = 1;
    int32_t lc = 56;          // LDS list entries per thread of the scanning passes (32 or 56)
    int32_t persistent = 0;   // neighbour launches of device-filling workgroups looping over tiles
};
bool valid_shape(const LaunchShape &ls);

// Batched (many-scene) forms of the step's launches; n of c is the total.
void launch_predict(const SolverConsts &c, const DomMany &d, const LaunchShape &ls, const float4 *pos,
                    const float4 *vel, float4 *rec, uint32_t *keys, hipStream_t st);
void launch_findcells(const SolverConsts &c, const DomMany &d, const LaunchShape &ls, const float4 *sorted,
                      int32_t *start, unsigned long long *end64, const uint32_t *epoch, hipStream_t st);
void launch_neighbourcells(const SolverConsts &c, const DomMany &d, const LaunchShape &ls, const float4 *sorted,
                           const int32_t *start, const unsigned long long *end64, int32_t *nbr9, hipStream_t st,
                           const DenseList *dl = nullptr);
void launch_lambda(const SolverConsts &c, const DomMany &d, const LaunchShape &ls, const float4 *rec,
                   const uint2 *q, const int32_t *nbr9, float *lambda, float4 *nrec, const PairList &pl, hipStream_t st,
                   const DenseList *dl = nullptr);
void launch_updatepos(const SolverConsts &c, const DomMany &d, const LaunchShape &ls, const float4 *rec_in,
                      const float *lambda, const float4 *nrec, const int32_t *nbr9, const PairList &pl, float4 *rec_out,
                      uint2 *q_out, hipStream_t st, const DenseList *dl = nullptr, const StepUpdate *up = nullptr);
void launch_update(const SolverConsts &c, const DomMany &d, const LaunchShape &ls, const float4 *rec, float4 *pos,
                   float4 *vel, float4 *vel_sorted, hipStream_t st);
void launch_vorticity_a(const SolverConsts &c, const DomMany &d, const LaunchShape &ls, const float4 *rec,
                        const uint2 *q, const int32_t *nbr9, const float4 *vel_sorted, const float4 *vel_id,
                        float4 *vxsph, float4 *omega, float *vort, float4 *nrec, const PairList &pl, hipStream_t st,
                        const DenseList *dl = nullptr);
void launch_vorticity_b(const SolverConsts &c, const DomMany &d, const LaunchShape &ls, const float4 *rec,
                        const int32_t *nbr9, const PairList &pl, const float4 *vxsph, const float4 *omega,
                        const float4 *nrec, float4 *vel_id, hipStream_t st, const DenseList *dl = nullptr);
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
// keys is clobbered. The permutation is written to perm (unless null), which
// must not alias the scratch buffers. With gather_dst, the last pass also
// writes gather_dst[k] = gather_src[perm[k]] (the sorted records), in place of
// a separate gather, and with q_dst, their quantized positions (scale qscale).
// With tiny_dst, the flag word of the sorted records (note_tiny).
void radix_sort_perm(RadixScratch &s, uint32_t *keys, int32_t n, int key_bits, uint32_t *perm, hipStream_t st,
                     const float4 *gather_src = nullptr, float4 *gather_dst = nullptr, uint2 *q_dst = nullptr,
                     float qscale = 0.0f, uint32_t *tiny_dst = nullptr);

// ------------------------------------------------------------ errors and helpers
void set_error(const std::string &msg);
int hip_fail(hipError_t e, const char *what);
int check_params(const pbf_params *p);
int key_bits_of(const int32_t g[3]);

inline int invalid(const std::string &msg) {
    set_error(msg);
    return PBF_ERR_INVALID;
}

inline bool valid_grid(int32_t gx, int32_t gy, int32_t gz) {
    return gx > 0 && gy > 0 && gz > 0 && (uint64_t)gx * gy * gz < (1ull << 31);
}

template <class T>
hipError_t dalloc(T **p, size_t count) {
    return hipMalloc((void **)p, (count ? count : 1) * sizeof(T));
}

template <class T>
void dfree(T *&p) {
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

}  // namespace pbf

#define PBF_HIP(call)                                            \
    do {                                                         \
        hipError_t e_ = (call);                                  \
        if (e_ != hipSuccess) return pbf::hip_fail(e_, #call);   \
    } while (0)
