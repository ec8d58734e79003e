// gfx950 kernels of the pbf simulation step (everything except the sort).
//
// Built with -ffp-contract=off: every float expression below is evaluated
// with IEEE round-to-nearest in the written order, and division / sqrtf are
// the correctly rounded sequences. The order of operations is the one the
// reference's GLSL has when compiled for llvmpipe (Mesa 23.2.1): dot products
// reduced from the last component, r*r of r = sqrt(sq) folded to sq,
// 315/(64 pi)*t*t*t evaluated as (t*t)*(t*315/(64 pi)), and acc + (a - b) in
// the vorticity sum as (acc + a) - b. The oracle restates the same order and
// is pinned bit-exact to llvmpipe runs of the reference's shaders.
//
// Layouts (HBM): records are float4 {x, y, z, id bits} (ParticleKey,
// predictpos.glsl:3-6); id-order state is float4 pos/vel with w = 0
// (src/SPH.cpp:119-133); the neighbour table is int32[9][n] (row-major over
// the reference's 9 (dy,dz) rows, each start + (count << 24)); the end table
// is uint64 {step tag, gid} so the newest step's highest gid wins.
#include <type_traits>

#include "pbf_internal.h"
#include "pbf_math.h"

namespace pbf {
namespace {

constexpr float POLY6 = 1.56668147106f;                // calclambda.glsl:46
constexpr float SPIKY_GRAD = -3.0f * 4.774648292756860f; // calclambda.glsl:63, folded in FP32
constexpr int BLOCK = 256;

__device__ __forceinline__ float sq_len(float x, float y, float z) { return (z * z + y * y) + x * x; }

// Wpoly6 (calclambda.glsl:41-47) of r = sqrt(sq) for a pair known to lie
// within h (r <= h: the callers test it first, or replay a list recorded at
// the same positions, which only holds such pairs), so the shader's
// `r > h` branch is not taken. H1: h == 1, where the reference's division by
// h^9 is a division by 1.
template <bool H1>
__device__ __forceinline__ float wpoly6_within(const SolverConsts &c, float sq) {
    const float t = c.hh - sq;
    const float w = (t * t) * (t * POLY6);
    return H1 ? w : w / c.h9;
}

// gradWspiky (calclambda.glsl:57-64) for l = sqrt(sq) with 0 < l <= h:
// (-3 * 15/pi * t * t) * r / (l * h*h*h*h*h*h), t = h - l; the three
// quotients correctly rounded (pbf_math.h). With h = 1 the denominator l is at
// most 1 and the numerators at most 14.33 (1 + 2^-22) in magnitude (each
// component of r within l up to rounding), inside the fast upper bounds.
// A component whose numerator is zero (the two positions share that
// coordinate) may come out +0 where the reference has -0: every use of a
// gradient component, alone or in a product, is added to a sum that starts
// at +0 and is never -0 (x + y is -0 in round-to-nearest only when both are
// -0), and adding a zero of either sign to such a sum leaves it unchanged, so
// no result depends on that sign.
template <bool H1>
__device__ __forceinline__ void grad_spiky(const SolverConsts &c, float dx, float dy, float dz, float l, float &gx,
                                           float &gy, float &gz, bool tame = false) {
    const float t = c.h - l;
    const float s = (SPIKY_GRAD * t) * t;
    const float den = H1 ? l : (((((l * c.h) * c.h) * c.h) * c.h) * c.h) * c.h;
    if (H1 && tame) {
        // every operand inside the fast ranges (positions_tame): no range tests
        const float y = rcp_rn_fast(den);
        gx = div_markstein<false>(s * dx, den, y);
        gy = div_markstein<false>(s * dy, den, y);
        gz = div_markstein<false>(s * dz, den, y);
        return;
    }
    div3_rn<H1, false>(s * dx, s * dy, s * dz, den, gx, gy, gz);
}

// The positions read by a launch are tame (h = 1 and no tiny record, see
// note_tiny): grad_spiky may skip the division's range tests.
template <bool H1>
__device__ __forceinline__ bool positions_tame(const uint32_t *flag) {
    return H1 && flag && *flag == 0u;
}

__device__ __forceinline__ float clamp_glsl(float x, float lo, float hi) {
    float m = x < lo ? lo : x;
    return m > hi ? hi : m;
}

// Clamp state of one axis for the adjoint's tape: 1 inside the walls, 8 at
// the upper wall, 0 at the lower one (shifted by the axis).
__device__ __forceinline__ uint32_t clamp_state(float x, float lo, float hi, int axis) {
    return (((x >= lo) & (x <= hi)) ? 1u : (x > hi ? 8u : 0u)) << axis;
}

// GetHash, counting.glsl:53-57: FP32 dot of the clamped integer cell.
__device__ __forceinline__ uint32_t cell_hash(const Seg &s, float x, float y, float z) {
    int32_t ix = (int32_t)clamp_glsl(x, 0.0f, (float)s.gx);
    int32_t iy = (int32_t)clamp_glsl(y, 0.0f, (float)s.gy);
    int32_t iz = (int32_t)clamp_glsl(z, 0.0f, (float)s.gz);
    const float w1 = (float)(s.gx * s.gz), w2 = (float)s.gx;
    float d = (float)ix * 1.0f + (float)iy * w1 + (float)iz * w2;
    return (uint32_t)d;
}

__device__ __forceinline__ int32_t id_of(const float4 &r) { return __float_as_int(r.w); }

// ----------------------------------------------------------------- predict
// predictpos.glsl:18-38 fused with the key of counting.glsl:70. Records carry
// the scene-local id; keys put the scene number above the cell key, so one
// stable sort orders every scene by its own key and keeps scenes apart.
template <class Dom>
__device__ __forceinline__ void predict_one(const SolverConsts &c, const Dom &dom, int32_t i,
                                            const float4 *__restrict__ pos, const float4 *__restrict__ vel,
                                            float4 *__restrict__ rec, uint32_t *__restrict__ keys) {
    const Seg sg = dom.at(i);
    const float4 p = pos[i];
    const float4 v = vel[i];
    float vx = v.x, vy = v.y, vz = v.z;
    if (c.extforce && p.z > (float)sg.gz / 2.0f) {
        const float g2 = 2.0f * c.gravity;
        vx = vx + (g2 * 0.0f) * c.dt;
        vy = vy + (g2 * 0.0f) * c.dt;
        vz = vz + (g2 * -1.0f) * c.dt;
    }
    vx = vx + (c.gravity * sg.gdir[0]) * c.dt;
    vy = vy + (c.gravity * sg.gdir[1]) * c.dt;
    vz = vz + (c.gravity * sg.gdir[2]) * c.dt;
    const float x = p.x + c.dt * vx, y = p.y + c.dt * vy, z = p.z + c.dt * vz;
    rec[i] = make_float4(x, y, z, __int_as_float(i - sg.off));
    if (keys) keys[i] = sg.key_hi | (cell_hash(sg, x, y, z) & sg.key_mask);
}

// Streaming kernels take any workgroup size and ppt particles per thread
// (particle i = workgroup base + k * blockDim + thread, k < ppt).
#define PBF_STREAM_FOR(i)                                                                     \
    for (int32_t k_ = 0, i = (int32_t)(blockIdx.x * blockDim.x) * ppt + (int32_t)threadIdx.x; \
         k_ < ppt && i < c.n; k_++, i += (int32_t)blockDim.x)

// hl (one scene, or null): the highlight words by id, of which the step keeps
// only bit 0 (clearhighlight.glsl:5-8, imageAtomicAnd (highlight[gid], 1) over
// every index; no other pass of the step runs in between, so a plain
// read-modify-write of one's own word is the same).
template <class Dom>
__global__ __launch_bounds__(1024) void k_predict(SolverConsts c, Dom dom, int32_t ppt, const float4 *__restrict__ pos,
                                                  const float4 *__restrict__ vel, float4 *__restrict__ rec,
                                                  uint32_t *__restrict__ keys, uint32_t *__restrict__ hl,
                                                  uint32_t *__restrict__ tiny, int32_t ntiny) {
    PBF_STREAM_FOR(i) {
        predict_one(c, dom, i, pos, vel, rec, keys);
        if (hl) hl[i] = hl[i] & 1u;
        for (int32_t k = i; k < ntiny; k += c.n) tiny[k] = 0u;  // the step's tiny-coordinate flags
    }
}

template <class Dom>
__global__ __launch_bounds__(BLOCK) void k_keys(SolverConsts c, Dom dom, const float4 *__restrict__ rec,
                                                uint32_t *__restrict__ keys) {
    const int32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= c.n) return;
    const Seg sg = dom.at(i);
    const float4 r = rec[i];
    keys[i] = sg.key_hi | (cell_hash(sg, r.x, r.y, r.z) & sg.key_mask);
}

// globalsort.glsl:64 as a gather: out[g] = in[perm[g]] (bit-preserving).
__global__ __launch_bounds__(BLOCK) void k_gather(int32_t n, const uint32_t *__restrict__ perm,
                                                  const float4 *__restrict__ in, float4 *__restrict__ out) {
    const int32_t g = blockIdx.x * BLOCK + threadIdx.x;
    if (g >= n) return;
    out[g] = in[perm[g]];
}

__global__ __launch_bounds__(BLOCK) void k_gather_by_id(int32_t n, const float4 *__restrict__ rec,
                                                        const float4 *__restrict__ src, float4 *__restrict__ out) {
    const int32_t g = blockIdx.x * BLOCK + threadIdx.x;
    if (g >= n) return;
    out[g] = src[id_of(rec[g])];
}

// Neighbour records of a replaying pass: (x, y, z of rec[g], v[g]).
__global__ __launch_bounds__(BLOCK) void k_pack_records(int32_t n, const float4 *__restrict__ rec,
                                                        const float *__restrict__ v, float4 *__restrict__ out) {
    const int32_t g = blockIdx.x * BLOCK + threadIdx.x;
    if (g >= n) return;
    const float4 r = rec[g];
    out[g] = make_float4(r.x, r.y, r.z, v[g]);
}

__global__ __launch_bounds__(BLOCK) void k_fill_i32(int32_t *__restrict__ p, int32_t v, size_t count) {
    const size_t stride = (size_t)gridDim.x * BLOCK * 4;
    for (size_t i = ((size_t)blockIdx.x * BLOCK + threadIdx.x) * 4; i < count; i += stride) {
        if (i + 4 <= count) {
            *reinterpret_cast<int4 *>(p + i) = make_int4(v, v, v, v);
        } else {
            for (size_t k = i; k < count; k++) p[k] = v;
        }
    }
}

// The start table cleared to -1, the tag of this step's end writes advanced
// by one (read by the cell finder that follows on the stream), and the dense
// level counts emptied.
__global__ __launch_bounds__(BLOCK) void k_begin_cells(int32_t *__restrict__ start, size_t count,
                                                       uint32_t *__restrict__ epoch, DenseList dl) {
    if (blockIdx.x == 0) {
        if (threadIdx.x == 0) epoch[0] = epoch[0] + 1u;
        if (dl.levels)
            for (int k = (int)threadIdx.x; k < kDenseLevels * kDenseCopies; k += BLOCK) dl.levels[k] = 0u;
    }
    const size_t stride = (size_t)gridDim.x * BLOCK * 4;
    for (size_t i = ((size_t)blockIdx.x * BLOCK + threadIdx.x) * 4; i < count; i += stride) {
        if (i + 4 <= count) {
            *reinterpret_cast<int4 *>(start + i) = make_int4(-1, -1, -1, -1);
        } else {
            for (size_t k = i; k < count; k++) start[k] = -1;
        }
    }
}

// ----------------------------------------------------------------- cells
__device__ __forceinline__ int64_t clamped_cell_index(const Seg &s, const float4 &r) {
    const int32_t x = (int32_t)clamp_glsl(r.x, 0.0f, (float)s.gx);
    const int32_t y = (int32_t)clamp_glsl(r.y, 0.0f, (float)s.gy);
    const int32_t z = (int32_t)clamp_glsl(r.z, 0.0f, (float)s.gz);
    if (x >= s.gx || y >= s.gy || z >= s.gz) return -1;  // imageStore out of range: dropped
    return (int64_t)s.cell_base + (int64_t)x + (int64_t)s.gx * ((int64_t)y + (int64_t)s.gy * z);
}

__device__ __forceinline__ bool same_cell(const Seg &s, const float4 &a, const float4 &b) {
    return (int32_t)clamp_glsl(a.x, 0.0f, (float)s.gx) == (int32_t)clamp_glsl(b.x, 0.0f, (float)s.gx) &&
           (int32_t)clamp_glsl(a.y, 0.0f, (float)s.gy) == (int32_t)clamp_glsl(b.y, 0.0f, (float)s.gy) &&
           (int32_t)clamp_glsl(a.z, 0.0f, (float)s.gz) == (int32_t)clamp_glsl(b.z, 0.0f, (float)s.gz);
}

// findcells.glsl:34-53. The reference writes in dispatch order, so with
// colliding keys the highest gid is the last writer: atomicMax reproduces it
// (start is cleared to -1 first; end carries the step tag in its high word).
// gid is the scene-local sorted index.
template <class Dom>
__device__ __forceinline__ void findcells_one(const SolverConsts &c, const Dom &dom, int32_t g,
                                              const float4 *__restrict__ sorted, int32_t *__restrict__ start,
                                              unsigned long long *__restrict__ end64, uint32_t ep) {
    const Seg sg = dom.at(g);
    const int32_t gl = g - sg.off;
    if (gl == 0) {
        atomicMax(&start[sg.cell_base], 0);  // imageStore (gridtexture, ivec3 (0, 0, 0), 0)
        return;
    }
    const float4 a = sorted[g], b = sorted[g - 1];
    if (same_cell(sg, a, b)) return;
    const int64_t ca = clamped_cell_index(sg, a), cb = clamped_cell_index(sg, b);
    if (ca >= 0) atomicMax(&start[ca], gl);
    if (cb >= 0) atomicMax(&end64[cb], ((unsigned long long)ep << 32) | (unsigned long long)(uint32_t)gl);
}

template <class Dom>
__global__ __launch_bounds__(1024) void k_findcells(SolverConsts c, Dom dom, int32_t ppt,
                                                    const float4 *__restrict__ sorted, int32_t *__restrict__ start,
                                                    unsigned long long *__restrict__ end64,
                                                    const uint32_t *__restrict__ epoch) {
    const uint32_t ep = epoch[0];
    PBF_STREAM_FOR(g) findcells_one(c, dom, g, sorted, start, end64, ep);
}

__device__ __forceinline__ bool in_grid(const Seg &s, int32_t x, int32_t y, int32_t z) {
    return x >= 0 && y >= 0 && z >= 0 && x < s.gx && y < s.gy && z < s.gz;
}

// neighbourcells.glsl:52-92: per particle, 9 (dy,dz) rows merging x-1..x+1;
// texel fetches outside the grid read 0 (as llvmpipe does).
template <class Dom>
__device__ __forceinline__ int32_t neighbourcells_one(const SolverConsts &c, const Dom &dom, int32_t i,
                                                      const float4 *__restrict__ sorted,
                                                      const int32_t *__restrict__ start,
                                                      const unsigned long long *__restrict__ end64,
                                                      int32_t *__restrict__ nbr9, const float4 &p,
                                                      int32_t words[9]) {
    int32_t cand = 0;
    const Seg sg = dom.at(i);
    const int32_t px = (int32_t)p.x, py = (int32_t)p.y, pz = (int32_t)p.z;  // ivec3 (pos): truncation
    // The 27 cells' table words in two rounds of loads in flight together
    // (a load under a per-lane condition is waited for before the next):
    // every start word (index 0 outside the grid), then the end words of the
    // cells with a start, the others reading one shared word.
    uint32_t idx[27];
    int32_t s[27];
#pragma unroll
    for (int q = 0; q < 27; q++) {
        const int32_t x = px + (q % 3 - 1), y = py + (q / 9 - 1), z = pz + ((q / 3) % 3 - 1);
        const bool inb = in_grid(sg, x, y, z);
        idx[q] = inb ? (uint32_t)sg.cell_base + (uint32_t)x + (uint32_t)sg.gx * ((uint32_t)y + (uint32_t)sg.gy * (uint32_t)z)
                     : 0xFFFFFFFFu;
        s[q] = start[idx[q] == 0xFFFFFFFFu ? 0u : idx[q]];
    }
    int32_t e[27];
#pragma unroll
    for (int q = 0; q < 27; q++) {
        const bool inb = idx[q] != 0xFFFFFFFFu;
        s[q] = inb ? s[q] : 0;
        const bool used = inb && s[q] != -1;
        e[q] = (int32_t)(uint32_t)(end64[used ? idx[q] : 0u] & 0xFFFFFFFFull);
        e[q] = used ? e[q] : 0;
    }
#pragma unroll
    for (int o = 0; o < 9; o++) {
        int32_t cell = -1, entries = 0;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const int q = o * 3 + k;  // (dy, dz) row o = (o / 3 - 1, o % 3 - 1), dx = k - 1
            if (cell == -1) cell = s[q];
            entries += s[q] != -1 ? e[q] - s[q] : 0;
        }
        const int32_t word = (int32_t)((uint32_t)cell + ((uint32_t)entries << 24));
        nbr9[(size_t)o * c.n + i] = word;
        words[o] = word;
        cand += max(word >> 24, 0);
    }
    return cand;
}

// highlight.glsl:16-30 for sorted slot i with record p and row words w: when
// the particle's word has bit 0 set, every candidate of its nine rows but its
// own slot gets bit 1 (no distance test); an index past the scene's n reads
// the zero record, whose id is 0. The pass only sets bit 1 and only tests bit
// 0, so it needs no ordering against itself; the clear ran in the predict pass.
__device__ __forceinline__ void highlight_one(const Seg &sg, int32_t i, const float4 &p, const int32_t w[9],
                                              const float4 *__restrict__ sorted, uint32_t *__restrict__ hl) {
    if ((hl[sg.off + id_of(p)] & 1u) != 1u) return;
    const int32_t li = i - sg.off;
    for (int o = 0; o < 9; o++) {
        const int32_t s0 = w[o] & 0xFFFFFF, e = w[o] >> 24;
        for (int32_t j = s0; j < s0 + e; j++) {
            if (j == li) continue;
            const int32_t jid = j < sg.n ? id_of(sorted[sg.off + j]) : 0;
            atomicOr(&hl[sg.off + jid], 2u);
        }
    }
}

// Candidate levels of the dense-particle selection.
__constant__ int32_t kDenseLevelValue[kDenseLevels] = {24, 32, 48, 64, 96, 128, 192, 256};

// Also keeps each particle's candidate count and counts the particles above
// each dense level: per workgroup in LDS, then one atomic per level into one
// of kDenseCopies words (few workgroups per word: atomics on one word
// serialize).
// hl (one scene, or null): also runs the highlight pass (highlight_one).
template <class Dom>
__global__ __launch_bounds__(1024) void k_neighbourcells(SolverConsts c, Dom dom, int32_t ppt,
                                                         const float4 *__restrict__ sorted,
                                                         const int32_t *__restrict__ start,
                                                         const unsigned long long *__restrict__ end64,
                                                         int32_t *__restrict__ nbr9, DenseList dl,
                                                         uint32_t *__restrict__ hl) {
    __shared__ uint32_t above[kDenseLevels];
    if (threadIdx.x < kDenseLevels) above[threadIdx.x] = 0u;
    __syncthreads();
    PBF_STREAM_FOR(i) {
        const float4 p = sorted[i];
        int32_t w[9];
        const int32_t cand = neighbourcells_one(c, dom, i, sorted, start, end64, nbr9, p, w);
        if (hl) highlight_one(dom.at(i), i, p, w, sorted, hl);
        if (dl.cand) {
            dl.cand[i] = (uint16_t)cand;  // at most 9 * 127
            const int lane = (int)(threadIdx.x & 63u);
            const int leader = __builtin_ctzll(__ballot(1));
#pragma unroll
            for (int l = 0; l < kDenseLevels; l++) {
                const uint64_t m = __ballot(cand > kDenseLevelValue[l]);
                if (lane == leader && m) atomicAdd(&above[l], (uint32_t)__popcll(m));
            }
        }
    }
    __syncthreads();
    if (dl.cand && threadIdx.x < kDenseLevels && above[threadIdx.x])
        atomicAdd(&dl.levels[threadIdx.x * kDenseCopies + blockIdx.x % kDenseCopies], above[threadIdx.x]);
}

// The step's dense threshold: the lowest level with at most kDenseBudget
// particles above it, else the highest level. One workgroup of
// kDenseSelectThreads sums the copies of each level.
constexpr int kDenseSelectThreads = 1024;
__global__ __launch_bounds__(kDenseSelectThreads) void k_dense_select(DenseList dl) {
    constexpr int per = kDenseSelectThreads / kDenseLevels;
    static_assert(kDenseCopies % per == 0, "copies split evenly over the threads of a level");
    __shared__ uint32_t part[kDenseLevels][per];
    __shared__ uint32_t above[kDenseLevels];
    const int t = threadIdx.x;
    const int l = t / per, k0 = t % per;
    uint32_t sum = 0;
    for (int k = k0; k < kDenseCopies; k += per) sum += dl.levels[l * kDenseCopies + k];
    part[l][k0] = sum;
    __syncthreads();
    if (t < kDenseLevels) {
        uint32_t tot = 0;
        for (int k = 0; k < per; k++) tot += part[t][k];
        above[t] = tot;
    }
    __syncthreads();
    if (t == 0) {
        int level = kDenseLevels - 1;
        for (int l = 0; l < kDenseLevels; l++) {
            if (above[l] <= kDenseBudget) {
                level = l;
                break;
            }
        }
        dl.threshold[0] = kDenseLevelValue[level];
        dl.threshold[1] = (int32_t)above[level];  // particles above it
    }
}

// ----------------------------------------------------------------- solver
// Neighbour records j (scene-local) of the packed ranges; indices past the
// scene's n (a wrapped count) read a zero record (all four components, so a
// packed operand in w reads 0 too), as the reference's robust buffer access
// does.
// PAD: the arrays hold at least kPad zero entries after the last particle
// (one scene), so indices past n read the zero record without a test; the
// index is non-negative and below 2^28, so its byte offset is formed in 32
// bits and the load takes the uniform base in scalar registers (no 64-bit
// address arithmetic per neighbour).
// Without padding the load is of a clamped index and the zero made after it
// by masking the bits (+0.0): a load under the condition would be a branch,
// which keeps the loads of a batch from being in flight together (and the
// compiler turns a select of the loaded value back into such a branch).
__device__ __forceinline__ float mask_bits(float v, int32_t m) { return __int_as_float(__float_as_int(v) & m); }
template <bool PAD>
__device__ __forceinline__ float4 rec_at(const float4 *__restrict__ rec, const Seg &sg, int32_t j) {
    if (PAD) return *reinterpret_cast<const float4 *>(reinterpret_cast<const char *>(rec) + (uint32_t)j * 16u);
    const float4 r = rec[sg.off + min(j, sg.n - 1)];
    const int32_t m = j < sg.n ? -1 : 0;
    return make_float4(mask_bits(r.x, m), mask_bits(r.y, m), mask_bits(r.z, m), mask_bits(r.w, m));
}

// Neighbour iteration in two phases. FOR_EACH_NEIGHBOUR (foreachneighbour.glsl:
// 1-10) visits the particles of 27 cells, of which about one in six lies
// within h. Evaluating the kernels under a per-candidate branch runs them for
// the whole wave on nearly every candidate (some lane is almost always in
// range), so each thread first scans its candidates with a cheap conservative
// distance test, kBatch loads in flight at a time (phase 1), and appends the
// survivors, in the reference's visiting order, to its own column of an LDS
// list as 16-bit (row, offset) codes; phase 2 runs the exact arithmetic over
// the list. A column that fills up is evaluated early and reused. Sums keep the
// reference's order, and candidates with fl(sqrt(sq)) > h add exact zeros
// there, so skipping them changes no bit of the result.
// LC entries per column: the LDS of a workgroup is (LC + 1) * 2 + 45 bytes
// per thread, and LDS is what limits the resident waves per CU. The scanning
// passes (lambda, vorticity A) keep LCAP; the replaying passes (position
// update, vorticity B) scan only for particles whose pair list overflowed and
// keep kReplayLcap, so more of their waves fit on a CU.
#ifndef PBF_LIST_LDS
#define PBF_LIST_LDS 56
#endif
#ifndef PBF_REPLAY_LDS
#define PBF_REPLAY_LDS 16
#endif
constexpr int LCAP = PBF_LIST_LDS;
constexpr int kReplayLcap = PBF_REPLAY_LDS;
static_assert(LCAP % 8 == 0 && kReplayLcap % 8 == 0, "list column of whole batches");
// Candidates (and listed neighbours) whose loads are issued together: a
// particle in a dense clump walks hundreds of them serially, and its wave
// alone can set the kernel's duration, so memory-level parallelism per lane
// matters more than registers here.
#ifndef PBF_BATCH
#define PBF_BATCH 8
#endif
constexpr int kBatch = PBF_BATCH;
static_assert(LCAP >= 2 * kBatch && kReplayLcap >= 2 * kBatch, "list column must hold at least two batches");
static_assert(kBatch % 2 == 0, "codes are written in pairs");
static_assert(kListCap + kBatch <= kListRows, "a replay's next-batch rows stay inside the list");
static_assert(126 + kBatch <= kPad, "batches past a row stay inside the zero padding");

template <int NB, int LC = LCAP>
struct NbrLds {
    static constexpr int kCap = LC;
    uint16_t list[(LC + 1) * NB];  // row LC: sink of the branch-free appends
    int32_t rowstart[9 * NB];
    int8_t rowcount[9 * NB];  // start + (count << 24) >> 24: in [-128, 127]
};

// True when the candidate may satisfy sq <= h2cut: the fused multiply-add
// estimate is compared with h2loose = h2cut (1 + 2^-20), which bounds its error.
__device__ __forceinline__ bool maybe_near(const SolverConsts &c, const float4 &pi, const float4 &pj) {
    const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
    return fmaf(dx, dx, fmaf(dy, dy, dz * dz)) <= c.h2loose;
}

// True when the candidate may satisfy sq <= h2cut, from quantized positions
// (quantize_pos): each axis difference modulo 2^16 as a signed 16-bit value,
// the squares summed by two saturating dot products. A pair within h differs
// by at most h qscale (1 + 1e-5) < 4097 units per axis, so its differences
// are exact up to the truncations (under one unit per coordinate: within
// 2 sqrt(3) of the scaled distance), and qcut = ceil((h qscale (1 + 2e-5)
// + 4)^2) keeps it. Far pairs may alias to near ones and are rejected by the
// exact test.
typedef short short2_v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ bool maybe_near_q(const SolverConsts &c, const uint2 &qi, const uint2 &qj) {
    const short2_v dxy = __builtin_bit_cast(short2_v, qi.x) - __builtin_bit_cast(short2_v, qj.x);
    const short2_v dz = __builtin_bit_cast(short2_v, qi.y) - __builtin_bit_cast(short2_v, qj.y);
    return __builtin_amdgcn_sdot2(dxy, dxy, __builtin_amdgcn_sdot2(dz, dz, 0, true), true) <= c.qcut;
}

// Row word o of particle i: start + (count << 24). The count is the
// reference's wrapped one, read back with an arithmetic shift.
__device__ __forceinline__ int32_t row_count(int32_t word) { return word >> 24; }
__device__ __forceinline__ int32_t row_start(int32_t word) { return word & 0xFFFFFF; }

// The nine row words of slot i, loaded together (one round trip rather than
// one per row ahead of that row's candidates) into the thread's LDS columns;
// returns the number of candidates (the sum of the positive counts).
template <int NB, class L>
__device__ __forceinline__ int32_t load_rows(const SolverConsts &c, const int32_t *__restrict__ nbr9, int32_t i,
                                             L &lds) {
    const int t = threadIdx.x;
    int32_t d[9];
#pragma unroll
    for (int o = 0; o < 9; o++) d[o] = nbr9[(size_t)o * c.n + i];
    int32_t cand = 0;
#pragma unroll
    for (int o = 0; o < 9; o++) {
        lds.rowstart[o * NB + t] = row_start(d[o]);
        lds.rowcount[o * NB + t] = (int8_t)row_count(d[o]);
        cand += max(row_count(d[o]), 0);
    }
    return cand;
}

// Full candidate scan of global slot i (scene-local index li) over the rows
// that load_rows put in LDS; body(j, code, rec[j], extra(j)) for every
// candidate j (scene-local) that may lie within h, in visiting order. Appends
// go to the slot after the last kept one whatever the test gives, and the
// count only advances for kept candidates. The kept candidates are evaluated
// kBatch at a time with their loads issued together, so a particle with
// hundreds of neighbours waits for one round trip per batch rather than per
// neighbour.
// QUANT: the distance test reads the quantized positions qpos (8 bytes per
// candidate instead of 16, four integer operations instead of six float
// ones); else the records themselves.
template <int NB, bool PAD, bool QUANT, class L, class Extra, class Body>
__device__ __forceinline__ void scan_neighbours(const SolverConsts &c, const Seg &sg, const float4 *__restrict__ rec,
                                                const uint2 *__restrict__ qpos, int32_t i, const float4 &pi, L &lds,
                                                Extra &&extra, Body &&body) {
    constexpr int LC = L::kCap;
    const int t = threadIdx.x;
    const int32_t li = i - sg.off;
    const float4 *__restrict__ srec = rec + sg.off;
    const uint2 *__restrict__ sqpos = QUANT ? qpos + sg.off : nullptr;
    uint2 qi = make_uint2(0u, 0u);
    if constexpr (QUANT) qi = qpos[i];
    using X = decltype(extra(0));
    int cnt = 0;
    auto flush = [&]() {
        for (int k = 0; k < cnt; k += kBatch) {
            uint32_t e[kBatch];
            int32_t j[kBatch];
            float4 pj[kBatch];
            X x[kBatch];
#pragma unroll
            for (int u = 0; u < kBatch; u++) e[u] = lds.list[min(k + u, cnt - 1) * NB + t];
#pragma unroll
            for (int u = 0; u < kBatch; u++) j[u] = lds.rowstart[(e[u] >> 8) * NB + t] + (int32_t)(e[u] & 255u);
#pragma unroll
            for (int u = 0; u < kBatch; u++) {
                pj[u] = rec_at<PAD>(rec, sg, j[u]);
                x[u] = extra(j[u]);
            }
            // every load of the batch is issued before any arithmetic on them
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int u = 0; u < kBatch; u++)
                if (k + u < cnt) body(j[u], e[u], pj[u], x[u]);
        }
        cnt = 0;
    };
    // Appends go through boff, the byte offset in the list of the column's
    // next entry (2 (cnt NB + t)), advanced by 2 NB per kept candidate; the
    // codes of two consecutive candidates share one register (low and high
    // halves), since the code is the same in every lane.
    char *const lbase = reinterpret_cast<char *>(lds.list);
    auto scan_row = [&](int o, int32_t s0, int32_t e, int32_t selfq, auto check_self) {
        constexpr bool SELF = decltype(check_self)::value;
        // Loads stay in bounds: with PAD, a row's batches end at most
        // 126 + kBatch entries past n (start < n, count < 128), inside the
        // zero padding; without, indices past the row or past the scene
        // repeat the last valid one, and candidates past the scene are the
        // zero record.
        const int32_t last = max(min(s0 + e, sg.n) - 1, 0);
        const float4 *__restrict__ row = srec + s0;
        const uint2 *__restrict__ qrow = QUANT ? sqpos + s0 : nullptr;
        uint32_t boff = 2u * (uint32_t)(cnt * NB + t);
        const uint32_t bfull = 2u * (uint32_t)((LC - kBatch) * NB + t);
        for (int32_t q = 0; q < e; q += kBatch) {
            // All lanes flush together when one column is full: a per-lane
            // branch would run the flush once for every subset of lanes that
            // fills up at a different candidate, which in a dense clump is
            // most of the wave's time. Each lane still evaluates its own list
            // in order.
            if (wave_any(boff > bfull)) {
                cnt = (int)(boff / 2u - (uint32_t)t) / NB;
                flush();
                boff = 2u * (uint32_t)t;
            }
            const int32_t j0 = s0 + q;
            float4 p[kBatch];
            uint2 pq[kBatch];
#pragma unroll
            for (int u = 0; u < kBatch; u++) {
                if constexpr (QUANT)
                    pq[u] = PAD ? qrow[q + u] : sqpos[min(j0 + u, last)];
                else
                    p[u] = PAD ? row[q + u] : srec[min(j0 + u, last)];
            }
            const uint32_t code0 = (uint32_t)((o << 8) | q);
#pragma unroll
            for (int u = 0; u < kBatch; u += 2) {
                const uint32_t two_s = (code0 + (uint32_t)u) | ((code0 + (uint32_t)u + 1u) << 16);
                // one register move for both codes; the high half is stored
                // from the same register (ds_write_b16_d16_hi)
                uint32_t two;
                asm("v_mov_b32 %0, %1" : "=v"(two) : "s"(two_s));
#pragma unroll
                for (int v = 0; v < 2; v++) {
                    // past the scene: the zero record, whose quantized form is zero
                    const bool inside = PAD || j0 + u + v < sg.n;
                    bool near;
                    if constexpr (QUANT)
                        near = maybe_near_q(c, qi, inside ? pq[u + v] : make_uint2(0u, 0u));
                    else
                        near = maybe_near(c, pi, inside ? p[u + v] : make_float4(0.0f, 0.0f, 0.0f, 0.0f));
                    bool f = (q + u + v < e) & near;
                    if (SELF) f = f & (q + u + v != selfq);
                    *reinterpret_cast<uint16_t *>(lbase + boff) = (uint16_t)(v ? two >> 16 : two);
                    boff += f ? 2u * NB : 0u;
                }
            }
        }
        cnt = (int)(boff / 2u - (uint32_t)t) / NB;
    };
    for (int o = 0; o < 9; o++) {
        const int32_t s0 = lds.rowstart[o * NB + t], e = (int32_t)lds.rowcount[o * NB + t];
        // The particle's own slot, when this row holds it (the centre row,
        // or any row in the reference's boundary cases): the one candidate
        // skipped besides the out-of-range ones. Rows without it in any lane
        // of the wave skip the test.
        const int32_t selfq = (li >= s0 && li - s0 < e) ? li - s0 : -1;
        if (wave_any(selfq >= 0))
            scan_row(o, s0, e, selfq, std::true_type());
        else
            scan_row(o, s0, e, selfq, std::false_type());
    }
    flush();
}

// Per-particle list of the neighbours within h at one set of positions, in
// visiting order, as (row, offset) codes: entries[k * n + i], k < count[i].
// The lambda pass (and vorticity pass A) writes it while scanning; the
// position update (and vorticity pass B), which reads the same positions,
// replays it instead of scanning. count LIST_NONE: more than cap neighbours,
// the replaying pass scans.
constexpr uint16_t LIST_NONE = kListNone;

// The next entry's address is kept and advanced by n per entry (an add, not
// a 64-bit multiply per entry).
struct PairListOut {
    uint16_t *entries;
    uint16_t *count;
    int32_t cap;
    int32_t cnt;
    uint16_t *next;  // &entries[cnt * n + i]: entries + i to begin with
    __device__ __forceinline__ void add(int32_t, int32_t n, uint32_t code) {
        if (entries) {
            if (cnt < cap) *next = (uint16_t)code;
            next += n;
            cnt++;
        }
    }
    __device__ __forceinline__ void finish(int32_t i) {
        if (entries) count[i] = cnt <= cap ? (uint16_t)cnt : LIST_NONE;
    }
};

// body(j, rec[j], extra(j), listed) for i's neighbours within h (listed: the
// recording pass saw these exact positions, so the range tests are known to
// pass) or for candidates that may be within h (scanned); j is scene-local;
// extra(j) gathers any further per-neighbour operand, so that on the list path
// the loads of a batch of entries are in flight together. The position update
// and vorticity B pass the packed neighbour records as rec (the operand in w)
// and no extra.
template <int NB, bool PAD, class L, class Extra, class Body>
__device__ __forceinline__ void replay_neighbours(const SolverConsts &c, const Seg &sg, const float4 *__restrict__ rec,
                                                  const uint16_t *__restrict__ entries,
                                                  const uint16_t *__restrict__ count, int32_t i, const float4 &pi,
                                                  L &lds, Extra &&extra, Body &&body) {
    const int t = threadIdx.x;
    const uint32_t cnt = entries ? count[i] : LIST_NONE;
    if (cnt == LIST_NONE) {
        scan_neighbours<NB, PAD, false>(c, sg, rec, nullptr, i, pi, lds, extra,
                                        [&](int32_t j, uint32_t, const float4 &pj, auto x) { body(j, pj, x, false); });
        return;
    }
    using X = decltype(extra(0));
    // The entries of the next batch are requested before this batch's
    // neighbours, so each batch waits for one round trip rather than two.
    // Every load is unconditional, so that a batch's loads are in flight
    // together (a load under a per-lane condition is a branch, and the
    // compiler waits for each before the next): the rows of the next batch
    // are requested whether they hold entries or not (the lists have
    // kListRows >= kListCap + kBatch rows), and codes past the list are
    // replaced by the batch's first, so every code and neighbour index is a
    // valid one. The row address advances by n per row (no multiply).
    const size_t stride = (size_t)c.n;
    const uint16_t *__restrict__ rp = entries + i;  // row k of the list
    uint32_t next[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; u++) next[u] = rp[u * stride];
    for (uint32_t k = 0; k < cnt; k += kBatch) {
        uint32_t e[kBatch];
        int32_t j[kBatch];
        float4 pj[kBatch];
        X x[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; u++) e[u] = k + u < cnt ? next[u] : next[0];
        rp += kBatch * stride;
#pragma unroll
        for (int u = 0; u < kBatch; u++) next[u] = rp[u * stride];
#pragma unroll
        for (int u = 0; u < kBatch; u++) j[u] = lds.rowstart[(e[u] >> 8) * NB + t] + (int32_t)(e[u] & 255u);
#pragma unroll
        for (int u = 0; u < kBatch; u++) {
            pj[u] = rec_at<PAD>(rec, sg, j[u]);
            x[u] = extra(j[u]);
        }
#pragma unroll
        for (int u = 0; u < kBatch; u++)
            if (k + u < cnt) body(j[u], pj[u], x[u], true);
    }
}

// ----------------------------------------------------------------- dense particles
// A particle in a dense clump has hundreds of candidates, and its thread walks
// them serially: one such wave, running alone at the end of a launch, sets the
// launch's duration (tens of microseconds while the rest of the launch takes
// about ten). The step's particles with the most candidates (DenseList) are
// skipped by the one-thread path, and the first workgroups of the same launch
// give each a group of kDenseGroup lanes, so that they run beside the bulk
// rather than after it: the group walks the particle's
// candidates in visiting order, kDenseSlots per round (rows flattened, so short
// rows do not cost a round each), evaluates the pair terms in parallel with the
// same arithmetic as the one-thread passes, and one lane per accumulator then
// adds the round's in-range terms in visiting order. Every sum is the same
// sequence of the same operations, so results equal the one-thread passes bit
// for bit. The pair lists are recorded and replayed in the same format.
#ifndef PBF_DENSE_GROUP
#define PBF_DENSE_GROUP 16
#endif
constexpr int kDenseGroup = PBF_DENSE_GROUP;
static_assert(kDenseGroup == 8 || kDenseGroup == 16, "dense group width (a round's slots fit a 64-bit mask)");
constexpr int kDenseBatch = 4;
constexpr int kDenseSlots = kDenseGroup * kDenseBatch;
constexpr int kDenseGroups = 64 / kDenseGroup;
// Waves of a neighbour launch that work through the dense particles (in
// whole workgroups, a multiple of 8 of them, so that the rest keep their XCD
// round robin).
#ifndef PBF_DENSE_WAVES
#define PBF_DENSE_WAVES 2048
#endif
constexpr uint32_t kDenseWaves = PBF_DENSE_WAVES;

// Per group (W waves of kDenseGroups groups). Terms are stored
// accumulator-major, so an accumulator's lane reads a round's terms four at a
// time.
template <int NP, int NM, int W>
struct DenseLds {
    int32_t s0[W * kDenseGroups][9];
    int32_t cum[W * kDenseGroups][10];
    int32_t total[W * kDenseGroups];
    float4 p[W * kDenseGroups][NP][kDenseSlots / 4];
    float4 m[W * kDenseGroups][NM > 0 ? NM : 1][kDenseSlots / 4];
    float acc[W * kDenseGroups][NP];
};

// Orders the LDS accesses of one wave's lanes (the dense groups of a wave
// exchange data through LDS; waves of a workgroup do not wait for each other).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Group-local bits of a wave ballot.
__device__ __forceinline__ uint64_t group_bits(uint64_t ballot, int g) {
    return (ballot >> (kDenseGroup * g)) & ((1ull << kDenseGroup) - 1ull);
}

// Drives one pass over the dense particles with workgroup wg of wgs, each of
// W waves. Pass
// supplies NP accumulators (the last NM of them of the form
// acc = (acc + p) - m, the rest acc + p), RECORD (the pass scans and records the pair list) or replays
// the recorded list, begin(i, sg), extra(sg, j), eval(pj, x, listed, P, M)
// -> in range, and finish(i, sg, acc). Waves synchronize on their own, and
// every synchronization is reached by all lanes of the wave.
template <class Pass, bool PAD, int W, class Dom>
__device__ __forceinline__ void dense_pass(const SolverConsts &c, const Dom &dom, const float4 *__restrict__ rec,
                                           const int32_t *__restrict__ nbr9, const DenseList &dl,
                                           uint16_t *__restrict__ list_entries, uint16_t *__restrict__ list_count,
                                           int32_t list_cap, Pass &ps, DenseLds<Pass::NP, Pass::NM, W> &lds,
                                           uint32_t wg, uint32_t wgs) {
    constexpr int NP = Pass::NP, NM = Pass::NM;
    const int lane = (int)(threadIdx.x & 63u);
    const int wave = (int)(threadIdx.x >> 6);
    const int lg = lane % kDenseGroup;
    const int g = wave * kDenseGroups + lane / kDenseGroup;  // this lane's group in the workgroup
    const int g0 = wave * kDenseGroups;                      // the wave's first group
    const uint32_t block = wg * W + (uint32_t)wave, blocks = wgs * W;
    const int32_t th = dl.threshold[0];
    if (dl.threshold[1] == 0) return;  // no particle above the threshold
    // Each wave works alone. Its share: particles block, block + blocks, ...
    // (clumps, contiguous in the sorted order, spread over all dense waves),
    // examined 64 at a time; its groups take the dense ones four at a time.
    for (uint64_t m0 = 0; m0 * blocks < (uint64_t)c.n; m0 += 64) {
        const uint64_t ii = (uint64_t)block + (uint64_t)blocks * (m0 + (uint64_t)lane);
        uint64_t pending = __ballot(ii < (uint64_t)c.n && (int32_t)dl.cand[ii] > th);
        while (pending) {
            uint64_t mine = pending;
            for (int q = 0; q < g - g0; q++) mine &= mine - 1ull;
            for (int q = 0; q < kDenseGroups; q++) pending &= pending - 1ull;
            const bool act = mine != 0ull;
            const uint64_t slot = m0 + (uint64_t)(act ? __builtin_ctzll(mine) : 0);
            const int32_t i = act ? (int32_t)((uint64_t)block + (uint64_t)blocks * slot) : 0;
            const Seg sg = dom.at(i);
            const int32_t li = i - sg.off;
            for (int o = lg; o < 9; o += kDenseGroup) {
                const int32_t d = act ? nbr9[(size_t)o * c.n + i] : -1;  // -1: no candidates
                lds.s0[g][o] = row_start(d);
                lds.cum[g][o + 1] = max(row_count(d), 0);
            }
            wave_sync();
            if (lg == 0) {
                int32_t run = 0;
                lds.cum[g][0] = 0;
                for (int o = 0; o < 9; o++) {
                    run += lds.cum[g][o + 1];
                    lds.cum[g][o + 1] = run;
                }
            }
            wave_sync();
            uint32_t lc = LIST_NONE;
            if (!Pass::RECORD && act && list_entries) lc = list_count[i];
            const bool listed = !Pass::RECORD && lc != LIST_NONE;
            if (lg == 0) lds.total[g] = !act ? 0 : (listed ? (int32_t)lc : lds.cum[g][9]);
            wave_sync();
            const int32_t cn = lds.total[g];
            int32_t rounds = 0;
#pragma unroll
            for (int q = 0; q < kDenseGroups; q++)
                rounds = max(rounds, (lds.total[g0 + q] + kDenseSlots - 1) / kDenseSlots);
            ps.begin(i, sg);
            float acc = 0.0f;
            int32_t kept = 0;
            int o = 0;
            using X = decltype(ps.extra(sg, 0));
            for (int r = 0; r < rounds; r++) {
                int32_t j[kDenseBatch];
                uint32_t code[kDenseBatch];
                bool valid[kDenseBatch];
#pragma unroll
                for (int b = 0; b < kDenseBatch; b++) {
                    const int32_t k = r * kDenseSlots + b * kDenseGroup + lg;
                    valid[b] = k < cn;
                    if (listed) {
                        // unconditional load (a listed group's particle is real;
                        // the index stays inside its list), selected after
                        const uint32_t raw = list_entries[(size_t)max(min(k, cn - 1), 0) * c.n + i];
                        code[b] = valid[b] ? raw : 0u;
                        j[b] = lds.s0[g][code[b] >> 8] + (int32_t)(code[b] & 255u);
                    } else {
                        if (valid[b])
                            while (o < 8 && k >= lds.cum[g][o + 1]) o++;
                        const int32_t q = k - lds.cum[g][o];
                        j[b] = lds.s0[g][o] + q;
                        code[b] = ((uint32_t)o << 8) | (uint32_t)(q & 255);
                        valid[b] = valid[b] & (j[b] != li);
                    }
                }
                float4 pj[kDenseBatch];
                X x[kDenseBatch];
#pragma unroll
                for (int b = 0; b < kDenseBatch; b++) {
                    // invalid slots load the particle itself (a valid index), so
                    // that the loads are unconditional; only valid slots are used
                    const int32_t jl = valid[b] ? j[b] : li;
                    pj[b] = rec_at<PAD>(rec, sg, jl);
                    x[b] = ps.extra(sg, jl);
                }
                uint64_t mask = 0;
                bool in[kDenseBatch];
#pragma unroll
                for (int b = 0; b < kDenseBatch; b++) {
                    float P[NP], M[NM > 0 ? NM : 1];  // M[a]: accumulator NP - NM + a
#pragma unroll
                    for (int a = 0; a < NP; a++) P[a] = 0.0f;
#pragma unroll
                    for (int a = 0; a < (NM > 0 ? NM : 1); a++) M[a] = 0.0f;
                    in[b] = valid[b] && ps.eval(pj[b], x[b], listed, P, M);
                    const int slot = b * kDenseGroup + lg;
#pragma unroll
                    for (int a = 0; a < NP; a++)
                        reinterpret_cast<float *>(lds.p[g][a])[slot] = in[b] ? P[a] : 0.0f;
                    if (NM > 0) {
#pragma unroll
                        for (int a = 0; a < NM; a++)
                            reinterpret_cast<float *>(lds.m[g][a])[slot] = in[b] ? M[a] : 0.0f;
                    }
                    mask |= group_bits(__ballot(in[b]), g - g0) << (kDenseGroup * b);
                }
                if (Pass::RECORD && list_entries) {
#pragma unroll
                    for (int b = 0; b < kDenseBatch; b++) {
                        const int slot = b * kDenseGroup + lg;
                        const int32_t pos = kept + __popcll(mask & ((1ull << slot) - 1ull));
                        if (in[b] && pos < list_cap) list_entries[(size_t)pos * c.n + i] = (uint16_t)code[b];
                    }
                }
                wave_sync();
                // Slots without an in-range pair hold +0.0, and an accumulator
                // that starts at +0.0 is never -0.0 (round to nearest), so adding
                // them changes nothing: every slot up to the round's last
                // candidate is added in order.
                const int used = min(kDenseSlots, cn - r * kDenseSlots);
                if (lg < NP && used > 0) {
                    for (int q = 0; q < (used + 3) / 4; q++) {
                        const float4 v = lds.p[g][lg][q];
                        if (NM > 0) {
                            // the first NP - NM accumulators subtract +0.0
                            const float4 w0 = lds.m[g][max(lg - (NP - NM), 0)][q];
                            const bool has_m = lg >= NP - NM;
                            const float4 w = has_m ? w0 : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                            acc = (acc + v.x) - w.x;
                            acc = (acc + v.y) - w.y;
                            acc = (acc + v.z) - w.z;
                            acc = (acc + v.w) - w.w;
                        } else {
                            acc = acc + v.x;
                            acc = acc + v.y;
                            acc = acc + v.z;
                            acc = acc + v.w;
                        }
                    }
                }
                kept += __popcll(mask);
                wave_sync();
            }
            if (lg < NP) lds.acc[g][lg] = acc;
            wave_sync();
            if (act && lg == 0) {
                ps.finish(i, sg, lds.acc[g]);
                if (Pass::RECORD && list_entries) list_count[i] = kept <= list_cap ? (uint16_t)kept : LIST_NONE;
            }
            wave_sync();
        }
    }
}

// calclambda.glsl:66-103 for the dense particles (terms of k_lambda).
template <bool H1, bool PAD>
struct LambdaPass {
    static constexpr int NP = 5, NM = 0;
    static constexpr bool RECORD = true;
    const SolverConsts &c;
    const float4 *__restrict__ rec;
    float *__restrict__ lambda;
    float4 *__restrict__ nrec;  // neighbour records of the position update
    SolverTape tp;
    bool tame;  // positions_tame: the division's range tests can be skipped
    float4 pi;
    __device__ __forceinline__ void begin(int32_t i, const Seg &) { pi = rec[i]; }
    __device__ __forceinline__ int extra(const Seg &, int32_t) const { return 0; }
    __device__ __forceinline__ bool eval(const float4 &pj, int, bool, float *P, float *) const {
        const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        const float sq = sq_len(dx, dy, dz);
        if (sq > c.h2cut) return false;
        const float r = sqrt_rn_dist(sq, tame);
        if (r > c.h) return false;
        P[0] = wpoly6_within<H1>(c, sq);
        if (r == 0.0f) return true;
        float gx, gy, gz;
        grad_spiky<H1>(c, dx, dy, dz, r, gx, gy, gz, tame);
        gx = gx * c.inv_rho0;
        gy = gy * c.inv_rho0;
        gz = gz * c.inv_rho0;
        P[1] = sq_len(gx, gy, gz);
        P[2] = gx;
        P[3] = gy;
        P[4] = gz;
        return true;
    }
    __device__ __forceinline__ void finish(int32_t i, const Seg &, const float *a) const {
        const float sk = a[1] + sq_len(a[2], a[3], a[4]);
        const float ci = a[0] * c.inv_rho0 - 1.0f;
        const float l = -ci / (sk + c.eps);
        lambda[i] = l;
        nrec[i] = make_float4(pi.x, pi.y, pi.z, l);
        if (tp.sden) {
            tp.sden[i] = sk + c.eps;
            tp.G[i] = make_float4(a[2], a[3], a[4], 0.0f);
        }
    }
};

// updatepos.glsl:42-105 for the dense particles (terms of k_updatepos).
// update.glsl:16-28 for sorted slot i with final record r (scene sg).
__device__ __forceinline__ void update_one(const SolverConsts &c, const Seg &sg, int32_t i, const float4 &r,
                                           const StepUpdate &up) {
    const int32_t id = sg.off + id_of(r);
    const float4 old = up.pos[id];
    const float4 v = make_float4((r.x - old.x) / c.dt, (r.y - old.y) / c.dt, (r.z - old.z) / c.dt, 0.0f);
    up.pos[id] = make_float4(r.x, r.y, r.z, 0.0f);
    up.vel[id] = v;
    if (up.vel_sorted) up.vel_sorted[i] = v;
}

template <bool H1, bool PAD>
struct UpdatePosPass {
    static constexpr int NP = 3, NM = 0;
    static constexpr bool RECORD = false;
    const SolverConsts &c;
    const float4 *__restrict__ rec_in;
    const float *__restrict__ lambda;
    float4 *__restrict__ rec_out;
    uint2 *__restrict__ q_out;  // quantized positions of rec_out, or none
    StepUpdate up;  // the step's closing update (last iteration), or none
    SolverTape tp;
    bool tame;                  // positions_tame of rec_in
    uint32_t *__restrict__ tiny_out;  // note_tiny flag of rec_out, or none
    float4 pi;
    float li;
    __device__ __forceinline__ void begin(int32_t i, const Seg &) {
        pi = rec_in[i];
        li = lambda[i];
    }
    // neighbours come from the lambda pass's records: lambda_j in pj.w
    __device__ __forceinline__ int extra(const Seg &, int32_t) const { return 0; }
    __device__ __forceinline__ bool eval(const float4 &pj, int, bool listed, float *P, float *) const {
        const float lj = pj.w;
        const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        const float sq = sq_len(dx, dy, dz);
        if (!listed && sq > c.h2cut) return false;
        const float r = sqrt_rn_dist(sq, tame);
        if ((!listed && r > c.h) || r == 0.0f) return false;
        float sc = c.scale * wpoly6_within<H1>(c, sq);  // r <= h here, listed or tested
        sc *= sc;
        sc *= sc;
        sc = -c.k * sc;
        float gx, gy, gz;
        grad_spiky<H1>(c, dx, dy, dz, r, gx, gy, gz, tame);
        const float f = li + lj + sc;
        P[0] = f * gx;
        P[1] = f * gy;
        P[2] = f * gz;
        return true;
    }
    __device__ __forceinline__ void finish(int32_t i, const Seg &sg, const float *a) const {
        float x = pi.x + c.inv_rho0 * a[0], y = pi.y + c.inv_rho0 * a[1], z = pi.z + c.inv_rho0 * a[2];
        if (tp.clamp)
            tp.clamp[i] = (uint8_t)(clamp_state(x, sg.lo[0], sg.hi[0], 0) | clamp_state(y, sg.lo[1], sg.hi[1], 1) |
                                    clamp_state(z, sg.lo[2], sg.hi[2], 2));
        x = clamp_glsl(x, sg.lo[0], sg.hi[0]);
        y = clamp_glsl(y, sg.lo[1], sg.hi[1]);
        z = clamp_glsl(z, sg.lo[2], sg.hi[2]);
        const float4 r = make_float4(x, y, z, pi.w);
        rec_out[i] = r;
        if (q_out) q_out[i] = quantize_pos(r, c.qscale);
        note_tiny(tiny_out, r);
        if (up.pos) update_one(c, sg, i, r, up);
    }
};

// vorticity.glsl:34-64 for the dense particles (terms of k_vorticity_a).
template <bool H1, bool PAD>
struct VorticityAPass {
    static constexpr int NP = 6, NM = 3;
    static constexpr bool RECORD = true;
    const SolverConsts &c;
    const float4 *__restrict__ rec;
    const float4 *__restrict__ vel_sorted;
    const float4 *__restrict__ vel_id;
    float4 *__restrict__ vxsph;
    float4 *__restrict__ omega;
    float *__restrict__ vort;   // |omega| by sorted slot, or none
    float4 *__restrict__ nrec;  // neighbour records of the confinement pass
    bool tame;
    float4 pi, vi;
    __device__ __forceinline__ void begin(int32_t i, const Seg &) {
        pi = rec[i];
        vi = vel_sorted[i];
    }
    __device__ __forceinline__ float4 extra(const Seg &sg, int32_t j) const {
        return *(j < sg.n ? vel_sorted + sg.off + j : vel_id + sg.off);  // one load of either
    }
    __device__ __forceinline__ bool eval(const float4 &pj, const float4 &vj, bool, float *P, float *M) const {
        const float px = pi.x - pj.x, py = pi.y - pj.y, pz = pi.z - pj.z;
        const float sq = sq_len(px, py, pz);
        if (sq > c.h2cut) return false;
        const float r = sqrt_rn_dist(sq, tame);
        if (r > c.h) return false;
        const float ux = vj.x - vi.x, uy = vj.y - vi.y, uz = vj.z - vi.z;
        const float w = wpoly6_within<H1>(c, sq);
        P[0] = ux * w;
        P[1] = uy * w;
        P[2] = uz * w;
        if (r == 0.0f) return true;
        float gx, gy, gz;
        grad_spiky<H1>(c, px, py, pz, r, gx, gy, gz, tame);
        P[3] = uy * gz;
        M[0] = uz * gy;
        P[4] = uz * gx;
        M[1] = ux * gz;
        P[5] = ux * gy;
        M[2] = uy * gx;
        return true;
    }
    __device__ __forceinline__ void finish(int32_t i, const Seg &sg, const float *a) const {
        vxsph[i] = make_float4(vi.x + sg.xsph_c * a[0], vi.y + sg.xsph_c * a[1], vi.z + sg.xsph_c * a[2], 0.0f);
        omega[i] = make_float4(a[3], a[4], a[5], 0.0f);
        const float o = sqrtf(sq_len(a[3], a[4], a[5]));
        if (vort) vort[i] = o;
        nrec[i] = make_float4(pi.x, pi.y, pi.z, o);
    }
};

// vorticity.glsl:65-85 for the dense particles (terms of k_vorticity_b).
template <bool H1, bool PAD>
struct VorticityBPass {
    static constexpr int NP = 3, NM = 0;
    static constexpr bool RECORD = false;
    const SolverConsts &c;
    const float4 *__restrict__ rec;
    const float4 *__restrict__ vxsph;
    const float4 *__restrict__ omega;
    float4 *__restrict__ vel_id;
    SolverTape tp;
    bool tame;
    float4 pi;
    __device__ __forceinline__ void begin(int32_t i, const Seg &) { pi = rec[i]; }
    // neighbours come from the first half's records: |omega_j| in pj.w
    __device__ __forceinline__ int extra(const Seg &, int32_t) const { return 0; }
    __device__ __forceinline__ bool eval(const float4 &pj, int, bool listed, float *P, float *) const {
        const float om = pj.w;
        const float px = pi.x - pj.x, py = pi.y - pj.y, pz = pi.z - pj.z;
        const float sq = sq_len(px, py, pz);
        if (!listed && sq > c.h2cut) return false;
        const float r = sqrt_rn_dist(sq, tame);
        if ((!listed && r > c.h) || r == 0.0f) return false;
        float gx, gy, gz;
        grad_spiky<H1>(c, px, py, pz, r, gx, gy, gz, tame);
        P[0] = om * gx;
        P[1] = om * gy;
        P[2] = om * gz;
        return true;
    }
    __device__ __forceinline__ void finish(int32_t i, const Seg &sg, const float *a) const {
        float nx = a[0], ny = a[1], nz = a[2];
        if (tp.eta) tp.eta[i] = make_float4(nx, ny, nz, 0.0f);
        const float l = sqrtf(sq_len(nx, ny, nz));
        if (l > 0.0f) {
            nx = nx / l;
            ny = ny / l;
            nz = nz / l;
        }
        const float4 w = omega[i];
        const float4 v = vxsph[i];
        const float sc = c.dt * c.vort_eps;
        const float ox = v.x + sc * (ny * w.z - nz * w.y);
        const float oy = v.y + sc * (nz * w.x - nx * w.z);
        const float oz = v.z + sc * (nx * w.y - ny * w.x);
        vel_id[sg.off + id_of(pi)] = make_float4(ox, oy, oz, 0.0f);
    }
};

// True for a particle on the step's dense list (more candidates than the
// step's threshold), which the launch's dense workgroups handle.
__device__ __forceinline__ bool leave_dense(int32_t cand, const DenseList &dl) {
    return dl.threshold && cand > dl.threshold[0];
}

// calclambda.glsl:66-103; records the neighbours within h for the position update.
template <bool H1, int NB, bool PAD, class Dom, int LC = LCAP>
__global__ __launch_bounds__(NB) void k_lambda(SolverConsts c, Dom dom, const float4 *__restrict__ rec,
                                                     const uint2 *__restrict__ qpos,
                                                     const int32_t *__restrict__ nbr9, float *__restrict__ lambda,
                                                     float4 *__restrict__ nrec,
                                                     uint16_t *__restrict__ list_entries,
                                                     uint16_t *__restrict__ list_count, int32_t list_cap,
                                                     DenseList dl, uint32_t dense_blocks, SolverTape tp, uint32_t ntiles,
                                                     const uint32_t *__restrict__ tiny) {
    __shared__ union {
        NbrLds<NB, LC> light;
        DenseLds<5, 0, NB / 64> dense;
    } u;
    const bool tame = positions_tame<H1>(tiny);
    // tile bid of ntiles (one per workgroup, or several per workgroup of a
    // persistent launch, whose grid is a multiple of 8 so that tile bid runs on
    // XCD bid % 8 as xcd_block assumes)
    auto tile = [&](const uint32_t bid) {
        {
            if (bid < dense_blocks) {
                LambdaPass<H1, PAD> ps{c, rec, lambda, nrec, tp, tame, {}};
                dense_pass<LambdaPass<H1, PAD>, PAD, NB / 64>(c, dom, rec, nbr9, dl, list_entries, list_count, list_cap, ps,
                                                     u.dense, bid, dense_blocks);
                return;
            }
        }
        NbrLds<NB, LC> &lds = u.light;
        const int32_t i = xcd_block<NB>(bid - dense_blocks, ntiles - dense_blocks) * NB + threadIdx.x;
        if (i >= c.n) return;
        if (leave_dense(load_rows<NB>(c, nbr9, i, lds), dl)) return;
        const Seg sg = dom.at(i);
        const float4 pi = rec[i];
        float rho = 0.0f, sk = 0.0f, gix = 0.0f, giy = 0.0f, giz = 0.0f;
        PairListOut out = {list_entries, list_count, list_cap, 0, list_entries ? list_entries + i : nullptr};
        scan_neighbours<NB, PAD, true>(c, sg, rec, qpos, i, pi, lds, [](int32_t) { return 0; },
                                 [&](int32_t, uint32_t code, const float4 &pj, int) {
            const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
            const float sq = sq_len(dx, dy, dz);
            if (sq > c.h2cut) return;
            const float r = sqrt_rn_dist(sq, tame);
            if (r > c.h) return;
            out.add(i, c.n, code);
            rho += wpoly6_within<H1>(c, sq);
            if (r == 0.0f) return;
            float gx, gy, gz;
            grad_spiky<H1>(c, dx, dy, dz, r, gx, gy, gz, tame);
            gx = gx * c.inv_rho0;
            gy = gy * c.inv_rho0;
            gz = gz * c.inv_rho0;
            sk += sq_len(gx, gy, gz);
            gix += gx;
            giy += gy;
            giz += gz;
        });
        out.finish(i);
        sk += sq_len(gix, giy, giz);
        const float ci = rho * c.inv_rho0 - 1.0f;
        const float l = -ci / (sk + c.eps);
        lambda[i] = l;
        nrec[i] = make_float4(pi.x, pi.y, pi.z, l);
        if (tp.sden) {
            tp.sden[i] = sk + c.eps;
            tp.G[i] = make_float4(gix, giy, giz, 0.0f);
        }
    };
    if (std::is_same<Dom, DomOne>::value || gridDim.x >= ntiles) {
        tile(blockIdx.x);  // one tile per workgroup (always so for one scene)
        return;
    }
    for (uint32_t b = blockIdx.x; b < ntiles; b += gridDim.x) {
        tile(b);
        __syncthreads();  // the next tile reuses the LDS
    }
}

// updatepos.glsl:42-105, race-free: neighbours read from the lambda pass's
// records nrec (the positions of rec_in with lambda in w, one load per
// neighbour), result to rec_out; replays the lambda pass's neighbour list
// (same positions).
template <bool H1, int NB, bool PAD, class Dom>
__global__ __launch_bounds__(NB) void k_updatepos(SolverConsts c, Dom dom, const float4 *__restrict__ rec_in,
                                                        const float *__restrict__ lambda,
                                                        const float4 *__restrict__ nrec,
                                                        const int32_t *__restrict__ nbr9,
                                                        const uint16_t *__restrict__ list_entries,
                                                        const uint16_t *__restrict__ list_count,
                                                        float4 *__restrict__ rec_out, uint2 *__restrict__ q_out,
                                                        DenseList dl, uint32_t dense_blocks, StepUpdate up,
                                                        SolverTape tp, uint32_t ntiles,
                                                        const uint32_t *__restrict__ tiny_in,
                                                        uint32_t *__restrict__ tiny_out) {
    __shared__ union {
        NbrLds<NB, kReplayLcap> light;
        DenseLds<3, 0, NB / 64> dense;
    } u;
    const bool tame = positions_tame<H1>(tiny_in);
    // tile bid of ntiles (one per workgroup, or several per workgroup of a
    // persistent launch, whose grid is a multiple of 8 so that tile bid runs on
    // XCD bid % 8 as xcd_block assumes)
    auto tile = [&](const uint32_t bid) {
        {
            if (bid < dense_blocks) {
                UpdatePosPass<H1, PAD> ps{c, rec_in, lambda, rec_out, q_out, up, tp, tame, tiny_out, {}, 0.0f};
                dense_pass<UpdatePosPass<H1, PAD>, PAD, NB / 64>(c, dom, nrec, nbr9, dl, const_cast<uint16_t *>(list_entries),
                                                        const_cast<uint16_t *>(list_count), 0, ps, u.dense, bid,
                                                        dense_blocks);
                return;
            }
        }
        NbrLds<NB, kReplayLcap> &lds = u.light;
        const int32_t i = xcd_block<NB>(bid - dense_blocks, ntiles - dense_blocks) * NB + threadIdx.x;
        if (i >= c.n) return;
        if (leave_dense(load_rows<NB>(c, nbr9, i, lds), dl)) return;
        const Seg sg = dom.at(i);
        const float4 pi = rec_in[i];
        const float li = lambda[i];
        float dpx = 0.0f, dpy = 0.0f, dpz = 0.0f;
        replay_neighbours<NB, PAD>(c, sg, nrec, list_entries, list_count, i, pi, lds, [](int32_t) { return 0; },
                          [&](int32_t, const float4 &pj, int, bool listed) {
            const float lj = pj.w;
            const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
            const float sq = sq_len(dx, dy, dz);
            if (!listed && sq > c.h2cut) return;
            const float r = sqrt_rn_dist(sq, tame);
            if ((!listed && r > c.h) || r == 0.0f) return;
            float sc = c.scale * wpoly6_within<H1>(c, sq);  // r <= h here, listed or tested
            sc *= sc;
            sc *= sc;
            sc = -c.k * sc;
            float gx, gy, gz;
            grad_spiky<H1>(c, dx, dy, dz, r, gx, gy, gz, tame);
            const float f = li + lj + sc;
            dpx += f * gx;
            dpy += f * gy;
            dpz += f * gz;
        });
        float x = pi.x + c.inv_rho0 * dpx, y = pi.y + c.inv_rho0 * dpy, z = pi.z + c.inv_rho0 * dpz;
        if (tp.clamp)
            tp.clamp[i] = (uint8_t)(clamp_state(x, sg.lo[0], sg.hi[0], 0) | clamp_state(y, sg.lo[1], sg.hi[1], 1) |
                                    clamp_state(z, sg.lo[2], sg.hi[2], 2));
        x = clamp_glsl(x, sg.lo[0], sg.hi[0]);
        y = clamp_glsl(y, sg.lo[1], sg.hi[1]);
        z = clamp_glsl(z, sg.lo[2], sg.hi[2]);
        const float4 r = make_float4(x, y, z, pi.w);
        rec_out[i] = r;
        if (q_out) q_out[i] = quantize_pos(r, c.qscale);
        note_tiny(tiny_out, r);
        if (up.pos) update_one(c, sg, i, r, up);
    };
    if (std::is_same<Dom, DomOne>::value || gridDim.x >= ntiles) {
        tile(blockIdx.x);  // one tile per workgroup (always so for one scene)
        return;
    }
    for (uint32_t b = blockIdx.x; b < ntiles; b += gridDim.x) {
        tile(b);
        __syncthreads();  // the next tile reuses the LDS
    }
}

// update.glsl:16-28; also keeps the velocities in sorted order for the
// vorticity passes.
template <class Dom>
__global__ __launch_bounds__(1024) void k_update(SolverConsts c, Dom dom, int32_t ppt, const float4 *__restrict__ rec,
                                                 StepUpdate up) {
    PBF_STREAM_FOR(i) {
        update_one(c, dom.at(i), i, rec[i], up);
    }
}

// vorticity.glsl:34-60 (first half): XSPH and omega from the velocities before
// this pass; |omega| to the vorticity buffer (vorticities[], :60). Records the
// neighbours within h for the second half.
template <bool H1, int NB, bool PAD, class Dom, int LC = LCAP>
__global__ __launch_bounds__(NB) void k_vorticity_a(SolverConsts c, Dom dom, const float4 *__restrict__ rec,
                                                          const uint2 *__restrict__ qpos,
                                                          const int32_t *__restrict__ nbr9,
                                                          const float4 *__restrict__ vel_sorted,
                                                          const float4 *__restrict__ vel_id,
                                                          float4 *__restrict__ vxsph, float4 *__restrict__ omega,
                                                          float *__restrict__ vort, float4 *__restrict__ nrec,
                                                          uint16_t *__restrict__ list_entries,
                                                          uint16_t *__restrict__ list_count, int32_t list_cap,
                                                          DenseList dl, uint32_t dense_blocks, uint32_t ntiles,
                                                          const uint32_t *__restrict__ tiny) {
    __shared__ union {
        NbrLds<NB, LC> light;
        DenseLds<6, 3, NB / 64> dense;
    } u;
    const bool tame = positions_tame<H1>(tiny);
    // tile bid of ntiles (one per workgroup, or several per workgroup of a
    // persistent launch, whose grid is a multiple of 8 so that tile bid runs on
    // XCD bid % 8 as xcd_block assumes)
    auto tile = [&](const uint32_t bid) {
        {
            if (bid < dense_blocks) {
                VorticityAPass<H1, PAD> ps{c, rec, vel_sorted, vel_id, vxsph, omega, vort, nrec, tame, {}, {}};
                dense_pass<VorticityAPass<H1, PAD>, PAD, NB / 64>(c, dom, rec, nbr9, dl, list_entries, list_count, list_cap, ps,
                                                         u.dense, bid, dense_blocks);
                return;
            }
        }
        NbrLds<NB, LC> &lds = u.light;
        const int32_t i = xcd_block<NB>(bid - dense_blocks, ntiles - dense_blocks) * NB + threadIdx.x;
        if (i >= c.n) return;
        if (leave_dense(load_rows<NB>(c, nbr9, i, lds), dl)) return;
        const Seg sg = dom.at(i);
        const float4 pi = rec[i];
        const float4 vi = vel_sorted[i];
        float sx = 0.0f, sy = 0.0f, sz = 0.0f, wx = 0.0f, wy = 0.0f, wz = 0.0f;
        PairListOut out = {list_entries, list_count, list_cap, 0, list_entries ? list_entries + i : nullptr};
        // a zero record past the scene has id 0: imageLoad (velocitytexture, 0)
        scan_neighbours<NB, PAD, true>(c, sg, rec, qpos, i, pi, lds,
                                 [&](int32_t j) { return *(j < sg.n ? vel_sorted + sg.off + j : vel_id + sg.off); },
                                 [&](int32_t, uint32_t code, const float4 &pj, const float4 &vj) {
            const float px = pi.x - pj.x, py = pi.y - pj.y, pz = pi.z - pj.z;
            const float sq = sq_len(px, py, pz);
            if (sq > c.h2cut) return;
            const float r = sqrt_rn_dist(sq, tame);
            if (r > c.h) return;
            out.add(i, c.n, code);
            const float ux = vj.x - vi.x, uy = vj.y - vi.y, uz = vj.z - vi.z;
            const float w = wpoly6_within<H1>(c, sq);
            sx += ux * w;
            sy += uy * w;
            sz += uz * w;
            if (r == 0.0f) return;
            float gx, gy, gz;
            grad_spiky<H1>(c, px, py, pz, r, gx, gy, gz, tame);
            wx = (wx + uy * gz) - uz * gy;
            wy = (wy + uz * gx) - ux * gz;
            wz = (wz + ux * gy) - uy * gx;
        });
        out.finish(i);
        vxsph[i] = make_float4(vi.x + sg.xsph_c * sx, vi.y + sg.xsph_c * sy, vi.z + sg.xsph_c * sz, 0.0f);
        omega[i] = make_float4(wx, wy, wz, 0.0f);
        const float o = sqrtf(sq_len(wx, wy, wz));
        if (vort) vort[i] = o;
        nrec[i] = make_float4(pi.x, pi.y, pi.z, o);
    };
    if (std::is_same<Dom, DomOne>::value || gridDim.x >= ntiles) {
        tile(blockIdx.x);  // one tile per workgroup (always so for one scene)
        return;
    }
    for (uint32_t b = blockIdx.x; b < ntiles; b += gridDim.x) {
        tile(b);
        __syncthreads();  // the next tile reuses the LDS
    }
}

// vorticity.glsl:65-85 (second half): confinement from this step's |omega|,
// read with the neighbour's position from the first half's records nrec;
// replays the first half's neighbour list.
template <bool H1, int NB, bool PAD, class Dom>
__global__ __launch_bounds__(NB) void k_vorticity_b(SolverConsts c, Dom dom, const float4 *__restrict__ rec,
                                                          const int32_t *__restrict__ nbr9,
                                                          const uint16_t *__restrict__ list_entries,
                                                          const uint16_t *__restrict__ list_count,
                                                          const float4 *__restrict__ vxsph,
                                                          const float4 *__restrict__ omega,
                                                          const float4 *__restrict__ nrec,
                                                          float4 *__restrict__ vel_id, DenseList dl,
                                                          uint32_t dense_blocks, SolverTape tp, uint32_t ntiles,
                                                          const uint32_t *__restrict__ tiny) {
    __shared__ union {
        NbrLds<NB, kReplayLcap> light;
        DenseLds<3, 0, NB / 64> dense;
    } u;
    const bool tame = positions_tame<H1>(tiny);
    // tile bid of ntiles (one per workgroup, or several per workgroup of a
    // persistent launch, whose grid is a multiple of 8 so that tile bid runs on
    // XCD bid % 8 as xcd_block assumes)
    auto tile = [&](const uint32_t bid) {
        {
            if (bid < dense_blocks) {
                VorticityBPass<H1, PAD> ps{c, rec, vxsph, omega, vel_id, tp, tame, {}};
                dense_pass<VorticityBPass<H1, PAD>, PAD, NB / 64>(c, dom, nrec, nbr9, dl, const_cast<uint16_t *>(list_entries),
                                                         const_cast<uint16_t *>(list_count), 0, ps, u.dense, bid,
                                                         dense_blocks);
                return;
            }
        }
        NbrLds<NB, kReplayLcap> &lds = u.light;
        const int32_t i = xcd_block<NB>(bid - dense_blocks, ntiles - dense_blocks) * NB + threadIdx.x;
        if (i >= c.n) return;
        if (leave_dense(load_rows<NB>(c, nbr9, i, lds), dl)) return;
        const Seg sg = dom.at(i);
        const float4 pi = rec[i];
        float nx = 0.0f, ny = 0.0f, nz = 0.0f;
        replay_neighbours<NB, PAD>(c, sg, nrec, list_entries, list_count, i, pi, lds, [](int32_t) { return 0; },
                          [&](int32_t, const float4 &pj, int, bool listed) {
            const float om = pj.w;
            const float px = pi.x - pj.x, py = pi.y - pj.y, pz = pi.z - pj.z;
            const float sq = sq_len(px, py, pz);
            if (!listed && sq > c.h2cut) return;
            const float r = sqrt_rn_dist(sq, tame);
            if ((!listed && r > c.h) || r == 0.0f) return;
            float gx, gy, gz;
            grad_spiky<H1>(c, px, py, pz, r, gx, gy, gz, tame);
            nx += om * gx;
            ny += om * gy;
            nz += om * gz;
        });
        if (tp.eta) tp.eta[i] = make_float4(nx, ny, nz, 0.0f);
        const float l = sqrtf(sq_len(nx, ny, nz));
        if (l > 0.0f) {
            nx = nx / l;
            ny = ny / l;
            nz = nz / l;
        }
        const float4 w = omega[i];
        const float4 v = vxsph[i];
        const float sc = c.dt * c.vort_eps;
        const float ox = v.x + sc * (ny * w.z - nz * w.y);
        const float oy = v.y + sc * (nz * w.x - nx * w.z);
        const float oz = v.z + sc * (nx * w.y - ny * w.x);
        vel_id[sg.off + id_of(pi)] = make_float4(ox, oy, oz, 0.0f);
    };
    if (std::is_same<Dom, DomOne>::value || gridDim.x >= ntiles) {
        tile(blockIdx.x);  // one tile per workgroup (always so for one scene)
        return;
    }
    for (uint32_t b = blockIdx.x; b < ntiles; b += gridDim.x) {
        tile(b);
        __syncthreads();  // the next tile reuses the LDS
    }
}

__global__ __launch_bounds__(BLOCK) void k_quantize(float qscale, const float4 *__restrict__ rec,
                                                   uint2 *__restrict__ q, int32_t count) {
    const int32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < count) q[i] = quantize_pos(rec[i], qscale);
}

// ----------------------------------------------------------------- layout conversions
__global__ __launch_bounds__(BLOCK) void k_nbr9_to_12(int32_t n, const int32_t *__restrict__ nbr9,
                                                      int32_t *__restrict__ nbr12) {
    const int32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    for (int r = 0; r < 3; r++) {
        int4 v = make_int4(nbr9[(size_t)(3 * r) * n + i], nbr9[(size_t)(3 * r + 1) * n + i],
                           nbr9[(size_t)(3 * r + 2) * n + i], 0);
        reinterpret_cast<int4 *>(nbr12)[(size_t)3 * i + r] = v;
    }
}

__global__ __launch_bounds__(BLOCK) void k_nbr12_to_9(int32_t n, const int32_t *__restrict__ nbr12,
                                                      int32_t *__restrict__ nbr9) {
    const int32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    for (int r = 0; r < 3; r++) {
        const int4 v = reinterpret_cast<const int4 *>(nbr12)[(size_t)3 * i + r];
        nbr9[(size_t)(3 * r) * n + i] = v.x;
        nbr9[(size_t)(3 * r + 1) * n + i] = v.y;
        nbr9[(size_t)(3 * r + 2) * n + i] = v.z;
    }
}

__global__ __launch_bounds__(BLOCK) void k_end64_to_i32(size_t cells, const unsigned long long *__restrict__ e64,
                                                        int32_t *__restrict__ e) {
    for (size_t k = (size_t)blockIdx.x * BLOCK + threadIdx.x; k < cells; k += (size_t)gridDim.x * BLOCK)
        e[k] = (int32_t)(uint32_t)(e64[k] & 0xFFFFFFFFull);
}

__global__ __launch_bounds__(BLOCK) void k_i32_to_end64(size_t cells, const int32_t *__restrict__ e,
                                                        unsigned long long *__restrict__ e64) {
    for (size_t k = (size_t)blockIdx.x * BLOCK + threadIdx.x; k < cells; k += (size_t)gridDim.x * BLOCK)
        e64[k] = (unsigned long long)(uint32_t)e[k];
}

inline unsigned blocks_for(size_t n) { return (unsigned)((n + BLOCK - 1) / BLOCK); }
inline unsigned grid_stride_blocks(size_t n) {
    size_t b = (n + BLOCK - 1) / BLOCK;
    return (unsigned)(b > 65536 ? 65536 : (b == 0 ? 1 : b));
}

}  // namespace

// ----------------------------------------------------------------- launchers
Seg seg_of(const SolverConsts &c) {
    Seg s;
    s.off = 0;
    s.n = c.n;
    s.gx = c.gx;
    s.gy = c.gy;
    s.gz = c.gz;
    s.cell_base = 0;
    s.key_mask = c.key_mask;
    s.key_hi = 0;
    for (int a = 0; a < 3; a++) {
        s.lo[a] = c.lo[a];
        s.hi[a] = c.hi[a];
        s.gdir[a] = c.gdir[a];
    }
    s.xsph_c = c.xsph_c;
    return s;
}

bool valid_shape(const LaunchShape &ls) {
    const bool nb = ls.nb == 64 || ls.nb == 128 || ls.nb == 256 || ls.nb == 512;
    const bool sb = ls.sb >= 64 && ls.sb <= 1024 && (ls.sb & (ls.sb - 1)) == 0;
    const bool lc = ls.lc == 32 || ls.lc == LCAP;
    return nb && sb && lc && ls.ppt >= 1 && ls.ppt <= 4 && (ls.persistent == 0 || ls.persistent == 1);
}

namespace {
DomOne one(const SolverConsts &c) { return DomOne{seg_of(c)}; }

const LaunchShape kDefaultShape;

unsigned stream_blocks(int32_t n, const LaunchShape &ls) {
    const int64_t per = (int64_t)ls.sb * ls.ppt;
    return (unsigned)((n + per - 1) / per);
}
unsigned nb_blocks(int32_t n, int nb) { return (unsigned)((n + nb - 1) / nb); }

// Calls f with the neighbour workgroup size and the padding flag as
// compile-time constants: one scene runs PBF_NB_ONE-wide, padded when its arrays are
// (the SPH handle's are, caller buffers of the stage entry points are not);
// many scenes take the tuned width, unpadded.
#ifndef PBF_NB_ONE
#define PBF_NB_ONE 64
#endif
#ifndef PBF_NB_REPLAY
#define PBF_NB_REPLAY 64
#endif
static_assert(PBF_NB_ONE == 64 || PBF_NB_ONE == 128 || PBF_NB_ONE == 256, "neighbour workgroup width");
static_assert(PBF_NB_REPLAY == 64 || PBF_NB_REPLAY == 128 || PBF_NB_REPLAY == 256, "neighbour workgroup width");
// REPLAY: the replaying passes' width for one scene (their LDS is smaller).
template <class Dom, bool REPLAY = false, class F>
void with_nb(int nb, bool pad, F &&f) {
    if constexpr (std::is_same<Dom, DomOne>::value) {
        constexpr int W = REPLAY ? PBF_NB_REPLAY : PBF_NB_ONE;
        if (pad)
            f(std::integral_constant<int, W>(), std::true_type());
        else
            f(std::integral_constant<int, W>(), std::false_type());
    } else {
        switch (nb) {
            case 64: f(std::integral_constant<int, 64>(), std::false_type()); break;
            case 128: f(std::integral_constant<int, 128>(), std::false_type()); break;
            case 512: f(std::integral_constant<int, 512>(), std::false_type()); break;
            default: f(std::integral_constant<int, 256>(), std::false_type()); break;
        }
    }
}

template <class Dom>
void predict_impl(const SolverConsts &c, const Dom &d, const LaunchShape &ls, const float4 *pos, const float4 *vel,
                  float4 *rec, uint32_t *keys, hipStream_t st, uint32_t *hl = nullptr, uint32_t *tiny = nullptr,
                  int32_t ntiny = 0) {
    if (c.n > 0)
        k_predict<Dom><<<stream_blocks(c.n, ls), ls.sb, 0, st>>>(c, d, ls.ppt, pos, vel, rec, keys, hl, tiny, ntiny);
}
template <class Dom>
void findcells_impl(const SolverConsts &c, const Dom &d, const LaunchShape &ls, const float4 *sorted, int32_t *start,
                    unsigned long long *end64, const uint32_t *epoch, hipStream_t st) {
    if (c.n > 0)
        k_findcells<Dom><<<stream_blocks(c.n, ls), ls.sb, 0, st>>>(c, d, ls.ppt, sorted, start, end64, epoch);
}
template <class Dom>
void neighbourcells_impl(const SolverConsts &c, const Dom &d, const LaunchShape &ls, const float4 *sorted,
                         const int32_t *start, const unsigned long long *end64, int32_t *nbr9, hipStream_t st,
                         const DenseList *dl = nullptr, uint32_t *hl = nullptr) {
    if (c.n > 0)
        k_neighbourcells<Dom><<<stream_blocks(c.n, ls), ls.sb, 0, st>>>(c, d, ls.ppt, sorted, start, end64, nbr9,
                                                                         dl ? *dl : DenseList(), hl);
    if (dl && c.n > 0) launch_dense_select(*dl, st);
}

// Grid of a neighbour launch of ntiles workgroup tiles: one workgroup per
// tile, or for a persistent launch as many workgroups as fit on the device at
// once (a multiple of 8), each looping over tiles.
template <class K>
unsigned neighbour_grid(K kernel, int nb, unsigned ntiles, bool persistent) {
    if (!persistent) return ntiles;
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, nb, 0) != hipSuccess || per <= 0)
        return ntiles;
    const unsigned cap = (unsigned)(cus * per) / 8u * 8u;
    return cap >= 8u && cap < ntiles ? cap : ntiles;
}

// Calls f with the scanning passes' LDS list depth as a compile-time constant:
// the default for one scene; 32 or 56 entries for many (a tuning choice).
template <class Dom, class F>
void with_lc(int lc, F &&f) {
    if constexpr (std::is_same<Dom, DomOne>::value) {
        f(std::integral_constant<int, LCAP>());
    } else {
        if (lc == 32)
            f(std::integral_constant<int, 32>());
        else
            f(std::integral_constant<int, LCAP>());
    }
}

// Workgroups of a neighbour launch given to the dense list (none without one).
unsigned dense_blocks_of(const DenseList *dl, int32_t n, int nb) {
    if (!dl) return 0u;
    const int64_t waves = std::min<int64_t>(kDenseWaves, ((int64_t)n + kDenseGroups - 1) / kDenseGroups);
    const int64_t wgs = (waves + nb / 64 - 1) / (nb / 64);
    return (unsigned)((wgs + 7) / 8 * 8);
}
template <class Dom>
void update_impl(const SolverConsts &c, const Dom &d, const LaunchShape &ls, const float4 *rec, float4 *pos,
                 float4 *vel, float4 *vel_sorted, hipStream_t st) {
    StepUpdate up;
    up.pos = pos;
    up.vel = vel;
    up.vel_sorted = vel_sorted;
    if (c.n > 0) k_update<Dom><<<stream_blocks(c.n, ls), ls.sb, 0, st>>>(c, d, ls.ppt, rec, up);
}
template <class Dom>
void lambda_impl(const SolverConsts &c, const Dom &d, const LaunchShape &ls, bool pad, const float4 *rec,
                 const uint2 *q, const int32_t *nbr9, float *lambda, float4 *nrec, const PairList &pl, hipStream_t st,
                 const DenseList *dl = nullptr, const SolverTape *tape = nullptr, const uint32_t *tiny = nullptr) {
    if (c.n <= 0) return;
    const DenseList dlv = dl ? *dl : DenseList();
    const SolverTape tp = tape ? *tape : SolverTape();
    with_nb<Dom>(ls.nb, pad, [&](auto nbc, auto padc) {
        constexpr int NB = decltype(nbc)::value;
        constexpr bool PAD = decltype(padc)::value;
        const unsigned db = dense_blocks_of(dl, c.n, NB);
        const unsigned ntiles = nb_blocks(c.n, NB) + db;
        with_lc<Dom>(ls.lc, [&](auto lcc) {
            constexpr int LC = decltype(lcc)::value;
            if (c.h_is_one) {
                auto k = k_lambda<true, NB, PAD, Dom, LC>;
                k<<<neighbour_grid(k, NB, ntiles, ls.persistent), NB, 0, st>>>(
                    c, d, rec, q, nbr9, lambda, nrec, pl.entries, pl.count, pl.cap, dlv, db, tp, ntiles, tiny);
            } else {
                auto k = k_lambda<false, NB, PAD, Dom, LC>;
                k<<<neighbour_grid(k, NB, ntiles, ls.persistent), NB, 0, st>>>(
                    c, d, rec, q, nbr9, lambda, nrec, pl.entries, pl.count, pl.cap, dlv, db, tp, ntiles, tiny);
            }
        });
    });
}
template <class Dom>
void updatepos_impl(const SolverConsts &c, const Dom &d, const LaunchShape &ls, bool pad, const float4 *rec_in,
                    const float *lambda, const float4 *nrec, const int32_t *nbr9, const PairList &pl, float4 *rec_out, uint2 *q_out,
                    hipStream_t st, const DenseList *dl = nullptr, const StepUpdate *up = nullptr,
                    const SolverTape *tape = nullptr, const uint32_t *tiny_in = nullptr,
                    uint32_t *tiny_out = nullptr) {
    if (c.n <= 0) return;
    const DenseList dlv = dl ? *dl : DenseList();
    const StepUpdate upv = up ? *up : StepUpdate();
    const SolverTape tp = tape ? *tape : SolverTape();
    with_nb<Dom, true>(ls.nb, pad, [&](auto nbc, auto padc) {
        constexpr int NB = decltype(nbc)::value;
        constexpr bool PAD = decltype(padc)::value;
        const unsigned db = dense_blocks_of(dl, c.n, NB);
        const unsigned ntiles = nb_blocks(c.n, NB) + db;
        if (c.h_is_one) {
            auto k = k_updatepos<true, NB, PAD, Dom>;
            k<<<neighbour_grid(k, NB, ntiles, ls.persistent), NB, 0, st>>>(
                c, d, rec_in, lambda, nrec, nbr9, pl.entries, pl.count, rec_out, q_out, dlv, db, upv, tp, ntiles, tiny_in,
                tiny_out);
        } else {
            auto k = k_updatepos<false, NB, PAD, Dom>;
            k<<<neighbour_grid(k, NB, ntiles, ls.persistent), NB, 0, st>>>(
                c, d, rec_in, lambda, nrec, nbr9, pl.entries, pl.count, rec_out, q_out, dlv, db, upv, tp, ntiles, tiny_in,
                tiny_out);
        }
    });
}
template <class Dom>
void vorticity_a_impl(const SolverConsts &c, const Dom &d, const LaunchShape &ls, bool pad, const float4 *rec,
                      const uint2 *q, const int32_t *nbr9, const float4 *vel_sorted, const float4 *vel_id, float4 *vxsph,
                      float4 *omega, float *vort, float4 *nrec, const PairList &pl, hipStream_t st,
                      const DenseList *dl = nullptr, const uint32_t *tiny = nullptr) {
    if (c.n <= 0) return;
    const DenseList dlv = dl ? *dl : DenseList();
    with_nb<Dom>(ls.nb, pad, [&](auto nbc, auto padc) {
        constexpr int NB = decltype(nbc)::value;
        constexpr bool PAD = decltype(padc)::value;
        const unsigned db = dense_blocks_of(dl, c.n, NB);
        const unsigned ntiles = nb_blocks(c.n, NB) + db;
        with_lc<Dom>(ls.lc, [&](auto lcc) {
            constexpr int LC = decltype(lcc)::value;
            if (c.h_is_one) {
                auto k = k_vorticity_a<true, NB, PAD, Dom, LC>;
                k<<<neighbour_grid(k, NB, ntiles, ls.persistent), NB, 0, st>>>(
                    c, d, rec, q, nbr9, vel_sorted, vel_id, vxsph, omega, vort, nrec, pl.entries, pl.count, pl.cap, dlv, db,
                    ntiles, tiny);
            } else {
                auto k = k_vorticity_a<false, NB, PAD, Dom, LC>;
                k<<<neighbour_grid(k, NB, ntiles, ls.persistent), NB, 0, st>>>(
                    c, d, rec, q, nbr9, vel_sorted, vel_id, vxsph, omega, vort, nrec, pl.entries, pl.count, pl.cap, dlv, db,
                    ntiles, tiny);
            }
        });
    });
}
template <class Dom>
void vorticity_b_impl(const SolverConsts &c, const Dom &d, const LaunchShape &ls, bool pad, const float4 *rec,
                      const int32_t *nbr9, const PairList &pl, const float4 *vxsph, const float4 *omega,
                      const float4 *nrec, float4 *vel_id, hipStream_t st, const DenseList *dl = nullptr,
                      const SolverTape *tape = nullptr, const uint32_t *tiny = nullptr) {
    if (c.n <= 0) return;
    const DenseList dlv = dl ? *dl : DenseList();
    const SolverTape tp = tape ? *tape : SolverTape();
    with_nb<Dom, true>(ls.nb, pad, [&](auto nbc, auto padc) {
        constexpr int NB = decltype(nbc)::value;
        constexpr bool PAD = decltype(padc)::value;
        const unsigned db = dense_blocks_of(dl, c.n, NB);
        const unsigned ntiles = nb_blocks(c.n, NB) + db;
        if (c.h_is_one) {
            auto k = k_vorticity_b<true, NB, PAD, Dom>;
            k<<<neighbour_grid(k, NB, ntiles, ls.persistent), NB, 0, st>>>(
                c, d, rec, nbr9, pl.entries, pl.count, vxsph, omega, nrec, vel_id, dlv, db, tp, ntiles, tiny);
        } else {
            auto k = k_vorticity_b<false, NB, PAD, Dom>;
            k<<<neighbour_grid(k, NB, ntiles, ls.persistent), NB, 0, st>>>(
                c, d, rec, nbr9, pl.entries, pl.count, vxsph, omega, nrec, vel_id, dlv, db, tp, ntiles, tiny);
        }
    });
}
}  // namespace

// one scene
void launch_predict(const SolverConsts &c, const float4 *pos, const float4 *vel, float4 *rec, uint32_t *keys,
                    hipStream_t st, uint32_t *highlight, uint32_t *tiny, int32_t ntiny) {
    predict_impl(c, one(c), kDefaultShape, pos, vel, rec, keys, st, highlight, tiny, ntiny);
}
void launch_keys_of_records(const SolverConsts &c, const float4 *rec, uint32_t *keys, hipStream_t st) {
    if (c.n > 0) k_keys<DomOne><<<blocks_for(c.n), BLOCK, 0, st>>>(c, one(c), rec, keys);
}
void launch_gather_records(int32_t n, const uint32_t *perm, const float4 *in, float4 *out, hipStream_t st) {
    if (n > 0) k_gather<<<blocks_for(n), BLOCK, 0, st>>>(n, perm, in, out);
}
void launch_gather_by_id(int32_t n, const float4 *rec, const float4 *src, float4 *out, hipStream_t st) {
    if (n > 0) k_gather_by_id<<<blocks_for(n), BLOCK, 0, st>>>(n, rec, src, out);
}
void launch_pack_records(int32_t n, const float4 *rec, const float *v, float4 *out, hipStream_t st) {
    if (n > 0) k_pack_records<<<blocks_for(n), BLOCK, 0, st>>>(n, rec, v, out);
}
void launch_fill_i32(int32_t *p, int32_t value, size_t count, hipStream_t st) {
    if (count > 0) k_fill_i32<<<grid_stride_blocks((count + 3) / 4), BLOCK, 0, st>>>(p, value, count);
}
void launch_begin_cells(int32_t *start, size_t cells, uint32_t *epoch, hipStream_t st, const DenseList *dl) {
    k_begin_cells<<<grid_stride_blocks((cells + 3) / 4), BLOCK, 0, st>>>(start, cells, epoch, dl ? *dl : DenseList());
}
void launch_dense_select(const DenseList &dl, hipStream_t st) {
    k_dense_select<<<1, kDenseSelectThreads, 0, st>>>(dl);
}
void launch_findcells(const SolverConsts &c, const float4 *sorted, int32_t *start, unsigned long long *end64,
                      const uint32_t *epoch, hipStream_t st) {
    findcells_impl(c, one(c), kDefaultShape, sorted, start, end64, epoch, st);
}
void launch_neighbourcells(const SolverConsts &c, const float4 *sorted, const int32_t *start,
                           const unsigned long long *end64, int32_t *nbr9, hipStream_t st, const DenseList *dl,
                           uint32_t *highlight) {
    neighbourcells_impl(c, one(c), kDefaultShape, sorted, start, end64, nbr9, st, dl, highlight);
}
void launch_lambda(const SolverConsts &c, const float4 *rec, const uint2 *q, const int32_t *nbr9, float *lambda,
                   float4 *nrec, const PairList &pl, hipStream_t st, bool padded, const DenseList *dl,
                   const SolverTape *tape, const uint32_t *tiny) {
    lambda_impl(c, one(c), kDefaultShape, padded, rec, q, nbr9, lambda, nrec, pl, st, dl, tape, tiny);
}
void launch_updatepos(const SolverConsts &c, const float4 *rec_in, const float *lambda, const float4 *nrec,
                      const int32_t *nbr9, const PairList &pl, float4 *rec_out, uint2 *q_out, hipStream_t st,
                      bool padded, const DenseList *dl, const StepUpdate *up, const SolverTape *tape,
                      const uint32_t *tiny_in, uint32_t *tiny_out) {
    updatepos_impl(c, one(c), kDefaultShape, padded, rec_in, lambda, nrec, nbr9, pl, rec_out, q_out, st, dl, up, tape,
                   tiny_in, tiny_out);
}
void launch_update(const SolverConsts &c, const float4 *rec, float4 *pos, float4 *vel, float4 *vel_sorted,
                   hipStream_t st) {
    update_impl(c, one(c), kDefaultShape, rec, pos, vel, vel_sorted, st);
}
void launch_vorticity_a(const SolverConsts &c, const float4 *rec, const uint2 *q, const int32_t *nbr9,
                        const float4 *vel_sorted, const float4 *vel_id, float4 *vxsph, float4 *omega, float *vort,
                        float4 *nrec, const PairList &pl, hipStream_t st, bool padded, const DenseList *dl,
                        const uint32_t *tiny) {
    vorticity_a_impl(c, one(c), kDefaultShape, padded, rec, q, nbr9, vel_sorted, vel_id, vxsph, omega, vort, nrec, pl,
                     st, dl, tiny);
}
void launch_quantize(const SolverConsts &c, const float4 *rec, uint2 *q, int32_t count, hipStream_t st) {
    if (count > 0) k_quantize<<<blocks_for(count), BLOCK, 0, st>>>(c.qscale, rec, q, count);
}
void launch_vorticity_b(const SolverConsts &c, const float4 *rec, const int32_t *nbr9, const PairList &pl,
                        const float4 *vxsph, const float4 *omega, const float4 *nrec, float4 *vel_id, hipStream_t st,
                        bool padded, const DenseList *dl, const SolverTape *tape, const uint32_t *tiny) {
    vorticity_b_impl(c, one(c), kDefaultShape, padded, rec, nbr9, pl, vxsph, omega, nrec, vel_id, st, dl, tape, tiny);
}

// many scenes
void launch_predict(const SolverConsts &c, const DomMany &d, const LaunchShape &ls, const float4 *pos,
                    const float4 *vel, float4 *rec, uint32_t *keys, hipStream_t st) {
    predict_impl(c, d, ls, pos, vel, rec, keys, st);
}
void launch_findcells(const SolverConsts &c, const DomMany &d, const LaunchShape &ls, const float4 *sorted,
                      int32_t *start, unsigned long long *end64, const uint32_t *epoch, hipStream_t st) {
    findcells_impl(c, d, ls, sorted, start, end64, epoch, st);
}
void launch_neighbourcells(const SolverConsts &c, const DomMany &d, const LaunchShape &ls, const float4 *sorted,
                           const int32_t *start, const unsigned long long *end64, int32_t *nbr9, hipStream_t st,
                           const DenseList *dl) {
    neighbourcells_impl(c, d, ls, sorted, start, end64, nbr9, st, dl);
}
void launch_lambda(const SolverConsts &c, const DomMany &d, const LaunchShape &ls, const float4 *rec,
                   const uint2 *q, const int32_t *nbr9, float *lambda, float4 *nrec, const PairList &pl, hipStream_t st,
                   const DenseList *dl) {
    lambda_impl(c, d, ls, false, rec, q, nbr9, lambda, nrec, pl, st, dl);
}
void launch_updatepos(const SolverConsts &c, const DomMany &d, const LaunchShape &ls, const float4 *rec_in,
                      const float *lambda, const float4 *nrec, const int32_t *nbr9, const PairList &pl, float4 *rec_out,
                      uint2 *q_out, hipStream_t st, const DenseList *dl, const StepUpdate *up) {
    updatepos_impl(c, d, ls, false, rec_in, lambda, nrec, nbr9, pl, rec_out, q_out, st, dl, up);
}
void launch_update(const SolverConsts &c, const DomMany &d, const LaunchShape &ls, const float4 *rec, float4 *pos,
                   float4 *vel, float4 *vel_sorted, hipStream_t st) {
    update_impl(c, d, ls, rec, pos, vel, vel_sorted, st);
}
void launch_vorticity_a(const SolverConsts &c, const DomMany &d, const LaunchShape &ls, const float4 *rec,
                        const uint2 *q, const int32_t *nbr9, const float4 *vel_sorted, const float4 *vel_id,
                        float4 *vxsph, float4 *omega, float *vort, float4 *nrec, const PairList &pl, hipStream_t st,
                        const DenseList *dl) {
    vorticity_a_impl(c, d, ls, false, rec, q, nbr9, vel_sorted, vel_id, vxsph, omega, vort, nrec, pl, st, dl);
}
void launch_vorticity_b(const SolverConsts &c, const DomMany &d, const LaunchShape &ls, const float4 *rec,
                        const int32_t *nbr9, const PairList &pl, const float4 *vxsph, const float4 *omega,
                        const float4 *nrec, float4 *vel_id, hipStream_t st, const DenseList *dl) {
    vorticity_b_impl(c, d, ls, false, rec, nbr9, pl, vxsph, omega, nrec, vel_id, st, dl);
}
void launch_nbr9_to_12(int32_t n, const int32_t *nbr9, int32_t *nbr12, hipStream_t st) {
    if (n > 0) k_nbr9_to_12<<<blocks_for(n), BLOCK, 0, st>>>(n, nbr9, nbr12);
}
void launch_nbr12_to_9(int32_t n, const int32_t *nbr12, int32_t *nbr9, hipStream_t st) {
    if (n > 0) k_nbr12_to_9<<<blocks_for(n), BLOCK, 0, st>>>(n, nbr12, nbr9);
}
void launch_end64_to_i32(size_t cells, const unsigned long long *end64, int32_t *end, hipStream_t st) {
    if (cells > 0) k_end64_to_i32<<<grid_stride_blocks(cells), BLOCK, 0, st>>>(cells, end64, end);
}
void launch_i32_to_end64(size_t cells, const int32_t *end, unsigned long long *end64, hipStream_t st) {
    if (cells > 0) k_i32_to_end64<<<grid_stride_blocks(cells), BLOCK, 0, st>>>(cells, end, end64);
}

}  // namespace pbf
