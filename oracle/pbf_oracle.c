/* CPU restatement of the reference's simulation step: TEST INFRASTRUCTURE ONLY.
 * See pbf_oracle.h for who may use it. Built with -ffp-contract=off.
 *
 * Parity pins (tests/test_oracle_golden.py and tests/test_oracle_invariants.py):
 *   predict, keys, sort, findcells, neighbourcells, lambda, update: bit-exact
 *   against llvmpipe runs of the reference's shaders (fixtures in tests/golden);
 *   updatepos and vorticity: bit-exact against the same runs in the
 *   reference-order emulation; race-free order pinned by invariants
 *   (momentum, lattice closed form) and by sharing the per-particle code. */
#include "pbf_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------- FP32 */
#define REAL float
#define SFX(x) x##_f32
#define LIT(x) x##f
#define SQRT sqrtf
#include "pbf_oracle_impl.inc"
#undef REAL
#undef SFX
#undef LIT
#undef SQRT

/* ---------------------------------------------------------------- FP64 */
#define REAL double
#define SFX(x) x##_f64
#define LIT(x) x
#define SQRT sqrt
#include "pbf_oracle_impl.inc"
#undef REAL
#undef SFX
#undef LIT
#undef SQRT

int pbfo_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

static inline int32_t cells_of(const pbfo_cfg *c) { return c->grid[0] * c->grid[1] * c->grid[2]; }

/* GLSL clamp (x, lo, hi) = min (max (x, lo), hi) */
static inline float clampf_glsl(float x, float lo, float hi) {
    float m = x < lo ? lo : x;
    return m > hi ? hi : m;
}

/* GetHash, counting.glsl:53-57 = globalsort.glsl:50-55: the integer cell is
 * converted back to float for dot() (GLSL has no integer dot), summed in FP32
 * in component order and truncated to uint. */
static inline uint32_t hash_f32(const pbfo_cfg *c, float x, float y, float z) {
    const float gx = (float)c->grid[0], gy = (float)c->grid[1], gz = (float)c->grid[2];
    int32_t ix = (int32_t)clampf_glsl(x, 0.0f, gx);
    int32_t iy = (int32_t)clampf_glsl(y, 0.0f, gy);
    int32_t iz = (int32_t)clampf_glsl(z, 0.0f, gz);
    const float w1 = (float)(c->grid[0] * c->grid[2]), w2 = (float)c->grid[0];
    float d = (float)ix * 1.0f + (float)iy * w1 + (float)iz * w2;
    return (uint32_t)d;
}

static inline uint32_t hash_f64(const pbfo_cfg *c, double x, double y, double z) {
    const double gx = c->grid[0], gy = c->grid[1], gz = c->grid[2];
    int64_t ix = (int64_t)(x < 0 ? 0 : (x > gx ? gx : x));
    int64_t iy = (int64_t)(y < 0 ? 0 : (y > gy ? gy : y));
    int64_t iz = (int64_t)(z < 0 ? 0 : (z > gz ? gz : z));
    return (uint32_t)(ix + iy * (int64_t)c->grid[0] * c->grid[2] + iz * (int64_t)c->grid[0]);
}

static inline uint32_t key_mask(const pbfo_cfg *c) {
    return c->key_bits >= 32 ? 0xFFFFFFFFu : ((1u << c->key_bits) - 1u);
}

/* Stable sort of indices by key: the closed form of the reference's LSD radix
 * sort (RadixSort::Run, src/RadixSort.cpp:124-133; counting.glsl ranks by the
 * exclusive count of equal digits earlier in the block, globalsort.glsl:64
 * scatters digit-major), which is a stable sort of the records by the key bits
 * [0, key_bits); predictpos writes record id at index id, so ties keep ascending
 * id. Here: LSD radix with 8-bit digits. perm[g] = input index at sorted slot g. */
static void stable_sort(const uint32_t *keys_in, int32_t n, int bits, int32_t *perm) {
    uint32_t *k = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)n);
    uint32_t *k2 = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)n);
    int32_t *p2 = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    for (int32_t i = 0; i < n; i++) {
        k[i] = keys_in[i];
        perm[i] = i;
    }
    for (int shift = 0; shift < bits; shift += 8) {
        size_t cnt[257];
        memset(cnt, 0, sizeof cnt);
        for (int32_t i = 0; i < n; i++) cnt[((k[i] >> shift) & 0xFF) + 1]++;
        for (int d = 0; d < 256; d++) cnt[d + 1] += cnt[d];
        for (int32_t i = 0; i < n; i++) {
            size_t dst = cnt[(k[i] >> shift) & 0xFF]++;
            k2[dst] = k[i];
            p2[dst] = perm[i];
        }
        memcpy(k, k2, sizeof(uint32_t) * (size_t)n);
        memcpy(perm, p2, sizeof(int32_t) * (size_t)n);
    }
    free(k);
    free(k2);
    free(p2);
}

void pbfo_predict(const pbfo_cfg *c, const float *pos4, const float *vel4, float *rec4) {
    predict_f32(c, pos4, vel4, (rec_f32 *)rec4);
}

void pbfo_keys(const pbfo_cfg *c, const float *rec4, uint32_t *keys) {
    const rec_f32 *r = (const rec_f32 *)rec4;
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < c->n; i++) keys[i] = hash_f32(c, r[i].x, r[i].y, r[i].z);
}

void pbfo_sort(const pbfo_cfg *c, const float *rec4_in, float *rec4_out, int32_t *perm) {
    const int32_t n = c->n;
    uint32_t *keys = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)n);
    int32_t *p = perm ? perm : (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    pbfo_keys(c, rec4_in, keys);
    const uint32_t m = key_mask(c);
    for (int32_t i = 0; i < n; i++) keys[i] &= m;
    stable_sort(keys, n, c->key_bits, p);
    const rec_f32 *in = (const rec_f32 *)rec4_in;
    rec_f32 *out = (rec_f32 *)rec4_out;
    for (int32_t g = 0; g < n; g++) out[g] = in[p[g]]; /* bit-preserving move, globalsort.glsl:64 */
    if (!perm) free(p);
    free(keys);
}

/* ivec3 (clamp (p, 0, GRID_SIZE)) of findcells.glsl:45-46 as a texel index, or
 * -1 when the clamped cell equals GRID_SIZE in some axis (store dropped). */
static inline int64_t cell_index_clamped(const pbfo_cfg *c, int32_t ix, int32_t iy, int32_t iz) {
    if (ix < 0 || iy < 0 || iz < 0 || ix >= c->grid[0] || iy >= c->grid[1] || iz >= c->grid[2]) return -1;
    return (int64_t)ix + (int64_t)c->grid[0] * ((int64_t)iy + (int64_t)c->grid[1] * iz);
}

static inline void clamped_cell(const pbfo_cfg *c, double x, double y, double z, int32_t *ix, int32_t *iy,
                                int32_t *iz, int is_f32) {
    if (is_f32) {
        *ix = (int32_t)clampf_glsl((float)x, 0.0f, (float)c->grid[0]);
        *iy = (int32_t)clampf_glsl((float)y, 0.0f, (float)c->grid[1]);
        *iz = (int32_t)clampf_glsl((float)z, 0.0f, (float)c->grid[2]);
    } else {
        *ix = (int32_t)(x < 0 ? 0 : (x > c->grid[0] ? c->grid[0] : x));
        *iy = (int32_t)(y < 0 ? 0 : (y > c->grid[1] ? c->grid[1] : y));
        *iz = (int32_t)(z < 0 ? 0 : (z > c->grid[2] ? c->grid[2] : z));
    }
}

/* NeighbourCellFinder::FindNeighbourCells clear (src/NeighbourCellFinder.cpp:115-126)
 * + findcells.glsl:34-53 in dispatch order (the highest gid writes last). The
 * end table is the persistent state: it is never cleared. */
static void findcells_xyz(const pbfo_cfg *c, const double *xyz, int is_f32, int32_t *start, int32_t *end) {
    const int32_t n = c->n, C = cells_of(c);
    for (int32_t k = 0; k < C; k++) start[k] = -1;
    start[0] = 0; /* gid 0: imageStore (gridtexture, ivec3 (0, 0, 0), 0) */
    int32_t px, py, pz;
    clamped_cell(c, xyz[0], xyz[1], xyz[2], &px, &py, &pz, is_f32);
    for (int32_t g = 1; g < n; g++) {
        int32_t qx, qy, qz;
        clamped_cell(c, xyz[3 * (size_t)g], xyz[3 * (size_t)g + 1], xyz[3 * (size_t)g + 2], &qx, &qy, &qz, is_f32);
        if (qx != px || qy != py || qz != pz) {
            int64_t a = cell_index_clamped(c, qx, qy, qz), b = cell_index_clamped(c, px, py, pz);
            if (a >= 0) start[a] = g;
            if (b >= 0) end[b] = g;
        }
        px = qx;
        py = qy;
        pz = qz;
    }
}

/* texelFetch of a 3D table; out-of-range coordinates read 0 on llvmpipe. */
static inline int32_t fetch3d(const pbfo_cfg *c, const int32_t *t, int32_t x, int32_t y, int32_t z) {
    if (x < 0 || y < 0 || z < 0 || x >= c->grid[0] || y >= c->grid[1] || z >= c->grid[2]) return 0;
    return t[(size_t)x + (size_t)c->grid[0] * ((size_t)y + (size_t)c->grid[1] * (size_t)z)];
}

/* neighbourcells.glsl:52-92: for the 9 (dy,dz) rows merge x-1..x+1 into one
 * range; pack start + (count << 24) with int32 wrap-around; 3 ivec4 per particle. */
static void neighbourcells_xyz(const pbfo_cfg *c, const double *xyz, int is_f32, const int32_t *start,
                               const int32_t *end, int32_t *nbr12) {
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < c->n; i++) {
        int32_t gx, gy, gz;
        if (is_f32) {
            gx = (int32_t)(float)xyz[3 * (size_t)i];
            gy = (int32_t)(float)xyz[3 * (size_t)i + 1];
            gz = (int32_t)(float)xyz[3 * (size_t)i + 2];
        } else {
            gx = (int32_t)xyz[3 * (size_t)i];
            gy = (int32_t)xyz[3 * (size_t)i + 1];
            gz = (int32_t)xyz[3 * (size_t)i + 2];
        }
        int32_t cells[9];
        for (int o = 0; o < 9; o++) {
            const int32_t dy = o / 3 - 1, dz = o % 3 - 1;
            int32_t cell = -1, entries = 0;
            for (int32_t dx = -1; dx <= 1; dx++) {
                int32_t s = fetch3d(c, start, gx + dx, gy + dy, gz + dz);
                if (cell == -1) cell = s;
                if (s != -1) entries += fetch3d(c, end, gx + dx, gy + dy, gz + dz) - s;
            }
            cells[o] = (int32_t)((uint32_t)cell + ((uint32_t)entries << 24));
        }
        int32_t *dst = nbr12 + 12 * (size_t)i;
        for (int r = 0; r < 3; r++) {
            dst[4 * r + 0] = cells[3 * r + 0];
            dst[4 * r + 1] = cells[3 * r + 1];
            dst[4 * r + 2] = cells[3 * r + 2];
            dst[4 * r + 3] = 0;
        }
    }
}

static double *xyz_of_f32(const pbfo_cfg *c, const float *rec4) {
    double *xyz = (double *)malloc(sizeof(double) * 3 * (size_t)c->n);
    for (int32_t i = 0; i < c->n; i++) {
        xyz[3 * (size_t)i] = rec4[4 * (size_t)i];
        xyz[3 * (size_t)i + 1] = rec4[4 * (size_t)i + 1];
        xyz[3 * (size_t)i + 2] = rec4[4 * (size_t)i + 2];
    }
    return xyz;
}

void pbfo_findcells(const pbfo_cfg *c, const float *sorted4, int32_t *start, int32_t *end) {
    double *xyz = xyz_of_f32(c, sorted4);
    findcells_xyz(c, xyz, 1, start, end);
    free(xyz);
}

void pbfo_neighbourcells(const pbfo_cfg *c, const float *sorted4, const int32_t *start, const int32_t *end,
                         int32_t *nbr12) {
    double *xyz = xyz_of_f32(c, sorted4);
    neighbourcells_xyz(c, xyz, 1, start, end, nbr12);
    free(xyz);
}

void pbfo_lambda(const pbfo_cfg *c, const float *rec4, const int32_t *nbr12, float *lambda) {
    lambda_f32(c, (const rec_f32 *)rec4, nbr12, lambda);
}

void pbfo_updatepos(const pbfo_cfg *c, const float *rec4_in, const float *lambda, const int32_t *nbr12,
                    float *rec4_out, int order) {
    updatepos_f32(c, (const rec_f32 *)rec4_in, lambda, nbr12, (rec_f32 *)rec4_out, order);
}

void pbfo_update(const pbfo_cfg *c, const float *rec4, float *pos4, float *vel4) {
    update_f32(c, (const rec_f32 *)rec4, pos4, vel4);
}

void pbfo_vorticity(const pbfo_cfg *c, const float *rec4, const int32_t *nbr12, float *vel4, float *vort,
                    int order) {
    vorticity_f32(c, (const rec_f32 *)rec4, nbr12, vel4, vort, order);
}

/* clearhighlight.glsl:5-8, then highlight.glsl:16-30 (dispatched by SPH::Run
 * at src/SPH.cpp:288-296, after the neighbour cells and before the solver):
 * every flag keeps only bit 0 (imageAtomicAnd by index); then each sorted
 * particle gid whose id has bit 0 set ORs 2 into the flags of the ids of every
 * candidate j != gid of its nine packed rows. There is no distance test, and
 * an index past n reads the zero record, whose id is 0. The passes only ever
 * set bit 1, which the test of bit 0 does not read, so the result does not
 * depend on the order in which the invocations run. */
void pbfo_highlight(const pbfo_cfg *c, const float *sorted4, const int32_t *nbr12, uint32_t *hl) {
    const int32_t n = c->n;
    const rec_f32 *r = (const rec_f32 *)sorted4;
    for (int32_t i = 0; i < n; i++) hl[i] &= 1u;
    for (int32_t g = 0; g < n; g++) {
        if ((hl[r[g].id] & 1u) != 1u) continue;
        for (int o = 0; o < 3; o++)
            for (int k = 0; k < 3; k++) {
                const int32_t d = nbr12[(size_t)g * 12 + o * 4 + k];
                const int32_t e = d >> 24, s0 = d & 0xFFFFFF;
                for (int32_t j = s0; j < s0 + e; j++) {
                    if (j == g) continue;
                    const int32_t id = (j >= 0 && j < n) ? r[j].id : 0;
                    hl[id] |= 2u;
                }
            }
    }
}

void pbfo_step(const pbfo_cfg *c, float *pos4, float *vel4, int32_t *end, float *vort, int order,
               pbfo_stages *st) {
    pbfo_step_hl(c, pos4, vel4, end, vort, order, st, NULL);
}

void pbfo_step_hl(const pbfo_cfg *c, float *pos4, float *vel4, int32_t *end, float *vort, int order,
                  pbfo_stages *st, uint32_t *highlight) {
    const int32_t n = c->n, C = cells_of(c);
    float *rec = (float *)malloc(sizeof(float) * 4 * (size_t)n);
    float *sorted = (float *)malloc(sizeof(float) * 4 * (size_t)n);
    float *tmp = (float *)malloc(sizeof(float) * 4 * (size_t)n);
    int32_t *start = (int32_t *)malloc(sizeof(int32_t) * (size_t)C);
    int32_t *nbr = (int32_t *)malloc(sizeof(int32_t) * 12 * (size_t)n);
    float *lambda = (float *)malloc(sizeof(float) * (size_t)n);

    pbfo_predict(c, pos4, vel4, rec);
    if (st && st->predict) memcpy(st->predict, rec, sizeof(float) * 4 * (size_t)n);
    pbfo_sort(c, rec, sorted, NULL);
    if (st && st->sorted) memcpy(st->sorted, sorted, sizeof(float) * 4 * (size_t)n);
    pbfo_findcells(c, sorted, start, end);
    if (st && st->start) memcpy(st->start, start, sizeof(int32_t) * (size_t)C);
    pbfo_neighbourcells(c, sorted, start, end, nbr);
    if (st && st->nbr12) memcpy(st->nbr12, nbr, sizeof(int32_t) * 12 * (size_t)n);
    if (highlight) pbfo_highlight(c, sorted, nbr, highlight);
    for (int it = 0; it < c->iterations; it++) {
        pbfo_lambda(c, sorted, nbr, lambda);
        if (st && st->lambda) memcpy(st->lambda + (size_t)it * n, lambda, sizeof(float) * (size_t)n);
        pbfo_updatepos(c, sorted, lambda, nbr, tmp, order);
        float *t = sorted;
        sorted = tmp;
        tmp = t;
        if (st && st->solved) memcpy(st->solved + (size_t)it * 4 * n, sorted, sizeof(float) * 4 * (size_t)n);
    }
    pbfo_update(c, sorted, pos4, vel4);
    if (c->vorticity) pbfo_vorticity(c, sorted, nbr, vel4, vort, order);
    free(rec);
    free(sorted);
    free(tmp);
    free(start);
    free(nbr);
    free(lambda);
}

/* FP64 replay: the same step with every float operation in double. */
static void step_f64_core(const pbfo_cfg *c, double *pos4, double *vel4, const int32_t *perm_fixed,
                          const int32_t *nbr_fixed, int32_t *end, double *vort) {
    const int32_t n = c->n, C = cells_of(c);
    rec_f64 *rec = (rec_f64 *)malloc(sizeof(rec_f64) * (size_t)n);
    rec_f64 *sorted = (rec_f64 *)malloc(sizeof(rec_f64) * (size_t)n);
    rec_f64 *tmp = (rec_f64 *)malloc(sizeof(rec_f64) * (size_t)n);
    double *lambda = (double *)malloc(sizeof(double) * (size_t)n);
    int32_t *nbr = NULL;
    double *vbuf = vort;
    predict_f64(c, pos4, vel4, rec);
    if (perm_fixed) {
        for (int32_t g = 0; g < n; g++) sorted[g] = rec[perm_fixed[g]];
    } else {
        uint32_t *keys = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)n);
        int32_t *perm = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
        const uint32_t m = key_mask(c);
        for (int32_t i = 0; i < n; i++) keys[i] = hash_f64(c, rec[i].x, rec[i].y, rec[i].z) & m;
        stable_sort(keys, n, c->key_bits, perm);
        for (int32_t g = 0; g < n; g++) sorted[g] = rec[perm[g]];
        free(keys);
        free(perm);
    }
    if (nbr_fixed) {
        nbr = (int32_t *)nbr_fixed;
    } else {
        nbr = (int32_t *)malloc(sizeof(int32_t) * 12 * (size_t)n);
        int32_t *start = (int32_t *)malloc(sizeof(int32_t) * (size_t)C);
        double *xyz = (double *)malloc(sizeof(double) * 3 * (size_t)n);
        for (int32_t g = 0; g < n; g++) {
            xyz[3 * (size_t)g] = sorted[g].x;
            xyz[3 * (size_t)g + 1] = sorted[g].y;
            xyz[3 * (size_t)g + 2] = sorted[g].z;
        }
        findcells_xyz(c, xyz, 0, start, end);
        neighbourcells_xyz(c, xyz, 0, start, end, nbr);
        free(xyz);
        free(start);
    }
    for (int it = 0; it < c->iterations; it++) {
        lambda_f64(c, sorted, nbr, lambda);
        updatepos_f64(c, sorted, lambda, nbr, tmp, PBFO_ORDER_JACOBI);
        rec_f64 *t = sorted;
        sorted = tmp;
        tmp = t;
    }
    update_f64(c, sorted, pos4, vel4);
    if (c->vorticity) {
        if (!vbuf) vbuf = (double *)malloc(sizeof(double) * (size_t)n);
        vorticity_f64(c, sorted, nbr, vel4, vbuf, PBFO_ORDER_JACOBI);
        if (vbuf != vort) free(vbuf);
    }
    if (!nbr_fixed) free(nbr);
    free(rec);
    free(sorted);
    free(tmp);
    free(lambda);
}

void pbfo_step_f64(const pbfo_cfg *c, double *pos4, double *vel4, int32_t *end, double *vort) {
    step_f64_core(c, pos4, vel4, NULL, NULL, end, vort);
}

void pbfo_step_fixed_f64(const pbfo_cfg *c, double *pos4, double *vel4, const int32_t *perm, const int32_t *nbr12) {
    step_f64_core(c, pos4, vel4, perm, nbr12, NULL, NULL);
}
