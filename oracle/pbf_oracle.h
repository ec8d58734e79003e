/* CPU restatement of the reference's simulation step: TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * this library, and only as the checker (or the timed CPU baseline). The
 * product path (pbf_amd/, libpbf_amd.so) never calls it.
 *
 * Every function restates one reference shader or host routine op for op in
 * FP32 without contraction (the reference's shaders run on llvmpipe without
 * fused multiply-add); the *_f64 entry points replay the same algorithm in
 * FP64. Citations are to files of ekpyron/pbf.
 *
 * Record layout: 16 bytes per particle, {x, y, z, id} with the int id stored
 * in the fourth float slot by bit pattern (ParticleKey, predictpos.glsl:3-6).
 */
#ifndef PBF_ORACLE_H
#define PBF_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pbfo_cfg {
    int32_t n;
    int32_t grid[3];
    float h;
    /* one_over_rho_0, epsilon, gravity, timestep, tensile k, tensile scale,
     * xsph c, vorticity epsilon: sphparams_t, src/SPH.h:252-285 */
    float p[8];
    int32_t iterations;
    int32_t vorticity;
    int32_t extforce;
    float wall_lo[3]; /* clamp bounds of updatepos.glsl:98-100: lo = wall, hi = GRID - wall */
    float wall_hi[3];
    float gdir[3];    /* gravity direction, (0,-1,0) in predictpos.glsl:32 */
    int32_t key_bits; /* bits the sort orders by: 2*ceil(numbits/2), src/RadixSort.cpp:44,127 */
} pbfo_cfg;

/* Solver orders: race-free (Jacobi) or the reference's in-place dispatch order
 * as llvmpipe executes it with one thread (workgroups of 256 in order, 8-lane
 * SIMD chunks, loads of a chunk before its stores). */
enum { PBFO_ORDER_JACOBI = 0, PBFO_ORDER_REFERENCE = 1 };

void pbfo_predict(const pbfo_cfg *c, const float *pos4, const float *vel4, float *rec4);
void pbfo_keys(const pbfo_cfg *c, const float *rec4, uint32_t *keys);
void pbfo_sort(const pbfo_cfg *c, const float *rec4_in, float *rec4_out, int32_t *perm);
void pbfo_findcells(const pbfo_cfg *c, const float *sorted4, int32_t *start, int32_t *end);
void pbfo_neighbourcells(const pbfo_cfg *c, const float *sorted4, const int32_t *start, const int32_t *end,
                         int32_t *nbr12);
void pbfo_lambda(const pbfo_cfg *c, const float *rec4, const int32_t *nbr12, float *lambda);
void pbfo_updatepos(const pbfo_cfg *c, const float *rec4_in, const float *lambda, const int32_t *nbr12,
                    float *rec4_out, int order);
void pbfo_update(const pbfo_cfg *c, const float *rec4, float *pos4, float *vel4);
void pbfo_vorticity(const pbfo_cfg *c, const float *rec4, const int32_t *nbr12, float *vel4, float *vort,
                    int order);

/* The two highlight passes of SPH::Run (src/SPH.cpp:288-296): clear every flag
 * but bit 0, then mark (|= 2) the candidates of each particle with bit 0 set.
 * sorted4 and nbr12 are the step's sorted records and packed table; hl holds
 * one flag word per particle id. */
void pbfo_highlight(const pbfo_cfg *c, const float *sorted4, const int32_t *nbr12, uint32_t *hl);

/* One whole step (SPH::Run, src/SPH.cpp:246-334). pos4/vel4 are in id order;
 * end (grid cells) and vort (n) are the state the reference keeps between
 * steps. Optional stage outputs may be NULL. pbfo_step_hl also runs the
 * highlight passes on the flag words in highlight (n, by id) unless it is
 * NULL; pbfo_step leaves them out. */
typedef struct pbfo_stages {
    float *predict;    /* n*4 */
    float *sorted;     /* n*4 */
    int32_t *start;    /* cells */
    int32_t *nbr12;    /* n*12 */
    float *lambda;     /* iterations*n */
    float *solved;     /* iterations*n*4 */
} pbfo_stages;

void pbfo_step(const pbfo_cfg *c, float *pos4, float *vel4, int32_t *end, float *vort, int order,
               pbfo_stages *stages);
void pbfo_step_hl(const pbfo_cfg *c, float *pos4, float *vel4, int32_t *end, float *vort, int order,
                  pbfo_stages *stages, uint32_t *highlight);

/* FP64 replay of one step, race-free order, same neighbour-table semantics. */
void pbfo_step_f64(const pbfo_cfg *c, double *pos4, double *vel4, int32_t *end, double *vort);

/* FP64 replay with a caller-fixed neighbour structure: the sorted order perm
 * (sorted slot -> id) and the packed table nbr12 of one step, as used for the
 * finite-difference checks of the adjoint. */
void pbfo_step_fixed_f64(const pbfo_cfg *c, double *pos4, double *vel4, const int32_t *perm,
                         const int32_t *nbr12);

int pbfo_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
