/* pbf_amd: C ABI of the MI355X (gfx950) simulation step of ekpyron/pbf.
 *
 * Drop-in boundary for the reference's three step classes. Every entry point
 * names the reference interface it replaces (paths relative to the reference
 * checkout). Conventions:
 *   - every function returns PBF_OK (0) or a negative pbf_status; none throws;
 *     pbf_last_error() returns the thread's last message;
 *   - handles own their device memory; pointers they return are borrowed and
 *     stay valid until the handle's next run/step (as RadixSort::GetBuffer,
 *     src/RadixSort.h:45-53) or its destruction;
 *   - "dev" pointers are HIP device pointers, "host" pointers are host memory;
 *   - work is enqueued on the handle's stream (pbf_*_set_stream; default: a
 *     stream the handle creates), in order, asynchronously to the host unless
 *     a function says it synchronizes; pbf_*_set_stream waits for the work
 *     already queued on the old stream, so the order holds across a switch;
 *   - one host thread per handle at a time; distinct handles are independent.
 *
 * Arithmetic: FP32 with the reference shaders' operation order and IEEE
 * rounding, no fused multiply-add (the library is built with
 * -ffp-contract=off); keys, sort permutation and neighbour tables are integer
 * exact. Solver passes are race-free (Jacobi): the reference updates positions
 * in place while neighbours read them (updatepos.glsl:53-54, :104).
 */
#ifndef PBF_H
#define PBF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum pbf_status {
    PBF_OK = 0,
    PBF_ERR_INVALID = -1,   /* bad argument (size, grid, null pointer) */
    PBF_ERR_HIP = -2,       /* a HIP runtime call failed */
    PBF_ERR_NOMEM = -3,     /* device allocation failed */
    PBF_ERR_LOGIC = -4,     /* e.g. an odd sort block size, src/RadixSort.cpp:41-42 */
    PBF_ERR_NODEVICE = -5   /* no HIP device */
} pbf_status;

/* Solver parameters: the eight floats of the reference's uniform block
 * (sphparams_t, src/SPH.h:252-285, defaults src/SPH.cpp:137-144), the host
 * fields of SPH (iterations src/SPH.h:197, vorticity flag :220, external force
 * src/SPH.cpp:242-244), the header constant h (src/SPH.cpp:58), and the wall
 * and gravity direction that the reference hard-codes (updatepos.glsl:98-100,
 * predictpos.glsl:32). pbf_params_default() returns the reference's values. */
typedef struct pbf_params {
    float one_over_rho_0;
    float epsilon;
    float gravity;
    float timestep;
    float tensile_instability_k;
    float tensile_instability_scale;
    float xsph_viscosity_c;
    float vorticity_epsilon;
    uint32_t iterations;
    int32_t vorticity;
    int32_t extforce;
    float h;
    float wall_lo[3];    /* clamp lower bound; reference (16, 0, 16) */
    float wall_hi[3];    /* clamp upper bound; reference GRID_SIZE - (16, 0, 16) */
    float gravity_dir[3];/* reference (0, -1, 0) */
} pbf_params;

/* Five phases of SPH::OutputTiming (src/SPH.cpp:218-240), milliseconds. */
typedef struct pbf_timing {
    double predict_ms;
    double sort_ms;
    double neighbour_cells_ms;
    double solver_ms;
    double vorticity_ms; /* XSPH/vorticity passes. The reference's last query also
                          * holds update.glsl; here the last position update of a
                          * step performs the update per particle, so its time is in
                          * solver_ms (vorticity_ms holds it only with 0 iterations) */
} pbf_timing;

typedef struct pbf_sph pbf_sph;
typedef struct pbf_radixsort pbf_radixsort;
typedef struct pbf_ncf pbf_ncf;

/* ---------------------------------------------------------------- library */
const char *pbf_last_error(void);
const char *pbf_version(void);
int pbf_device_count(void);
/* Copy between host and device memory on a stream (kind: 0 host->device,
 * 1 device->host, 2 device->device) and wait for it; for callers that hold
 * the library's device pointers but no HIP binding of their own. */
int pbf_memcpy(void *dst, const void *src, size_t bytes, int kind, void *stream);
int pbf_synchronize(void *stream);
/* Reference defaults for a grid: src/SPH.cpp:137-144 plus h = 2 (:58), wall
 * (16,0,16) (updatepos.glsl:98), gravity (0,-1,0) (predictpos.glsl:32). */
void pbf_params_default(pbf_params *out, int32_t gx, int32_t gy, int32_t gz);
/* SPH::Wpoly6, src/SPH.cpp:159-164. */
float pbf_wpoly6(float r, float h);
/* Bits the reference's sort orders by for a grid: 2*ceil(count_sortbits(C-1)/2),
 * src/RadixSort.cpp:24-30, 44, 127. */
int pbf_key_bits(int32_t gx, int32_t gy, int32_t gz);

/* Self-test of the fast correctly rounded FP32 primitives of the solver kernels
 * against the general IEEE sequences on the current device (no reference
 * counterpart): reciprocal over every float of [2^-60, 2^60) of either sign,
 * square root over every float of [2^-96, 2^8), division over div_samples
 * pseudo-random operand pairs. out[0..2]: mismatch counts, out[3..5]: cases. */
int pbf_selftest_math(uint64_t div_samples, uint64_t seed, uint64_t out[6]);
/* Self-test of the neighbour passes' workgroup renumbering (no reference
 * counterpart): launches nblocks workgroups and writes, for workgroup b, the
 * block of particles the passes of the given width (64, 128, 256, 512) give it
 * (out[2b]) and the XCD that ran it (out[2b+1]); out holds 2*nblocks words. */
int pbf_selftest_xcd_map(int32_t width, uint32_t nblocks, uint32_t *out);

/* ---------------------------------------------------------------- SPH */
/* SPH::SPH (src/SPH.h:40, src/SPH.cpp:24-151). numparticles must be a positive
 * multiple of 512, as the reference's dispatch sizes require (src/Simulation.cpp:202). */
int pbf_sph_create(pbf_sph **out, uint32_t numparticles, int32_t gx, int32_t gy, int32_t gz);
/* SPH::~SPH (src/SPH.cpp:153-157). */
int pbf_sph_destroy(pbf_sph *s);
/* SPH::Set* / UploadSPHParams (src/SPH.cpp:166-216), SetNumSolverIterations
 * (src/SPH.h:197), SetVorticityConfinementEnabled (:220), SetExternalForce
 * (src/SPH.cpp:242-244), in one call. */
int pbf_sph_set_params(pbf_sph *s, const pbf_params *p);
/* SPH::Get* (src/SPH.h:57-222). */
int pbf_sph_get_params(const pbf_sph *s, pbf_params *out);
/* Replaces Simulation::ResetParticleBuffer's glCopyBufferSubData into the
 * position and velocity buffers (src/Simulation.cpp:248-268): float4 per
 * particle in id order (w ignored, stored 0). host != 0: host pointers. */
int pbf_sph_upload_state(pbf_sph *s, const float *pos4, const float *vel4, int host);
/* Reads the float4 id-order state back (synchronizes the handle's stream when host != 0). */
int pbf_sph_download_state(pbf_sph *s, float *pos4, float *vel4, int host);
/* SPH::Run (src/SPH.cpp:246-334): one step. */
int pbf_sph_run(pbf_sph *s);
/* nsteps consecutive SPH::Run calls. */
int pbf_sph_step(pbf_sph *s, int32_t nsteps);
/* No reference counterpart: enable != 0 makes every later step one replay of
 * a hipGraph holding the step's launches (captured by the first such step on
 * the handle's stream; recaptured after pbf_sph_set_params or a stream
 * change), which removes most of the per-launch overhead of a step. Results
 * are those of eager steps bit for bit. Phase timing (pbf_sph_timing) needs
 * eager steps; per-kernel timing switches replay off while it is on. */
int pbf_sph_set_graph(pbf_sph *s, int enable);
/* No reference counterpart: the candidate threshold of the last step above
 * which particles took the dense-particle workgroups, and how many did
 * (synchronizes the handle's stream). Results do not depend on either. */
int pbf_sph_dense_info(pbf_sph *s, int32_t *threshold, int32_t *count);
/* SPH::GetPositionBuffer (src/SPH.h:49-51): float4 positions in id order,
 * stride 16 bytes, as PointSprite::SetPositionBuffer consumes them
 * (src/PointSprite.cpp:64-73). */
int pbf_sph_positions_device(pbf_sph *s, void **dev, size_t *stride);
/* SPH::GetVelocityBuffer (src/SPH.h:181-183). */
int pbf_sph_velocities_device(pbf_sph *s, void **dev, size_t *stride);
/* Render feed of the positions (the vertex data PointSprite::SetPositionBuffer,
 * src/PointSprite.cpp:64-73, and Selection::GetParticle, src/Selection.cpp:
 * 55-85, read from the position buffer): the first `components` (3 or 4)
 * floats of every particle's position, in id order, to dst with dst_stride
 * bytes between particles (12 for packed xyz, 16 for the float4 layout), e.g.
 * into a GL buffer mapped through HIP-GL interop. host != 0: dst is host
 * memory and the call synchronizes; otherwise it is enqueued on the stream. */
int pbf_sph_export_positions(pbf_sph *s, void *dst, size_t dst_stride, int32_t components, int host);
/* SPH::GetHighlightBuffer (src/SPH.h:205-207): one uint32 word per particle id
 * (stride 4), cleared to zero at creation (src/SPH.cpp:106-109). The caller
 * sets bit 0 to select a particle; every step keeps only bit 0 of each word
 * (clearhighlight.glsl:5-8) and sets bit 1 on every neighbour candidate of a
 * selected particle (highlight.glsl:16-30), as SPH::Run does at
 * src/SPH.cpp:288-296; the renderer reads the words (PointSprite::
 * SetHighlightBuffer, src/Simulation.cpp:79). */
int pbf_sph_highlight_device(pbf_sph *s, void **dev, size_t *stride);
/* The highlight words read into (get) or written from (set) uint32[n]; set with
 * src == NULL clears them all, as Simulation::ResetParticleBuffer does with
 * glClearBufferData (src/Simulation.cpp:270-272). host: synchronizes. */
int pbf_sph_get_highlight(pbf_sph *s, uint32_t *dst, int host);
int pbf_sph_set_highlight(pbf_sph *s, const uint32_t *src, int host);
/* The selection toggle of Simulation::OnMouseDown (src/Simulation.cpp:163-193):
 * a word above zero becomes 0, otherwise 1; *now (if not null) receives the
 * new word (synchronizes the handle's stream). */
int pbf_sph_toggle_highlight(pbf_sph *s, uint32_t id, uint32_t *now);
/* SPH::OutputTiming (src/SPH.cpp:218-240): phase times of the last step
 * (synchronizes the handle's stream). */
int pbf_sph_timing(pbf_sph *s, pbf_timing *out);
/* Per-kernel HIP-event timing of the handle's launches (no reference
 * counterpart; finer than OutputTiming's five phases). While enabled, every
 * kernel launch of pbf_sph_run is bracketed by two events on the handle's
 * stream; pbf_sph_kernel_times synchronizes, and returns and resets per-kernel
 * totals. Kernel ids: see pbf_kernel_name. The sort's time includes the
 * sorted records (its last pass writes them; "gather" has no launches), and
 * the last position update of a step includes the closing update ("update"
 * has launches only in steps without solver iterations). */
#define PBF_NUM_KERNEL_IDS 11
typedef struct pbf_kernel_time {
    double total_ms;
    uint32_t launches;
} pbf_kernel_time;
int pbf_sph_set_kernel_timing(pbf_sph *s, int enable);
int pbf_sph_kernel_times(pbf_sph *s, pbf_kernel_time out[PBF_NUM_KERNEL_IDS]);
const char *pbf_kernel_name(int id);
/* Stream of the handle (hipStream_t); NULL restores the handle's own stream. */
int pbf_sph_set_stream(pbf_sph *s, void *stream);
void *pbf_sph_get_stream(const pbf_sph *s);
/* Persistent end table of the step's NeighbourCellFinder (never cleared,
 * src/NeighbourCellFinder.cpp:113-126): read into or write from int32[gx*gy*gz]. */
int pbf_sph_get_end_table(pbf_sph *s, int32_t *dst, int host);
int pbf_sph_set_end_table(pbf_sph *s, const int32_t *src, int host);
/* Neighbour table of the last step in the reference's layout, 3 ivec4 per
 * particle in sorted order (NeighbourCellFinder::GetResult, src/NeighbourCellFinder.h:58). */
int pbf_sph_neighbour_table(pbf_sph *s, int32_t *dst12, int host);
/* Sorted particle records of the last step, {x, y, z, id bits} per particle
 * (RadixSort::GetBuffer after SPH::Run, src/RadixSort.cpp:118-122). */
int pbf_sph_sorted_records(pbf_sph *s, float *dst4, int host);

/* ---------------------------------------------------------------- RadixSort */
/* RadixSort::RadixSort (src/RadixSort.h:41, src/RadixSort.cpp:32-109): sorts
 * blocksize*numblocks float4 records {x, y, z, id bits} by the grid-cell key of
 * their position. An odd blocksize fails with PBF_ERR_LOGIC like the reference. */
int pbf_radixsort_create(pbf_radixsort **out, uint32_t blocksize, uint32_t numblocks, int32_t gx, int32_t gy,
                         int32_t gz);
int pbf_radixsort_destroy(pbf_radixsort *r);
/* RadixSort::GetBuffer (src/RadixSort.cpp:118-122): the record buffer to fill
 * before run and holding the sorted records after it. */
int pbf_radixsort_buffer(pbf_radixsort *r, void **dev);
/* RadixSort::Run (src/RadixSort.cpp:124-133). */
int pbf_radixsort_run(pbf_radixsort *r);
/* Sorted-slot -> input-index permutation of the last run (int32[n], device). */
int pbf_radixsort_permutation(pbf_radixsort *r, void **dev);
int pbf_radixsort_set_stream(pbf_radixsort *r, void *stream);
void *pbf_radixsort_get_stream(const pbf_radixsort *r);

/* ---------------------------------------------------------------- NeighbourCellFinder */
/* NeighbourCellFinder::NeighbourCellFinder (src/NeighbourCellFinder.h:42,
 * src/NeighbourCellFinder.cpp:24-100). */
int pbf_ncf_create(pbf_ncf **out, uint32_t numparticles, int32_t gx, int32_t gy, int32_t gz);
int pbf_ncf_destroy(pbf_ncf *f);
/* NeighbourCellFinder::FindNeighbourCells (src/NeighbourCellFinder.cpp:113-148):
 * from sorted float4 records (device). */
int pbf_ncf_find(pbf_ncf *f, const void *dev_sorted_records);
/* NeighbourCellFinder::GetResult (src/NeighbourCellFinder.cpp:108-111): 3
 * ivec4 per particle, each component start + (count << 24), w = 0 (device). */
int pbf_ncf_result(pbf_ncf *f, void **dev12);
/* The start table (cleared to -1 every call) and the persistent end table as
 * int32[gx*gy*gz]. */
int pbf_ncf_get_tables(pbf_ncf *f, int32_t *start, int32_t *end, int host);
int pbf_ncf_set_end_table(pbf_ncf *f, const int32_t *end, int host);
int pbf_ncf_set_stream(pbf_ncf *f, void *stream);
void *pbf_ncf_get_stream(const pbf_ncf *f);

/* ---------------------------------------------------------------- batched scenes
 * No reference counterpart: many independent scenes of uneven particle counts
 * advanced together, one launch per stage for all of them. Each scene has its
 * own grid, walls, gravity direction and XSPH coefficient; the other constants
 * are shared. Per scene, every result equals a pbf_sph handle created with that
 * scene's grid and parameters, bit for bit (same arithmetic, same neighbour
 * order): records carry scene-local ids and the tables scene-local indices.
 * State arrays hold the scenes one after another, each in id order. */
typedef struct pbf_scene_desc {
    uint32_t n;             /* particles of the scene, a positive multiple of 512 as for pbf_sph_create */
    int32_t grid[3];
    float wall_lo[3];
    float wall_hi[3];
    float gravity_dir[3];
    float xsph_viscosity_c;
} pbf_scene_desc;

/* Launch shapes the autotuner picks from: workgroup size of the neighbour
 * passes (64, 128, 256 or 512), and workgroup size (64 .. 1024) and particles per
 * thread (1 .. 4) of the streaming passes (predict, cell table, update); the
 * LDS list depth of the candidate-scanning passes (32 or 56 entries per
 * thread), and persistent neighbour launches (1: as many workgroups as fit on
 * the device, each looping over the tiles; 0: one workgroup per tile). */
typedef struct pbf_batch_tuning {
    int32_t neighbour_block;
    int32_t stream_block;
    int32_t stream_ppt;
    int32_t list_depth;
    int32_t persistent;
    double step_ms;         /* measured mean step time of this choice (autotune output) */
} pbf_batch_tuning;

typedef struct pbf_batch pbf_batch;

/* shared: h, the eight uniform-block floats except the XSPH coefficient,
 * iterations, vorticity and extforce flags; its walls, gravity direction and
 * XSPH coefficient are ignored (each scene has its own). At most 65536 scenes. */
int pbf_batch_create(pbf_batch **out, uint32_t nscenes, const pbf_scene_desc *scenes, const pbf_params *shared);
int pbf_batch_destroy(pbf_batch *b);
/* Total particle count and, when offsets is not null, the first index of each
 * scene in the state arrays (nscenes + 1 entries, the last one the total). */
int pbf_batch_layout(const pbf_batch *b, uint32_t *total, uint32_t *offsets);
int pbf_batch_upload_state(pbf_batch *b, const float *pos4, const float *vel4, int host);
int pbf_batch_download_state(pbf_batch *b, float *pos4, float *vel4, int host);
/* nsteps steps of every scene, asynchronous on the batch's stream. */
int pbf_batch_step(pbf_batch *b, int32_t nsteps);
/* enable != 0: the first step after this call captures the step's launches in
 * a hipGraph, and every step replays it (one launch call per step). */
int pbf_batch_set_graph(pbf_batch *b, int enable);
int pbf_batch_set_tuning(pbf_batch *b, const pbf_batch_tuning *t);
int pbf_batch_get_tuning(const pbf_batch *b, pbf_batch_tuning *out);
/* Times launch shapes on the current state (steps_per_trial steps each, state
 * restored afterwards) and keeps the fastest: first every neighbour shape
 * (workgroup size x list depth x persistent) at the current streaming shape,
 * then every streaming shape (workgroup size x particles per thread) with the
 * best neighbour shape. results, when not null, receives up to max_results
 * trials; *count their number. */
int pbf_batch_autotune(pbf_batch *b, int32_t steps_per_trial, pbf_batch_tuning *best, pbf_batch_tuning *results,
                       int32_t max_results, int32_t *count);
/* The scenes' persistent end tables, one after another (int32 per cell). */
int pbf_batch_get_end_tables(pbf_batch *b, int32_t *dst, int host);
int pbf_batch_set_end_tables(pbf_batch *b, const int32_t *src, int host);
int pbf_batch_set_stream(pbf_batch *b, void *stream);
void *pbf_batch_get_stream(const pbf_batch *b);

/* ---------------------------------------------------------------- adjoint
 * No reference counterpart: reverse mode through whole steps (predict, the
 * solver iterations, update, XSPH and vorticity confinement) for gradients of
 * a loss on the final positions with respect to the initial velocities, the
 * XSPH coefficient and the vorticity epsilon. Each step's sort order and
 * neighbour table are held fixed: they are computed by the plain step's
 * kernels from the state and not differentiated. The clamp to the walls has a
 * zero derivative where it clamps; N = eta/|eta| a zero derivative at eta = 0.
 * States are checkpointed every checkpoint_every steps and each segment's
 * states and intermediate values are recomputed from its checkpoint during the
 * reverse pass, except the last segment's, which the forward pass keeps (one
 * segment: nothing is recomputed). fp64 != 0: every
 * state, tape and adjoint in double (for finite-difference checks), and the
 * element type of every state array below is double instead of float. Arrays
 * hold 4 elements per particle (x, y, z, unused) in id order. */
typedef struct pbf_adjoint pbf_adjoint;
int pbf_adjoint_create(pbf_adjoint **out, uint32_t numparticles, int32_t gx, int32_t gy, int32_t gz,
                       const pbf_params *p, int32_t steps, int32_t checkpoint_every, int32_t fp64);
int pbf_adjoint_destroy(pbf_adjoint *a);
/* Values of the two differentiated coefficients for the next forward (defaults
 * from the parameters given at creation). */
int pbf_adjoint_set_coefficients(pbf_adjoint *a, double xsph_c, double vorticity_epsilon);
/* steps steps from (pos4, vel4), keeping the checkpoints and every step's
 * sort order, neighbour table, clamp states and the flags of the particles whose
 * rows are exactly their cells' contents (41 + iterations bytes per particle
 * and step), which the reverse pass and replays hold fixed. */
int pbf_adjoint_forward(pbf_adjoint *a, const void *pos4, const void *vel4, int host);
int pbf_adjoint_positions(pbf_adjoint *a, void *pos4, int host);
int pbf_adjoint_velocities(pbf_adjoint *a, void *vel4, int host);
/* From dL/d(final positions) to dL/d(initial velocities) and the two
 * coefficient derivatives (synchronizes the handle's stream). */
int pbf_adjoint_backward(pbf_adjoint *a, const void *grad_pos4, int host, void *grad_vel4, double *grad_xsph_c,
                         double *grad_vorticity_epsilon);
/* steps steps from (pos4, vel4) with the structure kept by the last forward
 * (the map whose derivative backward computes); final positions to pos_out4. */
int pbf_adjoint_replay(pbf_adjoint *a, const void *pos4, const void *vel4, int host, void *pos_out4);
/* pbf_adjoint_replay with every clamp to the walls (updatepos.glsl:98-100)
 * taking the decision the last forward took at that step, iteration, particle
 * and axis (inside, lower wall, upper wall): the smooth piece of the map that
 * backward differentiates. clamp_flips: how many of those decisions the clamp
 * would have taken otherwise on this replay's path (synchronizes). */
int pbf_adjoint_replay_frozen(pbf_adjoint *a, const void *pos4, const void *vel4, int host, void *pos_out4,
                              uint64_t *clamp_flips);
/* FP64 handles: the directional derivative of pbf_adjoint_replay's final
 * positions along (dvel4, dxsph_c, dvorticity_epsilon), by forward-mode
 * differentiation (dual numbers) of the same forward kernels, into dpos_out4
 * (and the positions into pos_out4 unless null): a second derivative of the
 * map, independent of backward's hand-written reverse kernels. */
int pbf_adjoint_jvp(pbf_adjoint *a, const void *pos4, const void *vel4, const void *dvel4, double dxsph_c,
                    double dvorticity_epsilon, int host, void *pos_out4, void *dpos_out4);
int pbf_adjoint_set_stream(pbf_adjoint *a, void *stream);
void *pbf_adjoint_get_stream(const pbf_adjoint *a);

/* ---------------------------------------------------------------- stages
 * Single shader dispatches of SPH::Run on caller memory (device pointers,
 * reference layouts), so each stage can be checked alone. n is a multiple of
 * 256 (the reference's BLOCKSIZE dispatch); nbr12 is the GetResult layout.
 * The neighbour stages allocate their scratch (the [9][n] table, quantized
 * positions, the packed neighbour records the pair passes read) with
 * hipMalloc and free it before returning, which synchronizes the stream;
 * pbf_stage_vorticity_noalloc below does neither. */
/* predictpos.glsl:18-38: rec4_out[i] = {x + dt v', i}. */
int pbf_stage_predict(uint32_t n, const int32_t grid[3], const pbf_params *p, const void *pos4, const void *vel4,
                      void *rec4_out, void *stream);
/* calclambda.glsl:66-103. */
int pbf_stage_lambda(uint32_t n, const int32_t grid[3], const pbf_params *p, const void *rec4, const void *nbr12,
                     void *lambda_out, void *stream);
/* updatepos.glsl:42-105, race-free: reads rec4_in, writes rec4_out. */
int pbf_stage_updatepos(uint32_t n, const int32_t grid[3], const pbf_params *p, const void *rec4_in,
                        const void *lambda, const void *nbr12, void *rec4_out, void *stream);
/* update.glsl:16-28: pos4/vel4 (id order) updated in place. */
int pbf_stage_update(uint32_t n, const pbf_params *p, const void *rec4, void *pos4, void *vel4, void *stream);
/* vorticity.glsl:34-86 as two race-free passes: vel4 (id order) updated in
 * place, |omega| per sorted slot to vort_out. */
int pbf_stage_vorticity(uint32_t n, const int32_t grid[3], const pbf_params *p, const void *rec4,
                        const void *nbr12, void *vel4, void *vort_out, void *stream);

/* No counterpart in the reference: pbf_stage_vorticity for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc). Same arguments and results; it allocates
 * and frees nothing and does not synchronize the stream. Its scratch (the
 * [9][n] neighbour table, four float4 arrays and the quantized positions,
 * 108 bytes per particle) is
 * one fixed device array of the library per GPU, held until the process ends,
 * so n is at most PBF_NOALLOC_MAX_N (PBF_ERR_INVALID beyond). Calls that
 * overlap on one GPU share that array and race; order them on one stream. */
#define PBF_NOALLOC_MAX_N 131072u
int pbf_stage_vorticity_noalloc(uint32_t n, const int32_t grid[3], const pbf_params *p, const void *rec4,
                                const void *nbr12, void *vel4, void *vort_out, void *stream);

#ifdef __cplusplus
}
#endif
#endif
