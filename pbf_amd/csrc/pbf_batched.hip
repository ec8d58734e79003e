// Batched form of the step (include/pbf.h, pbf_batch_*): many independent
// scenes in one set of arrays, one launch per stage for all of them.
//
// Layout: scene s owns slots [off_s, off_s + n_s) of every per-particle array,
// in id order for the state and in sorted order for the records, and cells
// [cell_base_s, cell_base_s + C_s) of the cell tables. Keys are
// (s << key_bits_max) | the scene's own cell key, so one stable LSD sort
// orders every scene by its own key while keeping scenes apart (the input is
// grouped by scene already). Records keep scene-local ids and the tables
// scene-local indices, which makes every array restricted to one scene equal
// to the plain handle's array for that scene.
//
// A whole step can be captured once in a hipGraph and replayed: all launch
// arguments are fixed per step (the end-table tag is a device word advanced by
// the clear kernel), and the step records no memset or host synchronisation.
#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "pbf_internal.h"

using namespace pbf;

struct pbf_batch {
    uint32_t nscenes = 0;
    uint32_t n = 0;               // total particles
    size_t cells = 0;             // total cells
    int key_bits = 0;             // composite key bits
    std::vector<pbf_scene_desc> desc;
    std::vector<uint32_t> offsets;  // nscenes + 1
    std::vector<size_t> cell_base;  // nscenes + 1
    pbf_params params;
    SolverConsts consts;
    LaunchShape shape;
    hipStream_t own_stream = nullptr, stream = nullptr;
    Seg *segs = nullptr;
    uint16_t *scene_of = nullptr;
    float4 *pos = nullptr, *vel = nullptr, *xs = nullptr, *rec0 = nullptr, *rec_a = nullptr, *rec_b = nullptr;
    uint2 *qpos = nullptr;  // quantized positions of the current sorted records
    float4 *vel_sorted = nullptr, *vxsph = nullptr, *omega = nullptr;
    float *lambda = nullptr;
    float4 *nrec = nullptr;  // neighbour records: position and lambda, then |omega|
    uint32_t *keys = nullptr, *perm = nullptr;
    uint16_t *list_entries = nullptr, *list_count = nullptr;
    int32_t *start = nullptr, *nbr9 = nullptr;
    unsigned long long *end64 = nullptr;
    uint32_t *epoch = nullptr;
    DenseList dense;  // the step's particles with the most candidates (all scenes)
    RadixScratch rs;
    bool use_graph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;

    DomMany dom() const { return DomMany{segs, scene_of}; }

    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    void release() {
        drop_graph();
        dfree(segs);
        dfree(scene_of);
        dfree(pos);
        dfree(vel);
        dfree(xs);
        dfree(rec0);
        dfree(rec_a);
        dfree(rec_b);
        dfree(qpos);
        dfree(vel_sorted);
        dfree(vxsph);
        dfree(omega);
        dfree(lambda);
        dfree(nrec);
        dfree(keys);
        dfree(perm);
        dfree(list_entries);
        dfree(list_count);
        dfree(start);
        dfree(nbr9);
        dfree(end64);
        dfree(epoch);
        dfree(dense.cand);
        dfree(dense.levels);
        dfree(dense.threshold);
        radix_free(rs);
        if (own_stream) (void)hipStreamDestroy(own_stream);
    }

    // The launches of one step, in SPH::Run's order (src/SPH.cpp:246-334).
    void enqueue_step(hipStream_t st) {
        const SolverConsts &c = consts;
        const DomMany d = dom();
        const int32_t nn = (int32_t)n;
        launch_predict(c, d, shape, pos, vel, xs, keys, st);
        // the last pass writes the records and their quantized positions
        radix_sort_perm(rs, keys, nn, key_bits, nullptr, st, xs, rec0, qpos, c.qscale);
        launch_begin_cells(start, cells, epoch, st, &dense);
        launch_findcells(c, d, shape, rec0, start, end64, epoch, st);
        launch_neighbourcells(c, d, shape, rec0, start, end64, nbr9, st, &dense);
        PairList pl;
        pl.entries = list_entries;
        pl.count = list_count;
        pl.cap = kListCap;
        // the last position update also performs the closing update
        StepUpdate up;
        up.pos = pos;
        up.vel = vel;
        up.vel_sorted = params.vorticity ? vel_sorted : nullptr;
        const float4 *cur = rec0;
        for (uint32_t it = 0; it < params.iterations; it++) {
            float4 *dst = (it & 1) ? rec_b : rec_a;
            launch_lambda(c, d, shape, cur, qpos, nbr9, lambda, nrec, pl, st, &dense);
            launch_updatepos(c, d, shape, cur, lambda, nrec, nbr9, pl, dst, qpos, st, &dense,
                             it + 1 == params.iterations ? &up : nullptr);
            cur = dst;
        }
        if (params.iterations == 0) launch_update(c, d, shape, cur, pos, vel, up.vel_sorted, st);
        if (params.vorticity) {
            launch_vorticity_a(c, d, shape, cur, qpos, nbr9, vel_sorted, vel, vxsph, omega, nullptr, nrec, pl, st,
                               &dense);
            launch_vorticity_b(c, d, shape, cur, nbr9, pl, vxsph, omega, nrec, vel, st, &dense);
        }
    }
};

namespace {

int check_desc(const pbf_scene_desc &s) {
    if (s.n == 0 || s.n % 512 != 0 || s.n >= (1u << 24)) return invalid("scene n must be a positive multiple of 512 below 2^24");
    if (!valid_grid(s.grid[0], s.grid[1], s.grid[2])) return invalid("bad scene grid");
    return PBF_OK;
}

int bits_for(uint32_t v) {  // bits to hold 0 .. v-1
    int b = 0;
    while (b < 32 && (1ull << b) < v) b++;
    return b;
}

}  // namespace

extern "C" {

int pbf_batch_create(pbf_batch **out, uint32_t nscenes, const pbf_scene_desc *scenes, const pbf_params *shared) {
    if (!out) return invalid("null output handle");
    *out = nullptr;
    if (nscenes == 0 || nscenes > 65536 || !scenes) return invalid("need 1 to 65536 scene descriptions");
    int rc = check_params(shared);
    if (rc) return rc;
    uint64_t total = 0, cells = 0;
    int kb = 0;
    for (uint32_t s = 0; s < nscenes; s++) {
        if ((rc = check_desc(scenes[s]))) return rc;
        total += scenes[s].n;
        cells += (uint64_t)scenes[s].grid[0] * scenes[s].grid[1] * scenes[s].grid[2];
        kb = std::max(kb, key_bits_of(scenes[s].grid));
    }
    if (total >= (1u << 31)) return invalid("total particle count must stay below 2^31");
    if (cells >= (1ull << 31)) return invalid("total cell count must stay below 2^31");
    const int sb = bits_for(nscenes);
    if (kb + sb > 32) return invalid("scene count and grids need more than 32 key bits");
    if (pbf_device_count() < 1) {
        set_error("no HIP device");
        return PBF_ERR_NODEVICE;
    }
    pbf_batch *b = new (std::nothrow) pbf_batch();
    if (!b) return invalid("host allocation failed");
    b->nscenes = nscenes;
    b->n = (uint32_t)total;
    b->cells = (size_t)cells;
    b->key_bits = kb + sb;
    b->desc.assign(scenes, scenes + nscenes);
    b->params = *shared;
    b->consts = make_consts(b->n, scenes[0].grid, *shared, kb);
    b->offsets.resize(nscenes + 1);
    b->cell_base.resize(nscenes + 1);
    std::vector<Seg> segs(nscenes);
    std::vector<uint16_t> scene_of(b->n);
    uint32_t off = 0;
    size_t cb = 0;
    for (uint32_t s = 0; s < nscenes; s++) {
        const pbf_scene_desc &d = scenes[s];
        b->offsets[s] = off;
        b->cell_base[s] = cb;
        Seg &g = segs[s];
        g.off = (int32_t)off;
        g.n = (int32_t)d.n;
        g.gx = d.grid[0];
        g.gy = d.grid[1];
        g.gz = d.grid[2];
        g.cell_base = (int32_t)cb;
        const int kbs = key_bits_of(d.grid);
        g.key_mask = kbs >= 32 ? 0xFFFFFFFFu : ((1u << kbs) - 1u);
        g.key_hi = (uint32_t)((uint64_t)s << kb);
        for (int a = 0; a < 3; a++) {
            g.lo[a] = d.wall_lo[a];
            g.hi[a] = d.wall_hi[a];
            g.gdir[a] = d.gravity_dir[a];
        }
        g.xsph_c = d.xsph_viscosity_c;
        std::fill(scene_of.begin() + off, scene_of.begin() + off + d.n, (uint16_t)s);
        off += d.n;
        cb += (size_t)d.grid[0] * d.grid[1] * d.grid[2];
    }
    b->offsets[nscenes] = off;
    b->cell_base[nscenes] = cb;
    const size_t n = b->n;
    hipError_t e = hipSuccess;
#define TRY(x)                              \
    if ((e = (x)) != hipSuccess) {          \
        int r_ = hip_fail(e, #x);           \
        b->release();                       \
        delete b;                           \
        return r_;                          \
    }
    TRY(hipStreamCreateWithFlags(&b->own_stream, hipStreamNonBlocking));
    b->stream = b->own_stream;
    TRY(dalloc(&b->segs, nscenes));
    TRY(dalloc(&b->scene_of, n));
    TRY(dalloc(&b->pos, n));
    TRY(dalloc(&b->vel, n));
    TRY(dalloc(&b->xs, n));
    TRY(dalloc(&b->rec0, n));
    TRY(dalloc(&b->rec_a, n));
    TRY(dalloc(&b->rec_b, n));
    TRY(dalloc(&b->qpos, n));
    TRY(dalloc(&b->vel_sorted, n));
    TRY(dalloc(&b->vxsph, n));
    TRY(dalloc(&b->omega, n));
    TRY(dalloc(&b->lambda, n));
    TRY(dalloc(&b->nrec, n));
    TRY(dalloc(&b->keys, n));
    TRY(dalloc(&b->perm, n));
    TRY(dalloc(&b->list_entries, (size_t)kListRows * n));
    TRY(dalloc(&b->list_count, n));
    TRY(dalloc(&b->start, b->cells));
    TRY(dalloc(&b->nbr9, (size_t)9 * n));
    TRY(dalloc(&b->end64, b->cells));
    TRY(dalloc(&b->epoch, 1));
    TRY(dalloc(&b->dense.cand, n));
    TRY(dalloc(&b->dense.levels, (size_t)kDenseLevels * kDenseCopies));
    TRY(dalloc(&b->dense.threshold, 2));
    TRY(hipMemset(b->dense.levels, 0, (size_t)kDenseLevels * kDenseCopies * sizeof(uint32_t)));
    if (!radix_alloc(b->rs, (int32_t)n)) TRY(hipErrorOutOfMemory);
    TRY(hipMemcpy(b->segs, segs.data(), nscenes * sizeof(Seg), hipMemcpyHostToDevice));
    TRY(hipMemcpy(b->scene_of, scene_of.data(), n * sizeof(uint16_t), hipMemcpyHostToDevice));
    TRY(hipMemset(b->pos, 0, n * sizeof(float4)));
    TRY(hipMemset(b->vel, 0, n * sizeof(float4)));
    TRY(hipMemset(b->end64, 0, b->cells * sizeof(unsigned long long)));  // as a fresh handle's end texture
    TRY(hipMemset(b->epoch, 0, sizeof(uint32_t)));
#undef TRY
    *out = b;
    return PBF_OK;
}

int pbf_batch_destroy(pbf_batch *b) {
    if (!b) return PBF_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    b->release();
    delete b;
    return PBF_OK;
}

int pbf_batch_layout(const pbf_batch *b, uint32_t *total, uint32_t *offsets) {
    if (!b) return invalid("null batch handle");
    if (total) *total = b->n;
    if (offsets) std::copy(b->offsets.begin(), b->offsets.end(), offsets);
    return PBF_OK;
}

int pbf_batch_upload_state(pbf_batch *b, const float *pos4, const float *vel4, int host) {
    if (!b || !pos4 || !vel4) return invalid("null argument");
    const size_t bytes = (size_t)b->n * sizeof(float4);
    const hipMemcpyKind k = host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
    PBF_HIP(hipMemcpyAsync(b->pos, pos4, bytes, k, b->stream));
    PBF_HIP(hipMemcpyAsync(b->vel, vel4, bytes, k, b->stream));
    if (host) PBF_HIP(hipStreamSynchronize(b->stream));
    return PBF_OK;
}

int pbf_batch_download_state(pbf_batch *b, float *pos4, float *vel4, int host) {
    if (!b) return invalid("null batch handle");
    const size_t bytes = (size_t)b->n * sizeof(float4);
    const hipMemcpyKind k = host ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    if (pos4) PBF_HIP(hipMemcpyAsync(pos4, b->pos, bytes, k, b->stream));
    if (vel4) PBF_HIP(hipMemcpyAsync(vel4, b->vel, bytes, k, b->stream));
    if (host) PBF_HIP(hipStreamSynchronize(b->stream));
    return PBF_OK;
}

int pbf_batch_step(pbf_batch *b, int32_t nsteps) {
    if (!b) return invalid("null batch handle");
    if (nsteps < 0) return invalid("negative step count");
    if (nsteps == 0) return PBF_OK;
    if (!b->use_graph) {
        for (int32_t k = 0; k < nsteps; k++) b->enqueue_step(b->stream);
        PBF_HIP(hipGetLastError());
        return PBF_OK;
    }
    if (!b->exec) {
        PBF_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        b->enqueue_step(b->stream);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (e != hipSuccess) return hip_fail(e, "pbf_batch_step: capture");
        b->graph = g;
        e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            b->drop_graph();
            return hip_fail(e, "pbf_batch_step: instantiate");
        }
    }
    for (int32_t k = 0; k < nsteps; k++) PBF_HIP(hipGraphLaunch(b->exec, b->stream));
    return PBF_OK;
}

int pbf_batch_set_graph(pbf_batch *b, int enable) {
    if (!b) return invalid("null batch handle");
    if (b->stream) PBF_HIP(hipStreamSynchronize(b->stream));
    b->drop_graph();
    b->use_graph = enable != 0;
    return PBF_OK;
}

int pbf_batch_set_tuning(pbf_batch *b, const pbf_batch_tuning *t) {
    if (!b || !t) return invalid("null argument");
    LaunchShape ls;
    ls.nb = t->neighbour_block;
    ls.sb = t->stream_block;
    ls.ppt = t->stream_ppt;
    ls.lc = t->list_depth;
    ls.persistent = t->persistent;
    if (!valid_shape(ls))
        return invalid("tuning: neighbour_block 64/128/256/512, stream_block a power of two in 64..1024, ppt 1..4, "
                       "list_depth 32 or 56, persistent 0 or 1");
    if (b->stream) PBF_HIP(hipStreamSynchronize(b->stream));
    b->drop_graph();
    b->shape = ls;
    return PBF_OK;
}

int pbf_batch_get_tuning(const pbf_batch *b, pbf_batch_tuning *out) {
    if (!b || !out) return invalid("null argument");
    out->neighbour_block = b->shape.nb;
    out->stream_block = b->shape.sb;
    out->stream_ppt = b->shape.ppt;
    out->list_depth = b->shape.lc;
    out->persistent = b->shape.persistent;
    out->step_ms = 0.0;
    return PBF_OK;
}

int pbf_batch_autotune(pbf_batch *b, int32_t steps_per_trial, pbf_batch_tuning *best, pbf_batch_tuning *results,
                       int32_t max_results, int32_t *count) {
    if (!b) return invalid("null batch handle");
    if (steps_per_trial < 1) return invalid("steps_per_trial must be at least 1");
    hipStream_t st = b->stream;
    const size_t n = b->n;
    // Snapshot of everything a step reads and writes persistently: state, end
    // table and its tag. Every trial starts from it; the snapshot is restored
    // at the end, so tuning leaves the simulation where it was.
    float4 *snap = nullptr;
    unsigned long long *snap_end = nullptr;
    uint32_t *snap_epoch = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = dalloc(&snap, 2 * n);
    if (e == hipSuccess) e = dalloc(&snap_end, b->cells);
    if (e == hipSuccess) e = dalloc(&snap_epoch, 1);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    auto cleanup = [&]() {
        dfree(snap);
        dfree(snap_end);
        dfree(snap_epoch);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    };
    auto save = [&]() -> hipError_t {
        hipError_t r;
        if ((r = hipMemcpyAsync(snap, b->pos, n * sizeof(float4), hipMemcpyDeviceToDevice, st))) return r;
        if ((r = hipMemcpyAsync(snap + n, b->vel, n * sizeof(float4), hipMemcpyDeviceToDevice, st))) return r;
        if ((r = hipMemcpyAsync(snap_end, b->end64, b->cells * 8, hipMemcpyDeviceToDevice, st))) return r;
        return hipMemcpyAsync(snap_epoch, b->epoch, 4, hipMemcpyDeviceToDevice, st);
    };
    auto restore = [&]() -> hipError_t {
        hipError_t r;
        if ((r = hipMemcpyAsync(b->pos, snap, n * sizeof(float4), hipMemcpyDeviceToDevice, st))) return r;
        if ((r = hipMemcpyAsync(b->vel, snap + n, n * sizeof(float4), hipMemcpyDeviceToDevice, st))) return r;
        if ((r = hipMemcpyAsync(b->end64, snap_end, b->cells * 8, hipMemcpyDeviceToDevice, st))) return r;
        return hipMemcpyAsync(b->epoch, snap_epoch, 4, hipMemcpyDeviceToDevice, st);
    };
    if (e == hipSuccess) e = save();
    if (e != hipSuccess) {
        cleanup();
        return hip_fail(e, "pbf_batch_autotune");
    }
    const bool graph = b->use_graph;
    b->drop_graph();
    b->use_graph = false;
    const LaunchShape keep = b->shape;
    LaunchShape best_shape = keep;
    double best_ms = 1e300;
    int32_t k = 0;
    // one trial: a warm-up step of the shape, then steps_per_trial timed steps,
    // each from the snapshot
    auto trial = [&](const LaunchShape &ls) {
        b->shape = ls;
        float ms = 0.0f;
        if (e == hipSuccess) e = restore();
        if (e == hipSuccess) {
            b->enqueue_step(st);
            e = restore();
        }
        if (e == hipSuccess) e = hipEventRecord(e0, st);
        if (e == hipSuccess) {
            for (int32_t s = 0; s < steps_per_trial; s++) b->enqueue_step(st);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipEventRecord(e1, st);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e != hipSuccess) return;
        const double per = ms / steps_per_trial;
        if (results && k < max_results) {
            results[k].neighbour_block = ls.nb;
            results[k].stream_block = ls.sb;
            results[k].stream_ppt = ls.ppt;
            results[k].list_depth = ls.lc;
            results[k].persistent = ls.persistent;
            results[k].step_ms = per;
        }
        k++;
        if (per < best_ms) {
            best_ms = per;
            best_shape = ls;
        }
    };
    static const int nbs[4] = {64, 128, 256, 512};
    static const int lcs[2] = {32, 56};
    static const int sbs[5] = {64, 128, 256, 512, 1024};
    // the neighbour passes' shape at the current streaming shape, then the
    // streaming shape with the best of those
    for (int nb : nbs)
        for (int lc : lcs)
            for (int persistent = 0; persistent <= 1 && e == hipSuccess; persistent++) {
                LaunchShape ls = keep;
                ls.nb = nb;
                ls.lc = lc;
                ls.persistent = persistent;
                trial(ls);
            }
    const LaunchShape nbest = best_shape;
    for (int sb : sbs)
        for (int ppt = 1; ppt <= 4 && e == hipSuccess; ppt++) {
            LaunchShape ls = nbest;
            ls.sb = sb;
            ls.ppt = ppt;
            if (ls.sb == nbest.sb && ls.ppt == nbest.ppt) continue;  // timed in the first round
            trial(ls);
        }
    if (e == hipSuccess) e = restore();
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    cleanup();
    b->use_graph = graph;
    if (e != hipSuccess) {
        b->shape = keep;
        return hip_fail(e, "pbf_batch_autotune");
    }
    b->shape = best_shape;
    if (count) *count = k;
    if (best) {
        best->neighbour_block = best_shape.nb;
        best->stream_block = best_shape.sb;
        best->stream_ppt = best_shape.ppt;
        best->list_depth = best_shape.lc;
        best->persistent = best_shape.persistent;
        best->step_ms = best_ms;
    }
    return PBF_OK;
}

int pbf_batch_get_end_tables(pbf_batch *b, int32_t *dst, int host) {
    if (!b || !dst) return invalid("null argument");
    int32_t *tmp = nullptr;
    PBF_HIP(dalloc(&tmp, b->cells));
    launch_end64_to_i32(b->cells, b->end64, tmp, b->stream);
    hipError_t e = hipMemcpyAsync(dst, tmp, b->cells * 4, host ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice,
                                  b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    (void)hipFree(tmp);
    if (e != hipSuccess) return hip_fail(e, "pbf_batch_get_end_tables");
    return PBF_OK;
}

int pbf_batch_set_end_tables(pbf_batch *b, const int32_t *src, int host) {
    if (!b || !src) return invalid("null argument");
    int32_t *tmp = nullptr;
    PBF_HIP(dalloc(&tmp, b->cells));
    hipError_t e = hipMemcpyAsync(tmp, src, b->cells * 4, host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice,
                                  b->stream);
    if (e == hipSuccess) {
        launch_i32_to_end64(b->cells, tmp, b->end64, b->stream);
        e = hipStreamSynchronize(b->stream);
    }
    (void)hipFree(tmp);
    if (e != hipSuccess) return hip_fail(e, "pbf_batch_set_end_tables");
    return PBF_OK;
}

int pbf_batch_set_stream(pbf_batch *b, void *stream) {
    if (!b) return invalid("null batch handle");
    PBF_HIP(hipStreamSynchronize(b->stream));
    b->drop_graph();
    b->stream = stream ? (hipStream_t)stream : b->own_stream;
    return PBF_OK;
}

void *pbf_batch_get_stream(const pbf_batch *b) { return b ? (void *)b->stream : nullptr; }

}  // extern "C"
