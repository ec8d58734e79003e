// C ABI of libpbf_amd (declared in include/pbf.h): handles for SPH, RadixSort
// and NeighbourCellFinder, plus single-stage entry points. Host-side
// orchestration only; the kernels live in pbf_kernels.hip and pbf_sort.hip.
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "pbf_internal.h"

namespace pbf {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string &msg) { g_last_error = msg; }
const char *last_error() { return g_last_error.c_str(); }

int hip_fail(hipError_t e, const char *what) {
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? PBF_ERR_NOMEM : PBF_ERR_HIP;
}

static uint32_t count_sortbits(uint64_t v) {  // src/RadixSort.cpp:24-30
    uint32_t r = 1;
    while (v >>= 1) r++;
    return r;
}

int key_bits_of(const int32_t g[3]) {
    const uint32_t nb = count_sortbits((uint64_t)g[0] * (uint64_t)g[1] * (uint64_t)g[2] - 1);
    return (int)(2 * ((nb + 1) >> 1));
}


int check_params(const pbf_params *p) {
    if (!p) return invalid("null parameters");
    if (!(p->h > 0.0f) || !std::isfinite(p->h)) return invalid("h must be positive and finite");
    if (!(p->timestep != 0.0f) || !std::isfinite(p->timestep)) return invalid("timestep must be nonzero and finite");
    if (p->iterations > (uint32_t)kMaxIterations) return invalid("iterations above 1000");
    return PBF_OK;
}

SolverConsts make_consts(uint32_t n, const int32_t grid[3], const pbf_params &p, int key_bits) {
    SolverConsts c;
    std::memset(&c, 0, sizeof c);
    c.n = (int32_t)n;
    c.gx = grid[0];
    c.gy = grid[1];
    c.gz = grid[2];
    const float h = p.h;
    c.h = h;
    c.hh = h * h;
    c.h9 = h * h * h * h * h * h * h * h * h;
    c.h_is_one = h == 1.0f;
    // Any float sq above h2cut has fl(sqrt(sq)) > h: h^2 (1 + 1e-5), rounded up.
    c.h2cut = std::nextafter((float)((double)h * (double)h * (1.0 + 1e-5)), INFINITY);
    c.h2loose = std::nextafter((float)((double)c.h2cut * (1.0 + 0x1p-20)), INFINITY);
    c.inv_rho0 = p.one_over_rho_0;
    c.eps = p.epsilon;
    c.gravity = p.gravity;
    c.dt = p.timestep;
    c.k = p.tensile_instability_k;
    c.scale = p.tensile_instability_scale;
    c.xsph_c = p.xsph_viscosity_c;
    c.vort_eps = p.vorticity_epsilon;
    for (int a = 0; a < 3; a++) {
        c.lo[a] = p.wall_lo[a];
        c.hi[a] = p.wall_hi[a];
        c.gdir[a] = p.gravity_dir[a];
    }
    c.extforce = p.extforce;
    c.key_mask = key_bits >= 32 ? 0xFFFFFFFFu : ((1u << key_bits) - 1u);
    // qscale = 2^(12 - e) for h = m 2^e, m in [0.5, 1): h qscale in [2048, 4096)
    int e = 0;
    (void)std::frexp((double)h, &e);
    const double qs = std::ldexp(1.0, 12 - e);
    c.qscale = (float)qs;
    const double qb = (double)h * qs * (1.0 + 2e-5) + 4.0;
    c.qcut = (int32_t)std::ceil(qb * qb);
    return c;
}

}  // namespace pbf

using namespace pbf;

namespace {

hipMemcpyKind to_dev(int host) { return host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice; }
hipMemcpyKind from_dev(int host) { return host ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice; }

// Packed-range neighbour table and cell tables of one NeighbourCellFinder.
struct CellTables {
    size_t cells = 0;
    int32_t *start = nullptr;
    unsigned long long *end64 = nullptr;
    int32_t *nbr9 = nullptr;
    uint32_t *epoch = nullptr;  // device word: tag of the current step's end writes

    hipError_t alloc(uint32_t n, size_t c) {
        cells = c;
        hipError_t e;
        if ((e = dalloc(&start, cells)) != hipSuccess) return e;
        if ((e = dalloc(&end64, cells)) != hipSuccess) return e;
        if ((e = dalloc(&nbr9, (size_t)9 * n)) != hipSuccess) return e;
        if ((e = dalloc(&epoch, 1)) != hipSuccess) return e;
        if ((e = hipMemset(epoch, 0, sizeof(uint32_t))) != hipSuccess) return e;
        // The reference's end texture starts as zeros (glTexImage3D with NULL
        // data reads 0 on llvmpipe) and is never cleared afterwards.
        return hipMemset(end64, 0, cells * sizeof(unsigned long long));
    }
    void release() {
        dfree(start);
        dfree(end64);
        dfree(nbr9);
        dfree(epoch);
    }
    // FindNeighbourCells (src/NeighbourCellFinder.cpp:113-148).
    void find(const SolverConsts &c, const float4 *sorted, hipStream_t st) {
        launch_begin_cells(start, cells, epoch, st);
        launch_findcells(c, sorted, start, end64, epoch, st);
        launch_neighbourcells(c, sorted, start, end64, nbr9, st);
    }
};

enum KernelId {
    KID_PREDICT = 0,
    KID_SORT,
    KID_GATHER,
    KID_CLEAR_START,
    KID_FINDCELLS,
    KID_NEIGHBOURCELLS,
    KID_LAMBDA,
    KID_UPDATEPOS,
    KID_UPDATE,
    KID_VORTICITY_A,
    KID_VORTICITY_B,
    KID_COUNT
};
static_assert(KID_COUNT == PBF_NUM_KERNEL_IDS, "kernel id table");

const char *const kKernelNames[KID_COUNT] = {
    "predict",  "radix_sort", "gather_records", "clear_start", "findcells",  "neighbourcells",
    "lambda",   "updatepos",  "update",         "vorticity_a", "vorticity_b"};

// Event brackets around kernel launches, folded into per-kernel totals.
struct KernelTimer {
    bool on = false;
    std::vector<hipEvent_t> pool;
    size_t used = 0;
    std::vector<int> ids;
    double total_ms[KID_COUNT] = {};
    uint32_t launches[KID_COUNT] = {};

    hipError_t begin(int id, hipStream_t st) {
        if (!on) return hipSuccess;
        hipError_t e;
        if (used + 2 > pool.size()) {
            if ((e = fold()) != hipSuccess) return e;
            while (pool.size() < used + 2) {
                hipEvent_t ev;
                if ((e = hipEventCreate(&ev)) != hipSuccess) return e;
                pool.push_back(ev);
            }
        }
        ids.push_back(id);
        return hipEventRecord(pool[used++], st);
    }
    hipError_t end(hipStream_t st) {
        if (!on) return hipSuccess;
        return hipEventRecord(pool[used++], st);
    }
    hipError_t fold() {
        hipError_t e;
        for (size_t k = 0; k < ids.size(); k++) {
            if ((e = hipEventSynchronize(pool[2 * k + 1])) != hipSuccess) return e;
            float ms = 0.0f;
            if ((e = hipEventElapsedTime(&ms, pool[2 * k], pool[2 * k + 1])) != hipSuccess) return e;
            total_ms[ids[k]] += ms;
            launches[ids[k]]++;
        }
        ids.clear();
        used = 0;
        return hipSuccess;
    }
    void reset() {
        for (int k = 0; k < KID_COUNT; k++) {
            total_ms[k] = 0.0;
            launches[k] = 0;
        }
    }
    void release() {
        for (auto ev : pool) (void)hipEventDestroy(ev);
        pool.clear();
        ids.clear();
        used = 0;
    }
};

}  // namespace

struct pbf_sph {
    uint32_t n = 0;
    int32_t grid[3] = {0, 0, 0};
    size_t cells = 0;
    int key_bits = 0;
    pbf_params params;
    SolverConsts consts;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    float4 *pos = nullptr, *vel = nullptr;       // id-order state
    float4 *xs = nullptr;                        // predicted records, id order
    float4 *rec0 = nullptr;                      // sorted records at x*
    float4 *rec_a = nullptr, *rec_b = nullptr;   // sorted records of the solver iterations (Jacobi)
    uint2 *qpos = nullptr;                       // quantized positions of the current sorted records
    uint32_t *tiny = nullptr;                    // note_tiny flags of x* and of each iteration's positions
    float4 *rec = nullptr;                       // sorted records after the last iteration (borrowed)
    uint16_t *list_entries = nullptr, *list_count = nullptr;  // neighbours within h, lambda -> update
    DenseList dense;                             // the step's particles with the most candidates
    float4 *vel_sorted = nullptr, *vxsph = nullptr, *omega = nullptr;
    float *lambda = nullptr;
    float4 *nrec = nullptr;                      // neighbour records: position and lambda, then |omega|
    uint32_t *keys = nullptr, *perm = nullptr;
    uint32_t *highlight = nullptr;               // highlight words by id (SPH::GetHighlightBuffer)
    CellTables cells_t;
    RadixScratch rs;
    hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    bool stepped = false;
    bool graph_step = false;                     // the last step ran as a graph replay
    KernelTimer kt;
    // Graph replay of one step (pbf_sph_set_graph): captured on first use on
    // graph_stream, dropped when parameters or the stream change.
    bool use_graph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t graph_stream = nullptr;

    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        graph_stream = nullptr;
    }

    void release() {
        drop_graph();
        kt.release();
        dfree(pos);
        dfree(vel);
        dfree(xs);
        dfree(rec0);
        dfree(rec_a);
        dfree(rec_b);
        dfree(qpos);
        dfree(tiny);
        rec = nullptr;
        dfree(list_entries);
        dfree(list_count);
        dfree(dense.cand);
        dfree(dense.levels);
        dfree(dense.threshold);
        dfree(vel_sorted);
        dfree(vxsph);
        dfree(omega);
        dfree(lambda);
        dfree(nrec);
        dfree(keys);
        dfree(perm);
        dfree(highlight);
        cells_t.release();
        radix_free(rs);
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
        if (own_stream) (void)hipStreamDestroy(own_stream);
    }
};

struct pbf_radixsort {
    uint32_t n = 0;
    int32_t grid[3] = {0, 0, 0};
    int key_bits = 0;
    hipStream_t own_stream = nullptr, stream = nullptr;
    float4 *buf = nullptr, *buf_alt = nullptr;
    uint32_t *keys = nullptr, *perm = nullptr;
    RadixScratch rs;
};

struct pbf_ncf {
    uint32_t n = 0;
    int32_t grid[3] = {0, 0, 0};
    hipStream_t own_stream = nullptr, stream = nullptr;
    CellTables t;
    int32_t *nbr12 = nullptr;
};

extern "C" {

const char *pbf_last_error(void) { return pbf::last_error(); }
const char *pbf_version(void) { return "pbf_amd 0.1 (gfx950)"; }

int pbf_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int pbf_memcpy(void *dst, const void *src, size_t bytes, int kind, void *stream) {
    if (!dst || !src) return invalid("null pointer");
    static const hipMemcpyKind kinds[3] = {hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice};
    if (kind < 0 || kind > 2) return invalid("memcpy kind must be 0, 1 or 2");
    PBF_HIP(hipMemcpyAsync(dst, src, bytes, kinds[kind], (hipStream_t)stream));
    PBF_HIP(hipStreamSynchronize((hipStream_t)stream));
    return PBF_OK;
}

int pbf_synchronize(void *stream) {
    PBF_HIP(hipStreamSynchronize((hipStream_t)stream));
    return PBF_OK;
}

void pbf_params_default(pbf_params *out, int32_t gx, int32_t gy, int32_t gz) {
    if (!out) return;
    std::memset(out, 0, sizeof *out);
    out->one_over_rho_0 = 1.0f;  // src/SPH.cpp:137-144
    out->epsilon = 5.0f;
    out->gravity = 10.0f;
    out->timestep = 0.016f;
    out->tensile_instability_k = 0.1f;
    out->tensile_instability_scale = 1.0f / pbf_wpoly6(0.2f, 2.0f);
    out->xsph_viscosity_c = 0.01f;
    out->vorticity_epsilon = 5.0f;
    out->iterations = 5;     // src/SPH.cpp:26
    out->vorticity = 0;      // src/SPH.cpp:25
    out->extforce = 0;
    out->h = 2.0f;           // src/SPH.cpp:58
    const float wall[3] = {16.0f, 0.0f, 16.0f};  // updatepos.glsl:98
    const int32_t g[3] = {gx, gy, gz};
    for (int a = 0; a < 3; a++) {
        out->wall_lo[a] = wall[a];
        out->wall_hi[a] = (float)g[a] - wall[a];
    }
    out->gravity_dir[0] = 0.0f;
    out->gravity_dir[1] = -1.0f;
    out->gravity_dir[2] = 0.0f;
}

float pbf_wpoly6(float r, float h) {  // src/SPH.cpp:159-164
    if (r > h) return 0.0f;
    float tmp = h * h - r * r;
    return 1.56668147106f * tmp * tmp * tmp / (h * h * h * h * h * h * h * h * h);
}

int pbf_key_bits(int32_t gx, int32_t gy, int32_t gz) {
    if (!valid_grid(gx, gy, gz)) return invalid("grid dimensions must be positive and below 2^31 cells");
    const int32_t g[3] = {gx, gy, gz};
    return key_bits_of(g);
}

// ---------------------------------------------------------------- SPH
int pbf_sph_create(pbf_sph **out, uint32_t n, int32_t gx, int32_t gy, int32_t gz) {
    if (!out) return invalid("pbf_sph_create: null output handle");
    *out = nullptr;
    if (n == 0 || n % 512 != 0) return invalid("numparticles must be a positive multiple of 512");
    if (n >= (1u << 24)) return invalid("numparticles must be below 2^24 (24-bit start of the packed ranges)");
    if (!valid_grid(gx, gy, gz)) return invalid("grid dimensions must be positive and below 2^31 cells");
    if (pbf_device_count() < 1) {
        set_error("no HIP device");
        return PBF_ERR_NODEVICE;
    }
    pbf_sph *s = new (std::nothrow) pbf_sph();
    if (!s) return invalid("host allocation failed");
    s->n = n;
    s->grid[0] = gx;
    s->grid[1] = gy;
    s->grid[2] = gz;
    s->cells = (size_t)gx * gy * gz;
    s->key_bits = key_bits_of(s->grid);
    pbf_params_default(&s->params, gx, gy, gz);
    s->consts = make_consts(n, s->grid, s->params, s->key_bits);
    hipError_t e = hipSuccess;
#define TRY(x)                                \
    if ((e = (x)) != hipSuccess) {            \
        int rc = hip_fail(e, #x);             \
        s->release();                         \
        delete s;                             \
        return rc;                            \
    }
    TRY(hipStreamCreateWithFlags(&s->own_stream, hipStreamNonBlocking));
    s->stream = s->own_stream;
    TRY(dalloc(&s->pos, n));
    TRY(dalloc(&s->vel, n));
    TRY(dalloc(&s->xs, n));
    // sorted records, lambda and the neighbour records keep kPad zero entries
    // after the n particles (the solver kernels read them for indices past n)
    TRY(dalloc(&s->rec0, (size_t)n + kPad));
    TRY(dalloc(&s->rec_a, (size_t)n + kPad));
    TRY(dalloc(&s->rec_b, (size_t)n + kPad));
    TRY(dalloc(&s->qpos, (size_t)n + kPad));
    TRY(dalloc(&s->tiny, (size_t)kMaxIterations + 1));
    s->rec = s->rec0;
    TRY(dalloc(&s->list_entries, (size_t)kListRows * n));
    TRY(dalloc(&s->list_count, n));
    TRY(dalloc(&s->dense.cand, n));
    TRY(dalloc(&s->dense.levels, (size_t)kDenseLevels * kDenseCopies));
    TRY(dalloc(&s->dense.threshold, 2));
    TRY(hipMemset(s->dense.levels, 0, (size_t)kDenseLevels * kDenseCopies * sizeof(uint32_t)));
    TRY(dalloc(&s->vel_sorted, n));
    TRY(dalloc(&s->vxsph, n));
    TRY(dalloc(&s->omega, n));
    TRY(dalloc(&s->lambda, (size_t)n + kPad));
    TRY(dalloc(&s->nrec, (size_t)n + kPad));
    TRY(dalloc(&s->keys, n));
    TRY(dalloc(&s->perm, n));
    TRY(dalloc(&s->highlight, n));
    // cleared like the reference's glClearBufferData (src/SPH.cpp:106-109)
    TRY(hipMemset(s->highlight, 0, (size_t)n * sizeof(uint32_t)));
    TRY(s->cells_t.alloc(n, s->cells));
    if (!radix_alloc(s->rs, (int32_t)n)) TRY(hipErrorOutOfMemory);
    for (auto &ev : s->ev) TRY(hipEventCreate(&ev));
    TRY(hipMemset(s->pos, 0, (size_t)n * sizeof(float4)));
    TRY(hipMemset(s->vel, 0, (size_t)n * sizeof(float4)));
    TRY(hipMemset(s->omega, 0, (size_t)n * sizeof(float4)));
    TRY(hipMemset(s->rec0, 0, ((size_t)n + kPad) * sizeof(float4)));
    TRY(hipMemset(s->rec_a, 0, ((size_t)n + kPad) * sizeof(float4)));
    TRY(hipMemset(s->rec_b, 0, ((size_t)n + kPad) * sizeof(float4)));
    TRY(hipMemset(s->qpos, 0, ((size_t)n + kPad) * sizeof(uint2)));
    TRY(hipMemset(s->lambda, 0, ((size_t)n + kPad) * sizeof(float)));
    TRY(hipMemset(s->nrec, 0, ((size_t)n + kPad) * sizeof(float4)));
    TRY(hipDeviceSynchronize());
#undef TRY
    *out = s;
    return PBF_OK;
}

int pbf_sph_destroy(pbf_sph *s) {
    if (!s) return PBF_OK;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    s->release();
    delete s;
    return PBF_OK;
}

int pbf_sph_set_params(pbf_sph *s, const pbf_params *p) {
    if (!s) return invalid("null SPH handle");
    int rc = check_params(p);
    if (rc) return rc;
    // the captured launches carry the old constants; a replay may still run
    // on the stream it was captured on
    if (s->exec) PBF_HIP(hipStreamSynchronize(s->graph_stream));
    s->drop_graph();
    s->params = *p;
    s->consts = make_consts(s->n, s->grid, s->params, s->key_bits);
    return PBF_OK;
}

int pbf_sph_get_params(const pbf_sph *s, pbf_params *out) {
    if (!s || !out) return invalid("null argument");
    *out = s->params;
    return PBF_OK;
}

int pbf_sph_upload_state(pbf_sph *s, const float *pos4, const float *vel4, int host) {
    if (!s || !pos4 || !vel4) return invalid("null argument");
    const size_t bytes = (size_t)s->n * sizeof(float4);
    PBF_HIP(hipMemcpyAsync(s->pos, pos4, bytes, to_dev(host), s->stream));
    PBF_HIP(hipMemcpyAsync(s->vel, vel4, bytes, to_dev(host), s->stream));
    if (host) PBF_HIP(hipStreamSynchronize(s->stream));
    return PBF_OK;
}

int pbf_sph_download_state(pbf_sph *s, float *pos4, float *vel4, int host) {
    if (!s) return invalid("null SPH handle");
    const size_t bytes = (size_t)s->n * sizeof(float4);
    if (pos4) PBF_HIP(hipMemcpyAsync(pos4, s->pos, bytes, from_dev(host), s->stream));
    if (vel4) PBF_HIP(hipMemcpyAsync(vel4, s->vel, bytes, from_dev(host), s->stream));
    if (host) PBF_HIP(hipStreamSynchronize(s->stream));
    return PBF_OK;
}

namespace {
// One step's launches on st (SPH::Run, src/SPH.cpp:246-334); with events, the
// five phase events and the kernel timer's brackets are recorded too.
int sph_enqueue(pbf_sph *s, hipStream_t st, bool events) {
    const SolverConsts &c = s->consts;
    const int32_t n = (int32_t)s->n;
    KernelTimer &kt = s->kt;
#define TIMED(id, launch)                    \
    do {                                     \
        if (events) PBF_HIP(kt.begin(id, st)); \
        launch;                              \
        if (events) PBF_HIP(kt.end(st));     \
    } while (0)
#define PHASE(k)                                              \
    do {                                                      \
        if (events) PBF_HIP(hipEventRecord(s->ev[k], st));    \
    } while (0)
    PHASE(0);
    // predict positions (src/SPH.cpp:247-261) + cell keys
    // with the clear of the highlight words (clearhighlight.glsl, src/SPH.cpp:290-293)
    const uint32_t iters = s->params.iterations;
    TIMED(KID_PREDICT, launch_predict(c, s->pos, s->vel, s->xs, s->keys, st, s->highlight, s->tiny, (int32_t)iters + 1));
    PHASE(1);
    // sort (src/SPH.cpp:263-268)
    // the sort's last pass writes the sorted records (globalsort.glsl:64 as a
    // gather) and their quantized positions
    TIMED(KID_SORT,
          radix_sort_perm(s->rs, s->keys, n, s->key_bits, nullptr, st, s->xs, s->rec0, s->qpos, c.qscale, s->tiny));
    PHASE(2);
    // neighbour cells (src/SPH.cpp:270-275; src/NeighbourCellFinder.cpp:113-148)
    CellTables &t = s->cells_t;
    TIMED(KID_CLEAR_START, launch_begin_cells(t.start, t.cells, t.epoch, st, &s->dense));
    TIMED(KID_FINDCELLS, launch_findcells(c, s->rec0, t.start, t.end64, t.epoch, st));
    // with the highlight pass (highlight.glsl, src/SPH.cpp:294-296), which reads
    // the same sorted records and rows
    TIMED(KID_NEIGHBOURCELLS,
          launch_neighbourcells(c, s->rec0, t.start, t.end64, t.nbr9, st, &s->dense, s->highlight));
    PHASE(3);
    // solver iterations (src/SPH.cpp:303-311), race-free: x* -> a -> b -> a ...
    PairList pl;
    pl.entries = s->list_entries;
    pl.count = s->list_count;
    pl.cap = kListCap;
    // The last position update also performs the step's closing update
    // (src/SPH.cpp:315-318) for each particle as its final position is known.
    StepUpdate up;
    up.pos = s->pos;
    up.vel = s->vel;
    up.vel_sorted = s->params.vorticity ? s->vel_sorted : nullptr;
    const float4 *cur = s->rec0;
    for (uint32_t it = 0; it < iters; it++) {
        float4 *dst = (it & 1) ? s->rec_b : s->rec_a;
        TIMED(KID_LAMBDA,
              launch_lambda(c, cur, s->qpos, t.nbr9, s->lambda, s->nrec, pl, st, true, &s->dense, nullptr,
                            s->tiny + it));
        TIMED(KID_UPDATEPOS, launch_updatepos(c, cur, s->lambda, s->nrec, t.nbr9, pl, dst, s->qpos, st, true,
                                              &s->dense, it + 1 == iters ? &up : nullptr, nullptr, s->tiny + it,
                                              s->tiny + it + 1));
        cur = dst;
    }
    s->rec = const_cast<float4 *>(cur);
    PHASE(4);
    // update (when no iteration ran) + vorticity (src/SPH.cpp:315-333)
    if (iters == 0) TIMED(KID_UPDATE, launch_update(c, s->rec, s->pos, s->vel, up.vel_sorted, st));
    if (s->params.vorticity) {
        TIMED(KID_VORTICITY_A,
              launch_vorticity_a(c, s->rec, s->qpos, t.nbr9, s->vel_sorted, s->vel, s->vxsph, s->omega, nullptr,
                                 s->nrec, pl, st, true, &s->dense, s->tiny + iters));
        TIMED(KID_VORTICITY_B,
              launch_vorticity_b(c, s->rec, t.nbr9, pl, s->vxsph, s->omega, s->nrec, s->vel, st, true, &s->dense,
                                 nullptr, s->tiny + iters));
    }
#undef TIMED
    PHASE(5);
#undef PHASE
    PBF_HIP(hipGetLastError());
    return PBF_OK;
}
}  // namespace

int pbf_sph_run(pbf_sph *s) {
    if (!s) return invalid("null SPH handle");
    if (!s->use_graph || s->kt.on) {
        const int rc = sph_enqueue(s, s->stream, true);
        if (rc) return rc;
        s->stepped = true;
        s->graph_step = false;
        return PBF_OK;
    }
    if (!s->exec || s->graph_stream != s->stream) {
        s->drop_graph();
        PBF_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
        const int rc = sph_enqueue(s, s->stream, false);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(s->stream, &g);
        if (rc || e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return rc ? rc : hip_fail(e, "pbf_sph_run: capture");
        }
        s->graph = g;
        const hipError_t ei = hipGraphInstantiate(&s->exec, g, nullptr, nullptr, 0);
        if (ei != hipSuccess) {
            s->drop_graph();
            return hip_fail(ei, "pbf_sph_run: instantiate");
        }
        s->graph_stream = s->stream;
    }
    PBF_HIP(hipGraphLaunch(s->exec, s->stream));
    s->stepped = true;
    s->graph_step = true;
    return PBF_OK;
}

int pbf_sph_dense_info(pbf_sph *s, int32_t *threshold, int32_t *count) {
    if (!s || !threshold || !count) return invalid("null argument");
    int32_t v[2] = {0, 0};
    if (s->stepped) {
        PBF_HIP(hipStreamSynchronize(s->stream));
        PBF_HIP(hipMemcpy(v, s->dense.threshold, sizeof v, hipMemcpyDeviceToHost));
    }
    *threshold = v[0];
    *count = v[1];
    return PBF_OK;
}

int pbf_sph_set_graph(pbf_sph *s, int enable) {
    if (!s) return invalid("null SPH handle");
    if (s->exec) PBF_HIP(hipStreamSynchronize(s->graph_stream));
    s->drop_graph();
    s->use_graph = enable != 0;
    return PBF_OK;
}

int pbf_sph_step(pbf_sph *s, int32_t nsteps) {
    if (!s) return invalid("null SPH handle");
    if (nsteps < 0) return invalid("negative step count");
    for (int32_t k = 0; k < nsteps; k++) {
        int rc = pbf_sph_run(s);
        if (rc) return rc;
    }
    return PBF_OK;
}

int pbf_sph_positions_device(pbf_sph *s, void **dev, size_t *stride) {
    if (!s || !dev) return invalid("null argument");
    *dev = s->pos;
    if (stride) *stride = sizeof(float4);
    return PBF_OK;
}

int pbf_sph_velocities_device(pbf_sph *s, void **dev, size_t *stride) {
    if (!s || !dev) return invalid("null argument");
    *dev = s->vel;
    if (stride) *stride = sizeof(float4);
    return PBF_OK;
}

int pbf_sph_export_positions(pbf_sph *s, void *dst, size_t dst_stride, int32_t components, int host) {
    if (!s || !dst) return invalid("null argument");
    if (components != 3 && components != 4) return invalid("components must be 3 or 4");
    const size_t width = (size_t)components * sizeof(float);
    if (dst_stride < width) return invalid("destination stride below the exported width");
    PBF_HIP(hipMemcpy2DAsync(dst, dst_stride, s->pos, sizeof(float4), width, s->n, from_dev(host), s->stream));
    if (host) PBF_HIP(hipStreamSynchronize(s->stream));
    return PBF_OK;
}

int pbf_sph_highlight_device(pbf_sph *s, void **dev, size_t *stride) {
    if (!s || !dev) return invalid("null argument");
    *dev = s->highlight;
    if (stride) *stride = sizeof(uint32_t);
    return PBF_OK;
}

int pbf_sph_get_highlight(pbf_sph *s, uint32_t *dst, int host) {
    if (!s || !dst) return invalid("null argument");
    PBF_HIP(hipMemcpyAsync(dst, s->highlight, (size_t)s->n * sizeof(uint32_t), from_dev(host), s->stream));
    if (host) PBF_HIP(hipStreamSynchronize(s->stream));
    return PBF_OK;
}

int pbf_sph_set_highlight(pbf_sph *s, const uint32_t *src, int host) {
    if (!s) return invalid("null SPH handle");
    const size_t bytes = (size_t)s->n * sizeof(uint32_t);
    if (src) {
        PBF_HIP(hipMemcpyAsync(s->highlight, src, bytes, to_dev(host), s->stream));
        if (host) PBF_HIP(hipStreamSynchronize(s->stream));
    } else {
        PBF_HIP(hipMemsetAsync(s->highlight, 0, bytes, s->stream));
    }
    return PBF_OK;
}

int pbf_sph_toggle_highlight(pbf_sph *s, uint32_t id, uint32_t *now) {
    if (!s) return invalid("null SPH handle");
    if (id >= s->n) return invalid("particle id out of range");
    uint32_t w = 0;
    PBF_HIP(hipMemcpyAsync(&w, s->highlight + id, sizeof w, hipMemcpyDeviceToHost, s->stream));
    PBF_HIP(hipStreamSynchronize(s->stream));
    w = w > 0 ? 0u : 1u;  // src/Simulation.cpp:181-185
    PBF_HIP(hipMemcpyAsync(s->highlight + id, &w, sizeof w, hipMemcpyHostToDevice, s->stream));
    PBF_HIP(hipStreamSynchronize(s->stream));
    if (now) *now = w;
    return PBF_OK;
}

int pbf_sph_timing(pbf_sph *s, pbf_timing *out) {
    if (!s || !out) return invalid("null argument");
    std::memset(out, 0, sizeof *out);
    if (!s->stepped) return PBF_OK;
    if (s->graph_step) return invalid("phase timing needs an eager step (graph replay is on)");
    PBF_HIP(hipEventSynchronize(s->ev[5]));
    float ms[5];
    for (int k = 0; k < 5; k++) PBF_HIP(hipEventElapsedTime(&ms[k], s->ev[k], s->ev[k + 1]));
    out->predict_ms = ms[0];
    out->sort_ms = ms[1];
    out->neighbour_cells_ms = ms[2];
    out->solver_ms = ms[3];
    out->vorticity_ms = ms[4];
    return PBF_OK;
}

int pbf_sph_set_kernel_timing(pbf_sph *s, int enable) {
    if (!s) return invalid("null SPH handle");
    if (s->kt.on && !enable) PBF_HIP(s->kt.fold());
    s->kt.on = enable != 0;
    return PBF_OK;
}

int pbf_sph_kernel_times(pbf_sph *s, pbf_kernel_time out[PBF_NUM_KERNEL_IDS]) {
    if (!s || !out) return invalid("null argument");
    PBF_HIP(s->kt.fold());
    for (int k = 0; k < KID_COUNT; k++) {
        out[k].total_ms = s->kt.total_ms[k];
        out[k].launches = s->kt.launches[k];
    }
    s->kt.reset();
    return PBF_OK;
}

const char *pbf_kernel_name(int id) { return id >= 0 && id < KID_COUNT ? kKernelNames[id] : nullptr; }

// The work queued on the old stream (eager launches, or a graph replay on the
// stream it was captured on) finishes before the switch, so steps keep their
// order across streams; the graph, captured on the old stream, is dropped.
int pbf_sph_set_stream(pbf_sph *s, void *stream) {
    if (!s) return invalid("null SPH handle");
    const hipStream_t next = stream ? (hipStream_t)stream : s->own_stream;
    if (next == s->stream) return PBF_OK;
    PBF_HIP(hipStreamSynchronize(s->stream));
    if (s->graph_stream && s->graph_stream != s->stream) PBF_HIP(hipStreamSynchronize(s->graph_stream));
    s->drop_graph();
    s->stream = next;
    return PBF_OK;
}

void *pbf_sph_get_stream(const pbf_sph *s) { return s ? (void *)s->stream : nullptr; }

int pbf_sph_get_end_table(pbf_sph *s, int32_t *dst, int host) {
    if (!s || !dst) return invalid("null argument");
    int32_t *tmp = dst;
    if (host) PBF_HIP(dalloc(&tmp, s->cells));
    launch_end64_to_i32(s->cells, s->cells_t.end64, tmp, s->stream);
    if (host) {
        hipError_t e = hipMemcpyAsync(dst, tmp, s->cells * 4, hipMemcpyDeviceToHost, s->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
        (void)hipFree(tmp);
        if (e != hipSuccess) return hip_fail(e, "pbf_sph_get_end_table");
    }
    return PBF_OK;
}

int pbf_sph_set_end_table(pbf_sph *s, const int32_t *src, int host) {
    if (!s || !src) return invalid("null argument");
    int32_t *tmp = nullptr;
    PBF_HIP(dalloc(&tmp, s->cells));
    hipError_t e = hipMemcpyAsync(tmp, src, s->cells * 4, to_dev(host), s->stream);
    if (e == hipSuccess) {
        launch_i32_to_end64(s->cells, tmp, s->cells_t.end64, s->stream);
        e = hipStreamSynchronize(s->stream);
    }
    (void)hipFree(tmp);
    if (e != hipSuccess) return hip_fail(e, "pbf_sph_set_end_table");
    return PBF_OK;
}

int pbf_sph_neighbour_table(pbf_sph *s, int32_t *dst12, int host) {
    if (!s || !dst12) return invalid("null argument");
    int32_t *tmp = dst12;
    if (host) PBF_HIP(dalloc(&tmp, (size_t)12 * s->n));
    launch_nbr9_to_12((int32_t)s->n, s->cells_t.nbr9, tmp, s->stream);
    if (host) {
        hipError_t e = hipMemcpyAsync(dst12, tmp, (size_t)48 * s->n, hipMemcpyDeviceToHost, s->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
        (void)hipFree(tmp);
        if (e != hipSuccess) return hip_fail(e, "pbf_sph_neighbour_table");
    }
    return PBF_OK;
}

int pbf_sph_sorted_records(pbf_sph *s, float *dst4, int host) {
    if (!s || !dst4) return invalid("null argument");
    PBF_HIP(hipMemcpyAsync(dst4, s->rec, (size_t)s->n * sizeof(float4), from_dev(host), s->stream));
    if (host) PBF_HIP(hipStreamSynchronize(s->stream));
    return PBF_OK;
}

// ---------------------------------------------------------------- RadixSort
int pbf_radixsort_create(pbf_radixsort **out, uint32_t blocksize, uint32_t numblocks, int32_t gx, int32_t gy,
                         int32_t gz) {
    if (!out) return invalid("null output handle");
    *out = nullptr;
    if (blocksize & 1) {  // src/RadixSort.cpp:41-42
        set_error("The block size for sorting has to be even.");
        return PBF_ERR_LOGIC;
    }
    const uint64_t n = (uint64_t)blocksize * numblocks;
    if (n == 0 || n >= (1u << 31)) return invalid("blocksize * numblocks must be in [1, 2^31)");
    if (!valid_grid(gx, gy, gz)) return invalid("grid dimensions must be positive and below 2^31 cells");
    if (pbf_device_count() < 1) {
        set_error("no HIP device");
        return PBF_ERR_NODEVICE;
    }
    pbf_radixsort *r = new (std::nothrow) pbf_radixsort();
    if (!r) return invalid("host allocation failed");
    r->n = (uint32_t)n;
    r->grid[0] = gx;
    r->grid[1] = gy;
    r->grid[2] = gz;
    r->key_bits = key_bits_of(r->grid);
    hipError_t e;
    bool ok = (e = hipStreamCreateWithFlags(&r->own_stream, hipStreamNonBlocking)) == hipSuccess &&
              (e = dalloc(&r->buf, n)) == hipSuccess && (e = dalloc(&r->buf_alt, n)) == hipSuccess &&
              (e = dalloc(&r->keys, n)) == hipSuccess && (e = dalloc(&r->perm, n)) == hipSuccess;
    if (ok && !radix_alloc(r->rs, (int32_t)n)) {
        ok = false;
        e = hipErrorOutOfMemory;
    }
    if (!ok) {
        int rc = hip_fail(e, "pbf_radixsort_create");
        pbf_radixsort_destroy(r);
        return rc;
    }
    r->stream = r->own_stream;
    *out = r;
    return PBF_OK;
}

int pbf_radixsort_destroy(pbf_radixsort *r) {
    if (!r) return PBF_OK;
    if (r->stream) (void)hipStreamSynchronize(r->stream);
    dfree(r->buf);
    dfree(r->buf_alt);
    dfree(r->keys);
    dfree(r->perm);
    radix_free(r->rs);
    if (r->own_stream) (void)hipStreamDestroy(r->own_stream);
    delete r;
    return PBF_OK;
}

int pbf_radixsort_buffer(pbf_radixsort *r, void **dev) {
    if (!r || !dev) return invalid("null argument");
    *dev = r->buf;
    return PBF_OK;
}

int pbf_radixsort_run(pbf_radixsort *r) {
    if (!r) return invalid("null RadixSort handle");
    pbf_params p;
    pbf_params_default(&p, r->grid[0], r->grid[1], r->grid[2]);
    const SolverConsts c = make_consts(r->n, r->grid, p, r->key_bits);
    launch_keys_of_records(c, r->buf, r->keys, r->stream);
    radix_sort_perm(r->rs, r->keys, (int32_t)r->n, r->key_bits, r->perm, r->stream);
    launch_gather_records((int32_t)r->n, r->perm, r->buf, r->buf_alt, r->stream);
    std::swap(r->buf, r->buf_alt);
    PBF_HIP(hipGetLastError());
    return PBF_OK;
}

int pbf_radixsort_permutation(pbf_radixsort *r, void **dev) {
    if (!r || !dev) return invalid("null argument");
    *dev = r->perm;
    return PBF_OK;
}

int pbf_radixsort_set_stream(pbf_radixsort *r, void *stream) {
    if (!r) return invalid("null RadixSort handle");
    const hipStream_t next = stream ? (hipStream_t)stream : r->own_stream;
    if (next != r->stream) PBF_HIP(hipStreamSynchronize(r->stream));  // runs stay in order
    r->stream = next;
    return PBF_OK;
}

void *pbf_radixsort_get_stream(const pbf_radixsort *r) { return r ? (void *)r->stream : nullptr; }

// ---------------------------------------------------------------- NeighbourCellFinder
int pbf_ncf_create(pbf_ncf **out, uint32_t n, int32_t gx, int32_t gy, int32_t gz) {
    if (!out) return invalid("null output handle");
    *out = nullptr;
    if (n == 0 || n >= (1u << 24)) return invalid("numparticles must be in [1, 2^24)");
    if (!valid_grid(gx, gy, gz)) return invalid("grid dimensions must be positive and below 2^31 cells");
    if (pbf_device_count() < 1) {
        set_error("no HIP device");
        return PBF_ERR_NODEVICE;
    }
    pbf_ncf *f = new (std::nothrow) pbf_ncf();
    if (!f) return invalid("host allocation failed");
    f->n = n;
    f->grid[0] = gx;
    f->grid[1] = gy;
    f->grid[2] = gz;
    hipError_t e;
    bool ok = (e = hipStreamCreateWithFlags(&f->own_stream, hipStreamNonBlocking)) == hipSuccess &&
              (e = f->t.alloc(n, (size_t)gx * gy * gz)) == hipSuccess &&
              (e = dalloc(&f->nbr12, (size_t)12 * n)) == hipSuccess;
    if (!ok) {
        int rc = hip_fail(e, "pbf_ncf_create");
        pbf_ncf_destroy(f);
        return rc;
    }
    f->stream = f->own_stream;
    *out = f;
    return PBF_OK;
}

int pbf_ncf_destroy(pbf_ncf *f) {
    if (!f) return PBF_OK;
    if (f->stream) (void)hipStreamSynchronize(f->stream);
    f->t.release();
    dfree(f->nbr12);
    if (f->own_stream) (void)hipStreamDestroy(f->own_stream);
    delete f;
    return PBF_OK;
}

int pbf_ncf_find(pbf_ncf *f, const void *dev_sorted_records) {
    if (!f || !dev_sorted_records) return invalid("null argument");
    pbf_params p;
    pbf_params_default(&p, f->grid[0], f->grid[1], f->grid[2]);
    const SolverConsts c = make_consts(f->n, f->grid, p, key_bits_of(f->grid));
    f->t.find(c, (const float4 *)dev_sorted_records, f->stream);
    launch_nbr9_to_12((int32_t)f->n, f->t.nbr9, f->nbr12, f->stream);
    PBF_HIP(hipGetLastError());
    return PBF_OK;
}

int pbf_ncf_result(pbf_ncf *f, void **dev12) {
    if (!f || !dev12) return invalid("null argument");
    *dev12 = f->nbr12;
    return PBF_OK;
}

int pbf_ncf_get_tables(pbf_ncf *f, int32_t *start, int32_t *end, int host) {
    if (!f) return invalid("null NeighbourCellFinder handle");
    const size_t bytes = f->t.cells * 4;
    if (start) PBF_HIP(hipMemcpyAsync(start, f->t.start, bytes, from_dev(host), f->stream));
    if (end) {
        int32_t *tmp = end;
        if (host) PBF_HIP(dalloc(&tmp, f->t.cells));
        launch_end64_to_i32(f->t.cells, f->t.end64, tmp, f->stream);
        if (host) {
            hipError_t e = hipMemcpyAsync(end, tmp, bytes, hipMemcpyDeviceToHost, f->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(f->stream);
            (void)hipFree(tmp);
            if (e != hipSuccess) return hip_fail(e, "pbf_ncf_get_tables");
        }
    }
    if (host) PBF_HIP(hipStreamSynchronize(f->stream));
    return PBF_OK;
}

int pbf_ncf_set_end_table(pbf_ncf *f, const int32_t *end, int host) {
    if (!f || !end) return invalid("null argument");
    int32_t *tmp = nullptr;
    PBF_HIP(dalloc(&tmp, f->t.cells));
    hipError_t e = hipMemcpyAsync(tmp, end, f->t.cells * 4, to_dev(host), f->stream);
    if (e == hipSuccess) {
        launch_i32_to_end64(f->t.cells, tmp, f->t.end64, f->stream);
        e = hipStreamSynchronize(f->stream);
    }
    (void)hipFree(tmp);
    if (e != hipSuccess) return hip_fail(e, "pbf_ncf_set_end_table");
    return PBF_OK;
}

int pbf_ncf_set_stream(pbf_ncf *f, void *stream) {
    if (!f) return invalid("null NeighbourCellFinder handle");
    const hipStream_t next = stream ? (hipStream_t)stream : f->own_stream;
    if (next != f->stream) PBF_HIP(hipStreamSynchronize(f->stream));  // calls stay in order
    f->stream = next;
    return PBF_OK;
}

void *pbf_ncf_get_stream(const pbf_ncf *f) { return f ? (void *)f->stream : nullptr; }

// ---------------------------------------------------------------- stages
static int stage_check(uint32_t n, const int32_t *grid, const pbf_params *p) {
    if (n == 0 || n % 256 != 0 || n >= (1u << 24)) return invalid("n must be a positive multiple of 256 below 2^24");
    if (grid && !valid_grid(grid[0], grid[1], grid[2])) return invalid("bad grid");
    return check_params(p);
}

static const int32_t k_no_grid[3] = {1, 1, 1};

int pbf_stage_predict(uint32_t n, const int32_t grid[3], const pbf_params *p, const void *pos4, const void *vel4,
                      void *rec4_out, void *stream) {
    int rc = stage_check(n, grid, p);
    if (rc) return rc;
    if (!pos4 || !vel4 || !rec4_out) return invalid("null pointer");
    const SolverConsts c = make_consts(n, grid, *p, key_bits_of(grid));
    launch_predict(c, (const float4 *)pos4, (const float4 *)vel4, (float4 *)rec4_out, nullptr, (hipStream_t)stream);
    PBF_HIP(hipGetLastError());
    return PBF_OK;
}

// Converts a caller's GetResult-layout table into the kernels' [9][n] layout.
static int nbr_scratch(uint32_t n, const void *nbr12, int32_t **nbr9, hipStream_t st) {
    PBF_HIP(dalloc(nbr9, (size_t)9 * n));
    launch_nbr12_to_9((int32_t)n, (const int32_t *)nbr12, *nbr9, st);
    return PBF_OK;
}

// The two vorticity passes of the stage on caller memory, with the [9][n]
// table in nbr9 (already converted) and kVortStageBytes per particle of
// scratch in tmp: four float4 arrays (sorted velocities, XSPH velocities,
// omega, neighbour records), then the quantized positions.
static void launch_vorticity_stage(uint32_t n, const int32_t grid[3], const pbf_params &p, const void *rec4,
                                   void *vel4, void *vort_out, const int32_t *nbr9, float4 *tmp, hipStream_t st) {
    const SolverConsts c = make_consts(n, grid, p, key_bits_of(grid));
    const float4 *rec = (const float4 *)rec4;
    float4 *vel = (float4 *)vel4;
    float4 *vs = tmp, *vx = tmp + n, *om = tmp + 2 * (size_t)n, *nrec = tmp + 3 * (size_t)n;
    uint2 *q = reinterpret_cast<uint2 *>(tmp + 4 * (size_t)n);
    // sorted copy of the velocities, vs[g] = vel[id(rec[g])] (SPH::Run keeps
    // one from the update pass)
    launch_gather_by_id((int32_t)n, rec, vel, vs, st);
    launch_quantize(c, rec, q, (int32_t)n, st);
    launch_vorticity_a(c, rec, q, nbr9, vs, vel, vx, om, (float *)vort_out, nrec, PairList(), st);
    launch_vorticity_b(c, rec, nbr9, PairList(), vx, om, nrec, vel, st);
}

constexpr size_t kVortStageBytes = 4 * sizeof(float4) + sizeof(uint2);

static int release_scratch(void *p, hipStream_t st) {
    hipError_t e = hipStreamSynchronize(st);
    (void)hipFree(p);
    if (e != hipSuccess) return hip_fail(e, "stage");
    PBF_HIP(hipGetLastError());
    return PBF_OK;
}

int pbf_stage_lambda(uint32_t n, const int32_t grid[3], const pbf_params *p, const void *rec4, const void *nbr12,
                     void *lambda_out, void *stream) {
    int rc = stage_check(n, grid, p);
    if (rc) return rc;
    if (!rec4 || !nbr12 || !lambda_out) return invalid("null pointer");
    hipStream_t st = (hipStream_t)stream;
    int32_t *nbr9 = nullptr;
    if ((rc = nbr_scratch(n, nbr12, &nbr9, st))) return rc;
    const SolverConsts c = make_consts(n, grid, *p, key_bits_of(grid));
    uint2 *q = nullptr;        // quantized positions of the caller's records
    float4 *nrec = nullptr;    // neighbour records the pass writes beside lambda (unused here)
    hipError_t e = dalloc(&q, n);
    if (e == hipSuccess) e = dalloc(&nrec, n);
    if (e != hipSuccess) {
        (void)hipFree(nbr9);
        (void)hipFree(q);
        return hip_fail(e, "pbf_stage_lambda");
    }
    launch_quantize(c, (const float4 *)rec4, q, (int32_t)n, st);
    launch_lambda(c, (const float4 *)rec4, q, nbr9, (float *)lambda_out, nrec, PairList(), st);
    rc = release_scratch(nbr9, st);
    (void)hipFree(q);
    (void)hipFree(nrec);
    return rc;
}

int pbf_stage_updatepos(uint32_t n, const int32_t grid[3], const pbf_params *p, const void *rec4_in,
                        const void *lambda, const void *nbr12, void *rec4_out, void *stream) {
    int rc = stage_check(n, grid, p);
    if (rc) return rc;
    if (!rec4_in || !lambda || !nbr12 || !rec4_out) return invalid("null pointer");
    if (rec4_in == rec4_out) return invalid("updatepos is race-free: input and output must differ");
    hipStream_t st = (hipStream_t)stream;
    int32_t *nbr9 = nullptr;
    if ((rc = nbr_scratch(n, nbr12, &nbr9, st))) return rc;
    const SolverConsts c = make_consts(n, grid, *p, key_bits_of(grid));
    float4 *nrec = nullptr;  // the caller's positions with their lambda, read per neighbour
    const hipError_t e = dalloc(&nrec, n);
    if (e != hipSuccess) {
        (void)hipFree(nbr9);
        return hip_fail(e, "pbf_stage_updatepos");
    }
    launch_pack_records((int32_t)n, (const float4 *)rec4_in, (const float *)lambda, nrec, st);
    launch_updatepos(c, (const float4 *)rec4_in, (const float *)lambda, nrec, nbr9, PairList(), (float4 *)rec4_out,
                     nullptr, st);
    rc = release_scratch(nbr9, st);
    (void)hipFree(nrec);
    return rc;
}

int pbf_stage_update(uint32_t n, const pbf_params *p, const void *rec4, void *pos4, void *vel4, void *stream) {
    int rc = stage_check(n, nullptr, p);
    if (rc) return rc;
    if (!rec4 || !pos4 || !vel4) return invalid("null pointer");
    const SolverConsts c = make_consts(n, k_no_grid, *p, 1);
    launch_update(c, (const float4 *)rec4, (float4 *)pos4, (float4 *)vel4, nullptr, (hipStream_t)stream);
    PBF_HIP(hipGetLastError());
    return PBF_OK;
}

int pbf_stage_vorticity(uint32_t n, const int32_t grid[3], const pbf_params *p, const void *rec4,
                        const void *nbr12, void *vel4, void *vort_out, void *stream) {
    int rc = stage_check(n, grid, p);
    if (rc) return rc;
    if (!rec4 || !nbr12 || !vel4 || !vort_out) return invalid("null pointer");
    hipStream_t st = (hipStream_t)stream;
    int32_t *nbr9 = nullptr;
    float4 *tmp = nullptr;  // sorted velocities, XSPH velocities, omega, neighbour records, quantized positions
    if ((rc = nbr_scratch(n, nbr12, &nbr9, st))) return rc;
    hipError_t e = dalloc(&tmp, ((size_t)kVortStageBytes * n + sizeof(float4) - 1) / sizeof(float4));
    if (e != hipSuccess) {
        (void)hipFree(nbr9);
        return hip_fail(e, "pbf_stage_vorticity");
    }
    launch_vorticity_stage(n, grid, *p, rec4, vel4, vort_out, nbr9, tmp, st);
    e = hipStreamSynchronize(st);
    (void)hipFree(nbr9);
    (void)hipFree(tmp);
    if (e != hipSuccess) return hip_fail(e, "pbf_stage_vorticity");
    PBF_HIP(hipGetLastError());
    return PBF_OK;
}


// ------------------------------------------------------ vorticity, no allocation
// Scratch of pbf_stage_vorticity_noalloc: the [9][n] neighbour table (36 B per
// particle, then aligned) followed by the sorted, XSPH and omega velocities,
// the neighbour records and the quantized positions (72 B per particle), for at most
// PBF_NOALLOC_MAX_N particles. Each GPU the
// library runs on holds one copy until the process ends. In host code the
// name `scratch` is a host-side stand-in, not a device address; the address
// on a device comes from hipGetSymbolAddress, looked up once per device with
// that device current, and only that address is passed to kernels.
}  // extern "C"

static constexpr size_t kNoAllocTableBytes = (size_t)36 * PBF_NOALLOC_MAX_N;
static constexpr size_t kNoAllocBytes = kNoAllocTableBytes + kVortStageBytes * PBF_NOALLOC_MAX_N;
static_assert(kNoAllocTableBytes % 256 == 0, "the velocity arrays start 256-byte aligned");
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_dev[kMaxScratchDevices];  // device address of `scratch`, per device

// The current device's address of the scratch array (looked up on its first use).
int scratch_of_current_device(void **out) {
    int dev = 0;
    PBF_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxScratchDevices) return invalid("device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "pbf_stage_vorticity_noalloc: hipGetSymbolAddress(scratch)");
        if (!p) return invalid("pbf_stage_vorticity_noalloc: no device address for the scratch array");
        g_scratch_dev[dev] = p;
    }
    *out = g_scratch_dev[dev];
    return PBF_OK;
}
}  // namespace

extern "C" {

int pbf_stage_vorticity_noalloc(uint32_t n, const int32_t grid[3], const pbf_params *p, const void *rec4,
                                const void *nbr12, void *vel4, void *vort_out, void *stream) {
    int rc = stage_check(n, grid, p);
    if (rc) return rc;
    if (n > PBF_NOALLOC_MAX_N)
        return invalid("pbf_stage_vorticity_noalloc: n = " + std::to_string(n) + " exceeds PBF_NOALLOC_MAX_N = " +
                       std::to_string(PBF_NOALLOC_MAX_N) + " (the fixed scratch array); use pbf_stage_vorticity");
    if (!rec4 || !nbr12 || !vel4 || !vort_out) return invalid("null pointer");
    void *base = nullptr;
    if ((rc = scratch_of_current_device(&base))) return rc;
    hipStream_t st = (hipStream_t)stream;
    int32_t *nbr9 = (int32_t *)base;
    float4 *tmp = (float4 *)((unsigned char *)base + kNoAllocTableBytes);
    // every part of the scratch is written before it is read: nothing has
    // to start at zero, whatever an earlier call left there
    launch_nbr12_to_9((int32_t)n, (const int32_t *)nbr12, nbr9, st);
    launch_vorticity_stage(n, grid, *p, rec4, vel4, vort_out, nbr9, tmp, st);
    PBF_HIP(hipGetLastError());
    return PBF_OK;
}

}  // extern "C"
