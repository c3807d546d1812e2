// rz_api.hip -- host side of the C ABI declared in include/rz.h.
//
// A context owns one HIP stream and the device buffers of the pipeline.
// Draws are recorded on the host; the work runs as one job at resolve time
// (plain frame: one scene) or at rz_batch_submit (batch: many scenes), so
// the 4x MSAA colour and depth of each pixel stay in registers for the whole
// frame (no MSAA traffic to HBM). Buffers that a job outgrows are detected
// on the device (overflow bits); the host then grows them and replays the
// job, so no triangle is ever dropped.
//
// Meshes and textures live in a pool shared by every context created from
// the same root context (rz_create_shared), so several contexts -- several
// jobs in flight on their own streams -- can render the same resident data.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "rz_kernels.h"

using namespace rzk;

namespace {

constexpr uint32_t MAX_DIM = 16384;
// a coarse bin holds at most 256 16x16 tiles (fine binning keeps one running
// count per tile in 4 registers per lane)
constexpr uint32_t MAX_COARSE_PX = 256;
constexpr uint32_t MAX_DRAWS = 65536;   // draw index lives in 16 bits of a prim's flags
constexpr uint32_t MAX_SCENES = 65536;
constexpr uint32_t MAX_REPLAYS = 4;
constexpr uint64_t NOALLOC_MAX_TRIS = 262144;  // rz_render_noalloc: triangles per call

const char* kNullCtxMsg = "null context";

struct Texture {
    uint32_t* data = nullptr;
    uint32_t tw = 0, th = 0;
};

template <class T>
struct DBuf {
    T* p = nullptr;
    size_t cap = 0;  // elements
};

}  // namespace

// the persistent raster service of a context group (rz_service_start)
struct SvcState {
    bool on = false;                  // between rz_service_start and rz_service_stop
    hipStream_t stream = nullptr;     // CU-masked: a hardware queue of its own
    SvcEntry* ents = nullptr;         // SVC_MAX_ENTRIES entries
    SvcCtl* ctl = nullptr;
    SvcHost* host = nullptr;          // pinned
    uint32_t gen[SVC_MAX_ENTRIES] = {};  // last job number posted per entry
    uint32_t used = 0;                // entries held by attached contexts (bits)
    uint32_t grid = 0;
    unsigned long long idle_ticks = 0, life_ticks = 0;
    uint64_t launches = 0;
};

// resources shared by the contexts of one root (same device)
struct rz_pool {
    int device = 0;
    std::vector<Texture> textures;
    std::vector<rz_ctx*> ctxs;
    uint64_t resident_bytes = 0;  // device bytes of the group's meshes and textures
    SvcState svc;
};

struct rz_mesh {
    rz_pool* pool = nullptr;
    float* pos = nullptr;
    float* col = nullptr;
    float* uv = nullptr;
    uint32_t* idx = nullptr;
    uint32_t nv = 0;
    uint32_t ntri = 0;
    uint32_t pending = 0;  // recorded draws not yet finished, over all contexts
    uint64_t bytes = 0;    // device bytes of its streams
};

namespace {

struct PendingDraw {
    rz_mesh* mesh;
    uint32_t scene;
    uint32_t fs;
    uint32_t tex;  // texture index (RZ_FS_TEXTURE)
    Mat3 m;
};

struct PendingScene {
    uint32_t W, H;
    float clear_depth;
    uint32_t* out;
    uint32_t* msaa_color;
    float* msaa_depth;
};

// host-side layout of one job
struct JobPlan {
    uint32_t ndraws = 0, nscenes = 0;
    uint32_t ntri = 0, ntiles = 0, nbins = 0;
    uint32_t max_nb = 1, max_T = 1;
    uint64_t tris = 0;  // input triangles
};

enum class JobKind { None, Frame, Batch };

}  // namespace

struct rz_ctx {
    rz_pool* pool = nullptr;
    hipStream_t stream = nullptr;
    uint32_t W = 0, H = 0;  // plain-frame screen
    float clear_depth = FLT_MAX;
    float next_clear_depth = FLT_MAX;
    Mat3 uni{};
    uint32_t coarse_req = 0;  // plain frame: coarse bin side, 0 = automatic
    uint32_t fine_req = 32;   // fine-binning tile side asked for (8, 16 or 32; per scene where it fits)
    // batch tuning
    uint32_t bin_target = MAX_BINS;  // automatic bins: smallest power-of-2 side with at most this many bins
    uint32_t raster_per_cu = 0;      // persistent raster workgroups per CU (0: one workgroup per quadrant)
    bool use_graph = true;
    int n_cu = 256;

    // recorded work
    std::vector<PendingDraw> draws;
    std::vector<PendingScene> scenes;
    JobKind recording = JobKind::None;
    bool in_flight = false;
    JobPlan plan;

    // job tables: pinned staging block and its device copy (same layout)
    uint8_t* h_job = nullptr;
    uint8_t* d_job = nullptr;
    size_t job_bytes = 0;
    uint32_t dcap = 0, scap = 0;
    size_t off_draws = 0, off_mats = 0, off_scenes = 0;
    SceneRun* d_runs = nullptr;
    SceneRun* h_runs = nullptr;  // pinned
    uint32_t* h_sflags = nullptr;  // pinned: scene flags the raster service writes
    FrameCounters* d_cnt = nullptr;
    FrameCounters* h_cnt = nullptr;  // pinned

    // pipeline buffers
    DBuf<uint32_t> cls, nsub, suboff, parts;
    DBuf<unsigned long long> parts64;
    DBuf<uint32_t> clip_tri, clip_draw;
    DBuf<float> clip_out;  // MAXV * 10 floats per slot
    DBuf<PrimGeom> geom;
    DBuf<float> attrs;  // ATTR_STRIDE per prim
    DBuf<float> uvat;   // UV_STRIDE per prim
    DBuf<uint4> pbin;
    DBuf<uint32_t> bin_counts, bin_offsets;
    DBuf<uint4> list;
    DBuf<uint2> bin_rng;
    DBuf<uint4> tile_rec;
    DBuf<uint2> fchunk_rec;
    DBuf<uint32_t> tchunk_draw, pchunk_scene;
    DBuf<unsigned long long> rstat_parts;
    DBuf<unsigned long long> fpack, fpre;
    DBuf<uint32_t> fine_counts, fine_off;
    DBuf<uint2> flist;
    // plain-frame outputs
    DBuf<uint32_t> out;
    DBuf<uint32_t> msaa_color;
    DBuf<float> msaa_depth;
    uint64_t gen = 0;  // bumped whenever a buffer moves (invalidates the graph)

    // graph of the batch launch sequence
    hipGraphExec_t graph = nullptr;
    uint64_t graph_gen = ~0ull;
    uint32_t graph_dcap = 0, graph_scap = 0, graph_raster = 0, graph_tiles = 0;
    uint32_t tiles_cap = 0;  // raster grid of the graph: most tiles of any job so far
    bool graph_stats = false;

    hipEvent_t done = nullptr;
    uint32_t replays = 0, ovf_seen = 0;
    bool job_graph = false;  // the job in flight was launched from the graph

    // raster service (rz_service_attach): this context's entry, or -1
    int svc_entry = -1;
    int graph_svc = -1;    // the entry the captured graph posts to
    bool job_svc = false;  // the job in flight was posted to the service
    uint32_t job_gen = 0;  // its number

    uint32_t frame_flags = 0;
    uint64_t frame_stats[11] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    std::vector<uint32_t> scene_flags;
    std::vector<double> scene_density;   // build_job's working arrays (kept to avoid reallocation)
    std::vector<uint32_t> scene_order;
    bool raster_stats = false;  // RZ_RASTER_STATS=1 at rz_create
    bool clip_wide = false;     // RZ_CLIP_WIDE=1 at rz_create: one wave per clipped polygon
    bool exact_paths = false;   // RZ_EXACT_PATHS=1 at rz_create: no per-prim fast-path proofs

    bool prof = false;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    float prof_ms[RZ_NSTAGES] = {0, 0, 0};
    uint32_t prof_frames = 0;

    std::string err;
};

namespace {

int fail(rz_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

#define HIP_TRY(ctx, expr)                                                                      \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return fail((ctx), RZ_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class T>
int grow(rz_ctx* c, DBuf<T>& b, size_t need) {
    if (b.cap >= need && b.p) return RZ_OK;
    size_t cap = std::max<size_t>(need, 1024);
    cap = cap + cap / 4;
    // a buffer still in use by queued work must not be freed under it
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (b.p) HIP_TRY(c, hipFree(b.p));
    b.p = nullptr;
    b.cap = 0;
    hipError_t e = hipMalloc(&b.p, cap * sizeof(T));
    if (e != hipSuccess) {
        b.p = nullptr;
        return fail(c, RZ_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    b.cap = cap;
    c->gen++;
    return RZ_OK;
}

template <class T>
void release(DBuf<T>& b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
}

// coarse bin side for a W x H scene: req if given, else the smallest
// power-of-two multiple of 16 giving at most `target` bins (at most 256 px)
uint32_t coarse_side(uint32_t W, uint32_t H, uint32_t req, uint32_t target) {
    if (req) return req;
    uint32_t cb = 16;
    while (cb < MAX_COARSE_PX && ((W + cb - 1) / cb) * ((H + cb - 1) / cb) > target) cb *= 2;
    return cb;
}

uint32_t nbins_of(uint32_t W, uint32_t H, uint32_t cb) { return ((W + cb - 1) / cb) * ((H + cb - 1) / cb); }

// Coarse bin side and fine tile shift of a scene for the requested fine tile
// side: the fine tile must divide the coarse bin and a bin may hold at most
// MAX_BINS fine tiles. 8x8 tiles shrink an automatic bin to 128 px when the
// scene still fits MAX_BINS bins; where the request cannot be met the scene
// uses 16x16 tiles, which fit every bin side.
void scene_bins(uint32_t W, uint32_t H, uint32_t req_cb, uint32_t target, uint32_t fine_req, uint32_t& cb,
                uint32_t& fts) {
    cb = coarse_side(W, H, req_cb, target);
    fts = 4;
    if (fine_req == 8) {
        if (!req_cb && cb > 128 && nbins_of(W, H, 128) <= MAX_BINS) cb = 128;
        if (cb <= 128) fts = 3;
    } else if (fine_req == 32) {
        if (cb % 32 == 0) fts = 5;
    }
}

uint32_t clamp_u32(uint64_t v) { return v > 0xFFFFFFF0ull ? 0xFFFFFFF0u : (uint32_t)v; }

// table block layout at capacities (dcap draws, scap scenes)
int ensure_tables(rz_ctx* c, uint32_t nd, uint32_t ns) {
    if (nd <= c->dcap && ns <= c->scap && c->h_job) return RZ_OK;
    uint32_t dcap = std::max<uint32_t>(c->dcap, 16), scap = std::max<uint32_t>(c->scap, 4);
    while (dcap < nd) dcap *= 2;
    while (scap < ns) scap *= 2;
    auto al = [](size_t v) { return (v + 255) & ~(size_t)255; };
    size_t off_d = al(sizeof(JobHead));
    size_t off_m = al(off_d + sizeof(DrawDev) * dcap);
    size_t off_s = al(off_m + sizeof(Mat3) * dcap);
    size_t bytes = al(off_s + sizeof(SceneCfg) * scap);
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (c->h_job) (void)hipHostFree(c->h_job);
    if (c->d_job) (void)hipFree(c->d_job);
    if (c->h_runs) (void)hipHostFree(c->h_runs);
    if (c->d_runs) (void)hipFree(c->d_runs);
    if (c->h_sflags) (void)hipHostFree(c->h_sflags);
    c->h_job = nullptr;
    c->d_job = nullptr;
    c->h_runs = nullptr;
    c->d_runs = nullptr;
    c->h_sflags = nullptr;
    c->dcap = c->scap = 0;
    HIP_TRY(c, hipHostMalloc(&c->h_job, bytes, hipHostMallocDefault));
    HIP_TRY(c, hipMalloc(&c->d_job, bytes));
    HIP_TRY(c, hipHostMalloc(&c->h_runs, sizeof(SceneRun) * scap, hipHostMallocDefault));
    HIP_TRY(c, hipMalloc(&c->d_runs, sizeof(SceneRun) * scap));
    HIP_TRY(c, hipHostMalloc(&c->h_sflags, sizeof(uint32_t) * scap, hipHostMallocDefault));
    memset(c->h_job, 0, bytes);
    c->job_bytes = bytes;
    c->off_draws = off_d;
    c->off_mats = off_m;
    c->off_scenes = off_s;
    c->dcap = dcap;
    c->scap = scap;
    c->gen++;
    return RZ_OK;
}

// Lay out the recorded scenes and draws (draws must be grouped by scene,
// in order) into the pinned table block.
int build_job(rz_ctx* c, JobPlan& pl) {
    const uint32_t nd = (uint32_t)c->draws.size(), ns = (uint32_t)c->scenes.size();
    int rc;
    if ((rc = ensure_tables(c, nd, ns))) return rc;
    JobHead* head = reinterpret_cast<JobHead*>(c->h_job);
    DrawDev* dd = reinterpret_cast<DrawDev*>(c->h_job + c->off_draws);
    Mat3* mats = reinterpret_cast<Mat3*>(c->h_job + c->off_mats);
    SceneCfg* sc = reinterpret_cast<SceneCfg*>(c->h_job + c->off_scenes);
    pl = JobPlan{};
    pl.ndraws = nd;
    pl.nscenes = ns;
    uint64_t tri = 0, tiles = 0, bins = 0;
    uint32_t d = 0;
    std::vector<double>& dens = c->scene_density;  // input triangles per pixel, for the raster order
    dens.assign(ns, 0.0);
    for (uint32_t s = 0; s < ns; ++s) {
        const PendingScene& ps = c->scenes[s];
        SceneCfg& g = sc[s];
        memset(&g, 0, sizeof(g));
        g.W = ps.W;
        g.H = ps.H;
        g.clear_depth = ps.clear_depth;
        if (c->recording == JobKind::Frame)
            scene_bins(ps.W, ps.H, c->coarse_req, MAX_BINS, c->fine_req, g.cb, g.fts);
        else
            scene_bins(ps.W, ps.H, 0, c->bin_target, c->fine_req, g.cb, g.fts);
        g.nbx = (ps.W + g.cb - 1) / g.cb;
        g.nby = (ps.H + g.cb - 1) / g.cb;
        g.tile_base = (uint32_t)tiles;
        g.bin_base = (uint32_t)bins;
        g.out = ps.out;
        g.msaa_color = ps.msaa_color;
        g.msaa_depth = ps.msaa_depth;
        g.tri_begin = (uint32_t)tri;
        for (; d < nd && c->draws[d].scene == s; ++d) {
            const PendingDraw& pd = c->draws[d];
            DrawDev& h = dd[d];
            memset(&h, 0, sizeof(h));
            h.pos = pd.mesh->pos;
            // only the stream the fragment shader reads is fetched (the other
            // attribute channels never reach an output)
            h.col = pd.fs == RZ_FS_COLOR ? pd.mesh->col : nullptr;
            h.uv = pd.fs == RZ_FS_TEXTURE ? pd.mesh->uv : nullptr;
            h.idx = pd.mesh->idx;
            h.ntri = pd.mesh->ntri;
            h.fs = pd.fs;
            h.scene = s;
            h.tri_base = (uint32_t)tri;
            if (pd.fs == RZ_FS_TEXTURE) {
                const Texture& t = c->pool->textures[pd.tex];
                h.tex = t.data;
                h.tw = t.tw;
                h.th = t.th;
            }
            mats[d] = pd.m;
            tri += ((uint64_t)pd.mesh->ntri + TRI_ALIGN - 1) / TRI_ALIGN * TRI_ALIGN;
            pl.tris += pd.mesh->ntri;
            dens[s] += pd.mesh->ntri;
            if (tri > 0xFFFF0000ull) return fail(c, RZ_ERR_LIMIT, "more than 2^32 triangles in one job");
        }
        g.tri_end = (uint32_t)tri;
        tiles += (uint64_t)((ps.W + 15) / 16) * ((ps.H + 15) / 16);
        bins += (uint64_t)g.nbx * g.nby;
        pl.max_nb = std::max(pl.max_nb, g.nbx * g.nby);
        pl.max_T = std::max(pl.max_T, (g.cb >> g.fts) * (g.cb >> g.fts));
    }
    if (d != nd) return fail(c, RZ_ERR_STATE, "draws are not grouped by scene");
    if (tiles > 0xFFFF0000ull || bins > 0xFFFF0000ull) return fail(c, RZ_ERR_LIMIT, "job too large");
    // raster order: the scenes' tile ranges densest first (ties in scene
    // order), so that the longest tiles do not trail at the end of the launch
    std::vector<uint32_t>& order = c->scene_order;
    order.resize(ns);
    for (uint32_t s = 0; s < ns; ++s) {
        order[s] = s;
        dens[s] /= (double)sc[s].W * sc[s].H;
    }
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return dens[a] > dens[b]; });
    uint32_t rb = 0;
    for (uint32_t s : order) {
        sc[s].raster_base = rb;
        rb += ((sc[s].W + 15) / 16) * ((sc[s].H + 15) / 16);
    }
    pl.ntri = (uint32_t)tri;
    pl.ntiles = (uint32_t)tiles;
    pl.nbins = (uint32_t)bins;
    memset(head, 0, sizeof(*head));
    head->ndraws = nd;
    head->nscenes = ns;
    head->ntri = pl.ntri;
    head->ntiles = pl.ntiles;
    head->nbins = pl.nbins;
    return RZ_OK;
}

// capacities for a job (estimates; the device reports exact needs on overflow)
struct Needs {
    uint64_t prims = 0, clip = 0, pairs = 0, fine = 0, mat = 0, fmat = 0;
};

// capacity grow() gives a buffer of cap elements asked for need
size_t grown_cap(size_t cap, size_t need, bool allocated) {
    if (cap >= need && allocated) return cap;
    size_t n = std::max<size_t>(need, 1024);
    return n + n / 4;
}

// Grows every pipeline buffer to fit a job (Real), or only adds up the bytes
// the buffers would hold afterwards (Dry: nothing is allocated).
struct RealGrow {
    rz_ctx* c;
    int rc = RZ_OK;
    template <class T>
    size_t operator()(DBuf<T>& b, size_t need) {
        if (rc == RZ_OK) rc = grow(c, b, need);
        return b.cap;
    }
};
struct DryGrow {
    uint64_t bytes = 0;
    int rc = RZ_OK;
    template <class T>
    size_t operator()(DBuf<T>& b, size_t need) {
        const size_t cap = grown_cap(b.cap, need, b.p != nullptr);
        bytes += cap * sizeof(T);
        return cap;
    }
};

template <class G>
int size_buffers(rz_ctx* c, const JobPlan& pl, const Needs& n_in, G& g) {
    Needs n = n_in;
    const uint64_t NT = pl.ntri;
    n.prims = std::max<uint64_t>(n.prims, NT + NT / 4 + 4096 + (uint64_t)pl.nscenes * CHUNK);
    n.clip = std::max<uint64_t>(n.clip, NT / 8 + 1024);
    n.pairs = std::max<uint64_t>(n.pairs, n.prims * 2);
    n.fine = std::max<uint64_t>(n.fine, std::max<uint64_t>(c->flist.cap, n.pairs * 8));
    const uint64_t lim = 0xF0000000ull;
    if (n.prims > lim || n.pairs > lim || n.fine > lim || n.clip > lim)
        return fail(c, RZ_ERR_LIMIT, "job needs more than 2^32 pipeline entries");
    const size_t cls = g(c->cls, std::max<uint64_t>(NT, 256));
    g(c->nsub, cls);
    g(c->suboff, cls);
    const size_t geom = g(c->geom, n.prims);
    g(c->attrs, geom * ATTR_STRIDE);
    g(c->uvat, geom * UV_STRIDE);
    g(c->pbin, geom);
    g(c->tchunk_draw, cls / TRI_ALIGN + 1);
    g(c->pchunk_scene, geom / CHUNK + 1);
    if (c->raster_stats) g(c->rstat_parts, 5ull * (4ull * pl.ntiles + 512 + 32ull * (uint32_t)c->n_cu));
    const size_t clip = g(c->clip_tri, n.clip);
    g(c->clip_draw, clip);
    g(c->clip_out, clip * MAXV * 10);
    const size_t list = g(c->list, n.pairs);
    // coarse matrix: sum over scenes of bins * chunks <= max_nb * (slots / CHUNK)
    const uint64_t mat = std::max<uint64_t>(n.mat, (uint64_t)pl.max_nb * (geom / CHUNK + 1));
    const size_t mat_cap = g(c->bin_counts, mat);
    g(c->bin_offsets, mat_cap);
    const size_t bins = g(c->bin_rng, std::max<uint64_t>(pl.nbins, 1));
    g(c->tile_rec, 4 * std::max<uint64_t>(pl.ntiles, 1));
    g(c->fpack, bins);
    g(c->fpre, bins);
    g(c->flist, n.fine);
    // fine-binning chunks: sum over bins of ceil(len / CHUNK) <= entries / CHUNK + bins
    g(c->fchunk_rec, list / CHUNK + bins + 1);
    // fine matrix: sum over bins of T * ceil(len / CHUNK) <= max_T * (entries / CHUNK + bins)
    const uint64_t fmat = std::max<uint64_t>(n.fmat, (uint64_t)pl.max_T * (list / CHUNK + bins + 1));
    if (mat > lim || fmat > lim) return fail(c, RZ_ERR_LIMIT, "bin matrices exceed 2^32 entries");
    const size_t fmat_cap = g(c->fine_counts, fmat);
    g(c->fine_off, fmat_cap);
    const size_t scan_n = std::max({cls, mat_cap, fmat_cap});
    g(c->parts, scan_n / SCAN_TILE + 2);
    g(c->parts64, bins / SCAN_TILE + 2);
    return g.rc;
}

int ensure_buffers(rz_ctx* c, const JobPlan& pl, const Needs& n_in) {
    RealGrow g{c};
    return size_buffers(c, pl, n_in, g);
}

// Sizes of a job (no tables written): draw(i) gives the triangle count of
// draw i, scene(i, W, H) the size of scene i.
template <class DrawTris, class SceneSize>
void plan_sizes(const rz_ctx* c, bool frame, uint32_t nd, uint32_t ns, DrawTris draw, SceneSize scene, JobPlan& pl) {
    pl = JobPlan{};
    pl.ndraws = nd;
    pl.nscenes = ns;
    uint64_t tri = 0, tiles = 0, bins = 0;
    for (uint32_t i = 0; i < nd; ++i) {
        const uint32_t n = draw(i);
        tri += ((uint64_t)n + TRI_ALIGN - 1) / TRI_ALIGN * TRI_ALIGN;
        pl.tris += n;
    }
    for (uint32_t i = 0; i < ns; ++i) {
        uint32_t W, H, cb, fts;
        scene(i, W, H);
        if (frame)
            scene_bins(W, H, c->coarse_req, MAX_BINS, c->fine_req, cb, fts);
        else
            scene_bins(W, H, 0, c->bin_target, c->fine_req, cb, fts);
        const uint32_t nb = nbins_of(W, H, cb);
        tiles += (uint64_t)((W + 15) / 16) * ((H + 15) / 16);
        bins += nb;
        pl.max_nb = std::max(pl.max_nb, nb);
        pl.max_T = std::max(pl.max_T, (cb >> fts) * (cb >> fts));
    }
    pl.ntri = clamp_u32(tri);
    pl.ntiles = clamp_u32(tiles);
    pl.nbins = clamp_u32(bins);
}

// The plan of the recorded scenes and draws.
void plan_job(const rz_ctx* c, JobPlan& pl) {
    plan_sizes(
        c, c->recording == JobKind::Frame, (uint32_t)c->draws.size(), (uint32_t)c->scenes.size(),
        [&](uint32_t i) { return c->draws[i].mesh->ntri; },
        [&](uint32_t i, uint32_t& W, uint32_t& H) {
            W = c->scenes[i].W;
            H = c->scenes[i].H;
        },
        pl);
}

template <class T>
uint64_t held(const DBuf<T>& b) {
    return b.p ? (uint64_t)b.cap * sizeof(T) : 0;
}

// device bytes of the context's pipeline buffers, outputs and tables
uint64_t ctx_bytes(const rz_ctx* c) {
    uint64_t n = held(c->cls) + held(c->nsub) + held(c->suboff) + held(c->parts) + held(c->parts64) +
                 held(c->clip_tri) + held(c->clip_draw) + held(c->clip_out) + held(c->geom) +
                 held(c->attrs) + held(c->uvat) + held(c->pbin) + held(c->bin_counts) + held(c->bin_offsets) + held(c->list) +
                 held(c->bin_rng) + held(c->tile_rec) + held(c->fchunk_rec) + held(c->tchunk_draw) +
                 held(c->pchunk_scene) + held(c->rstat_parts) + held(c->fpack) + held(c->fpre) +
                 held(c->fine_counts) + held(c->fine_off) + held(c->flist) + held(c->out) + held(c->msaa_color) +
                 held(c->msaa_depth);
    n += c->d_job ? c->job_bytes : 0;
    n += c->d_runs ? sizeof(SceneRun) * c->scap : 0;
    n += c->d_cnt ? sizeof(FrameCounters) : 0;
    return n;
}

JobBufs job_bufs(const rz_ctx* c) {
    JobBufs b{};
    b.head = reinterpret_cast<const JobHead*>(c->d_job);
    b.draws = reinterpret_cast<DrawDev*>(c->d_job + c->off_draws);
    b.mats = reinterpret_cast<const Mat3*>(c->d_job + c->off_mats);
    b.scenes = reinterpret_cast<const SceneCfg*>(c->d_job + c->off_scenes);
    b.runs = c->d_runs;
    b.cnt = c->d_cnt;
    b.svc_flags = c->h_sflags;
    b.cls = c->cls.p;
    b.nsub = c->nsub.p;
    b.suboff = c->suboff.p;
    b.clip_tri = c->clip_tri.p;
    b.clip_draw = c->clip_draw.p;
    b.clip_out = c->clip_out.p;
    b.geom = c->geom.p;
    b.attrs = c->attrs.p;
    b.uvat = c->uvat.p;
    b.pbin = c->pbin.p;
    b.bin_counts = c->bin_counts.p;
    b.bin_offsets = c->bin_offsets.p;
    b.list = c->list.p;
    b.bin_rng = c->bin_rng.p;
    b.tile_rec = c->tile_rec.p;
    b.fchunk_rec = c->fchunk_rec.p;
    b.tchunk_draw = c->tchunk_draw.p;
    b.pchunk_scene = c->pchunk_scene.p;
    b.rstat_parts = c->rstat_parts.p;
    b.fpack = c->fpack.p;
    b.fpre = c->fpre.p;
    b.fine_counts = c->fine_counts.p;
    b.fine_off = c->fine_off.p;
    b.flist = c->flist.p;
    b.parts = c->parts.p;
    b.parts64 = c->parts64.p;
    b.tri_cap = clamp_u32(c->cls.cap);
    b.clip_cap = clamp_u32(c->clip_tri.cap);
    b.prim_cap = clamp_u32(c->geom.cap);
    b.pair_cap = clamp_u32(c->list.cap);
    b.mat_cap = clamp_u32(c->bin_counts.cap);
    b.fine_cap = clamp_u32(c->flist.cap);
    b.fmat_cap = clamp_u32(c->fine_counts.cap);
    b.bins_cap = clamp_u32(c->bin_rng.cap);
    b.tile_cap = clamp_u32(c->tile_rec.cap / 4);
    b.fchunk_cap = clamp_u32(c->fchunk_rec.cap);
    b.tchunk_cap = clamp_u32(c->tchunk_draw.cap);
    b.pchunk_cap = clamp_u32(c->pchunk_scene.cap);
    b.rstat_cap = clamp_u32(c->rstat_parts.cap / 5);
    b.clip_wide = c->clip_wide ? 1u : 0u;
    b.exact_paths = c->exact_paths ? 1u : 0u;
    return b;
}

// grids: exact where the host knows the work (plain launches), capacity
// bounds otherwise (graph launches serve any job that fits the buffers)
// ---- raster service ------------------------------------------------------
constexpr double SVC_LIFE_S = 4.0;  // a service launch exits after this long (the host relaunches it); job
                                   // numbers (16 bits) cannot wrap within one launch
constexpr double SVC_WAIT_S = 30.0;    // longest host wait for a posted job

hipError_t svc_launch(rz_pool* p) {
    SvcState& v = p->svc;
    __atomic_store_n(&v.host->stop, 0u, __ATOMIC_RELEASE);
    hipError_t e = hipMemsetAsync(v.ctl, 0, sizeof(SvcCtl), v.stream);
    if (e != hipSuccess) return e;
    launch_raster_service(v.stream, v.grid, v.ents, SVC_MAX_ENTRIES, v.ctl, v.host, v.idle_ticks, v.life_ticks);
    e = hipGetLastError();
    if (e == hipSuccess) v.launches++;
    return e;
}

// the service is running, or is relaunched (it exits when idle or old)
int svc_ensure(rz_ctx* c) {
    SvcState& v = c->pool->svc;
    if (!v.on) return fail(c, RZ_ERR_STATE, "the raster service is not running (rz_service_start)");
    const hipError_t q = hipStreamQuery(v.stream);
    if (q == hipErrorNotReady) return RZ_OK;
    if (q != hipSuccess) return fail(c, RZ_ERR_HIP, std::string("raster service: ") + hipGetErrorString(q));
    const hipError_t e = svc_launch(c->pool);
    if (e != hipSuccess) return fail(c, RZ_ERR_HIP, std::string("raster service launch: ") + hipGetErrorString(e));
    return RZ_OK;
}

bool svc_job_done(const rz_ctx* c) {
    const uint32_t d = __atomic_load_n(&c->pool->svc.host->done[c->svc_entry], __ATOMIC_ACQUIRE);
    return d == (c->job_gen & 0xFFFFu);
}

// wait for the service to finish the posted job (relaunching it if it
// exited meanwhile)
int svc_wait(rz_ctx* c) {
    const auto t0 = std::chrono::steady_clock::now();
    for (uint32_t i = 0;; ++i) {
        if (svc_job_done(c)) return RZ_OK;
        if ((i & 255u) == 255u) {
            int rc = svc_ensure(c);
            if (rc) return rc;
            if (std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() > SVC_WAIT_S)
                return fail(c, RZ_ERR_HIP, "the raster service did not finish a posted job in time");
        }
        std::this_thread::yield();
    }
}

void svc_free(rz_pool* p) {
    SvcState& v = p->svc;
    if (v.on && v.stream) {
        __atomic_store_n(&v.host->stop, 1u, __ATOMIC_RELEASE);
        (void)hipStreamSynchronize(v.stream);
        v.on = false;
    }
    if (v.ents) (void)hipFree(v.ents);
    if (v.ctl) (void)hipFree(v.ctl);
    if (v.host) (void)hipHostFree(v.host);
    if (v.stream) (void)hipStreamDestroy(v.stream);
    v = SvcState{};
}

// posts the job in flight to the service instead of running the raster
bool posts_to_service(const rz_ctx* c) { return c->svc_entry >= 0 && c->recording == JobKind::Batch; }

// before each launch of a posting job: its number in the job table, and a
// running service
int svc_prepare(rz_ctx* c, const JobPlan& pl) {
    c->job_svc = posts_to_service(c);
    if (!c->job_svc) return RZ_OK;
    if (pl.ntiles > SVC_MAX_TILES) return fail(c, RZ_ERR_LIMIT, "too many tiles for one raster service job");
    SvcState& v = c->pool->svc;
    c->job_gen = (v.gen[c->svc_entry] + 1u) & 0xFFFFu;
    if (c->job_gen == (__atomic_load_n(&v.host->done[c->svc_entry], __ATOMIC_ACQUIRE) & 0xFFFFu))
        c->job_gen = (c->job_gen + 1u) & 0xFFFFu;  // never the number already marked done
    v.gen[c->svc_entry] = c->job_gen;
    reinterpret_cast<JobHead*>(c->h_job)->svc_gen = c->job_gen;
    return svc_ensure(c);
}

JobGrid job_grid(const rz_ctx* c, const JobPlan& pl, bool for_graph) {
    JobGrid g{};
    const uint32_t nd = for_graph ? c->dcap : pl.ndraws;
    const uint32_t nt = for_graph ? (uint32_t)c->cls.cap : pl.ntri;
    g.draws = (nd + 255) / 256;
    g.tris = std::min<uint32_t>((nt + 255) / 256, 8192);
    g.clip = (uint32_t)std::min<size_t>(c->clip_tri.cap, 131072);  // slots; k_clip takes CLIP_GROUPS per wave
    const uint64_t chunks = c->geom.cap / CHUNK + (for_graph ? c->scap : pl.nscenes);
    g.bins = (uint32_t)std::min<uint64_t>(chunks, 8192);  // one workgroup per chunk
    const uint64_t fch = c->list.cap / CHUNK + c->bin_rng.cap;
    g.fine = (uint32_t)std::min<uint64_t>(fch, 8192);
    g.raster = c->raster_per_cu ? c->raster_per_cu * (uint32_t)c->n_cu : 0;
    g.ntiles = for_graph ? std::max(c->tiles_cap, pl.ntiles) : pl.ntiles;
    g.scenes = for_graph ? c->scap : pl.nscenes;
    return g;
}

int enqueue_eager(rz_ctx* c, const JobPlan& pl) {
    hipStream_t s = c->stream;
    int rc = svc_prepare(c, pl);
    if (rc) return rc;
    SvcState& v = c->pool->svc;
    if (c->prof) HIP_TRY(c, hipEventRecord(c->ev[0], s));
    HIP_TRY(c, hipMemcpyAsync(c->d_job, c->h_job, c->job_bytes, hipMemcpyHostToDevice, s));
    launch_job(s, job_bufs(c), job_grid(c, pl, false), c->raster_stats, c->prof ? c->ev[1] : nullptr,
               c->prof ? c->ev[2] : nullptr, c->job_svc ? v.ents + c->svc_entry : nullptr, v.host, v.ctl,
               (uint32_t)std::max(c->svc_entry, 0));
    if (c->prof) HIP_TRY(c, hipEventRecord(c->ev[3], s));
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipMemcpyAsync(c->h_cnt, c->d_cnt, sizeof(FrameCounters), hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipMemcpyAsync(c->h_runs, c->d_runs, sizeof(SceneRun) * std::max<uint32_t>(pl.nscenes, 1),
                              hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipEventRecord(c->done, s));
    c->job_graph = false;
    return RZ_OK;
}

// capture (once per buffer generation / table capacity / tuning) and launch
int enqueue_graph(rz_ctx* c, const JobPlan& pl) {
    hipStream_t s = c->stream;
    int rc = svc_prepare(c, pl);
    if (rc) return rc;
    SvcState& v = c->pool->svc;
    const int svc = c->job_svc ? c->svc_entry : -1;
    if (pl.ntiles > c->tiles_cap) c->tiles_cap = std::max<uint32_t>(pl.ntiles, c->tiles_cap + c->tiles_cap / 4);
    const JobGrid g = job_grid(c, pl, true);
    if (!c->graph || c->graph_gen != c->gen || c->graph_dcap != c->dcap || c->graph_scap != c->scap ||
        c->graph_raster != g.raster || c->graph_tiles != g.ntiles || c->graph_stats != c->raster_stats ||
        c->graph_svc != svc) {
        if (c->graph) {
            HIP_TRY(c, hipStreamSynchronize(s));
            (void)hipGraphExecDestroy(c->graph);
            c->graph = nullptr;
        }
        hipGraph_t graph = nullptr;
        HIP_TRY(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        hipError_t e = hipMemcpyAsync(c->d_job, c->h_job, c->job_bytes, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) {
            launch_job(s, job_bufs(c), g, c->raster_stats, nullptr, nullptr, svc >= 0 ? v.ents + svc : nullptr,
                       v.host, v.ctl, (uint32_t)std::max(svc, 0));
            e = hipGetLastError();
        }
        if (e == hipSuccess)
            e = hipMemcpyAsync(c->h_cnt, c->d_cnt, sizeof(FrameCounters), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(c->h_runs, c->d_runs, sizeof(SceneRun) * c->scap, hipMemcpyDeviceToHost, s);
        hipError_t e2 = hipStreamEndCapture(s, &graph);
        if (e != hipSuccess || e2 != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            return fail(c, RZ_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e != hipSuccess ? e : e2));
        }
        e = hipGraphInstantiate(&c->graph, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            c->graph = nullptr;
            return fail(c, RZ_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        }
        c->graph_gen = c->gen;
        c->graph_dcap = c->dcap;
        c->graph_scap = c->scap;
        c->graph_raster = g.raster;
        c->graph_tiles = g.ntiles;
        c->graph_stats = c->raster_stats;
        c->graph_svc = svc;
    }
    if (c->prof) HIP_TRY(c, hipEventRecord(c->ev[0], s));
    HIP_TRY(c, hipGraphLaunch(c->graph, s));
    if (c->prof) HIP_TRY(c, hipEventRecord(c->ev[3], s));
    HIP_TRY(c, hipEventRecord(c->done, s));
    c->job_graph = true;
    return RZ_OK;
}

int enqueue(rz_ctx* c, bool graph) {
    int rc = ensure_buffers(c, c->plan, Needs{});
    if (rc) return rc;
    return graph ? enqueue_graph(c, c->plan) : enqueue_eager(c, c->plan);
}

// the recorded draws hold their meshes (pending) until the job is done
void release_job(rz_ctx* c) {
    for (PendingDraw& d : c->draws) d.mesh->pending--;
    c->draws.clear();
    c->scenes.clear();
    c->recording = JobKind::None;
    c->in_flight = false;
}

// After the job's completion event: on overflow grow and replay
// (asynchronously); else record the results. Returns 1 when finished,
// 0 when a replay was queued, 2 while the raster service has not finished
// the posted job, < 0 on error.
int complete_job(rz_ctx* c) {
    const FrameCounters fc = *c->h_cnt;
    c->ovf_seen |= fc.overflow;
    if (fc.overflow) {
        if (fc.overflow & OVF_POLY) return fail(c, RZ_ERR_LIMIT, "a clipped polygon exceeded 16 vertices");
        if (++c->replays > MAX_REPLAYS) return fail(c, RZ_ERR_LIMIT, "job did not fit after repeated growth");
        Needs n;
        n.prims = std::max<uint64_t>(fc.nprims, c->geom.cap);
        n.clip = std::max<uint64_t>(fc.nclip, c->clip_tri.cap);
        // the coarse total is exact only when the coarse count pass ran (no
        // overflow before binning); otherwise the matrix holds stale counts
        const bool binned = !(fc.overflow & (OVF_PRIMS | OVF_CLIP | OVF_POLY | OVF_MATRIX));
        n.pairs = std::max<uint64_t>(binned ? fc.pairs_binned : 0u, c->list.cap);
        n.fine = std::max<uint64_t>(fc.fine_total, c->flist.cap);
        n.mat = std::max<uint64_t>(fc.mat_len, c->bin_counts.cap);
        n.fmat = std::max<uint64_t>(fc.fmat_len, c->fine_counts.cap);
        if (fc.overflow & OVF_PRIMS) n.prims = std::max<uint64_t>(n.prims, c->geom.cap * 2);
        if (fc.overflow & OVF_CLIP) n.clip = std::max<uint64_t>(n.clip, c->clip_tri.cap * 2);
        if (fc.overflow & OVF_PAIRS) n.pairs = std::max<uint64_t>(n.pairs, c->list.cap * 2);
        if (fc.overflow & OVF_FINE) n.fine = std::max<uint64_t>(n.fine, c->flist.cap * 2);
        int rc = ensure_buffers(c, c->plan, n);
        if (rc) return rc;
        rc = c->job_graph ? enqueue_graph(c, c->plan) : enqueue_eager(c, c->plan);
        return rc ? rc : 0;
    }
    const uint32_t ns = c->plan.nscenes;
    if (c->job_svc) {
        // the front end is done and the job posted; the service writes the
        // scenes' final flags to host memory before it marks the job done
        if (!svc_job_done(c)) return 2;
        for (uint32_t s = 0; s < ns; ++s) c->h_runs[s].flags = __atomic_load_n(&c->h_sflags[s], __ATOMIC_ACQUIRE);
        c->job_svc = false;
    }
    c->scene_flags.assign(ns, 0);
    uint32_t fl = 0;
    for (uint32_t s = 0; s < ns; ++s) {
        c->scene_flags[s] = c->h_runs[s].flags;
        fl |= c->h_runs[s].flags;
    }
    c->frame_flags = fl;
    c->frame_stats[0] = c->plan.tris;
    c->frame_stats[1] = fc.nclip;
    c->frame_stats[2] = fc.nprims_raw;
    c->frame_stats[3] = fc.pairs_binned;
    c->frame_stats[4] = c->replays;
    for (int i = 0; i < 5; ++i) c->frame_stats[5 + i] = fc.rstat[i];
    c->frame_stats[10] = c->ovf_seen;
    if (c->prof) {
        float a = 0, b = 0, d = 0;
        if (c->job_graph) {
            HIP_TRY(c, hipEventElapsedTime(&d, c->ev[0], c->ev[3]));
            c->prof_ms[RZ_STAGE_RASTER] += d;  // the whole graph
        } else {
            HIP_TRY(c, hipEventElapsedTime(&a, c->ev[0], c->ev[1]));
            HIP_TRY(c, hipEventElapsedTime(&b, c->ev[1], c->ev[2]));
            HIP_TRY(c, hipEventElapsedTime(&d, c->ev[2], c->ev[3]));
            c->prof_ms[RZ_STAGE_GEOMETRY] += a;
            c->prof_ms[RZ_STAGE_BINNING] += b;
            c->prof_ms[RZ_STAGE_RASTER] += d;
        }
        c->prof_frames++;
    }
    return 1;
}

// graph: replay the captured graph for jobs of at most graph_max_tris input
// triangles, launch directly otherwise
int start_job(rz_ctx* c, bool graph, uint64_t graph_max_tris = ~0ull) {
    int rc = build_job(c, c->plan);
    if (rc) return rc;
    c->replays = 0;
    c->ovf_seen = 0;
    c->in_flight = true;
    return enqueue(c, graph && c->plan.tris <= graph_max_tris);
}

// Frames of at most this many input triangles replay the captured graph: one
// submission instead of a launch per kernel, for frames whose time is the
// host's launches (a 12-triangle cube frame: mean 0.115 -> 0.103 ms, p99
// 0.19 -> 0.11 ms). Larger frames launch directly: the graph's grids, sized
// to the buffer capacities, made a 1M-triangle frame about 1% slower.
constexpr uint64_t FRAME_GRAPH_TRIS = 65536;

int wait_job(rz_ctx* c) {
    for (;;) {
        HIP_TRY(c, hipEventSynchronize(c->done));
        int rc = complete_job(c);
        if (rc == 2) {
            if ((rc = svc_wait(c))) return rc;
            continue;
        }
        if (rc < 0) return rc;
        if (rc == 1) return RZ_OK;
    }
}

int frame_impl(rz_ctx* c, uint32_t* host_out, uint32_t* dev_out, uint32_t* host_color4, float* host_depth4) {
    if (!c) return RZ_ERR_ARG;
    if (c->recording == JobKind::Batch || c->in_flight) return fail(c, RZ_ERR_STATE, "a batch is being recorded or in flight");
    int rc;
    const size_t npx = (size_t)c->W * c->H;
    uint32_t* target = dev_out;
    if (!target) {
        if ((rc = grow(c, c->out, npx))) return rc;
        target = c->out.p;
    }
    const bool want_msaa = host_color4 || host_depth4;
    if (want_msaa) {
        if ((rc = grow(c, c->msaa_color, npx * 4))) return rc;
        if ((rc = grow(c, c->msaa_depth, npx * 4))) return rc;
    }
    c->recording = JobKind::Frame;
    c->scenes.clear();
    c->scenes.push_back(PendingScene{c->W, c->H, c->clear_depth, target, want_msaa ? c->msaa_color.p : nullptr,
                                     want_msaa ? c->msaa_depth.p : nullptr});
    // small frames replay the captured graph; direct launches while
    // profiling, whose stage events sit between the kernels
    rc = start_job(c, c->use_graph && !c->prof, FRAME_GRAPH_TRIS);
    if (!rc) rc = wait_job(c);
    release_job(c);
    c->clear_depth = c->next_clear_depth;
    if (rc) return rc;
    // on the context's stream (non-blocking), so that a running raster
    // service, whose stream the null stream waits for, does not hold them up
    if (host_out)
        HIP_TRY(c, hipMemcpyAsync(host_out, target, npx * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    if (host_color4)
        HIP_TRY(c, hipMemcpyAsync(host_color4, c->msaa_color.p, npx * 4 * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                  c->stream));
    if (host_depth4)
        HIP_TRY(c, hipMemcpyAsync(host_depth4, c->msaa_depth.p, npx * 4 * sizeof(float), hipMemcpyDeviceToHost,
                                  c->stream));
    if (host_out || host_color4 || host_depth4) HIP_TRY(c, hipStreamSynchronize(c->stream));
    return RZ_OK;
}

int mesh_create(rz_ctx* c, const float* pos, const float* color, const float* uv, uint32_t nv, const uint32_t* idx,
                uint64_t ni, rz_mesh** out, hipMemcpyKind kind) {
    if (!c) return RZ_ERR_ARG;
    if (!out || !pos || (!idx && ni)) return fail(c, RZ_ERR_ARG, "null pointer");
    if (ni % 3 != 0) return fail(c, RZ_ERR_ARG, "index count is not a multiple of 3");
    if (ni / 3 > 0xFFFFFFFFull / 2) return fail(c, RZ_ERR_LIMIT, "too many triangles");
    if (kind == hipMemcpyHostToDevice) {
        // the reference indexes vertex_buf[idx] and would panic (render.rs:78-83)
        for (uint64_t i = 0; i < ni; ++i)
            if (idx[i] >= nv) return fail(c, RZ_ERR_ARG, "index out of range");
    }
    rz_mesh* m = new rz_mesh();
    m->pool = c->pool;
    m->nv = nv;
    m->ntri = (uint32_t)(ni / 3);
    auto up = [&](void** dst, const void* src, size_t bytes) -> int {
        if (!src || bytes == 0) return RZ_OK;
        hipError_t e = hipMalloc(dst, bytes);
        if (e != hipSuccess) return fail(c, RZ_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
        m->bytes += bytes;
        // host data: on the context's stream (not held up by a running
        // raster service); device data: on the null stream, ordered after
        // the caller's work there that produced it
        e = kind == hipMemcpyHostToDevice ? hipMemcpyAsync(*dst, src, bytes, kind, c->stream)
                                          : hipMemcpy(*dst, src, bytes, kind);
        if (e == hipSuccess && kind == hipMemcpyHostToDevice) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) return fail(c, RZ_ERR_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
        return RZ_OK;
    };
    int rc = up((void**)&m->pos, pos, (size_t)nv * 3 * sizeof(float));
    if (!rc) rc = up((void**)&m->col, color, (size_t)nv * 4 * sizeof(float));
    if (!rc) rc = up((void**)&m->uv, uv, (size_t)nv * 2 * sizeof(float));
    if (!rc) rc = up((void**)&m->idx, idx, (size_t)ni * sizeof(uint32_t));
    if (!rc && kind == hipMemcpyDeviceToDevice && ni) {
        // same check as the host path, on the device: largest index < nv
        uint32_t* dmax = nullptr;
        uint32_t hmax = 0;
        hipError_t e = hipMalloc(&dmax, sizeof(uint32_t));
        if (e == hipSuccess) e = hipMemset(dmax, 0, sizeof(uint32_t));
        if (e == hipSuccess) {
            launch_max_index(c->stream, m->idx, ni, dmax);
            e = hipMemcpyAsync(&hmax, dmax, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream);
        }
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (dmax) (void)hipFree(dmax);
        if (e != hipSuccess)
            rc = fail(c, RZ_ERR_HIP, std::string("index check: ") + hipGetErrorString(e));
        else if (hmax >= nv)
            rc = fail(c, RZ_ERR_ARG, "index out of range");
    }
    if (rc) {
        if (m->pos) (void)hipFree(m->pos);
        if (m->col) (void)hipFree(m->col);
        if (m->uv) (void)hipFree(m->uv);
        if (m->idx) (void)hipFree(m->idx);
        delete m;
        return rc;
    }
    c->pool->resident_bytes += m->bytes;
    *out = m;
    return RZ_OK;
}

int check_shaders(rz_ctx* c, int vs, int fs) {
    if (vs != RZ_VS_MVP) return fail(c, RZ_ERR_ARG, "unknown vertex shader");
    if (fs != RZ_FS_TEXTURE && fs != RZ_FS_COLOR && fs != RZ_FS_DEBUG)
        return fail(c, RZ_ERR_ARG, "unknown fragment shader");
    return RZ_OK;
}

int create_impl(rz_pool* pool, uint32_t width, uint32_t height, rz_ctx** out) {
    if (!out) return RZ_ERR_ARG;
    *out = nullptr;
    if (width == 0 || height == 0 || width > MAX_DIM || height > MAX_DIM) return RZ_ERR_ARG;
    // the screen must fit MAX_BINS coarse bins of at most MAX_COARSE_PX pixels
    if (((width + MAX_COARSE_PX - 1) / MAX_COARSE_PX) * ((height + MAX_COARSE_PX - 1) / MAX_COARSE_PX) > MAX_BINS)
        return RZ_ERR_LIMIT;
    rz_ctx* c = new rz_ctx();
    c->W = width;
    c->H = height;
    for (int i = 0; i < 16; ++i) {
        float v = (i % 5 == 0) ? 1.0f : 0.0f;  // Uniforms::new: identities (uniform.rs:17-27)
        c->uni.world[i] = c->uni.view[i] = c->uni.proj[i] = v;
    }
    const char* st = getenv("RZ_RASTER_STATS");
    c->raster_stats = st && st[0] == '1';
    const char* cw = getenv("RZ_CLIP_WIDE");
    c->clip_wide = cw && cw[0] == '1';
    const char* ex = getenv("RZ_EXACT_PATHS");
    c->exact_paths = ex && ex[0] == '1';
    if (pool) {
        c->pool = pool;
    } else {
        c->pool = new rz_pool();
        if (hipGetDevice(&c->pool->device) != hipSuccess) {
            delete c->pool;
            delete c;
            return RZ_ERR_HIP;
        }
    }
    c->pool->ctxs.push_back(c);
    auto bad = [&](hipError_t) {
        rz_destroy(c);
        return RZ_ERR_HIP;
    };
    hipError_t e;
    int dev = 0;
    if ((e = hipGetDevice(&dev)) != hipSuccess) return bad(e);
    if (dev != c->pool->device) {
        rz_destroy(c);
        return RZ_ERR_ARG;  // contexts sharing a pool must be on its device
    }
    if ((e = hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess)
        return bad(e);
    if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess) return bad(e);
    if ((e = hipMalloc(&c->d_cnt, sizeof(FrameCounters))) != hipSuccess) return bad(e);
    if ((e = hipHostMalloc(&c->h_cnt, sizeof(FrameCounters), hipHostMallocDefault)) != hipSuccess) return bad(e);
    for (int i = 0; i < 4; ++i)
        if ((e = hipEventCreate(&c->ev[i])) != hipSuccess) return bad(e);
    if ((e = hipEventCreateWithFlags(&c->done, hipEventDisableTiming)) != hipSuccess) return bad(e);
    memset(c->h_cnt, 0, sizeof(FrameCounters));
    if (ensure_tables(c, 16, 4) != RZ_OK) {
        rz_destroy(c);
        return RZ_ERR_NOMEM;
    }
    *out = c;
    return RZ_OK;
}

}  // namespace

// ===========================================================================
// Allocation-free frame (rz_render_noalloc)
//
// Every working buffer of the pipeline is carved out of `scratch`, one
// file-scope device array per GPU, held until the process ends. The job
// table is written by a kernel from its arguments, so the call makes no
// device allocation and no host-to-device copy; it reads back the frame
// counters (and, before that, the largest index) into host stack memory.
// Calls that overlap on one GPU share the array and race.
// ===========================================================================
namespace {

constexpr size_t NOALLOC_SCRATCH_BYTES = 128ull << 20;
constexpr int NOALLOC_MAX_DEVICES = 64;

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

// per device: the array's device address (hipGetSymbolAddress, taken once
// with that device current); the array's name is only a host-side stand-in
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

hipError_t scratch_for_device(int dev, unsigned char** out) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

// Bump allocator over [0, cap): offsets only (base may be null while sizing).
struct Carve {
    unsigned char* base;
    size_t off = 0, cap;
    template <class T>
    T* take(size_t n) {
        off = (off + 255) & ~(size_t)255;
        T* p = reinterpret_cast<T*>(reinterpret_cast<uintptr_t>(base) + off);
        off += n * sizeof(T);
        return p;
    }
};

// the one-scene, one-draw job table as written by k_write_words: JobHead at
// word 0, DrawDev at 64, Mat3 at 128, SceneCfg at 192 (256-byte offsets)
constexpr uint32_t NA_OFF_DRAW = 64, NA_OFF_MAT = 128, NA_OFF_SCENE = 192, NA_TABLE_WORDS = 212;
static_assert(sizeof(JobHead) <= 4 * NA_OFF_DRAW && sizeof(DrawDev) <= 4 * (NA_OFF_MAT - NA_OFF_DRAW) &&
                  sizeof(Mat3) <= 4 * (NA_OFF_SCENE - NA_OFF_MAT) &&
                  4 * NA_OFF_SCENE + sizeof(SceneCfg) <= 4 * NA_TABLE_WORDS && NA_TABLE_WORDS <= WORD_BLOCK,
              "allocation-free job table layout");

// Lays the working buffers of a W x H frame of ntri_pad (padded) triangles
// out in the scratch array: fixed-size buffers from the frame's shape, then
// the coarse and fine lists split what is left. Returns false if the fixed
// part does not fit.
bool noalloc_layout(unsigned char* base, uint32_t ntri_pad, const SceneCfg& sc, uint32_t ntiles, JobBufs& b,
                    uint32_t** maxidx, size_t* fixed_bytes) {
    Carve cv{base, 0, NOALLOC_SCRATCH_BYTES};
    uint32_t* table = cv.take<uint32_t>(NA_TABLE_WORDS);
    b = JobBufs{};
    b.head = reinterpret_cast<const JobHead*>(table);
    b.draws = reinterpret_cast<DrawDev*>(table + NA_OFF_DRAW);
    b.mats = reinterpret_cast<const Mat3*>(table + NA_OFF_MAT);
    b.scenes = reinterpret_cast<const SceneCfg*>(table + NA_OFF_SCENE);
    b.runs = cv.take<SceneRun>(1);
    b.cnt = cv.take<FrameCounters>(1);
    *maxidx = cv.take<uint32_t>(64);
    const uint64_t nbins = (uint64_t)sc.nbx * sc.nby, T = (uint64_t)(sc.cb >> sc.fts) * (sc.cb >> sc.fts);
    const uint64_t tri_cap = std::max<uint64_t>(ntri_pad, 256);
    const uint64_t prims = ntri_pad + ntri_pad / 4 + 4096 + CHUNK;
    const uint64_t clip = ntri_pad / 8 + 1024;
    const uint64_t mat = nbins * (prims / CHUNK + 1);
    b.cls = cv.take<uint32_t>(tri_cap);
    b.nsub = cv.take<uint32_t>(tri_cap);
    b.suboff = cv.take<uint32_t>(tri_cap);
    b.tchunk_draw = cv.take<uint32_t>(tri_cap / TRI_ALIGN + 1);
    b.geom = cv.take<PrimGeom>(prims);
    b.attrs = cv.take<float>(prims * ATTR_STRIDE);
    b.uvat = cv.take<float>(prims * UV_STRIDE);
    b.pbin = cv.take<uint4>(prims);
    b.pchunk_scene = cv.take<uint32_t>(prims / CHUNK + 1);
    b.clip_tri = cv.take<uint32_t>(clip);
    b.clip_draw = cv.take<uint32_t>(clip);
    b.clip_out = cv.take<float>(clip * MAXV * 10);
    b.bin_counts = cv.take<uint32_t>(mat);
    b.bin_offsets = cv.take<uint32_t>(mat);
    b.bin_rng = cv.take<uint2>(nbins);
    b.tile_rec = cv.take<uint4>(4 * (size_t)ntiles);
    b.fpack = cv.take<unsigned long long>(nbins);
    b.fpre = cv.take<unsigned long long>(nbins);
    b.parts64 = cv.take<unsigned long long>(nbins / SCAN_TILE + 2);
    // what is left: a third for the coarse list (16 B per entry, plus its
    // share of the fine-binning chunk table and fine count matrices), the
    // rest for the 8-byte fine entries and the scan partials
    const size_t slack = 16 * 256;
    *fixed_bytes = cv.off + slack;
    if (cv.off + slack >= NOALLOC_SCRATCH_BYTES) return false;
    const size_t left = NOALLOC_SCRATCH_BYTES - cv.off - slack;
    // bytes per coarse entry: 16 in the list, plus the fine-binning chunk
    // record (8 B) and two fine count matrices (8 B per tile) per CHUNK entries
    const size_t per_pair = 16 + (8 + 8 * T) / CHUNK + 1;
    uint64_t pairs = left / 3 / per_pair;
    const uint64_t fch = pairs / CHUNK + nbins + 1;
    const uint64_t fmat = T * fch;
    if (pairs < 1024) return false;
    b.list = cv.take<uint4>(pairs);
    b.fchunk_rec = cv.take<uint2>(fch);
    b.fine_counts = cv.take<uint32_t>(fmat);
    b.fine_off = cv.take<uint32_t>(fmat);
    const uint64_t scan_n = std::max<uint64_t>({tri_cap, mat, fmat});
    b.parts = cv.take<uint32_t>(scan_n / SCAN_TILE + 2);
    cv.off = (cv.off + 255) & ~(size_t)255;
    if (cv.off + 4096 >= NOALLOC_SCRATCH_BYTES) return false;
    const uint64_t fine = (NOALLOC_SCRATCH_BYTES - cv.off - 256) / sizeof(uint2);
    b.flist = cv.take<uint2>(fine);
    if (cv.off > NOALLOC_SCRATCH_BYTES) return false;
    const uint64_t lim = 0xF0000000ull;
    b.tri_cap = (uint32_t)std::min(tri_cap, lim);
    b.clip_cap = (uint32_t)clip;
    b.prim_cap = (uint32_t)prims;
    b.pair_cap = (uint32_t)pairs;
    b.mat_cap = (uint32_t)mat;
    b.fine_cap = (uint32_t)fine;
    b.fmat_cap = (uint32_t)fmat;
    b.bins_cap = (uint32_t)nbins;
    b.tile_cap = ntiles;
    b.fchunk_cap = (uint32_t)fch;
    b.tchunk_cap = (uint32_t)(tri_cap / TRI_ALIGN + 1);
    b.pchunk_cap = (uint32_t)(prims / CHUNK + 1);
    b.rstat_parts = nullptr;
    b.rstat_cap = 0;
    b.clip_wide = 0;
    b.exact_paths = 0;
    return true;
}

int noalloc_fail(const char** message, int code, const char* msg) {
    if (message) *message = msg;
    return code;
}

}  // namespace

extern "C" int rz_render_noalloc(void* stream, uint32_t width, uint32_t height, float clear_depth,
                                 const float world[16], const float view[16], const float proj[16],
                                 const float* pos_dev, const float* color_dev, const float* uv_dev, uint32_t nv,
                                 const uint32_t* idx_dev, uint64_t ni, const uint32_t* texture_dev, uint32_t tw,
                                 uint32_t th, int vertex_shader, int fragment_shader, uint32_t* argb_dev,
                                 uint32_t* flags_out, const char** message) {
    if (message) *message = "";
    if (flags_out) *flags_out = 0;
    // argument checks first: a rejected call touches no device
    if (width == 0 || height == 0 || width > MAX_DIM || height > MAX_DIM)
        return noalloc_fail(message, RZ_ERR_ARG, "width and height must be in [1, 16384]");
    if (((width + MAX_COARSE_PX - 1) / MAX_COARSE_PX) * ((height + MAX_COARSE_PX - 1) / MAX_COARSE_PX) > MAX_BINS)
        return noalloc_fail(message, RZ_ERR_LIMIT, "the screen needs more than 256 coarse bins");
    if (!world || !view || !proj) return noalloc_fail(message, RZ_ERR_ARG, "null matrix");
    if (!argb_dev) return noalloc_fail(message, RZ_ERR_ARG, "null output");
    if (!pos_dev || (!idx_dev && ni)) return noalloc_fail(message, RZ_ERR_ARG, "null pointer");
    if (ni % 3 != 0) return noalloc_fail(message, RZ_ERR_ARG, "index count is not a multiple of 3");
    if (vertex_shader != RZ_VS_MVP) return noalloc_fail(message, RZ_ERR_ARG, "unknown vertex shader");
    if (fragment_shader != RZ_FS_TEXTURE && fragment_shader != RZ_FS_COLOR && fragment_shader != RZ_FS_DEBUG)
        return noalloc_fail(message, RZ_ERR_ARG, "unknown fragment shader");
    if (fragment_shader == RZ_FS_TEXTURE && (!texture_dev || tw == 0 || th == 0))
        return noalloc_fail(message, RZ_ERR_STATE, "texture fragment shader with no texture");
    if (fragment_shader == RZ_FS_TEXTURE && (uint64_t)tw * th >= (1ull << 31))
        return noalloc_fail(message, RZ_ERR_LIMIT, "texture too large");
    const uint64_t ntri = ni / 3;
    if (ntri > NOALLOC_MAX_TRIS)
        return noalloc_fail(message, RZ_ERR_LIMIT, "more triangles than the allocation-free path takes (262144)");
    const uint32_t ntri_pad = (uint32_t)((ntri + TRI_ALIGN - 1) / TRI_ALIGN * TRI_ALIGN);

    // the job table (one scene, one draw), as rz_draw + rz_resolve_device build it
    WordBlock blk;
    memset(&blk, 0, sizeof(blk));
    JobHead* head = reinterpret_cast<JobHead*>(blk.w);
    DrawDev* dd = reinterpret_cast<DrawDev*>(blk.w + NA_OFF_DRAW);
    Mat3* mat = reinterpret_cast<Mat3*>(blk.w + NA_OFF_MAT);
    SceneCfg* sc = reinterpret_cast<SceneCfg*>(blk.w + NA_OFF_SCENE);
    sc->W = width;
    sc->H = height;
    sc->clear_depth = clear_depth;
    scene_bins(width, height, 0, MAX_BINS, 16, sc->cb, sc->fts);
    sc->nbx = (width + sc->cb - 1) / sc->cb;
    sc->nby = (height + sc->cb - 1) / sc->cb;
    sc->tri_begin = 0;
    sc->tri_end = ntri_pad;
    sc->out = argb_dev;
    const uint32_t ntiles = ((width + 15) / 16) * ((height + 15) / 16);
    dd->pos = pos_dev;
    dd->col = fragment_shader == RZ_FS_COLOR ? color_dev : nullptr;
    dd->uv = fragment_shader == RZ_FS_TEXTURE ? uv_dev : nullptr;
    dd->idx = idx_dev;
    dd->ntri = (uint32_t)ntri;
    dd->fs = (uint32_t)fragment_shader;
    if (fragment_shader == RZ_FS_TEXTURE) {
        dd->tex = texture_dev;
        dd->tw = tw;
        dd->th = th;
    }
    memcpy(mat->world, world, sizeof(mat->world));
    memcpy(mat->view, view, sizeof(mat->view));
    memcpy(mat->proj, proj, sizeof(mat->proj));
    head->ndraws = 1;
    head->nscenes = 1;
    head->ntri = ntri_pad;
    head->ntiles = ntiles;
    head->nbins = sc->nbx * sc->nby;

    JobBufs b;
    uint32_t* d_maxidx = nullptr;
    size_t fixed = 0;
    if (!noalloc_layout(nullptr, ntri_pad, *sc, ntiles, b, &d_maxidx, &fixed))
        return noalloc_fail(message, RZ_ERR_LIMIT,
                            "the frame's fixed buffers do not fit the 128 MiB scratch of the allocation-free path");

    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return noalloc_fail(message, RZ_ERR_HIP, "hipGetDevice failed");
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return noalloc_fail(message, RZ_ERR_LIMIT, "device ordinal above 63");
    unsigned char* base = nullptr;
    if (scratch_for_device(dev, &base) != hipSuccess)
        return noalloc_fail(message, RZ_ERR_HIP, "hipGetSymbolAddress of the scratch array failed");
    noalloc_layout(base, ntri_pad, *sc, ntiles, b, &d_maxidx, &fixed);
    hipStream_t s = static_cast<hipStream_t>(stream);

    // every index < nv (the reference indexes vertex_buf[idx], render.rs:78-83)
    if (ni) {
        uint32_t hmax = 0;
        launch_fill_u32(s, d_maxidx, 1, 0u);
        launch_max_index(s, idx_dev, ni, d_maxidx);
        if (hipMemcpyAsync(&hmax, d_maxidx, sizeof(uint32_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return noalloc_fail(message, RZ_ERR_HIP, "index check failed");
        if (hmax >= nv) return noalloc_fail(message, RZ_ERR_ARG, "index out of range");
    }

    launch_write_words(s, blk, NA_TABLE_WORDS, reinterpret_cast<uint32_t*>(const_cast<JobHead*>(b.head)));
    JobGrid g{};
    g.draws = 1;
    g.tris = std::min<uint32_t>((ntri_pad + 255) / 256, 8192);
    g.clip = std::min<uint32_t>(b.clip_cap, 131072);
    g.bins = std::min<uint32_t>(b.prim_cap / CHUNK + 1, 8192);
    g.fine = std::min<uint32_t>(b.pair_cap / CHUNK + b.bins_cap, 8192);
    g.raster = 0;
    g.ntiles = ntiles;
    g.scenes = 1;
    launch_job(s, b, g, false, nullptr, nullptr);
    if (hipGetLastError() != hipSuccess) return noalloc_fail(message, RZ_ERR_HIP, "kernel launch failed");
    FrameCounters fc;
    SceneRun run;
    if (hipMemcpyAsync(&fc, b.cnt, sizeof(fc), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(&run, b.runs, sizeof(run), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return noalloc_fail(message, RZ_ERR_HIP, "frame failed on the device");
    if (fc.overflow & OVF_POLY) return noalloc_fail(message, RZ_ERR_LIMIT, "a clipped polygon exceeded 16 vertices");
    // what outgrew the array (rz_resolve_device grows its buffers instead)
    if (fc.overflow & OVF_PRIMS)
        return noalloc_fail(message, RZ_ERR_LIMIT, "the frame's primitives do not fit the 128 MiB scratch");
    if (fc.overflow & OVF_CLIP)
        return noalloc_fail(message, RZ_ERR_LIMIT, "the frame's clipped triangles do not fit the 128 MiB scratch");
    if (fc.overflow & OVF_MATRIX)
        return noalloc_fail(message, RZ_ERR_LIMIT, "the coarse count matrix does not fit the 128 MiB scratch");
    if (fc.overflow & OVF_PAIRS)
        return noalloc_fail(message, RZ_ERR_LIMIT, "the coarse bin lists do not fit the 128 MiB scratch");
    if (fc.overflow)
        return noalloc_fail(message, RZ_ERR_LIMIT, "the tile bin lists do not fit the 128 MiB scratch");
    if (flags_out) *flags_out = run.flags;
    return RZ_OK;
}

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int rz_create(uint32_t width, uint32_t height, rz_ctx** out) { return create_impl(nullptr, width, height, out); }

int rz_create_shared(rz_ctx* parent, uint32_t width, uint32_t height, rz_ctx** out) {
    if (!parent || !out) return RZ_ERR_ARG;
    return create_impl(parent->pool, width, height, out);
}

int rz_destroy(rz_ctx* c) {
    if (!c) return RZ_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    // a job posted to the raster service still reads this context's buffers
    if (c->in_flight && c->job_svc && c->h_cnt && c->h_cnt->overflow == 0 && c->pool->svc.on && !svc_job_done(c))
        (void)svc_wait(c);
    if (c->svc_entry >= 0) c->pool->svc.used &= ~(1u << c->svc_entry);
    for (PendingDraw& d : c->draws) d.mesh->pending--;
    c->draws.clear();
    if (c->graph) (void)hipGraphExecDestroy(c->graph);
    release(c->cls);
    release(c->nsub);
    release(c->suboff);
    release(c->parts);
    release(c->parts64);
    release(c->clip_tri);
    release(c->clip_draw);
    release(c->clip_out);
    release(c->geom);
    release(c->attrs);
    release(c->uvat);
    release(c->pbin);
    release(c->bin_counts);
    release(c->bin_offsets);
    release(c->list);
    release(c->bin_rng);
    release(c->tile_rec);
    release(c->fchunk_rec);
    release(c->tchunk_draw);
    release(c->pchunk_scene);
    release(c->rstat_parts);
    release(c->fpack);
    release(c->fpre);
    release(c->fine_counts);
    release(c->fine_off);
    release(c->flist);
    release(c->out);
    release(c->msaa_color);
    release(c->msaa_depth);
    if (c->h_job) (void)hipHostFree(c->h_job);
    if (c->d_job) (void)hipFree(c->d_job);
    if (c->h_runs) (void)hipHostFree(c->h_runs);
    if (c->d_runs) (void)hipFree(c->d_runs);
    if (c->h_sflags) (void)hipHostFree(c->h_sflags);
    if (c->d_cnt) (void)hipFree(c->d_cnt);
    if (c->h_cnt) (void)hipHostFree(c->h_cnt);
    for (int i = 0; i < 4; ++i)
        if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);
    if (c->done) (void)hipEventDestroy(c->done);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    rz_pool* pool = c->pool;
    if (pool) {
        pool->ctxs.erase(std::remove(pool->ctxs.begin(), pool->ctxs.end(), c), pool->ctxs.end());
        if (pool->ctxs.empty()) {
            svc_free(pool);
            for (Texture& t : pool->textures)
                if (t.data) (void)hipFree(t.data);
            delete pool;
        }
    }
    delete c;
    return RZ_OK;
}

const char* rz_last_error(const rz_ctx* c) { return c ? c->err.c_str() : kNullCtxMsg; }

void* rz_stream(const rz_ctx* c) { return c ? (void*)c->stream : nullptr; }

int rz_set_clear_depth(rz_ctx* c, float depth) {
    if (!c) return RZ_ERR_ARG;
    c->next_clear_depth = depth;
    if (c->draws.empty() || c->recording != JobKind::Frame) c->clear_depth = depth;
    return RZ_OK;
}

int rz_set_uniforms(rz_ctx* c, const float world[16], const float view[16], const float proj[16]) {
    if (!c) return RZ_ERR_ARG;
    if (!world || !view || !proj) return fail(c, RZ_ERR_ARG, "null matrix");
    memcpy(c->uni.world, world, sizeof(c->uni.world));
    memcpy(c->uni.view, view, sizeof(c->uni.view));
    memcpy(c->uni.proj, proj, sizeof(c->uni.proj));
    return RZ_OK;
}

int rz_bind_texture(rz_ctx* c, uint32_t index, const uint8_t* rgba8, uint32_t tw, uint32_t th) {
    if (!c) return RZ_ERR_ARG;
    if (index != c->pool->textures.size())
        return fail(c, RZ_ERR_ARG, "texture index must equal the number of bound textures");
    if (!rgba8 || tw == 0 || th == 0) return fail(c, RZ_ERR_ARG, "empty texture");
    if ((uint64_t)tw * th >= (1ull << 31)) return fail(c, RZ_ERR_LIMIT, "texture too large");
    Texture t;
    t.tw = tw;
    t.th = th;
    size_t bytes = (size_t)tw * th * 4;
    HIP_TRY(c, hipMalloc(&t.data, bytes));
    hipError_t e = hipMemcpyAsync(t.data, rgba8, bytes, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        (void)hipFree(t.data);
        return fail(c, RZ_ERR_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
    }
    c->pool->textures.push_back(t);
    c->pool->resident_bytes += bytes;
    return RZ_OK;
}

uint32_t rz_texture_count(const rz_ctx* c) { return c ? (uint32_t)c->pool->textures.size() : 0u; }

int rz_mesh_create(rz_ctx* c, const float* pos, const float* color, const float* uv, uint32_t nv,
                   const uint32_t* idx, uint64_t ni, rz_mesh** out) {
    return mesh_create(c, pos, color, uv, nv, idx, ni, out, hipMemcpyHostToDevice);
}

int rz_mesh_create_device(rz_ctx* c, const float* pos, const float* color, const float* uv, uint32_t nv,
                          const uint32_t* idx, uint64_t ni, rz_mesh** out) {
    return mesh_create(c, pos, color, uv, nv, idx, ni, out, hipMemcpyDeviceToDevice);
}

int rz_mesh_destroy(rz_ctx* c, rz_mesh* m) {
    if (!c) return RZ_ERR_ARG;
    if (!m) return RZ_OK;
    if (m->pool != c->pool) return fail(c, RZ_ERR_ARG, "mesh belongs to another context pool");
    if (m->pending) return fail(c, RZ_ERR_STATE, "mesh is used by an unfinished frame or batch");
    for (rz_ctx* o : c->pool->ctxs) HIP_TRY(c, hipStreamSynchronize(o->stream));
    if (m->pos) (void)hipFree(m->pos);
    if (m->col) (void)hipFree(m->col);
    if (m->uv) (void)hipFree(m->uv);
    if (m->idx) (void)hipFree(m->idx);
    c->pool->resident_bytes -= m->bytes;
    delete m;
    return RZ_OK;
}

int rz_draw(rz_ctx* c, const rz_mesh* mesh_c, int vs, int fs) {
    if (!c) return RZ_ERR_ARG;
    rz_mesh* mesh = const_cast<rz_mesh*>(mesh_c);
    if (!mesh || mesh->pool != c->pool) return fail(c, RZ_ERR_ARG, "bad mesh");
    int rc = check_shaders(c, vs, fs);
    if (rc) return rc;
    if (c->recording == JobKind::Batch || c->in_flight)
        return fail(c, RZ_ERR_STATE, "a batch is being recorded or in flight");
    if (fs == RZ_FS_TEXTURE && c->pool->textures.empty())
        return fail(c, RZ_ERR_STATE, "texture fragment shader with no texture bound");
    if (c->draws.size() >= MAX_DRAWS) return fail(c, RZ_ERR_LIMIT, "too many draws in one frame");
    c->recording = JobKind::Frame;
    // uniforms.get_texture(0) (main.rs:69-71)
    c->draws.push_back(PendingDraw{mesh, 0u, (uint32_t)fs, 0u, c->uni});
    mesh->pending++;
    return RZ_OK;
}

int rz_resolve(rz_ctx* c, uint32_t* argb_out) {
    if (!c) return RZ_ERR_ARG;
    return frame_impl(c, argb_out, nullptr, nullptr, nullptr);
}

int rz_resolve_device(rz_ctx* c, uint32_t* argb_dev) {
    if (!c) return RZ_ERR_ARG;
    if (!argb_dev) return fail(c, RZ_ERR_ARG, "null output");
    return frame_impl(c, nullptr, argb_dev, nullptr, nullptr);
}

int rz_resolve_msaa(rz_ctx* c, uint32_t* argb_out, uint32_t* color4_out, float* depth4_out) {
    if (!c) return RZ_ERR_ARG;
    return frame_impl(c, argb_out, nullptr, color4_out, depth4_out);
}

uint32_t rz_frame_flags(const rz_ctx* c) { return c ? c->frame_flags : 0u; }

int rz_frame_stats(const rz_ctx* c, uint64_t out[11]) {
    if (!c || !out) return RZ_ERR_ARG;
    memcpy(out, c->frame_stats, sizeof(c->frame_stats));
    return RZ_OK;
}

int rz_set_coarse_bin(rz_ctx* c, uint32_t pixels) {
    if (!c) return RZ_ERR_ARG;
    if (pixels % 16 != 0 || pixels > MAX_COARSE_PX)
        return fail(c, RZ_ERR_ARG, "coarse bin must be a multiple of 16 and at most 256 pixels (0 = automatic)");
    if (!c->draws.empty()) return fail(c, RZ_ERR_STATE, "cannot change bins inside a frame");
    const uint32_t cb = coarse_side(c->W, c->H, pixels, MAX_BINS);
    if (((c->W + cb - 1) / cb) * ((c->H + cb - 1) / cb) > MAX_BINS)
        return fail(c, RZ_ERR_LIMIT, "more than 256 coarse bins");
    c->coarse_req = pixels;
    return RZ_OK;
}

// Bytes ctx would hold after running a job of plan pl (its buffers only grow).
uint64_t job_bytes_after(const rz_ctx* c, const JobPlan& pl) {
    rz_ctx* m = const_cast<rz_ctx*>(c);  // the dry sizing pass reads capacities only
    const std::string err = m->err;
    DryGrow g;
    const int rc = size_buffers(m, pl, Needs{}, g);
    m->err = err;
    const uint64_t now = ctx_bytes(c);
    if (rc != RZ_OK) return now;
    // the pipeline buffers as the job sizes them, plus what sizing leaves alone
    const uint64_t rest = held(c->out) + held(c->msaa_color) + held(c->msaa_depth) +
                          (c->raster_stats ? 0 : held(c->rstat_parts));
    uint64_t dcap = std::max<uint32_t>(c->dcap, 16), scap = std::max<uint32_t>(c->scap, 4);
    while (dcap < pl.ndraws) dcap *= 2;
    while (scap < pl.nscenes) scap *= 2;
    auto al = [](uint64_t v) { return (v + 255) & ~(uint64_t)255; };
    const uint64_t tables = al(sizeof(JobHead)) + al(sizeof(DrawDev) * dcap) + al(sizeof(Mat3) * dcap) +
                            al(sizeof(SceneCfg) * scap) + sizeof(SceneRun) * scap + sizeof(FrameCounters);
    return std::max(now, g.bytes + rest + tables);
}

int rz_memory(const rz_ctx* c, uint64_t out[3]) {
    if (!c || !out) return RZ_ERR_ARG;
    out[0] = ctx_bytes(c);
    out[1] = c->pool->resident_bytes;
    out[2] = out[0];
    if (c->recording != JobKind::None && !c->in_flight && !c->scenes.empty()) {
        JobPlan pl;
        plan_job(c, pl);
        out[2] = job_bytes_after(c, pl);
    }
    return RZ_OK;
}

int rz_trim(rz_ctx* c) {
    if (!c) return RZ_ERR_ARG;
    if (c->recording != JobKind::None || c->in_flight || !c->draws.empty())
        return fail(c, RZ_ERR_STATE, "not between jobs");
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (c->graph) {
        (void)hipGraphExecDestroy(c->graph);
        c->graph = nullptr;
    }
    release(c->cls);
    release(c->nsub);
    release(c->suboff);
    release(c->parts);
    release(c->parts64);
    release(c->clip_tri);
    release(c->clip_draw);
    release(c->clip_out);
    release(c->geom);
    release(c->attrs);
    release(c->uvat);
    release(c->pbin);
    release(c->bin_counts);
    release(c->bin_offsets);
    release(c->list);
    release(c->bin_rng);
    release(c->tile_rec);
    release(c->fchunk_rec);
    release(c->tchunk_draw);
    release(c->pchunk_scene);
    release(c->rstat_parts);
    release(c->fpack);
    release(c->fpre);
    release(c->fine_counts);
    release(c->fine_off);
    release(c->flist);
    release(c->out);
    release(c->msaa_color);
    release(c->msaa_depth);
    c->tiles_cap = 0;
    c->gen++;
    return RZ_OK;
}

int rz_batch_need(const rz_ctx* c, const rz_batch_item* items, uint32_t n, uint64_t* bytes) {
    if (!c || !bytes || (n && !items)) return RZ_ERR_ARG;
    for (uint32_t i = 0; i < n; ++i)
        if (!items[i].mesh || items[i].mesh->pool != c->pool || items[i].width == 0 || items[i].height == 0)
            return RZ_ERR_ARG;
    JobPlan pl;
    plan_sizes(
        c, false, n, n, [&](uint32_t i) { return items[i].mesh->ntri; },
        [&](uint32_t i, uint32_t& W, uint32_t& H) {
            W = items[i].width;
            H = items[i].height;
        },
        pl);
    *bytes = n ? job_bytes_after(c, pl) : ctx_bytes(c);
    return RZ_OK;
}

int rz_set_fine_tile(rz_ctx* c, uint32_t side) {
    if (!c) return RZ_ERR_ARG;
    if (side != 0 && side != 8 && side != 16 && side != 32)
        return fail(c, RZ_ERR_ARG, "fine tile side must be 8, 16 or 32 (0 = 32)");
    if (c->recording != JobKind::None || c->in_flight || !c->draws.empty())
        return fail(c, RZ_ERR_STATE, "not between jobs");
    c->fine_req = side ? side : 32;
    return RZ_OK;
}

uint32_t rz_get_fine_tile(const rz_ctx* c) { return c ? c->fine_req : 0u; }

int rz_profile(rz_ctx* c, int enable) {
    if (!c) return RZ_ERR_ARG;
    c->prof = enable != 0;
    for (float& v : c->prof_ms) v = 0;
    c->prof_frames = 0;
    return RZ_OK;
}

int rz_profile_read(const rz_ctx* c, float ms[RZ_NSTAGES], uint32_t* frames) {
    if (!c || !ms) return RZ_ERR_ARG;
    memcpy(ms, c->prof_ms, sizeof(c->prof_ms));
    if (frames) *frames = c->prof_frames;
    return RZ_OK;
}

// ------------------------------------------------------------------ batches

int rz_batch_begin(rz_ctx* c) {
    if (!c) return RZ_ERR_ARG;
    if (c->in_flight) return fail(c, RZ_ERR_STATE, "previous batch still in flight (rz_batch_wait)");
    if (c->recording != JobKind::None) return fail(c, RZ_ERR_STATE, "a frame or batch is being recorded");
    c->recording = JobKind::Batch;
    c->scenes.clear();
    c->draws.clear();
    return RZ_OK;
}

int rz_batch_scene(rz_ctx* c, uint32_t width, uint32_t height, float clear_depth, uint32_t* argb_dev,
                   uint32_t* scene_out) {
    if (!c) return RZ_ERR_ARG;
    if (c->recording != JobKind::Batch || c->in_flight) return fail(c, RZ_ERR_STATE, "no batch being recorded");
    if (!argb_dev) return fail(c, RZ_ERR_ARG, "null output");
    if (width == 0 || height == 0 || width > MAX_DIM || height > MAX_DIM) return fail(c, RZ_ERR_ARG, "bad size");
    if (((width + MAX_COARSE_PX - 1) / MAX_COARSE_PX) * ((height + MAX_COARSE_PX - 1) / MAX_COARSE_PX) > MAX_BINS)
        return fail(c, RZ_ERR_LIMIT, "scene needs more than 256 coarse bins");
    if (c->scenes.size() >= MAX_SCENES) return fail(c, RZ_ERR_LIMIT, "too many scenes in one batch");
    if (scene_out) *scene_out = (uint32_t)c->scenes.size();
    c->scenes.push_back(PendingScene{width, height, clear_depth, argb_dev, nullptr, nullptr});
    return RZ_OK;
}

int rz_batch_scene_msaa(rz_ctx* c, uint32_t scene, uint32_t* color4_dev, float* depth4_dev) {
    if (!c) return RZ_ERR_ARG;
    if (c->recording != JobKind::Batch || c->in_flight) return fail(c, RZ_ERR_STATE, "no batch being recorded");
    if (scene >= c->scenes.size()) return fail(c, RZ_ERR_ARG, "unknown scene");
    c->scenes[scene].msaa_color = color4_dev;
    c->scenes[scene].msaa_depth = depth4_dev;
    return RZ_OK;
}

int rz_batch_draw(rz_ctx* c, uint32_t scene, const rz_mesh* mesh_c, const float world[16], const float view[16],
                  const float proj[16], uint32_t texture, int vs, int fs) {
    if (!c) return RZ_ERR_ARG;
    if (c->recording != JobKind::Batch || c->in_flight) return fail(c, RZ_ERR_STATE, "no batch being recorded");
    rz_mesh* mesh = const_cast<rz_mesh*>(mesh_c);
    if (!mesh || mesh->pool != c->pool) return fail(c, RZ_ERR_ARG, "bad mesh");
    if (scene >= c->scenes.size()) return fail(c, RZ_ERR_ARG, "unknown scene");
    if (!world || !view || !proj) return fail(c, RZ_ERR_ARG, "null matrix");
    int rc = check_shaders(c, vs, fs);
    if (rc) return rc;
    if (fs == RZ_FS_TEXTURE && texture >= c->pool->textures.size())
        return fail(c, RZ_ERR_ARG, "texture index not bound");
    if (c->draws.size() >= MAX_DRAWS) return fail(c, RZ_ERR_LIMIT, "too many draws in one batch");
    PendingDraw d{mesh, scene, (uint32_t)fs, texture, Mat3{}};
    memcpy(d.m.world, world, sizeof(d.m.world));
    memcpy(d.m.view, view, sizeof(d.m.view));
    memcpy(d.m.proj, proj, sizeof(d.m.proj));
    c->draws.push_back(d);
    mesh->pending++;
    return RZ_OK;
}

int rz_batch_add(rz_ctx* c, const rz_batch_item* items, uint32_t n, uint32_t* first_scene) {
    if (!c) return RZ_ERR_ARG;
    if (c->recording != JobKind::Batch || c->in_flight) return fail(c, RZ_ERR_STATE, "no batch being recorded");
    if (n && !items) return fail(c, RZ_ERR_ARG, "null items");
    if (c->scenes.size() + n > MAX_SCENES) return fail(c, RZ_ERR_LIMIT, "too many scenes in one batch");
    if (c->draws.size() + n > MAX_DRAWS) return fail(c, RZ_ERR_LIMIT, "too many draws in one batch");
    for (uint32_t i = 0; i < n; ++i) {  // validate everything first: all or nothing
        const rz_batch_item& it = items[i];
        const std::string at = "item " + std::to_string(i) + ": ";
        if (!it.argb_dev) return fail(c, RZ_ERR_ARG, at + "null output");
        if (it.width == 0 || it.height == 0 || it.width > MAX_DIM || it.height > MAX_DIM)
            return fail(c, RZ_ERR_ARG, at + "bad size");
        if (((it.width + MAX_COARSE_PX - 1) / MAX_COARSE_PX) * ((it.height + MAX_COARSE_PX - 1) / MAX_COARSE_PX) >
            MAX_BINS)
            return fail(c, RZ_ERR_LIMIT, at + "scene needs more than 256 coarse bins");
        if (!it.mesh || it.mesh->pool != c->pool) return fail(c, RZ_ERR_ARG, at + "bad mesh");
        int rc = check_shaders(c, it.vertex_shader, it.fragment_shader);
        if (rc) return fail(c, rc, at + c->err);
        if (it.fragment_shader == RZ_FS_TEXTURE && it.texture >= c->pool->textures.size())
            return fail(c, RZ_ERR_ARG, at + "texture index not bound");
    }
    if (first_scene) *first_scene = (uint32_t)c->scenes.size();
    for (uint32_t i = 0; i < n; ++i) {
        const rz_batch_item& it = items[i];
        rz_mesh* mesh = const_cast<rz_mesh*>(it.mesh);
        const uint32_t scene = (uint32_t)c->scenes.size();
        c->scenes.push_back(PendingScene{it.width, it.height, it.clear_depth, it.argb_dev, nullptr, nullptr});
        PendingDraw d{mesh, scene, (uint32_t)it.fragment_shader, it.texture, Mat3{}};
        memcpy(d.m.world, it.world, sizeof(d.m.world));
        memcpy(d.m.view, it.view, sizeof(d.m.view));
        memcpy(d.m.proj, it.proj, sizeof(d.m.proj));
        c->draws.push_back(d);
        mesh->pending++;
    }
    return RZ_OK;
}

int rz_batch_submit(rz_ctx* c) {
    if (!c) return RZ_ERR_ARG;
    if (c->recording != JobKind::Batch || c->in_flight) return fail(c, RZ_ERR_STATE, "no batch being recorded");
    if (c->scenes.empty()) return fail(c, RZ_ERR_STATE, "empty batch");
    // group draws by scene, keeping each scene's submission order
    std::stable_sort(c->draws.begin(), c->draws.end(),
                     [](const PendingDraw& a, const PendingDraw& b) { return a.scene < b.scene; });
    int rc = start_job(c, c->use_graph);
    if (rc) {
        release_job(c);
        return rc;
    }
    return RZ_OK;
}

int rz_batch_query(rz_ctx* c) {
    if (!c) return RZ_ERR_ARG;
    if (!c->in_flight) return 1;
    hipError_t e = hipEventQuery(c->done);
    if (e == hipErrorNotReady) return 0;
    if (e != hipSuccess) {
        release_job(c);
        return fail(c, RZ_ERR_HIP, std::string("hipEventQuery: ") + hipGetErrorString(e));
    }
    int rc = complete_job(c);
    if (rc == 2) {
        rc = svc_ensure(c);  // relaunch a service that exited with the job posted
        if (rc == RZ_OK) return 0;
    }
    if (rc < 0 || rc == 1) release_job(c);
    return rc;
}

int rz_batch_wait(rz_ctx* c) {
    if (!c) return RZ_ERR_ARG;
    if (!c->in_flight) return RZ_OK;
    int rc = wait_job(c);
    release_job(c);
    return rc;
}

int rz_batch_scene_flags(const rz_ctx* c, uint32_t scene, uint32_t* flags) {
    if (!c || !flags) return RZ_ERR_ARG;
    if (scene >= c->scene_flags.size()) return RZ_ERR_ARG;
    *flags = c->scene_flags[scene];
    return RZ_OK;
}

int rz_batch_tuning(rz_ctx* c, uint32_t bin_target, uint32_t raster_per_cu, int use_graph) {
    if (!c) return RZ_ERR_ARG;
    if (bin_target == 0 || bin_target > MAX_BINS) return fail(c, RZ_ERR_ARG, "bin target must be in [1, 256]");
    if (raster_per_cu > 32) return fail(c, RZ_ERR_ARG, "at most 32 raster workgroups per CU");
    if (c->recording != JobKind::None || c->in_flight) return fail(c, RZ_ERR_STATE, "not between jobs");
    c->bin_target = bin_target;
    c->raster_per_cu = raster_per_cu;
    c->use_graph = use_graph != 0;
    return RZ_OK;
}

int rz_set_cu_mask(rz_ctx* c, const uint32_t* mask, uint32_t words) {
    if (!c) return RZ_ERR_ARG;
    if (words > 0 && !mask) return fail(c, RZ_ERR_ARG, "null mask");
    if (words > 1024) return fail(c, RZ_ERR_ARG, "mask longer than 1024 words");
    if (c->recording != JobKind::None || c->in_flight || !c->draws.empty())
        return fail(c, RZ_ERR_STATE, "not between jobs");
    if (words > 0) {
        bool any = false;
        for (uint32_t i = 0; i < (uint32_t)c->n_cu && i / 32 < words; ++i) any |= ((mask[i / 32] >> (i % 32)) & 1u) != 0;
        if (!any) return fail(c, RZ_ERR_ARG, "the mask selects no CU of the device");
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    hipStream_t ns = nullptr;
    const hipError_t e = words > 0 ? hipExtStreamCreateWithCUMask(&ns, words, mask)
                                   : hipStreamCreateWithFlags(&ns, hipStreamNonBlocking);
    if (e != hipSuccess) return fail(c, RZ_ERR_HIP, std::string("stream creation: ") + hipGetErrorString(e));
    (void)hipStreamDestroy(c->stream);
    c->stream = ns;
    return RZ_OK;
}

int rz_batch_get_tuning(const rz_ctx* c, uint32_t out[3]) {
    if (!c || !out) return RZ_ERR_ARG;
    out[0] = c->bin_target;
    out[1] = c->raster_per_cu;
    out[2] = c->use_graph ? 1u : 0u;
    return RZ_OK;
}

int rz_service_start(rz_ctx* c, uint32_t wg_per_cu, uint32_t idle_us) {
    if (!c) return RZ_ERR_ARG;
    SvcState& v = c->pool->svc;
    if (v.on) return fail(c, RZ_ERR_STATE, "the raster service is already running");
    if (wg_per_cu < 1 || wg_per_cu > 4) return fail(c, RZ_ERR_ARG, "service workgroups per CU must be in [1, 4]");
    if (idle_us < 1 || idle_us > 5000000) return fail(c, RZ_ERR_ARG, "service idle time must be in [1 us, 5 s]");
    int khz = 0;
    HIP_TRY(c, hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->pool->device));
    if (khz <= 0) return fail(c, RZ_ERR_HIP, "no wall clock rate");
    if (!v.ents) {
        hipError_t e = hipMalloc(&v.ents, sizeof(SvcEntry) * SVC_MAX_ENTRIES);
        if (e == hipSuccess) e = hipMemset(v.ents, 0, sizeof(SvcEntry) * SVC_MAX_ENTRIES);
        if (e == hipSuccess) e = hipMalloc(&v.ctl, sizeof(SvcCtl));
        if (e == hipSuccess) e = hipHostMalloc(&v.host, sizeof(SvcHost), hipHostMallocDefault);
        if (e == hipSuccess) {
            memset(v.host, 0, sizeof(SvcHost));
            std::vector<uint32_t> mask(((uint32_t)c->n_cu + 31) / 32, 0u);
            for (uint32_t i = 0; i < (uint32_t)c->n_cu; ++i) mask[i / 32] |= 1u << (i % 32);
            e = hipExtStreamCreateWithCUMask(&v.stream, (uint32_t)mask.size(), mask.data());
        }
        if (e != hipSuccess) {
            svc_free(c->pool);
            return fail(c, RZ_ERR_HIP, std::string("raster service set-up: ") + hipGetErrorString(e));
        }
    }
    v.grid = std::min<uint32_t>(wg_per_cu * (uint32_t)c->n_cu, SVC_MAX_GRID);
    v.idle_ticks = (unsigned long long)idle_us * (unsigned long long)khz / 1000ull;
    v.life_ticks = (unsigned long long)(SVC_LIFE_S * 1000.0) * (unsigned long long)khz;
    v.on = true;
    const hipError_t e = svc_launch(c->pool);
    if (e != hipSuccess) {
        v.on = false;
        return fail(c, RZ_ERR_HIP, std::string("raster service launch: ") + hipGetErrorString(e));
    }
    return RZ_OK;
}

int rz_service_stop(rz_ctx* c) {
    if (!c) return RZ_ERR_ARG;
    SvcState& v = c->pool->svc;
    if (!v.on) return RZ_OK;
    for (rz_ctx* o : c->pool->ctxs)
        if (o->svc_entry >= 0 && o->in_flight) return fail(c, RZ_ERR_STATE, "a context has a job in flight");
    __atomic_store_n(&v.host->stop, 1u, __ATOMIC_RELEASE);
    HIP_TRY(c, hipStreamSynchronize(v.stream));
    v.on = false;
    return RZ_OK;
}

int rz_service_attach(rz_ctx* c, int on) {
    if (!c) return RZ_ERR_ARG;
    if (c->recording != JobKind::None || c->in_flight) return fail(c, RZ_ERR_STATE, "not between jobs");
    SvcState& v = c->pool->svc;
    if (!on) {
        if (c->svc_entry >= 0) v.used &= ~(1u << c->svc_entry);
        c->svc_entry = -1;
        return RZ_OK;
    }
    if (c->svc_entry >= 0) return RZ_OK;
    if (!v.on) return fail(c, RZ_ERR_STATE, "the raster service is not running (rz_service_start)");
    if (c->raster_stats) return fail(c, RZ_ERR_STATE, "the counting raster (RZ_RASTER_STATS) cannot use the service");
    for (uint32_t i = 0; i < SVC_MAX_ENTRIES; ++i)
        if (!(v.used & (1u << i))) {
            v.used |= 1u << i;
            c->svc_entry = (int)i;
            return RZ_OK;
        }
    return fail(c, RZ_ERR_LIMIT, "every raster service entry is taken");
}

int rz_service_stats(const rz_ctx* c, uint64_t out[4]) {
    if (!c || !out) return RZ_ERR_ARG;
    const SvcState& v = c->pool->svc;
    out[0] = v.on ? 1u : 0u;
    out[1] = v.on && hipStreamQuery(v.stream) == hipErrorNotReady ? 1u : 0u;
    out[2] = v.launches;
    out[3] = v.grid;
    return RZ_OK;
}

}  // extern "C"
