// rz_api.hip -- host side of the C ABI declared in include/rz.h.
//
// A context owns one HIP stream, the frame's draw list, and the device
// buffers of the pipeline. Draws only enqueue the geometry stages; the
// frame is binned and rasterized at resolve time, so the 4x MSAA colour and
// depth of each pixel stay in registers for the whole frame (no MSAA
// traffic to HBM). Buffers that a frame outgrows are detected on the device
// (overflow bits); the resolve then grows them and replays the frame, so no
// triangle is ever dropped.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "rz_kernels.h"

using namespace rzk;

namespace {

constexpr uint32_t MAX_DIM = 16384;
// a coarse bin holds at most 256 16x16 tiles (fine binning keeps one running
// count per tile in 4 registers per lane)
constexpr uint32_t MAX_COARSE_PX = 256;
constexpr uint32_t MAX_DRAWS = 65536;

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

struct rz_mesh {
    rz_ctx* ctx = nullptr;
    float* pos = nullptr;
    float* col = nullptr;
    float* uv = nullptr;
    uint32_t* idx = nullptr;
    uint32_t nv = 0;
    uint32_t ntri = 0;
    uint32_t pending = 0;  // draws of the unresolved frame using this mesh
};

struct rz_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t W = 0, H = 0;
    float clear_depth = FLT_MAX;
    float next_clear_depth = FLT_MAX;
    Mat3 uni{};
    std::vector<Texture> textures;
    BinGeom bg{};
    uint32_t coarse_req = 0;

    // frame state
    std::vector<rz_mesh*> draw_mesh;
    DrawDev* d_draws = nullptr;  // MAX_DRAWS entries
    DrawDev* h_draws = nullptr;  // pinned mirror
    Mat3* h_mats = nullptr;      // pinned: uniforms of each draw
    FrameCounters* d_cnt = nullptr;
    FrameCounters* h_cnt = nullptr;  // pinned

    // per-draw scratch
    DBuf<uint32_t> cls, nsub, suboff, parts;
    // clipper
    DBuf<uint32_t> clip_tri, clip_n;
    DBuf<float> clip_out;  // MAXV * 10 floats per slot
    // rasterizer triangles
    DBuf<PrimGeom> geom;
    DBuf<float> attrs;  // ATTR_STRIDE per prim
    DBuf<uint4> pbin;
    // coarse bins
    DBuf<uint32_t> bin_counts, bin_offsets;
    DBuf<uint4> list;
    // fine bins (16x16 tiles)
    DBuf<uint32_t> chunk_base, fine_counts, fine_off;
    DBuf<uint4> flist;
    // outputs
    DBuf<uint32_t> out;
    DBuf<uint32_t> msaa_color;
    DBuf<float> msaa_depth;

    uint32_t frame_flags = 0;
    uint64_t frame_stats[11] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    bool raster_stats = false;  // RZ_RASTER_STATS=1 at rz_create
    uint64_t frame_tris = 0;

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
int grow(rz_ctx* c, DBuf<T>& b, size_t need, bool* grew = nullptr) {
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
    if (grew) *grew = true;
    return RZ_OK;
}

template <class T>
void release(DBuf<T>& b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
}

BinGeom choose_bins(uint32_t W, uint32_t H, uint32_t req) {
    BinGeom g{};
    uint32_t cb = req;
    if (cb == 0) {
        cb = 16;
        while (((W + cb - 1) / cb) * ((H + cb - 1) / cb) > MAX_BINS) cb *= 2;
    }
    g.cb = cb;
    g.nbx = (W + cb - 1) / cb;
    g.nby = (H + cb - 1) / cb;
    return g;
}

uint32_t bin_matrix_len(const rz_ctx* c) {
    uint32_t nb = c->bg.nbx * c->bg.nby;
    size_t nc = (c->geom.cap + CHUNK - 1) / CHUNK;
    return (uint32_t)(nb * nc);
}

// upper bound of coarse-list chunks: every bin adds at most one partial chunk
uint32_t nch_max(const rz_ctx* c) { return (uint32_t)(c->list.cap / CHUNK + MAX_BINS + 1); }

uint32_t fine_matrix_len(const rz_ctx* c) {
    uint32_t T = (c->bg.cb / 16) * (c->bg.cb / 16);
    return T * nch_max(c);
}

// capacities for the frame's rasterizer triangles, clip slots and bin entries
int ensure_frame_buffers(rz_ctx* c, uint64_t prims_need, uint64_t clip_need, uint64_t pairs_need,
                         uint64_t max_draw_tris, bool* grew_prims, uint64_t fine_need = 0) {
    if (fine_need == 0) fine_need = std::max<uint64_t>(c->flist.cap, pairs_need * 8);
    int rc;
    if ((rc = grow(c, c->cls, max_draw_tris))) return rc;
    if ((rc = grow(c, c->nsub, max_draw_tris))) return rc;
    if ((rc = grow(c, c->suboff, max_draw_tris))) return rc;
    bool g = false;
    if ((rc = grow(c, c->geom, prims_need, &g))) return rc;
    if ((rc = grow(c, c->attrs, c->geom.cap * ATTR_STRIDE))) return rc;
    if ((rc = grow(c, c->pbin, c->geom.cap))) return rc;
    if ((rc = grow(c, c->clip_tri, clip_need, &g))) return rc;
    if ((rc = grow(c, c->clip_n, c->clip_tri.cap))) return rc;
    if ((rc = grow(c, c->clip_out, c->clip_tri.cap * MAXV * 10))) return rc;
    if ((rc = grow(c, c->list, pairs_need))) return rc;
    size_t bm = bin_matrix_len(c);
    if ((rc = grow(c, c->bin_counts, bm))) return rc;
    if ((rc = grow(c, c->bin_offsets, bm))) return rc;
    if ((rc = grow(c, c->chunk_base, MAX_BINS + 1))) return rc;
    if ((rc = grow(c, c->flist, fine_need))) return rc;
    size_t fm = fine_matrix_len(c);
    if ((rc = grow(c, c->fine_counts, fm))) return rc;
    if ((rc = grow(c, c->fine_off, fm))) return rc;
    size_t scan_n = std::max<size_t>(std::max<size_t>(c->cls.cap, bm), fm);
    if ((rc = grow(c, c->parts, (scan_n + SCAN_TILE - 1) / SCAN_TILE + 1))) return rc;
    if (grew_prims) *grew_prims = g;
    return RZ_OK;
}

void enqueue_draw_geometry(rz_ctx* c, uint32_t d) {
    const DrawDev& h = c->h_draws[d];
    hipStream_t s = c->stream;
    launch_draw_reset(s, c->d_cnt);
    launch_classify(s, c->d_draws, d, h.ntri, c->d_cnt, c->cls.p, c->nsub.p, c->clip_tri.p,
                    (uint32_t)c->clip_tri.cap);
    launch_clip(s, c->d_draws, d, c->d_cnt, c->clip_tri.p, (uint32_t)c->clip_tri.cap, c->nsub.p, c->clip_out.p,
                c->clip_n.p);
    launch_scan(s, c->nsub.p, c->suboff.p, h.ntri, nullptr, 0, 1, c->parts.p, &c->d_cnt->nprims, nullptr);
    launch_setup(s, c->d_draws, d, h.ntri, c->d_cnt, c->cls.p, c->nsub.p, c->suboff.p, c->clip_out.p,
                 (uint32_t)c->clip_tri.cap, c->geom.p, c->attrs.p, c->pbin.p, (uint32_t)c->geom.cap, c->W, c->H,
                 c->bg);
}

// upload draw d's descriptor and uniforms, compute its MVP on the device
int enqueue_draw_desc(rz_ctx* c, uint32_t d) {
    HIP_TRY(c, hipMemcpyAsync(c->d_draws + d, c->h_draws + d, sizeof(DrawDev), hipMemcpyHostToDevice, c->stream));
    launch_mvp(c->stream, c->d_draws, d, c->h_mats[d]);
    return RZ_OK;
}

int replay_geometry(rz_ctx* c, uint32_t upto) {
    launch_frame_reset(c->stream, c->d_cnt);
    for (uint32_t d = 0; d < upto; ++d) {
        int rc = enqueue_draw_desc(c, d);
        if (rc) return rc;
        enqueue_draw_geometry(c, d);
    }
    return RZ_OK;
}

uint32_t max_draw_tris(const rz_ctx* c) {
    uint32_t m = 0;
    for (size_t d = 0; d < c->draw_mesh.size(); ++d) m = std::max(m, c->h_draws[d].ntri);
    return m;
}

void end_frame(rz_ctx* c) {
    for (rz_mesh* m : c->draw_mesh) m->pending--;
    c->draw_mesh.clear();
    c->frame_tris = 0;
    c->clear_depth = c->next_clear_depth;
}

int resolve_impl(rz_ctx* c, uint32_t* host_out, uint32_t* dev_out, uint32_t* host_color4, float* host_depth4) {
    if (!c) return RZ_ERR_ARG;
    int rc;
    hipStream_t s = c->stream;
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
    const uint32_t ndraws = (uint32_t)c->draw_mesh.size();
    if (ndraws == 0) {
        if ((rc = ensure_frame_buffers(c, 1024, 1024, 4096, 1024, nullptr))) return rc;
        launch_frame_reset(s, c->d_cnt);
        if (c->prof) {
            HIP_TRY(c, hipEventRecord(c->ev[0], s));
            HIP_TRY(c, hipEventRecord(c->ev[1], s));
        }
    }
    uint32_t replays = 0;
    uint32_t ovf_seen = 0;
    for (;;) {
        const uint32_t nb = c->bg.nbx * c->bg.nby;
        const uint32_t nc_max = (uint32_t)((c->geom.cap + CHUNK - 1) / CHUNK);
        const uint32_t bin_grid = std::min<uint32_t>((nc_max + 3) / 4, 4096);
        launch_bin(s, false, bin_grid, c->d_cnt, c->pbin.p, c->bg, c->bin_counts.p, nullptr, nullptr, 0);
        launch_scan(s, c->bin_counts.p, c->bin_offsets.p, nb * nc_max, &c->d_cnt->nprims, nb, CHUNK, c->parts.p,
                    nullptr, &c->d_cnt->pairs_binned);
        launch_bin(s, true, bin_grid, c->d_cnt, c->pbin.p, c->bg, nullptr, c->bin_offsets.p, c->list.p,
                   (uint32_t)c->list.cap);
        launch_fine(s, c->d_cnt, c->bin_offsets.p, c->list.p, c->chunk_base.p, c->bg, c->fine_counts.p,
                    c->fine_off.p, c->flist.p, (uint32_t)c->flist.cap, nch_max(c), c->parts.p);
        if (c->prof) HIP_TRY(c, hipEventRecord(c->ev[2], s));
        launch_raster(s, c->d_cnt, c->geom.p, c->attrs.p, c->d_draws, c->chunk_base.p, c->fine_off.p, c->flist.p,
                      (uint32_t)c->flist.cap, c->bg, c->W, c->H, c->clear_depth, target,
                      want_msaa ? c->msaa_color.p : nullptr, want_msaa ? c->msaa_depth.p : nullptr,
                      c->raster_stats);
        if (c->prof) HIP_TRY(c, hipEventRecord(c->ev[3], s));
        HIP_TRY(c, hipGetLastError());
        HIP_TRY(c, hipMemcpyAsync(c->h_cnt, c->d_cnt, sizeof(FrameCounters), hipMemcpyDeviceToHost, s));
        HIP_TRY(c, hipStreamSynchronize(s));
        const FrameCounters fc = *c->h_cnt;
        ovf_seen |= fc.overflow;
        if (fc.overflow == 0) break;
        if (fc.overflow & OVF_POLY) {
            end_frame(c);
            return fail(c, RZ_ERR_LIMIT, "a clipped polygon exceeded 16 vertices");
        }
        if (++replays > 4) {
            end_frame(c);
            return fail(c, RZ_ERR_LIMIT, "frame did not fit after repeated growth");
        }
        uint64_t prims_need = std::max<uint64_t>(fc.nprims, c->geom.cap);
        uint64_t clip_need = std::max<uint64_t>(fc.clip_need, c->clip_tri.cap);
        uint64_t pairs_need = std::max<uint64_t>(fc.npairs, c->list.cap);
        if (fc.overflow & OVF_PRIMS) prims_need = std::max<uint64_t>(fc.nprims, c->geom.cap * 2);
        if (fc.overflow & OVF_CLIP) clip_need = std::max<uint64_t>(fc.clip_need, c->clip_tri.cap * 2);
        if (fc.overflow & OVF_PAIRS) pairs_need = std::max<uint64_t>(fc.npairs, c->list.cap * 2);
        uint64_t fine_need = std::max<uint64_t>(fc.fine_total, c->flist.cap);
        if (fc.overflow & OVF_FINE) fine_need = std::max<uint64_t>(fc.fine_total, c->flist.cap * 2);
        if ((rc = ensure_frame_buffers(c, prims_need, clip_need, pairs_need, std::max(max_draw_tris(c), 1024u),
                                       nullptr, fine_need))) {
            end_frame(c);
            return rc;
        }
        if (c->prof) HIP_TRY(c, hipEventRecord(c->ev[0], s));
        if ((rc = replay_geometry(c, ndraws))) {
            end_frame(c);
            return rc;
        }
        if (c->prof) HIP_TRY(c, hipEventRecord(c->ev[1], s));
    }
    const FrameCounters fc = *c->h_cnt;
    c->frame_flags = fc.flags;
    c->frame_stats[0] = c->frame_tris;
    c->frame_stats[1] = fc.nclipped;
    c->frame_stats[2] = fc.nprims;
    c->frame_stats[3] = fc.pairs_binned;
    c->frame_stats[4] = replays;
    for (int i = 0; i < 5; ++i) c->frame_stats[5 + i] = fc.rstat[i];
    c->frame_stats[10] = ovf_seen;
    if (c->prof) {
        float a = 0, b = 0, d = 0;
        HIP_TRY(c, hipEventElapsedTime(&a, c->ev[0], c->ev[1]));
        HIP_TRY(c, hipEventElapsedTime(&b, c->ev[1], c->ev[2]));
        HIP_TRY(c, hipEventElapsedTime(&d, c->ev[2], c->ev[3]));
        c->prof_ms[RZ_STAGE_GEOMETRY] += a;
        c->prof_ms[RZ_STAGE_BINNING] += b;
        c->prof_ms[RZ_STAGE_RASTER] += d;
        c->prof_frames++;
    }
    end_frame(c);
    if (host_out) HIP_TRY(c, hipMemcpy(host_out, target, npx * sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (host_color4)
        HIP_TRY(c, hipMemcpy(host_color4, c->msaa_color.p, npx * 4 * sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (host_depth4)
        HIP_TRY(c, hipMemcpy(host_depth4, c->msaa_depth.p, npx * 4 * sizeof(float), hipMemcpyDeviceToHost));
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
    m->ctx = c;
    m->nv = nv;
    m->ntri = (uint32_t)(ni / 3);
    auto up = [&](void** dst, const void* src, size_t bytes) -> int {
        if (!src || bytes == 0) return RZ_OK;
        hipError_t e = hipMalloc(dst, bytes);
        if (e != hipSuccess) return fail(c, RZ_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
        e = hipMemcpy(*dst, src, bytes, kind);
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
    *out = m;
    return RZ_OK;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int rz_create(uint32_t width, uint32_t height, rz_ctx** out) {
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
    c->bg = choose_bins(width, height, 0);
    const char* st = getenv("RZ_RASTER_STATS");
    c->raster_stats = st && st[0] == '1';
    auto bad = [&](hipError_t e) {
        std::string m = hipGetErrorString(e);
        rz_destroy(c);
        return RZ_ERR_HIP;
    };
    hipError_t e;
    if ((e = hipGetDevice(&c->device)) != hipSuccess) return bad(e);
    if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess) return bad(e);
    if ((e = hipMalloc(&c->d_draws, sizeof(DrawDev) * MAX_DRAWS)) != hipSuccess) return bad(e);
    if ((e = hipHostMalloc(&c->h_draws, sizeof(DrawDev) * MAX_DRAWS, hipHostMallocDefault)) != hipSuccess)
        return bad(e);
    if ((e = hipHostMalloc(&c->h_mats, sizeof(Mat3) * MAX_DRAWS, hipHostMallocDefault)) != hipSuccess)
        return bad(e);
    if ((e = hipMalloc(&c->d_cnt, sizeof(FrameCounters))) != hipSuccess) return bad(e);
    if ((e = hipHostMalloc(&c->h_cnt, sizeof(FrameCounters), hipHostMallocDefault)) != hipSuccess) return bad(e);
    for (int i = 0; i < 4; ++i)
        if ((e = hipEventCreate(&c->ev[i])) != hipSuccess) return bad(e);
    memset(c->h_cnt, 0, sizeof(FrameCounters));
    launch_frame_reset(c->stream, c->d_cnt);
    if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return bad(e);
    *out = c;
    return RZ_OK;
}

int rz_destroy(rz_ctx* c) {
    if (!c) return RZ_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (Texture& t : c->textures)
        if (t.data) (void)hipFree(t.data);
    release(c->cls);
    release(c->nsub);
    release(c->suboff);
    release(c->parts);
    release(c->clip_tri);
    release(c->clip_n);
    release(c->clip_out);
    release(c->geom);
    release(c->attrs);
    release(c->pbin);
    release(c->bin_counts);
    release(c->bin_offsets);
    release(c->list);
    release(c->chunk_base);
    release(c->fine_counts);
    release(c->fine_off);
    release(c->flist);
    release(c->out);
    release(c->msaa_color);
    release(c->msaa_depth);
    if (c->d_draws) (void)hipFree(c->d_draws);
    if (c->h_draws) (void)hipHostFree(c->h_draws);
    if (c->h_mats) (void)hipHostFree(c->h_mats);
    if (c->d_cnt) (void)hipFree(c->d_cnt);
    if (c->h_cnt) (void)hipHostFree(c->h_cnt);
    for (int i = 0; i < 4; ++i)
        if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return RZ_OK;
}

const char* rz_last_error(const rz_ctx* c) { return c ? c->err.c_str() : kNullCtxMsg; }

void* rz_stream(const rz_ctx* c) { return c ? (void*)c->stream : nullptr; }

int rz_set_clear_depth(rz_ctx* c, float depth) {
    if (!c) return RZ_ERR_ARG;
    c->next_clear_depth = depth;
    if (c->draw_mesh.empty()) c->clear_depth = depth;
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
    if (index != c->textures.size())
        return fail(c, RZ_ERR_ARG, "texture index must equal the number of bound textures");
    if (!rgba8 || tw == 0 || th == 0) return fail(c, RZ_ERR_ARG, "empty texture");
    if ((uint64_t)tw * th >= (1ull << 31)) return fail(c, RZ_ERR_LIMIT, "texture too large");
    Texture t;
    t.tw = tw;
    t.th = th;
    size_t bytes = (size_t)tw * th * 4;
    HIP_TRY(c, hipMalloc(&t.data, bytes));
    hipError_t e = hipMemcpy(t.data, rgba8, bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(t.data);
        return fail(c, RZ_ERR_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
    }
    c->textures.push_back(t);
    return RZ_OK;
}

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
    if (m->ctx != c) return fail(c, RZ_ERR_ARG, "mesh belongs to another context");
    if (m->pending) return fail(c, RZ_ERR_STATE, "mesh is used by an unresolved frame");
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (m->pos) (void)hipFree(m->pos);
    if (m->col) (void)hipFree(m->col);
    if (m->uv) (void)hipFree(m->uv);
    if (m->idx) (void)hipFree(m->idx);
    delete m;
    return RZ_OK;
}

int rz_draw(rz_ctx* c, const rz_mesh* mesh_c, int vs, int fs) {
    if (!c) return RZ_ERR_ARG;
    rz_mesh* mesh = const_cast<rz_mesh*>(mesh_c);
    if (!mesh || mesh->ctx != c) return fail(c, RZ_ERR_ARG, "bad mesh");
    if (vs != RZ_VS_MVP) return fail(c, RZ_ERR_ARG, "unknown vertex shader");
    if (fs != RZ_FS_TEXTURE && fs != RZ_FS_COLOR && fs != RZ_FS_DEBUG)
        return fail(c, RZ_ERR_ARG, "unknown fragment shader");
    if (fs == RZ_FS_TEXTURE && c->textures.empty())
        return fail(c, RZ_ERR_STATE, "texture fragment shader with no texture bound");
    const uint32_t d = (uint32_t)c->draw_mesh.size();
    if (d >= MAX_DRAWS) return fail(c, RZ_ERR_LIMIT, "too many draws in one frame");
    int rc;
    if (d == 0) {
        // the previous frame's resolve has finished with the pinned mirrors
        launch_frame_reset(c->stream, c->d_cnt);
        if (c->prof) HIP_TRY(c, hipEventRecord(c->ev[0], c->stream));
    }
    // capacities: rasterizer triangles ~ input triangles (+ clip fans)
    uint64_t tris = c->frame_tris + mesh->ntri;
    uint64_t prims_est = tris + tris / 4 + 4096;
    uint64_t clip_est = std::max<uint64_t>(c->clip_tri.cap, mesh->ntri / 8 + 1024);
    uint64_t pairs_est = std::max<uint64_t>(c->list.cap, prims_est * 2);
    bool grew_prims = false;
    if ((rc = ensure_frame_buffers(c, std::max<uint64_t>(prims_est, c->geom.cap), clip_est, pairs_est,
                                   std::max(mesh->ntri, max_draw_tris(c)), &grew_prims)))
        return rc;
    DrawDev& h = c->h_draws[d];
    memset(&h, 0, sizeof(h));
    h.pos = mesh->pos;
    h.col = mesh->col;
    h.uv = mesh->uv;
    h.idx = mesh->idx;
    h.ntri = mesh->ntri;
    h.fs = (uint32_t)fs;
    if (fs == RZ_FS_TEXTURE) {  // uniforms.get_texture(0) (main.rs:69-71)
        h.tex = c->textures[0].data;
        h.tw = c->textures[0].tw;
        h.th = c->textures[0].th;
    }
    c->h_mats[d] = c->uni;
    c->draw_mesh.push_back(mesh);
    mesh->pending++;
    c->frame_tris = tris;
    if (grew_prims && d > 0) {
        // earlier draws' triangles lived in the old buffers: redo them
        if ((rc = replay_geometry(c, d + 1))) return rc;
    } else {
        if ((rc = enqueue_draw_desc(c, d))) return rc;
        enqueue_draw_geometry(c, d);
    }
    if (c->prof) HIP_TRY(c, hipEventRecord(c->ev[1], c->stream));
    HIP_TRY(c, hipGetLastError());
    return RZ_OK;
}

int rz_resolve(rz_ctx* c, uint32_t* argb_out) {
    if (!c) return RZ_ERR_ARG;
    return resolve_impl(c, argb_out, nullptr, nullptr, nullptr);
}

int rz_resolve_device(rz_ctx* c, uint32_t* argb_dev) {
    if (!c) return RZ_ERR_ARG;
    if (!argb_dev) return fail(c, RZ_ERR_ARG, "null output");
    return resolve_impl(c, nullptr, argb_dev, nullptr, nullptr);
}

int rz_resolve_msaa(rz_ctx* c, uint32_t* argb_out, uint32_t* color4_out, float* depth4_out) {
    if (!c) return RZ_ERR_ARG;
    return resolve_impl(c, argb_out, nullptr, color4_out, depth4_out);
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
    if (!c->draw_mesh.empty()) return fail(c, RZ_ERR_STATE, "cannot change bins inside a frame");
    BinGeom g = choose_bins(c->W, c->H, pixels);
    if (g.nbx * g.nby > MAX_BINS) return fail(c, RZ_ERR_LIMIT, "more than 256 coarse bins");
    c->coarse_req = pixels;
    c->bg = g;
    return RZ_OK;
}

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

}  // extern "C"
