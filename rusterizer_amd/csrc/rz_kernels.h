// rz_kernels.h -- shared layouts and kernel launchers of the gfx950 rasterizer.
//
// Unit of work: a job = S scenes (render targets) and D draws (meshes with
// their uniforms and fragment shader), each draw belonging to one scene. A
// plain rz_draw/rz_resolve frame is a job with one scene; a batch is a job
// with many. Every kernel reads its sizes from the device (job header and
// frame counters), so the same launch sequence serves any job that fits the
// buffers -- which is what lets a batch be captured once as a HIP graph.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/rz.h"

namespace rzk {

constexpr uint32_t CLEAR_COLOR = 0xFF191919u;  // rasterizer/buffers.rs:5
constexpr float FLT_MAX_F = 3.40282347e+38f;   // f32::MAX
constexpr uint32_t MAXV = 16;         // vertices kept per clipped polygon
constexpr uint32_t CLIP_GROUPS = 8;   // clipped triangles per k_clip wave (8 lanes each)
constexpr uint32_t CHUNK = 256;       // prims per binning chunk (one wave)
constexpr uint32_t MAX_BINS = 256;    // coarse bins per scene (one binning thread each)
constexpr uint32_t TRI_ALIGN = 256;   // draws start at multiples of this in the triangle index space
constexpr uint32_t ATTR_STRIDE = 12;  // floats of colour attributes per prim (3 x RGBA)
constexpr uint32_t UV_STRIDE = 6;     // floats of texture coordinates per prim (3 x uv)
constexpr int SCAN_ITEMS = 8;
constexpr uint32_t SCAN_TILE = 256 * SCAN_ITEMS;
constexpr uint32_t NO_TILE = 0xFFFFFFFFu;
constexpr uint32_t NO_PRIM = 0xFFFFFFFFu;

// PrimGeom.flags bits: texture coordinates not proven in bounds (see
// uv_in_bounds); barycentric arguments finite at covered samples (see finite_bary)
constexpr uint32_t PRIM_UV_CHECK = 8, PRIM_FINITE_BARY = 16;

// overflow bits: the job must be replayed with larger buffers
constexpr uint32_t OVF_PRIMS = 1, OVF_CLIP = 2, OVF_PAIRS = 4, OVF_POLY = 8, OVF_FINE = 16, OVF_MATRIX = 32;

struct FrameCounters {
    uint32_t nprims;        // prim slots of the job (each scene starts at a chunk boundary)
    uint32_t nprims_raw;    // rasterizer triangles of the job
    uint32_t nclip;         // clip slots asked for
    uint32_t overflow;      // OVF_* bits
    uint32_t nchunks;       // coarse-binning chunks (sum over scenes)
    uint32_t mat_len;       // coarse count matrix length (sum over scenes of bins * chunks)
    uint32_t pairs_binned;  // entries placed by the coarse bin scan
    uint32_t fine_total;    // entries placed by the fine bin scan
    uint32_t fine_chunks;   // chunks of the coarse lists (fine binning work items)
    uint32_t fmat_len;      // fine count matrix length (written with fine_chunks as one u64)
    uint32_t q[8];          // raster work queue heads, one per XCD band
    uint32_t pad[2];
    // raster work counters (stats variant only), per wave: list entries
    // scanned, candidates after the list culls, prims processed (some pixel
    // could take them); per sample: samples covered by processed prims; per
    // pixel: shading tasks (distinct surviving prims of a pixel)
    unsigned long long rstat[5];
};
static_assert(offsetof(FrameCounters, fine_chunks) % 8 == 0 &&
                  offsetof(FrameCounters, fmat_len) == offsetof(FrameCounters, fine_chunks) + 4,
              "fine_chunks | fmat_len << 32 is stored as one u64");

struct JobHead {
    uint32_t ndraws, nscenes;
    uint32_t ntri;    // padded triangle index space (sum of draws rounded up to TRI_ALIGN)
    uint32_t ntiles;  // 16x16 raster tiles of all scenes
    uint32_t nbins;   // coarse bins of all scenes
    uint32_t svc_gen;  // raster service: number of this job in its context (low 16 bits are posted)
    uint32_t pad[2];
};

// Per-draw state on the device (mesh streams, fragment shader, MVP).
struct DrawDev {
    float M[16];             // ((P*V)*W), written by k_job_init
    const float* pos;        // nv*3
    const float* col;        // nv*4 or null
    const float* uv;         // nv*2 or null
    const uint32_t* idx;     // ntri*3
    const uint32_t* tex;     // the draw's texture as RGBA8 words, for RZ_FS_TEXTURE
    uint32_t tw, th;
    uint32_t ntri;
    uint32_t fs;
    uint32_t scene;          // owning scene
    uint32_t tri_base;       // first index of the draw in the padded triangle space
};

struct Mat3 {
    float world[16], view[16], proj[16];
};

// Per-scene configuration, written by the host.
struct SceneCfg {
    uint32_t W, H;
    float clear_depth;
    uint32_t cb, nbx, nby;    // coarse bin side (multiple of 16, <= 256) and grid
    uint32_t tile_base;       // first global 16x16 raster tile of the scene
    uint32_t bin_base;        // first global coarse bin of the scene
    uint32_t tri_begin, tri_end;  // padded triangle range of the scene's draws
    uint32_t fts;             // fine-binning tile side = 1 << fts pixels (8, 16 or 32; divides cb)
    uint32_t raster_base;     // first raster slot of the scene's tiles (the host orders scenes densest first)
    uint32_t* out;            // resolved ARGB, W*H words (device)
    uint32_t* msaa_color;     // optional W*H*4 pre-resolve colours
    float* msaa_depth;        // optional W*H*4 pre-resolve depths
    uint64_t pad2;
};
static_assert(sizeof(SceneCfg) == 80, "SceneCfg layout");

// Per-scene values the device derives each frame.
struct SceneRun {
    uint32_t prim_off;    // rasterizer triangles of all earlier scenes
    uint32_t nprims;      // rasterizer triangles of this scene
    uint32_t prim_base;   // first prim slot (multiple of CHUNK)
    uint32_t chunk_base;  // prim_base / CHUNK
    uint32_t nch;         // ceil(nprims / CHUNK)
    uint32_t mat_base;    // first entry of the scene's coarse count matrix
    uint32_t flags;       // RZ_FLAG_* raised while rendering the scene
    uint32_t pad;
};

// One rasterizer triangle (RasterizerTriangle, rasterizer/mod.rs:177-184)
// plus exact-cull helpers, 112 bytes, read with scalar loads by the raster.
struct PrimGeom {
    float ax0, ay0, ax1, ay1, ax2, ay2;  // screen points
    float nx0, ny0, nx1, ny1, nx2, ny2;  // edge normals
    float inv_area;                      // 1 / (2 * area)
    float z0, z1, z2;                    // screen-space depth
    uint32_t flags;                      // tie-rule bit per edge (0..2) | PRIM_* bits | fs << 8 | draw << 16
    float w0, w1, w2;                    // clip-space w
    uint32_t bbx, bby;                   // clamped pixel bbox: lo | hi << 16
    float zlo;                           // lower bound of every interpolated depth
    float pz0, pzx, pzy;                 // depth lower-bound plane: pz0 + pzx*(x-bx0) + pzy*(y-by0)
    uint32_t pad0, pad1;
};
static_assert(sizeof(PrimGeom) == 112, "PrimGeom layout");

// Per-prim binning record and bin list entry: uint4 {prim, bbx, bby, zlo bits}
// (for the per-prim record .x is unused).

// Device pointers of one job's tables and working buffers.
struct JobBufs {
    const JobHead* head;
    DrawDev* draws;
    const Mat3* mats;
    const SceneCfg* scenes;
    SceneRun* runs;
    FrameCounters* cnt;
    // geometry
    uint32_t* cls;
    uint32_t* nsub;
    uint32_t* suboff;
    uint32_t* clip_tri;   // clipped triangle of each clip slot
    uint32_t* clip_draw;  // its draw
    float* clip_out;
    PrimGeom* geom;
    float* attrs;  // colour prims: ATTR_STRIDE floats each
    float* uvat;   // textured prims: UV_STRIDE floats each
    uint4* pbin;
    // coarse bins
    uint32_t* bin_counts;
    uint32_t* bin_offsets;
    uint4* list;
    // fine bins
    uint2* bin_rng;              // coarse list range of each global bin
    unsigned long long* fpack;   // per global bin: T * chunks << 32 | chunks
    unsigned long long* fpre;    // its exclusive scan
    uint32_t* fine_counts;
    uint32_t* fine_off;
    uint2* flist;
    uint4* tile_rec;    // per raster slot and 8x8 quadrant q, at 4 * slot + q: {scene, list begin, list end, x | y << 16}
    uint2* fchunk_rec;  // per fine-binning chunk: {global coarse bin, scene}
    uint32_t* tchunk_draw;   // per TRI_ALIGN-triangle chunk: its draw
    uint32_t* pchunk_scene;  // per CHUNK-prim binning chunk: its scene
    unsigned long long* rstat_parts;  // stats variant: per raster workgroup, its 5 work counts
    // scan partials
    uint32_t* parts;
    unsigned long long* parts64;
    // capacities (elements)
    uint32_t tri_cap, clip_cap, prim_cap, pair_cap, mat_cap, fine_cap, fmat_cap, bins_cap, fchunk_cap;
    uint32_t tile_cap;  // raster tiles tile_rec holds (4 records each)
    uint32_t tchunk_cap, pchunk_cap, rstat_cap;  // rstat_cap in workgroups
    uint32_t clip_wide;    // clip every polygon with a full wave
    uint32_t exact_paths;  // setup proves nothing: every prim takes the raster's exact paths
    uint32_t* svc_flags;   // raster service: pinned host array of the scenes' final flags
};

// Grid sizes of one launch sequence. Kernels loop over their work with the
// device counts, so any grid >= 1 is correct; larger grids only add idle
// workgroups.
struct JobGrid {
    uint32_t draws;      // draw blocks (k_job_init)
    uint32_t tris;       // k_classify / k_setup blocks
    uint32_t clip;       // k_clip waves
    uint32_t bins;       // k_bin blocks
    uint32_t fine;       // k_fine_bin blocks
    uint32_t raster;     // persistent raster workgroups; 0 = one one-wave workgroup per quadrant
    uint32_t ntiles;     // tiles the one-per-tile raster grid must cover (>= the job's)
    uint32_t scenes;     // scene count (k_job_init / k_scene_prep loops)
};

// A block of words passed to a kernel by value (kernel arguments hold up to
// 4 KiB): a small job table set up without any host-to-device copy.
constexpr uint32_t WORD_BLOCK = 256;
struct WordBlock {
    uint32_t w[WORD_BLOCK];
};

// ---- raster service ------------------------------------------------------
// A persistent raster kernel (k_raster_service) stays resident on a stream of
// its own and renders the tiles of jobs that the front ends of up to
// SVC_MAX_ENTRIES contexts post to it: each context owns one entry; its
// front end (every stage up to k_tile_prep) ends with k_svc_post, which
// publishes the job in the entry. A job's tiles are split into 8 contiguous
// bands, one per XCD, each with its own claim word (so claims stay in one
// XCD's L2 instead of bouncing one line between all of them); workgroups
// claim tiles one at a time from their XCD's band of the posted job with
// the fewest tiles left there, then from other XCDs' bands, and the
// workgroup that finishes a job's last tile writes the job's number to host
// memory.
constexpr uint32_t SVC_MAX_ENTRIES = 16;
constexpr uint32_t SVC_TILE_BITS = 24;
constexpr uint32_t SVC_TILE_MASK = (1u << SVC_TILE_BITS) - 1u;
// claims past a job's last tile (at most one per workgroup) must not carry
// into the tile-count field
constexpr uint32_t SVC_MAX_TILES = SVC_TILE_MASK - 8192u;
constexpr uint32_t SVC_MAX_GRID = 8192;

// what the raster of a posted job reads
struct SvcJob {
    uint32_t ntiles;
    const PrimGeom* geom;
    const float* attrs;
    const float* uvat;
    const uint2* flist;
    const uint4* tile_rec;
    const SceneCfg* scenes;
    const DrawDev* draws;
    SceneRun* runs;
    uint32_t* host_flags;  // pinned: each scene's RZ_FLAG_* bits once the job is done
    uint32_t nscenes;
};

// one XCD's claim word, alone in its cache line: job number (16 bits) << 48
// | the band's tiles << 24 | next unclaimed tile of the band. k_svc_post
// stores it; workgroups claim with one atomic add, and a claim returns the
// word as it was, so whoever adds owns the tile it names.
struct alignas(128) SvcBand {
    unsigned long long claim;
};

struct alignas(256) SvcEntry {
    SvcBand band[8];
    uint32_t finished;  // tiles of the posted job rendered
    uint32_t pad;
    SvcJob job;
};

struct alignas(256) SvcCtl {
    unsigned long long last_claim;  // wall clock of a recent claim by any workgroup
    uint32_t stop;                  // raised by workgroup 0 when the host asks the service to stop
    uint32_t posted;                // jobs posted so far (idle workgroups watch this one word)
};

// pinned host memory shared with the service
struct alignas(256) SvcHost {
    uint32_t done[SVC_MAX_ENTRIES];  // number of the last job each entry finished
    uint32_t stop;                   // the host asks the service to stop once idle
};

// ---- launchers (all asynchronous on `s`) ---------------------------------
void launch_max_index(hipStream_t s, const uint32_t* idx, uint64_t n, uint32_t* out);  // *out must start at 0
// dst[0, n) = blk.w[0, n), n <= WORD_BLOCK (one small kernel)
void launch_write_words(hipStream_t s, const WordBlock& blk, uint32_t n, uint32_t* dst);
// p[0, n) = v (one small kernel; used where a memset node would be captured)
void launch_fill_u32(hipStream_t s, uint32_t* p, uint32_t n, uint32_t v);
// the whole pipeline of one job, from counter reset to resolved outputs;
// with svc (the context's raster service entry) the raster is left to the
// service: the sequence ends by posting the job there
void launch_job(hipStream_t s, const JobBufs& b, const JobGrid& g, bool raster_stats, hipEvent_t ev_bin,
                hipEvent_t ev_raster, SvcEntry* svc = nullptr, SvcHost* svc_host = nullptr, SvcCtl* svc_ctl = nullptr,
                uint32_t svc_index = 0);
// the persistent raster service: grid workgroups (each CU's share of which
// must fit beside the front ends' kernels) that exit when the host raises
// svc_host->stop and they find no work, when no workgroup has claimed a tile
// for idle_ticks of the wall clock, or life_ticks after the launch
void launch_raster_service(hipStream_t s, uint32_t grid, SvcEntry* ents, uint32_t nent, SvcCtl* ctl,
                           SvcHost* host, unsigned long long idle_ticks, unsigned long long life_ticks);

}  // namespace rzk
