// rz_kernels.h -- shared layouts and kernel launchers of the gfx950 rasterizer.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/rz.h"

namespace rzk {

constexpr uint32_t CLEAR_COLOR = 0xFF191919u;  // rasterizer/buffers.rs:5
constexpr float FLT_MAX_F = 3.40282347e+38f;   // f32::MAX
constexpr uint32_t MAXV = 16;         // vertices kept per clipped polygon
constexpr uint32_t CLIP_LDS_V = 66;   // polygon vertices a clip wave can stage
constexpr uint32_t CHUNK = 256;       // prims per binning chunk (one wave)
constexpr uint32_t MAX_BINS = 256;    // coarse bins per frame (4 per lane)
constexpr uint32_t ATTR_STRIDE = 12;  // floats of attributes per prim
constexpr int SCAN_ITEMS = 8;
constexpr uint32_t SCAN_TILE = 256 * SCAN_ITEMS;

// overflow bits: the frame must be replayed with larger buffers
constexpr uint32_t OVF_PRIMS = 1, OVF_CLIP = 2, OVF_PAIRS = 4, OVF_POLY = 8, OVF_FINE = 16;

struct FrameCounters {
    uint32_t nprims;        // rasterizer triangles of the frame so far
    uint32_t nclip;         // clip slots used by the current draw
    uint32_t clip_need;     // most clip slots any draw of the frame asked for
    uint32_t overflow;      // OVF_* bits
    uint32_t flags;         // RZ_FLAG_* bits
    uint32_t nclipped;      // input triangles sent to the clipper
    unsigned long long npairs;  // coarse-bin entries counted by k_setup
    uint32_t pairs_binned;  // entries placed by the coarse bin scan
    uint32_t nch;           // chunks of the coarse lists (fine binning)
    uint32_t fine_total;    // entries placed by the fine bin scan
    uint32_t pad;
    // raster work counters (stats variant only): list entries scanned,
    // candidates after the bbox / depth / edge culls, hits processed,
    // hits with coverage, hits that shaded
    unsigned long long rstat[5];
};

// Per-draw state on the device (uniforms + mesh streams + fragment shader).
struct DrawDev {
    float M[16];             // ((P*V)*W), written by k_mvp
    const float* pos;        // nv*3
    const float* col;        // nv*4 or null
    const float* uv;         // nv*2 or null
    const uint32_t* idx;     // ntri*3
    const uint32_t* tex;     // texture 0 as RGBA8 words, for RZ_FS_TEXTURE
    uint32_t tw, th;
    uint32_t ntri;
    uint32_t fs;
};

// One rasterizer triangle (RasterizerTriangle, rasterizer/mod.rs:177-184)
// plus exact-cull helpers, 112 bytes, read with scalar loads by the raster.
struct PrimGeom {
    float ax0, ay0, ax1, ay1, ax2, ay2;  // screen points
    float nx0, ny0, nx1, ny1, nx2, ny2;  // edge normals
    float inv_area;                      // 1 / (2 * area)
    float z0, z1, z2;                    // screen-space depth
    float w0, w1, w2;                    // clip-space w
    uint32_t flags;                      // tie-rule bit per edge (0..2) | fs << 8 | draw << 16
    uint32_t bbx, bby;                   // clamped pixel bbox: lo | hi << 16
    float zlo;                           // lower bound of every interpolated depth
    float pz0, pzx, pzy;                 // depth lower-bound plane: pz0 + pzx*(x-bx0) + pzy*(y-by0)
    uint32_t pad0, pad1;
};
static_assert(sizeof(PrimGeom) == 112, "PrimGeom layout");

// Per-prim binning record and coarse-bin list entry: {prim, bbx, bby, zlo bits}.
// (uint4; for the per-prim record .x is unused.)

struct BinGeom {
    uint32_t cb;        // coarse bin side in pixels (multiple of 16)
    uint32_t nbx, nby;  // bins across / down
};

struct Mat3 {
    float world[16], view[16], proj[16];
};

// ---- launchers (all asynchronous on `s`) ---------------------------------
void launch_max_index(hipStream_t s, const uint32_t* idx, uint64_t n, uint32_t* out);  // *out must start at 0
void launch_frame_reset(hipStream_t s, FrameCounters* cnt);
void launch_draw_reset(hipStream_t s, FrameCounters* cnt);
void launch_mvp(hipStream_t s, DrawDev* draws, uint32_t d, const Mat3& m);
void launch_classify(hipStream_t s, const DrawDev* draws, uint32_t d, uint32_t ntri, FrameCounters* cnt,
                     uint32_t* cls, uint32_t* nsub, uint32_t* clip_tri, uint32_t clip_cap);
void launch_clip(hipStream_t s, const DrawDev* draws, uint32_t d, FrameCounters* cnt, const uint32_t* clip_tri,
                 uint32_t clip_cap, uint32_t* nsub, float* clip_out, uint32_t* clip_n);
// exclusive scan; with n_dev the length is mul * ceil(*n_dev / div), at most
// n_host (which also sizes the grid).
void launch_scan(hipStream_t s, const uint32_t* in, uint32_t* out, uint32_t n_host, const uint32_t* n_dev,
                 uint32_t mul, uint32_t div, uint32_t* parts, uint32_t* base_total, uint32_t* total2);
void launch_setup(hipStream_t s, const DrawDev* draws, uint32_t d, uint32_t ntri, FrameCounters* cnt,
                  const uint32_t* cls, const uint32_t* nsub, const uint32_t* sub_off, const float* clip_out,
                  uint32_t clip_cap, PrimGeom* geom, float* attrs, uint4* pbin, uint32_t prim_cap, uint32_t W,
                  uint32_t H, BinGeom bg);
void launch_bin(hipStream_t s, bool scatter, uint32_t grid, FrameCounters* cnt, const uint4* pbin, BinGeom bg,
                uint32_t* counts, const uint32_t* offsets, uint4* list, uint32_t pair_cap);
// fine binning of the coarse lists into 16x16-tile lists (nch_max bounds the chunk count)
void launch_fine(hipStream_t s, FrameCounters* cnt, const uint32_t* coff, const uint4* clist, uint32_t* chunk_base,
                 BinGeom bg, uint32_t* fcounts, uint32_t* foff, uint4* flist, uint32_t fcap, uint32_t nch_max,
                 uint32_t* parts);
void launch_raster(hipStream_t s, FrameCounters* cnt, const PrimGeom* geom, const float* attrs,
                   const DrawDev* draws, const uint32_t* chunk_base, const uint32_t* foff, const uint4* list,
                   uint32_t fcap, BinGeom bg, uint32_t W, uint32_t H, float clear_depth, uint32_t* out,
                   uint32_t* msaa_color, float* msaa_depth, bool stats);

}  // namespace rzk
