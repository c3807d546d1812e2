// rz_kernels.hip -- gfx950 kernels of the rasterizer (device side).
//
// Arithmetic contract: this file is compiled with -ffp-contract=off, IEEE
// f32 division and denormals kept, so each expression below produces the
// same bits as the reference's Rust expression it cites (no FMA fusion,
// same association, Rust's saturating float->int casts).
//
// Pipeline of one job (S scenes, D draws; all on one stream):
//   k_job_init (counter reset, MVPs) -> k_classify -> k_clip -> scan(nsub)
//   -> k_scene_prep -> k_setup                                  (geometry)
//   -> k_bin<count> -> scan -> k_bin<scatter>                   (coarse bins)
//   -> k_fine_prep -> scan64 -> k_fine_bin<count> -> scan
//   -> k_fine_bin<scatter>                      (8, 16 or 32 px fine tiles)
//   -> k_tile_prep -> k_raster_q (coverage, depth, shading, resolve)
// Rasterizer triangles ("prims") are stored per scene in submission order
// (draw, input triangle, clip-fan index), so every later stage keeps the
// reference's order per pixel without sorting. Each scene's prims start at a
// CHUNK boundary, so a binning chunk never mixes scenes.

#include "rz_kernels.h"

#include <hip/hip_runtime.h>

#include <algorithm>

namespace rzk {

// ---------------------------------------------------------------------------
// Rust scalar semantics on the device
// ---------------------------------------------------------------------------

// `f as u32`: saturating, NaN -> 0 (color.rs:15-20, bounding_box.rs:27-32).
__device__ __forceinline__ uint32_t sat_u32(float f) {
    if (!(f > 0.0f)) return 0u;
    if (f >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)f;
}

// f32::clamp(0, 1), NaN passes (mod.rs:102-106).
__device__ __forceinline__ float clamp_bary(float x) {
    if (x < 0.0f) x = 0.0f;
    if (x > 1.0f) x = 1.0f;
    return x;
}

__device__ __forceinline__ uint32_t lane_id() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}
// number of set bits of m below this lane
__device__ __forceinline__ uint32_t mbcnt(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
__device__ __forceinline__ uint32_t uniform(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}
__device__ __forceinline__ uint32_t lo32(unsigned long long v) { return (uint32_t)v; }
__device__ __forceinline__ uint32_t hi32(unsigned long long v) { return (uint32_t)(v >> 32); }

// ---------------------------------------------------------------------------
// Table lookups (binary searches). Called with a wave-uniform argument the
// compiler keeps the loads scalar; k_fine_prep calls find_scene_bin per lane.
// ---------------------------------------------------------------------------

// draw holding padded triangle t: largest d < nd with draws[d].tri_base <= t
__device__ __forceinline__ uint32_t find_draw(const DrawDev* __restrict__ draws, uint32_t nd, uint32_t t) {
    uint32_t lo = 0, hi = nd;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (draws[mid].tri_base <= t) lo = mid;
        else hi = mid;
    }
    return lo;
}
// scene holding global coarse-binning chunk c (scenes without prims share
// their successor's chunk_base; the largest index is the owner)
__device__ __forceinline__ uint32_t find_scene_chunk(const SceneRun* __restrict__ runs, uint32_t ns, uint32_t c) {
    uint32_t lo = 0, hi = ns;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (runs[mid].chunk_base <= c) lo = mid;
        else hi = mid;
    }
    return lo;
}
__device__ __forceinline__ uint32_t find_scene_tile(const SceneCfg* __restrict__ sc, uint32_t ns, uint32_t t) {
    uint32_t lo = 0, hi = ns;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (sc[mid].tile_base <= t) lo = mid;
        else hi = mid;
    }
    return lo;
}
__device__ __forceinline__ uint32_t find_scene_bin(const SceneCfg* __restrict__ sc, uint32_t ns, uint32_t g) {
    uint32_t lo = 0, hi = ns;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (sc[mid].bin_base <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}
// ---------------------------------------------------------------------------
// Vertex stage: ((P*V)*W) * (x, y, z, 1) with the reference's dot order
// ---------------------------------------------------------------------------

// vector.rs:17-23 / vector.rs:219-240: sum from 0.0, left to right.
__device__ __forceinline__ float4 mat_vec_p(const float* __restrict__ M, float x, float y, float z) {
    float4 r;
    r.x = (((0.0f + M[0] * x) + M[1] * y) + M[2] * z) + M[3] * 1.0f;
    r.y = (((0.0f + M[4] * x) + M[5] * y) + M[6] * z) + M[7] * 1.0f;
    r.z = (((0.0f + M[8] * x) + M[9] * y) + M[10] * z) + M[11] * 1.0f;
    r.w = (((0.0f + M[12] * x) + M[13] * y) + M[14] * z) + M[15] * 1.0f;
    return r;
}

// per-chunk owner tables, one thread per chunk (the searches run in parallel
// here instead of serially at the start of every workgroup that needs them):
// tchunk_map (in k_job_init): the draw of each TRI_ALIGN-triangle chunk (draws
// start at chunk boundaries); k_pchunk_map: the scene of each CHUNK-prim
// binning chunk
__device__ __forceinline__ void tchunk_map(const JobBufs& b) {
    const uint32_t nc = min((b.head->ntri + TRI_ALIGN - 1) / TRI_ALIGN, b.tchunk_cap), nd = b.head->ndraws;
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x)
        b.tchunk_draw[c] = find_draw(b.draws, nd, c * TRI_ALIGN);
}
__device__ __forceinline__ void pchunk_map(const JobBufs& b, uint32_t nchunks) {
    const uint32_t nc = min(nchunks, b.pchunk_cap), ns = b.head->nscenes;
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x)
        b.pchunk_scene[c] = find_scene_chunk(b.runs, ns, c);
}
__global__ __launch_bounds__(256) void k_pchunk_map(JobBufs b) {
    if (b.cnt->overflow != 0) return;
    pchunk_map(b, b.cnt->nchunks);
}

// main.rs:147-152: projection * view * world, evaluated left to right
// (matrix.rs:56-79). One thread per draw; written into the draw descriptor.
__device__ __forceinline__ void draw_mvp(const JobBufs& b) {
    const uint32_t nd = b.head->ndraws;
    for (uint32_t d = blockIdx.x * blockDim.x + threadIdx.x; d < nd; d += gridDim.x * blockDim.x) {
        const Mat3& m = b.mats[d];
        float pv[16], out[16];
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) {
                float s = 0.0f;
                for (int k = 0; k < 4; ++k) s = s + m.proj[i * 4 + k] * m.view[k * 4 + j];
                pv[i * 4 + j] = s;
            }
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) {
                float s = 0.0f;
                for (int k = 0; k < 4; ++k) s = s + pv[i * 4 + k] * m.world[k * 4 + j];
                out[i * 4 + j] = s;
            }
        for (int i = 0; i < 16; ++i) b.draws[d].M[i] = out[i];
    }
}

__device__ __forceinline__ void load_vertex(const DrawDev& dd, uint32_t vi, float4& p, float a[6]) {
    const float* pos = dd.pos + (size_t)vi * 3;
    p = mat_vec_p(dd.M, pos[0], pos[1], pos[2]);
    if (dd.col) {
        float4 c = *reinterpret_cast<const float4*>(dd.col + (size_t)vi * 4);
        a[0] = c.x; a[1] = c.y; a[2] = c.z; a[3] = c.w;
    } else {
        a[0] = a[1] = a[2] = a[3] = 0.0f;
    }
    if (dd.uv) {
        float2 t = *reinterpret_cast<const float2*>(dd.uv + (size_t)vi * 2);
        a[4] = t.x; a[5] = t.y;
    } else {
        a[4] = a[5] = 0.0f;
    }
}

// ---------------------------------------------------------------------------
// Job start: zero the counters and the per-scene flags (a kernel, not a
// memset, so the sequence can be captured in a graph).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void frame_reset(const JobBufs& b) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        FrameCounters* c = b.cnt;
        c->nprims = 0;
        c->nprims_raw = 0;
        c->nclip = 0;
        c->overflow = 0;
        c->nchunks = 0;
        c->mat_len = 0;
        c->pairs_binned = 0;
        c->fine_chunks = 0;
        c->fmat_len = 0;
        c->fine_total = 0;
        for (int i = 0; i < 8; ++i) c->q[i] = 0;
        for (int i = 0; i < 5; ++i) c->rstat[i] = 0;
    }
    const uint32_t ns = b.head->nscenes;
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < ns; s += gridDim.x * blockDim.x)
        b.runs[s].flags = 0;
}

// k_job_init: the job's three independent first passes in one launch (each a
// grid-stride loop, so any grid serves): the counter and flag reset, the
// draws' matrices and the triangle chunks' draws
__global__ __launch_bounds__(256) void k_job_init(JobBufs b) {
    frame_reset(b);
    draw_mvp(b);
    tchunk_map(b);
}

// ---------------------------------------------------------------------------
// k_classify: one thread per input triangle; a workgroup takes chunks of
// TRI_ALIGN triangles (each of one draw) in a grid-stride loop.
// render.rs:104-108 (VS), render.rs:75-96 (assembly), clipping.rs:62-112
// (degenerate cull, trivial reject / accept). Triangles that need the
// Sutherland-Hodgman clipper get a slot in the clip list.
// cls[t]: 0 = Outside (or padding), 1 = Inside, 2 + slot = clipped.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_classify(JobBufs b, const uint32_t* __restrict__ tchunk_draw,
                                                  const DrawDev* __restrict__ draws) {
    __shared__ uint32_t wcount[4], wbase[4];
    const uint32_t NT = b.head->ntri;
    const uint32_t wave = threadIdx.x >> 6;
    uint32_t wneed = 0;  // the wave's clipped triangles over all its chunks
    for (uint32_t blk = blockIdx.x; blk * 256u < NT; blk += gridDim.x) {
        const uint32_t t = blk * 256u + threadIdx.x;
        const uint32_t d = uniform(tchunk_draw[blk]);
        const DrawDev& dd = draws[d];
        const uint32_t lt = t - dd.tri_base;
        const bool live = lt < dd.ntri;
        uint32_t c = 0;
        if (live) {
            float4 v[3];
            for (int k = 0; k < 3; ++k) {
                uint32_t vi = dd.idx[(size_t)lt * 3 + k];
                const float* pos = dd.pos + (size_t)vi * 3;
                v[k] = mat_vec_p(dd.M, pos[0], pos[1], pos[2]);
            }
            // clipping.rs:63-65 with mod.rs:15-21 on clip-space xy
            float v10x = v[1].x - v[0].x, v10y = v[1].y - v[0].y;
            float v20x = v[2].x - v[0].x, v20y = v[2].y - v[0].y;
            float area2 = v10x * v20y - v20x * v10y;
            if (fabsf(area2) < 0.000001f) {
                c = 0;
            } else {
                bool in_all = true, out_any = false;
                bool ox0 = true, ox1 = true, oy0 = true, oy1 = true, oz0 = true, oz1 = true;
                for (int k = 0; k < 3; ++k) {
                    float x = v[k].x, y = v[k].y, z = v[k].z, w = v[k].w;
                    in_all = in_all && (x >= -w) && (y >= -w) && (z >= -w) && (x <= w) && (y <= w) && (z <= w);
                    ox0 = ox0 && (x < -w);
                    oy0 = oy0 && (y < -w);
                    oz0 = oz0 && (z < -w);
                    ox1 = ox1 && (x > w);
                    oy1 = oy1 && (y > w);
                    oz1 = oz1 && (z > w);
                }
                out_any = ox0 || oy0 || oz0 || ox1 || oy1 || oz1;
                c = out_any ? 0u : (in_all ? 1u : 2u);
            }
        }
        // pass 1 marks clipped triangles with 2; their slots come below
        const bool need = live && c == 2;
        wneed += (uint32_t)__popcll(__ballot(need));
        if (t < NT) {
            b.cls[t] = need ? 2u : c;
            b.nsub[t] = need ? 0u : c;  // Inside -> 1 rasterizer triangle, Outside / padding -> 0;
                                        // clipped: written by k_clip
        }
    }
    // one clip-list reservation per workgroup for all its chunks (a global
    // atomic per chunk would serialize thousands of workgroups on one counter)
    if (lane_id() == 0) wcount[wave] = wneed;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t base = 0;
        const uint32_t tot = wcount[0] + wcount[1] + wcount[2] + wcount[3];
        if (tot) base = atomicAdd(&b.cnt->nclip, tot);
        for (int w = 0; w < 4; ++w) {
            wbase[w] = base;
            base += wcount[w];
        }
    }
    __syncthreads();
    if (wneed == 0) return;
    // pass 2: each wave hands out its slots in triangle order, recording
    // each clipped triangle and its draw
    uint32_t next = wbase[wave];
    for (uint32_t blk = blockIdx.x; blk * 256u < NT; blk += gridDim.x) {
        const uint32_t t = blk * 256u + threadIdx.x;
        const bool need = t < NT && b.cls[t] == 2u;
        const uint64_t m = __ballot(need);
        if (m == 0) continue;
        const uint32_t d = uniform(tchunk_draw[blk]);
        if (need) {
            const uint32_t slot = next + mbcnt(m);
            b.cls[t] = 2u + slot;
            if (slot < b.clip_cap) {
                b.clip_tri[slot] = t;
                b.clip_draw[slot] = d;
            } else {
                atomicOr(&b.cnt->overflow, OVF_CLIP);
            }
        }
        next += (uint32_t)__popcll(m);
    }
}

// ---------------------------------------------------------------------------
// Clipping: Sutherland-Hodgman over the planes L, R, B, T, N, F
// (clipping.rs:114-177), one group of G lanes per clipped triangle, one lane
// per polygon vertex. Lane i of a group owns the edge (prev = i-1 mod n,
// cur = i) and emits 0, 1 or 2 vertices; a prefix count over the group
// places them, so the output keeps the reference's vertex order exactly.
// The final polygon (<= MAXV vertices) goes to its clip slot; the fan
// (clipping.rs:183-190) is formed in k_setup.
// k_clip runs CLIP_GROUPS groups of 64 / CLIP_GROUPS lanes per wave. A
// polygon that would need more lanes than its group has at some plane (a
// triangle clipped by all six planes can reach 9 vertices; rounding can add
// more) is clipped again from the start by the whole wave, once its groups
// are done.
// ---------------------------------------------------------------------------
struct ClipV {
    float x, y, z, w;
    float a[6];
};

// Returns true in the lanes of a group whose polygon needs more than G lanes
// (nothing is written for it then).
template <uint32_t G>
__device__ __forceinline__ bool clip_polygon(const JobBufs& b, uint32_t slot, bool have, float* __restrict__ sh) {
    const uint32_t lane = threadIdx.x & 63u, gl = lane & (G - 1), gb = lane - gl;
    const uint64_t gmask = G == 64 ? ~0ull : (((1ull << G) - 1) << gb);
    float* q = sh + gb * 10;  // this group's staging area: G vertices
    uint32_t t = 0, scene = 0;
    ClipV v;
    v.x = v.y = v.z = v.w = 0.0f;
    for (int j = 0; j < 6; ++j) v.a[j] = 0.0f;
    if (have) {
        t = b.clip_tri[slot];
        const uint32_t d = b.clip_draw[slot];
        const DrawDev& dd = b.draws[d];
        scene = dd.scene;
        if (gl < 3) {
            const uint32_t vi = dd.idx[(size_t)(t - dd.tri_base) * 3 + gl];
            float4 p;
            load_vertex(dd, vi, p, v.a);
            v.x = p.x; v.y = p.y; v.z = p.z; v.w = p.w;
        }
    }
    uint32_t n = have ? 3u : 0u;  // the same in every lane of the group
    bool big = false;             // needs more than G lanes
    for (int plane = 0; plane < 6; ++plane) {
        const bool live = n > 0 && !big;
        const bool act = live && gl < n;
        // clipping.rs:29-38
        float dcur;
        switch (plane) {
            case 0: dcur = v.w + v.x; break;
            case 1: dcur = v.w - v.x; break;
            case 2: dcur = v.w + v.y; break;
            case 3: dcur = v.w - v.y; break;
            case 4: dcur = v.w + v.z; break;
            default: dcur = v.w - v.z; break;
        }
        // a plane with every vertex of every polygon of the wave inside it
        // leaves the polygons unchanged (each edge emits its current vertex)
        if (__ballot(act && !(dcur >= 0.0f)) == 0) continue;
        const int prev = live ? (int)gb + (gl == 0 ? (int)n - 1 : (int)gl - 1) : (int)lane;
        ClipV pv;
        pv.x = __shfl(v.x, prev);
        pv.y = __shfl(v.y, prev);
        pv.z = __shfl(v.z, prev);
        pv.w = __shfl(v.w, prev);
        for (int j = 0; j < 6; ++j) pv.a[j] = __shfl(v.a[j], prev);
        const float dprev = __shfl(dcur, prev);
        const bool pin = dprev >= 0.0f, cin = dcur >= 0.0f;
        const uint32_t ncnt = act ? ((pin && cin) ? 1u : (pin ? 1u : (cin ? 2u : 0u))) : 0u;
        const uint64_t m1 = __ballot(ncnt >= 1) & gmask;
        const uint64_t m2 = __ballot(ncnt == 2) & gmask;
        const uint32_t pos = mbcnt(m1) + mbcnt(m2);
        const uint32_t newn = (uint32_t)__popcll(m1) + (uint32_t)__popcll(m2);
        ClipV o0 = v;
        if (act && (pin != cin)) {
            // clipping.rs:43-51 with p0 = prev, p1 = cur
            const float alpha = dprev / (dprev - dcur);
            const float s1 = 1.0f - alpha;
            o0.x = pv.x * s1 + v.x * alpha;
            o0.y = pv.y * s1 + v.y * alpha;
            o0.z = pv.z * s1 + v.z * alpha;
            o0.w = pv.w * s1 + v.w * alpha;
            // clipping.rs:149,161: (cur - prev) * alpha + prev
            for (int j = 0; j < 6; ++j) o0.a[j] = (v.a[j] - pv.a[j]) * alpha + pv.a[j];
        }
        __syncthreads();
        if (ncnt >= 1 && pos < G) {
            float* o = q + pos * 10;
            o[0] = o0.x; o[1] = o0.y; o[2] = o0.z; o[3] = o0.w;
            for (int j = 0; j < 6; ++j) o[4 + j] = o0.a[j];
        }
        if (ncnt == 2 && pos + 1 < G) {
            float* o = q + (pos + 1) * 10;
            o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
            for (int j = 0; j < 6; ++j) o[4 + j] = v.a[j];
        }
        __syncthreads();
        if (live) {
            n = newn;
            if (n > G) big = true;
            if (!big && gl < n) {
                const float* o = q + gl * 10;
                v.x = o[0]; v.y = o[1]; v.z = o[2]; v.w = o[3];
                for (int j = 0; j < 6; ++j) v.a[j] = o[4 + j];
            }
        }
    }
    if (!have) return false;
    if (big) {
        // more vertices than a wave's lanes: cannot continue
        if (G == 64 && gl == 0) atomicOr(&b.cnt->overflow, OVF_POLY);
        return G < 64;
    }
    uint32_t ntris = 0;
    if (n > MAXV) {
        if (gl == 0) atomicOr(&b.cnt->overflow, OVF_POLY);
    } else if (n >= 3) {
        ntris = n - 2;
        if (gl < n) {
            float* o = b.clip_out + ((size_t)slot * MAXV + gl) * 10;
            o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
            for (int j = 0; j < 6; ++j) o[4 + j] = v.a[j];
        }
    } else if (n > 0) {
        if (gl == 0) atomicOr(&b.runs[scene].flags, RZ_FLAG_CLIP_FEWER);
    }
    if (gl == 0) b.nsub[t] = ntris;
    return false;
}

// CLIP_GROUPS clipped triangles per wave, then with all 64 lanes each of them
// that needed more than its group (with clip_wide set, every one is clipped
// with all 64 lanes)
__global__ __launch_bounds__(64) void k_clip(JobBufs b) {
    __shared__ float sh[64 * 10];
    const uint32_t nslots = min(uniform(b.cnt->nclip), b.clip_cap);
    if (b.clip_wide) {
        for (uint32_t i = blockIdx.x; i < nslots; i += gridDim.x) (void)clip_polygon<64>(b, i, true, sh);
        return;
    }
    constexpr uint32_t G = 64 / CLIP_GROUPS;
    for (uint32_t base = blockIdx.x * CLIP_GROUPS; base < nslots; base += gridDim.x * CLIP_GROUPS) {
        const uint32_t slot = base + threadIdx.x / G;
        const bool big = clip_polygon<G>(b, slot, slot < nslots, sh);
        // one bit per group whose polygon was too big: lane G * k of group k
        uint64_t redo = __ballot(big && (threadIdx.x % G) == 0);
        while (redo != 0) {
            const uint32_t k = (uint32_t)__ffsll((unsigned long long)redo) - 1u;
            redo &= redo - 1;
            (void)clip_polygon<64>(b, base + k / G, true, sh);
        }
    }
}

// ---------------------------------------------------------------------------
// Exclusive scan of u32 or u64 (3 launches: tile sums, scan of sums, tile
// scan; one launch for short arrays, k_scan_one). The length is *n_dev when
// given (bounded by n_host, which also sizes the grid), else n_host. The
// grand total goes to *total (if given).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t scan_len(uint32_t n_host, const uint32_t* n_dev) {
    if (n_dev == nullptr) return n_host;
    uint32_t s = uniform(*n_dev);
    return s < n_host ? s : n_host;  // never past the buffers the host sized
}

template <class T>
__device__ __forceinline__ T block_exclusive_scan(T v, T* lds, T* total) {
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    T x = v;
    for (int o = 1; o < 64; o <<= 1) {
        T y = __shfl_up(x, o);
        if (lane >= (uint32_t)o) x += y;
    }
    if (lane == 63) lds[wave] = x;
    __syncthreads();
    uint32_t nw = blockDim.x >> 6;
    T pre = 0, tot = 0;
    for (uint32_t w = 0; w < nw; ++w) {
        T s = lds[w];
        if (w < wave) pre += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return pre + x - v;
}

template <class T>
__global__ __launch_bounds__(256) void k_scan_reduce(const T* __restrict__ in, uint32_t n_host,
                                                     const uint32_t* __restrict__ n_dev, T* __restrict__ parts) {
    __shared__ T lds[8];
    uint32_t n = scan_len(n_host, n_dev);
    uint32_t base = blockIdx.x * SCAN_TILE;
    if (base >= n) return;
    T s = 0;
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        uint32_t k = base + threadIdx.x * SCAN_ITEMS + i;
        if (k < n) s += in[k];
    }
    T tot;
    block_exclusive_scan<T>(s, lds, &tot);
    if (threadIdx.x == 0) parts[blockIdx.x] = tot;
}

template <class T>
__global__ __launch_bounds__(1024) void k_scan_top(T* __restrict__ parts, uint32_t n_host,
                                                   const uint32_t* __restrict__ n_dev, T* __restrict__ total) {
    __shared__ T lds[16];
    uint32_t n = scan_len(n_host, n_dev);
    uint32_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    T run = 0;
    for (uint32_t off = 0; off < nb; off += blockDim.x) {
        uint32_t k = off + threadIdx.x;
        T v = k < nb ? parts[k] : T(0);
        T tot;
        T ex = block_exclusive_scan<T>(v, lds, &tot);
        if (k < nb) parts[k] = run + ex;
        run += tot;
    }
    if (threadIdx.x == 0 && total) *total = run;
}

template <class T>
__global__ __launch_bounds__(256) void k_scan_down(const T* __restrict__ in, uint32_t n_host,
                                                   const uint32_t* __restrict__ n_dev, const T* __restrict__ parts,
                                                   T* __restrict__ out) {
    __shared__ T lds[8];
    uint32_t n = scan_len(n_host, n_dev);
    uint32_t base = blockIdx.x * SCAN_TILE;
    if (base >= n) return;
    T v[SCAN_ITEMS];
    T s = 0;
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        uint32_t k = base + threadIdx.x * SCAN_ITEMS + i;
        v[i] = k < n ? in[k] : T(0);
        s += v[i];
    }
    T tot;
    T ex = block_exclusive_scan<T>(s, lds, &tot) + parts[blockIdx.x];
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        uint32_t k = base + threadIdx.x * SCAN_ITEMS + i;
        if (k < n) out[k] = ex;
        ex += v[i];
    }
}

// The same scan by one workgroup, blockDim.x * SCAN_ITEMS elements per pass
// with the running total carried between passes: for short arrays, whose
// scan would otherwise be three launches of a few microseconds each (a small
// frame's job is mostly such launches).
constexpr uint32_t SCAN_ONE_MAX = 4u * 1024u * SCAN_ITEMS;

// (returns the total in every thread)
template <class T>
__device__ __forceinline__ T scan_one(const T* __restrict__ in, uint32_t n, T* __restrict__ out,
                                      T* __restrict__ total, T* lds) {
    T run = 0;
    for (uint32_t base = 0; base < n; base += blockDim.x * SCAN_ITEMS) {
        T v[SCAN_ITEMS];
        T s = 0;
        for (int i = 0; i < SCAN_ITEMS; ++i) {
            const uint32_t k = base + threadIdx.x * SCAN_ITEMS + i;
            v[i] = k < n ? in[k] : T(0);
            s += v[i];
        }
        T tot;
        T ex = run + block_exclusive_scan<T>(s, lds, &tot);
        for (int i = 0; i < SCAN_ITEMS; ++i) {
            const uint32_t k = base + threadIdx.x * SCAN_ITEMS + i;
            if (k < n) out[k] = ex;
            ex += v[i];
        }
        run += tot;
    }
    if (threadIdx.x == 0 && total) *total = run;
    return run;
}

template <class T>
__global__ __launch_bounds__(1024) void k_scan_one(const T* __restrict__ in, uint32_t n_host,
                                                   const uint32_t* __restrict__ n_dev, T* __restrict__ out,
                                                   T* __restrict__ total) {
    __shared__ T lds[16];
    (void)scan_one<T>(in, scan_len(n_host, n_dev), out, total, lds);
}

template <class T>
void launch_scan(hipStream_t s, const T* in, T* out, uint32_t n_host, const uint32_t* n_dev, T* parts, T* total) {
    if (n_host <= SCAN_ONE_MAX) {
        k_scan_one<T><<<1, 1024, 0, s>>>(in, n_host, n_dev, out, total);
        return;
    }
    uint32_t nb = (n_host + SCAN_TILE - 1) / SCAN_TILE;
    k_scan_reduce<T><<<nb, 256, 0, s>>>(in, n_host, n_dev, parts);
    k_scan_top<T><<<1, 1024, 0, s>>>(parts, n_host, n_dev, total);
    k_scan_down<T><<<nb, 256, 0, s>>>(in, n_host, n_dev, parts, out);
}

// ---------------------------------------------------------------------------
// k_scene_prep: one workgroup. From the scanned fan counts, each scene's
// rasterizer triangle range, its CHUNK-aligned prim slots, its binning
// chunks and the base of its coarse count matrix (bins x chunks).
// ---------------------------------------------------------------------------
// (blockDim.x threads; total = the job's rasterizer triangles; returns the
// job's binning chunks and coarse matrix length in every thread)
__device__ __forceinline__ uint2 scene_prep(const JobBufs& b, uint32_t total, uint32_t* l1, uint32_t* l2) {
    const uint32_t ns = b.head->nscenes, NT = b.head->ntri;
    uint32_t run_ch = 0, run_mat = 0;
    for (uint32_t off = 0; off < ns; off += blockDim.x) {
        const uint32_t s = off + threadIdx.x;
        uint32_t nch = 0, nmat = 0, beg = 0, np = 0;
        if (s < ns) {
            const SceneCfg& sc = b.scenes[s];
            beg = sc.tri_begin < NT ? b.suboff[sc.tri_begin] : total;
            const uint32_t end = sc.tri_end < NT ? b.suboff[sc.tri_end] : total;
            np = end - beg;
            nch = (np + CHUNK - 1) / CHUNK;
            nmat = sc.nbx * sc.nby * nch;
        }
        uint32_t tch, tmat;
        const uint32_t exc = block_exclusive_scan<uint32_t>(nch, l1, &tch);
        const uint32_t exm = block_exclusive_scan<uint32_t>(nmat, l2, &tmat);
        if (s < ns) {
            SceneRun& r = b.runs[s];
            r.prim_off = beg;
            r.nprims = np;
            r.chunk_base = run_ch + exc;
            r.prim_base = (run_ch + exc) * CHUNK;
            r.nch = nch;
            r.mat_base = run_mat + exm;
        }
        run_ch += tch;
        run_mat += tmat;
    }
    if (threadIdx.x == 0) {
        FrameCounters* c = b.cnt;
        c->nchunks = run_ch;
        c->nprims = run_ch * CHUNK;
        c->mat_len = run_mat;
        uint32_t o = 0;
        if ((uint64_t)run_ch * CHUNK > b.prim_cap) o |= OVF_PRIMS;
        if (run_mat > b.mat_cap) o |= OVF_MATRIX;
        if (o) atomicOr(&c->overflow, o);
    }
    return make_uint2(run_ch, run_mat);
}
__global__ __launch_bounds__(1024) void k_scene_prep(JobBufs b) {
    __shared__ uint32_t l1[16], l2[16];
    (void)scene_prep(b, b.cnt->nprims_raw, l1, l2);
}

// The scan of the prims per triangle, k_scene_prep and k_pchunk_map by one
// workgroup, for jobs of at most SCAN_ONE_MAX (padded) triangles: one launch
// instead of three. The totals each step needs pass in registers (every
// thread holds them); the barriers order each step's global writes before
// the next step's reads.
__global__ __launch_bounds__(1024) void k_scene_prep_one(JobBufs b) {
    __shared__ uint32_t l1[16], l2[16];
    const bool bad = b.cnt->overflow != 0;  // from the earlier kernels
    const uint32_t total = scan_one<uint32_t>(b.nsub, scan_len(b.tri_cap, &b.head->ntri), b.suboff,
                                              &b.cnt->nprims_raw, l1);
    __syncthreads();
    const uint2 cm = scene_prep(b, total, l1, l2);
    // the prim slots and the coarse matrix fit (the overflow bits scene_prep sets)
    const bool fits = (uint64_t)cm.x * CHUNK <= b.prim_cap && cm.y <= b.mat_cap;
    __syncthreads();
    if (!bad && fits) pchunk_map(b, cm.x);
}

// ---------------------------------------------------------------------------
// k_setup: one thread per input triangle; writes its rasterizer triangles
// at the scene's prim slots in fan order.
// mod.rs:284-313 (divide), mod.rs:315-345 (viewport), mod.rs:187-222 (edge
// setup), bounding_box.rs:13-42 + mod.rs:347-361 (clamped pixel bbox).
// Besides the reference's values each prim carries derived, exactly
// computed helpers for the raster: its tie-rule bits and a rigorous lower
// bound on any depth it can interpolate.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool tie_rule(float nx, float ny) {
    return (nx > 0.0f) || (!(nx < 0.0f) && (ny < 0.0f));  // mod.rs:159-168
}

// A plane that is <= the depth the reference interpolates (mod.rs:226-236)
// at every sample the prim can cover (samples inside its pixel bbox):
//   exact Z(p) = z2 + B0(p)(z0-z2) + B1(p)(z1-z2), B0 = inv*E1(p), B1 = inv*E2(p),
// computed in f64, minus a slack that bounds the f32 rounding of the edge
// functions (|dE| <= 4u*M), of e*inv, of the clamps (covered samples have
// barycentrics within that error of [0,1]), of b2 = (1-b0)-b1 and of the
// final three products and two sums, and the f32 evaluation of the plane
// itself -- all with a further 2x margin. u = 2^-24.
__device__ void depth_plane_bound(PrimGeom& o, uint32_t x0, uint32_t x1, uint32_t y0, uint32_t y1) {
    const double u = 0x1p-24;
    const double inv = o.inv_area;
    const double z0 = o.z0, z1 = o.z1, z2 = o.z2;
    const double dz0 = z0 - z2, dz1 = z1 - z2;
    const double n1x = o.nx1, n1y = o.ny1, n2x = o.nx2, n2y = o.ny2;
    const double a1x = o.ax1, a1y = o.ay1, a2x = o.ax2, a2y = o.ay2;
    const double bx = x0, by = y0, bw = (double)x1 - x0, bh = (double)y1 - y0;
    const double px = inv * (n1x * dz0 + n2x * dz1);
    const double py = inv * (n1y * dz0 + n2y * dz1);
    const double e1 = n1x * (bx - a1x) + n1y * (by - a1y);
    const double e2 = n2x * (bx - a2x) + n2y * (by - a2y);
    const double zb = z2 + inv * e1 * dz0 + inv * e2 * dz1;
    const double m1 = fabs(n1x) * fmax(fabs(bx - a1x), fabs(bx + bw - a1x)) +
                      fabs(n1y) * fmax(fabs(by - a1y), fabs(by + bh - a1y));
    const double m2 = fabs(n2x) * fmax(fabs(bx - a2x), fabs(bx + bw - a2x)) +
                      fabs(n2y) * fmax(fabs(by - a2y), fabs(by + bh - a2y));
    const double d = 10.0 * u * fmax(m1, m2) * fabs(inv) + 2.0 * u;
    const double slack = (3.0 * d + 16.0 * u) * (fabs(z0) + fabs(z1) + fabs(z2));
    const double ev = 8.0 * u * (fabs(zb) + fabs(px) * bw + fabs(py) * bh) + 1e-30;
    const double lo = zb - 2.0 * (slack + ev);
    const bool ok = (x0 < x1) && (y0 < y1) && isfinite(lo) && isfinite(px) && isfinite(py) &&
                    isfinite(zb) && fabs(px) < 1e30 && fabs(py) < 1e30;
    o.pz0 = ok ? (float)lo : -FLT_MAX_F;
    o.pzx = ok ? (float)px : 0.0f;
    o.pzy = ok ? (float)py : 0.0f;
}

// True when no fragment of the prim can address a texel past the texture,
// whatever its size: all w > 0 and every uv component finite and at most
// 1 - 2^-19. Proof sketch: a fragment is shaded at a covered sample (edge
// values >= 0) or, when all 4 samples pass, at the pixel centre, whose edge
// values are > 0 (the centre is the mean of the 4 samples and each sample's
// edge value errs by at most 4u*M, M <= 1.42*|n|*32768 < |n|/16, while the
// samples' spread along any edge normal is >= 0.28*|n|). So the perspective
// weights f_k = e_k / w_k are >= 0, u + v <= 1 + 2^-22, u + v + w <= 1 + 2^-22
// and tu <= max(a) * (1 + 2^-21); then tu * (tw - 1) < tw - 1, every texel
// coordinate is <= tw - 1 (th - 1 for v) and the index x + y*tw < tw*th.
// Negative and NaN coordinates clamp to texel 0. Prims that fail the test,
// and all prims of a texture whose texel index does not fit the raster's
// 32-bit fast path, have their fragments checked exactly in the raster.
__device__ __forceinline__ bool uv_in_bounds(const ClipV* tv, const float* w) {
    const float lim = 1.0f - 0x1p-19f;
    bool ok = (w[0] > 0.0f) & (w[1] > 0.0f) & (w[2] > 0.0f);
    for (int k = 0; k < 3; ++k)
        for (int j = 4; j < 6; ++j) {
            const float a = tv[k].a[j];
            ok &= (fabsf(a) <= FLT_MAX_F) & (a <= lim);
        }
    return ok;
}

// True when the edge values at every sample position are finite and e *
// inv_area is finite at every covered sample: 0 < inv_area < inf and the
// points and normals finite and below 2^24 (a covered sample's edge value is
// at most about 2A).
__device__ __forceinline__ bool finite_bary(const PrimGeom& o) {
    bool ok = (o.inv_area > 0.0f) & (o.inv_area <= FLT_MAX_F);
    const float v[12] = {o.ax0, o.ay0, o.ax1, o.ay1, o.ax2, o.ay2, o.nx0, o.ny0, o.nx1, o.ny1, o.nx2, o.ny2};
    for (int k = 0; k < 12; ++k) ok &= fabsf(v[k]) <= 0x1p24f;
    return ok;
}

__device__ __forceinline__ void setup_prim(const ClipV* tv, uint32_t W, uint32_t H, uint32_t draw, uint32_t fs,
                                           bool tex_idx32, bool proofs, PrimGeom* g, float* __restrict__ attr,
                                           float* __restrict__ uvat, uint4* __restrict__ pbin) {
    const float zmin = 0.0f, zmax = 1.0f;
    float sx[3], sy[3], sz[3], w[3];
    for (int k = 0; k < 3; ++k) {
        const ClipV& v = tv[k];
        float nx = v.x / v.w, ny = v.y / v.w, nz = v.z / v.w;  // mod.rs:284-313
        sx[k] = (float)W * (nx + 1.0f) / 2.0f;                  // mod.rs:323
        sy[k] = (float)H * (1.0f - (ny + 1.0f) / 2.0f);         // mod.rs:325
        sz[k] = (nz + 1.0f) * 0.5f * (zmax - zmin) + zmin;      // mod.rs:328
        w[k] = v.w;
    }
    PrimGeom o;
    o.ax0 = sx[0]; o.ay0 = sy[0]; o.ax1 = sx[1]; o.ay1 = sy[1]; o.ax2 = sx[2]; o.ay2 = sy[2];
    // mod.rs:199-204: v_k = p_{k+1} - p_k, n_k = (-v_k.y, v_k.x)
    o.nx0 = -(sy[1] - sy[0]); o.ny0 = sx[1] - sx[0];
    o.nx1 = -(sy[2] - sy[1]); o.ny1 = sx[2] - sx[1];
    o.nx2 = -(sy[0] - sy[2]); o.ny2 = sx[0] - sx[2];
    // mod.rs:206 with mod.rs:15-21
    float v10x = sx[1] - sx[0], v10y = sy[1] - sy[0];
    float v20x = sx[2] - sx[0], v20y = sy[2] - sy[0];
    o.inv_area = 1.0f / (v10x * v20y - v20x * v10y);
    o.z0 = sz[0]; o.z1 = sz[1]; o.z2 = sz[2];
    o.w0 = w[0]; o.w1 = w[1]; o.w2 = w[2];
    // bounding_box.rs:15-32: fold from (MAX, MIN); floor/ceil; saturating cast
    float a0 = FLT_MAX_F, a1 = -FLT_MAX_F, a2 = FLT_MAX_F, a3 = -FLT_MAX_F;
    for (int k = 0; k < 3; ++k) {
        a0 = fminf(a0, sx[k]);
        a1 = fmaxf(a1, sx[k]);
        a2 = fminf(a2, sy[k]);
        a3 = fmaxf(a3, sy[k]);
    }
    // mod.rs:355-360 clamp to the viewport (saturating at 32 bits is
    // equivalent: any value >= 2^32 clamps to W/H either way)
    uint32_t x0 = min(sat_u32(floorf(a0)), W), x1 = min(sat_u32(ceilf(a1)), W);
    uint32_t y0 = min(sat_u32(floorf(a2)), H), y1 = min(sat_u32(ceilf(a3)), H);
    o.bbx = x0 | (x1 << 16);
    o.bby = y0 | (y1 << 16);
    o.flags = (tie_rule(o.nx0, o.ny0) ? 1u : 0u) | (tie_rule(o.nx1, o.ny1) ? 2u : 0u) |
              (tie_rule(o.nx2, o.ny2) ? 4u : 0u) | (fs << 8) | (draw << 16);
    if (fs == RZ_FS_TEXTURE && !(proofs && tex_idx32 && uv_in_bounds(tv, w))) o.flags |= PRIM_UV_CHECK;
    if (proofs && finite_bary(o)) o.flags |= PRIM_FINITE_BARY;
    // Depth of a covered sample is ((b0*z0 + b1*z1) + b2*z2) with clamped
    // b_i in [0, 1] and 1 - 2^-23 <= b0+b1+b2 <= 2; three roundings later it
    // is >= zlo below, so a sample can pass LESS only where zbuf > zlo.
    float zm = fminf(fminf(sz[0], sz[1]), sz[2]);
    o.zlo = zm >= 0.0f ? zm * (1.0f - 0x1p-18f) - 0x1p-126f : zm * (2.0f + 0x1p-17f) - 0x1p-126f;
    depth_plane_bound(o, x0, x1, y0, y1);
    o.pad0 = 0;
    o.pad1 = 0;
    *g = o;
    *pbin = make_uint4(0u, o.bbx, o.bby, __float_as_uint(o.zlo));
    // vector stores of only the attributes the shader reads: the 3 colours
    // (48-byte records) or the 3 uv pairs (24-byte records in their own array)
    if (fs == RZ_FS_COLOR) {
        float4* a4 = reinterpret_cast<float4*>(attr);
        a4[0] = make_float4(tv[0].a[0], tv[0].a[1], tv[0].a[2], tv[0].a[3]);
        a4[1] = make_float4(tv[1].a[0], tv[1].a[1], tv[1].a[2], tv[1].a[3]);
        a4[2] = make_float4(tv[2].a[0], tv[2].a[1], tv[2].a[2], tv[2].a[3]);
    } else if (fs == RZ_FS_TEXTURE) {
        float2* a2 = reinterpret_cast<float2*>(uvat);
        a2[0] = make_float2(tv[0].a[4], tv[0].a[5]);
        a2[1] = make_float2(tv[1].a[4], tv[1].a[5]);
        a2[2] = make_float2(tv[2].a[4], tv[2].a[5]);
    }
}

// The prims of a workgroup's 256 triangles are one contiguous range; when it
// fits SETUP_STAGE records they are assembled in LDS and written with
// contiguous 16-byte stores (whole cache lines) instead of one scattered
// 112-byte record per thread.
constexpr uint32_t SETUP_STAGE = 336;

// The read-only inputs are __restrict__ kernel arguments, so that the
// compiler may read the uniform ones (draw, scene and chunk records) with
// scalar loads and keep them across the prim loop despite the record stores.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_setup(
    JobBufs b, const DrawDev* __restrict__ draws, const SceneCfg* __restrict__ scenes,
    const SceneRun* __restrict__ runs, const uint32_t* __restrict__ tchunk_draw,
    const uint32_t* __restrict__ suboff, const uint32_t* __restrict__ nsub, const uint32_t* __restrict__ cls,
    const float* __restrict__ clip_out) {
    __shared__ PrimGeom stage[SETUP_STAGE];
    __shared__ uint32_t poff[256], pcls[256];  // per triangle of the chunk: first prim (chunk-relative), class
    const uint32_t NT = b.head->ntri;
    const uint32_t prim_cap = b.prim_cap;
    // the chunk table words of the workgroup's next chunk are loaded one
    // chunk ahead (NT is a multiple of 256, so t0 + 255 is in range)
    uint32_t n_d = 0, n_s0 = 0, n_end = 0, n_sub = 0, n_cls = 0;
    auto fetch = [&](uint32_t c) {
        if (c * 256u < NT) {
            const uint32_t c0 = c * 256u;
            n_d = tchunk_draw[c];
            n_s0 = suboff[c0];
            n_end = suboff[c0 + 255u] + nsub[c0 + 255u];
            n_sub = suboff[c0 + threadIdx.x];
            n_cls = cls[c0 + threadIdx.x];
        }
    };
    fetch(blockIdx.x);
    for (uint32_t blk = blockIdx.x; blk * 256u < NT; blk += gridDim.x) {
        const uint32_t t0 = blk * 256u;
        const uint32_t t = t0 + threadIdx.x;
        const uint32_t d = uniform(n_d), s0 = uniform(n_s0), s_end = uniform(n_end), my_sub = n_sub, my_cls = n_cls;
        fetch(blk + gridDim.x);
        const DrawDev& dd = draws[d];
        const uint32_t scene = dd.scene;
        const SceneCfg& sc = scenes[scene];
        const SceneRun& run = runs[scene];
        const uint32_t W = sc.W, H = sc.H, fs = dd.fs;
        // texel indices of this draw's texture fit 32-bit (24-bit factor) arithmetic
        const bool tex_idx32 = dd.tw < (1u << 24) && dd.th < (1u << 24) && (uint64_t)dd.tw * dd.th <= 0xFFFFFFFFull;
        const bool proofs = b.exact_paths == 0;
        // the chunk's prim range (its 256 triangles belong to draw d: draws are 256-aligned)
        const uint32_t blk_first = uniform(run.prim_base + (s0 - run.prim_off));
        const uint32_t nblk = s_end - s0;
        const uint32_t blk_end = blk_first + nblk;
        const bool staged = nblk <= SETUP_STAGE && blk_end <= prim_cap;
        poff[threadIdx.x] = my_sub - s0;
        pcls[threadIdx.x] = (t - dd.tri_base < dd.ntri) ? my_cls : 0u;
        __syncthreads();
        // one thread per prim of the chunk: an input triangle (class 1) or one
        // fan triangle of a clipped polygon (class 2 + slot), so clipped
        // triangles' fans spread over the workgroup
        bool ovf = false;
        for (uint32_t pp = threadIdx.x; pp < nblk; pp += 256u) {
            // the chunk's triangle owning prim pp: the last one starting at or before it
            uint32_t j = 0;
            for (uint32_t step = 128; step >= 1; step >>= 1)
                if (poff[j + step] <= pp) j += step;
            const uint32_t c = pcls[j], k = pp - poff[j];
            const uint32_t prim = blk_first + pp;
            ClipV tv[3];
            if (c == 1) {
                const uint32_t lt = t0 + j - dd.tri_base;
                for (int q = 0; q < 3; ++q) {
                    const uint32_t vi = dd.idx[(size_t)lt * 3 + q];
                    float4 p;
                    load_vertex(dd, vi, p, tv[q].a);
                    tv[q].x = p.x; tv[q].y = p.y; tv[q].z = p.z; tv[q].w = p.w;
                }
            } else {
                // clipping.rs:185-190: fan triangle k is (v0, v_{k+1}, v_{k+2})
                const uint32_t slot = min(c - 2, b.clip_cap - 1);
                const float* poly = clip_out + (size_t)slot * MAXV * 10;
                const uint32_t vid[3] = {0u, k + 1, k + 2};
                for (int q = 0; q < 3; ++q) {
                    const float* pv = poly + min(vid[q], MAXV - 1) * 10;
                    tv[q].x = pv[0]; tv[q].y = pv[1]; tv[q].z = pv[2]; tv[q].w = pv[3];
                    for (int i = 0; i < 6; ++i) tv[q].a[i] = pv[4 + i];
                }
            }
            if (prim < prim_cap)
                setup_prim(tv, W, H, d, fs, tex_idx32, proofs, staged ? stage + pp : b.geom + prim,
                           b.attrs + (size_t)prim * ATTR_STRIDE, b.uvat + (size_t)prim * UV_STRIDE, b.pbin + prim);
            else
                ovf = true;
        }
        __syncthreads();
        if (staged) {
            const uint4* src = reinterpret_cast<const uint4*>(stage);
            uint4* dst = reinterpret_cast<uint4*>(b.geom + blk_first);
            const uint32_t n16 = nblk * (uint32_t)(sizeof(PrimGeom) / sizeof(uint4));
            for (uint32_t i = threadIdx.x; i < n16; i += 256u) dst[i] = src[i];
        }
        __syncthreads();  // the stage and the chunk tables are reused by the next chunk
        if (__ballot(ovf) && lane_id() == 0) atomicOr(&b.cnt->overflow, OVF_PRIMS);
    }
}

// ---------------------------------------------------------------------------
// XCD-aware work split: workgroups are dispatched round-robin over the 8
// XCDs (workgroup i runs on XCD i % 8), each with its own L2. Binning chunk
// c writes the cells of count matrices and the list entries next to those
// of chunks c - 1 and c + 1, so neighbouring chunks share cache lines. The
// chunks are dealt to the XCDs in runs of BAND_RUN consecutive chunks, so
// that a line written by several chunks is mostly merged in one L2 instead
// of being written back in parts from several, while every XCD still gets
// chunks from all over the job (uneven chunk costs stay balanced).
// ChunkBand walks this workgroup's chunks of [0, n).
// ---------------------------------------------------------------------------
constexpr uint32_t BAND_RUN = 16;

struct ChunkBand {
    uint32_t j, step, x, n;
    bool plain;
    __device__ ChunkBand(uint32_t n_) : n(n_) {
        const uint32_t G = gridDim.x;
        plain = G < 8;  // too few workgroups for one per XCD: plain grid stride
        x = blockIdx.x & 7u;
        j = plain ? blockIdx.x : blockIdx.x >> 3;
        step = plain ? G : (G - x + 7u) / 8u;  // workgroups on XCD x
    }
    // the j-th chunk of XCD x: run j / BAND_RUN of that XCD, offset j % BAND_RUN
    __device__ uint32_t chunk() const {
        return plain ? j : ((j / BAND_RUN) * 8u + x) * BAND_RUN + j % BAND_RUN;
    }
    __device__ bool more() const {
        return plain ? j < n : ((j / BAND_RUN) * 8u + x) * BAND_RUN < n;
    }
};

// ---------------------------------------------------------------------------
// k_bin: ordered coarse binning, one workgroup per chunk of CHUNK prims of one
// scene (thread t holds slot t). With r = the scene's SceneRun and lc the
// chunk's index in the scene:
// COUNT:   counts[r.mat_base + bin * r.nch + lc] = prims of the chunk overlapping bin
// SCATTER: writes {prim, bbox, zlo} to list position offsets[same index]
//          + rank, rank = number of earlier prims of the chunk in that bin.
// The chunk's bin membership is a bit matrix in LDS, bits[w][bin] bit k =
// slot 32w + k overlaps bin: each thread sets the bits of its own bins, so the
// work follows the bins a prim covers, not the bins of the scene. A bin's
// count is the popcount of its column; a slot's rank in a bin is the popcount
// of the column's bits below it, which keeps each bin's list in submission
// order. The bins of prims covering more than WIDE_BINS bins are instead
// spread over the whole workgroup: their (prim, bin) pairs are numbered
// slot by slot (an exclusive scan of their bin counts) and thread t takes
// pairs t, t + 256, ... (a few screen-sized triangles would otherwise leave
// single threads looping over hundreds of bins).
// ---------------------------------------------------------------------------
constexpr uint32_t BIN_WORDS = CHUNK / 32;
constexpr uint32_t WIDE_BINS = 16;
static_assert(MAX_BINS == 256 && CHUNK == 256, "one thread per bin and per chunk slot");

template <bool SCATTER>
__global__ __launch_bounds__(256) void k_bin(JobBufs b, const uint32_t* __restrict__ pchunk_scene,
                                             const SceneCfg* __restrict__ scenes, const SceneRun* __restrict__ runs,
                                             const uint4* __restrict__ pbin, const uint32_t* __restrict__ bin_offsets) {
    __shared__ uint32_t bits[BIN_WORDS * MAX_BINS];
    __shared__ uint32_t wbase[SCATTER ? BIN_WORDS * MAX_BINS : 1];
    __shared__ uint32_t wpre[CHUNK];  // first wide (prim, bin) pair of each slot
    __shared__ uint32_t scan_lds[CHUNK / 64];
    if (uniform(b.cnt->overflow) != 0) return;
    const uint32_t NC = uniform(b.cnt->nchunks);
    const uint32_t t = threadIdx.x, w = t >> 5, below = (1u << (t & 31)) - 1u;
    bool ovf = false;
    for (ChunkBand cb_(NC); cb_.more(); cb_.j += cb_.step) {
        const uint32_t chunk = cb_.chunk();
        if (chunk >= NC) continue;  // past the end inside this XCD's last run
        const uint32_t s = uniform(pchunk_scene[chunk]);
        const SceneCfg& sc = scenes[s];
        const SceneRun& r = runs[s];
        const uint32_t cb = sc.cb, nbx = sc.nbx, nb = nbx * sc.nby;
        const uint32_t nch = r.nch;
        const size_t mb = (size_t)r.mat_base + (chunk - r.chunk_base) + (size_t)t * nch;
        const uint32_t p = chunk * CHUNK + t;
        // both loads unconditional (clamped to the scene's first prim and the
        // chunk's first bin, in bounds whenever the chunk exists) so that they
        // are in flight together; slots past the scene's prims are masked by pv
        const bool pv = p < r.prim_base + r.nprims;
        const uint4 pb = pbin[pv ? p : r.prim_base];
        uint32_t run = 0;
        if (SCATTER) run = bin_offsets[t < nb ? mb : mb - (size_t)t * nch];
        if (t < nb)
            for (uint32_t k = 0; k < BIN_WORDS; ++k) bits[k * MAX_BINS + t] = 0u;
        const uint32_t x0 = pb.y & 0xFFFFu, x1 = pb.y >> 16, y0 = pb.z & 0xFFFFu, y1 = pb.z >> 16;
        const bool ne = pv && (x0 < x1) && (y0 < y1);
        const uint32_t cx0 = x0 / cb, cx1 = ne ? (x1 - 1) / cb : 0u;
        const uint32_t cy0 = y0 / cb, cy1 = ne ? (y1 - 1) / cb : 0u;
        const uint32_t nbin = ne ? (cx1 - cx0 + 1u) * (cy1 - cy0 + 1u) : 0u;
        const bool wide = nbin > WIDE_BINS;
        uint32_t* row = bits + w * MAX_BINS;
        const bool any_wide = __syncthreads_or(wide) != 0;
        if (ne && !wide)
            for (uint32_t by = cy0; by <= cy1; ++by)
                for (uint32_t bx = cx0; bx <= cx1; ++bx) atomicOr(&row[by * nbx + bx], 1u << (t & 31));
        // wide pair q: slot s = the last with wpre[s] <= q (slots of no wide
        // pairs have empty ranges), bin q - wpre[s] of its rectangle, row-major
        uint32_t npairs = 0;
        auto wide_pair = [&](uint32_t q, uint32_t& ws, uint4& wpb, uint32_t& wbi) {
            uint32_t k = 0;
            for (uint32_t step = CHUNK / 2; step >= 1; step >>= 1)
                if (wpre[k + step] <= q) k += step;
            ws = k;
            wpb = pbin[chunk * CHUNK + k];
            const uint32_t wx0 = (wpb.y & 0xFFFFu) / cb, wx1 = ((wpb.y >> 16) - 1u) / cb;
            const uint32_t wy0 = (wpb.z & 0xFFFFu) / cb, wbw = wx1 - wx0 + 1u, i = q - wpre[k];
            wbi = (wy0 + i / wbw) * nbx + wx0 + i % wbw;
        };
        if (any_wide) {
            wpre[t] = block_exclusive_scan<uint32_t>(wide ? nbin : 0u, scan_lds, &npairs);
            __syncthreads();
            for (uint32_t q = t; q < npairs; q += CHUNK) {
                uint32_t ws, wbi;
                uint4 wpb;
                wide_pair(q, ws, wpb, wbi);
                atomicOr(&bits[(ws >> 5) * MAX_BINS + wbi], 1u << (ws & 31));
            }
        }
        __syncthreads();
        if (t < nb) {
            for (uint32_t k = 0; k < BIN_WORDS; ++k) {
                const uint32_t c = (uint32_t)__popc(bits[k * MAX_BINS + t]);
                if (SCATTER) wbase[k * MAX_BINS + t] = run;
                run += c;
            }
            if (!SCATTER) b.bin_counts[mb] = run;
        }
        if (SCATTER) {
            __syncthreads();
            const uint4 en = make_uint4(p, pb.y, pb.z, pb.w);
            const uint32_t* wb = wbase + w * MAX_BINS;
            if (ne && !wide)
                for (uint32_t by = cy0; by <= cy1; ++by)
                    for (uint32_t bx = cx0; bx <= cx1; ++bx) {
                        const uint32_t bi = by * nbx + bx;
                        const uint32_t pos = wb[bi] + (uint32_t)__popc(row[bi] & below);
                        if (pos < b.pair_cap) b.list[pos] = en;
                        else ovf = true;
                    }
            for (uint32_t q = t; q < npairs; q += CHUNK) {
                uint32_t ws, wbi;
                uint4 wpb;
                wide_pair(q, ws, wpb, wbi);
                const uint32_t wi = (ws >> 5) * MAX_BINS + wbi;
                const uint32_t pos = wbase[wi] + (uint32_t)__popc(bits[wi] & ((1u << (ws & 31)) - 1u));
                if (pos < b.pair_cap) b.list[pos] = make_uint4(chunk * CHUNK + ws, wpb.y, wpb.z, wpb.w);
                else ovf = true;
            }
        }
        __syncthreads();  // bits and wbase are reused by the next chunk
    }
    if (SCATTER && __ballot(ovf) && lane_id() == 0) atomicOr(&b.cnt->overflow, OVF_PAIRS);
}

// ---------------------------------------------------------------------------
// Fine binning: each coarse bin's list is split again into its 16x16 screen
// tiles (the raster workgroups), keeping submission order.
// k_fine_prep, one thread per global coarse bin g: its list range and
//   fpack[g] = (T * n) << 32 | n, n = chunks of CHUNK entries, T = tiles of the bin.
// The exclusive scan fpre of fpack gives each bin its first fine-binning
// chunk (low word) and the base of its fine count matrix (high word); the
// matrix holds, for local tile lt and chunk c, the entry hi(fpre[g]) + lt*n + c.
// Its exclusive scan gives every tile one contiguous, ordered list.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t coarse_ofs(const JobBufs& b, uint32_t i, uint32_t mat_len, uint32_t tot) {
    return i < mat_len ? b.bin_offsets[i] : tot;
}

__device__ __forceinline__ void fine_prep(const JobBufs& b) {
    const uint32_t NB = b.head->nbins, ns = b.head->nscenes;
    const bool bad = b.cnt->overflow != 0;
    const uint32_t mat_len = b.cnt->mat_len, tot = b.cnt->pairs_binned;
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < NB; g += gridDim.x * blockDim.x) {
        uint32_t beg = 0, end = 0, T = 1;
        if (!bad) {
            const uint32_t s = find_scene_bin(b.scenes, ns, g);
            const SceneCfg& sc = b.scenes[s];
            const SceneRun& r = b.runs[s];
            const uint32_t bi = g - sc.bin_base;
            const uint32_t i0 = r.mat_base + bi * r.nch;
            beg = coarse_ofs(b, i0, mat_len, tot);
            end = coarse_ofs(b, i0 + r.nch, mat_len, tot);
            if (end < beg) end = beg;
            T = (sc.cb >> sc.fts) * (sc.cb >> sc.fts);
        }
        const uint32_t n = (end - beg + CHUNK - 1) / CHUNK;
        b.bin_rng[g] = make_uint2(beg, end);
        b.fpack[g] = ((unsigned long long)(T * n) << 32) | n;
    }
}
__global__ __launch_bounds__(256) void k_fine_prep(JobBufs b) { fine_prep(b); }

// A fine-list entry is 8 bytes: {prim, key}, key = the prim's depth lower
// bound zlo rounded toward -inf to a multiple of 2^QB ulps, with the low QB
// bits holding which 8x8 quadrants of the fine tile the prim's bbox overlaps:
// QB = 4 (bit 2*qy + qx of the 16x16 raster tile) for 8- and 16-pixel fine
// tiles, QB = 16 (bit 4*qy + qx of the 32x32 fine tile) for 32-pixel ones.
// The raster only compares the key against depths it stores, to skip prims,
// so a rounded-down bound skips no prim it would have kept.
__device__ __forceinline__ uint32_t quad_bits_of(uint32_t fts) { return fts == 5u ? 16u : 4u; }
__device__ __forceinline__ uint32_t zlo_key(uint32_t z, uint32_t qb) {
    // negative floats: round the magnitude up (a NaN stays a NaN, or -inf for
    // the lowest negative NaNs: either way no stored depth is above it)
    const uint32_t low = (1u << qb) - 1u;
    return (z & 0x80000000u) ? ((min(z, ~low) + low) & ~low) : (z & ~low);
}

// quadrants of the 16x16 tile at (ox, oy) that the bbox [x0,x1) x [y0,y1)
// overlaps, given that it overlaps the tile
__device__ __forceinline__ uint32_t quad_mask(uint32_t x0, uint32_t x1, uint32_t y0, uint32_t y1, uint32_t ox,
                                              uint32_t oy) {
    const uint32_t cx = (x0 < ox + 8 ? 1u : 0u) | (x1 > ox + 8 ? 2u : 0u);
    return ((y0 < oy + 8) ? cx : 0u) | ((y1 > oy + 8) ? cx << 2 : 0u);
}

// The entry's quadrant bits for a fine tile of side 1 << fts at (ox, oy),
// given that the bbox [x0,x1) x [y0,y1) overlaps it. 16: the quadrants of
// that raster tile; 8: the fine tile is one quadrant of a raster tile; 32:
// the 8x8 quadrants of the fine tile (a 4x4 grid) that the bbox overlaps.
__device__ __forceinline__ uint32_t fine_quad_bits(uint32_t fts, uint32_t x0, uint32_t x1, uint32_t y0, uint32_t y1,
                                                   uint32_t ox, uint32_t oy) {
    if (fts == 4u) return quad_mask(x0, x1, y0, y1, ox, oy);
    if (fts == 3u) return 1u << (((oy >> 2) & 2u) | ((ox >> 3) & 1u));
    const uint32_t qx0 = (max(x0, ox) - ox) >> 3, qx1 = (min(x1, ox + 32u) - 1u - ox) >> 3;
    const uint32_t qy0 = (max(y0, oy) - oy) >> 3, qy1 = (min(y1, oy + 32u) - 1u - oy) >> 3;
    const uint32_t row = (2u << qx1) - (1u << qx0);   // columns qx0..qx1 of one row (< 16)
    const uint32_t rows = (2u << qy1) - (1u << qy0);  // rows qy0..qy1
    // one bit per selected row at nibble 4*row, times the row's columns
    const uint32_t spread = (rows & 1u) | ((rows & 2u) << 3) | ((rows & 4u) << 6) | ((rows & 8u) << 9);
    return spread * row;
}

// k_fchunk_map, one thread per global coarse bin: writes {bin, scene} for
// each of the bin's fine-binning chunks, so a chunk's workgroup finds its bin
// with one load instead of two binary searches.
__device__ __forceinline__ void fchunk_map(const JobBufs& b) {
    if (b.cnt->overflow != 0) return;
    const uint32_t NB = b.head->nbins, ns = b.head->nscenes, cap = b.fchunk_cap;
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < NB; g += gridDim.x * blockDim.x) {
        const uint32_t first = lo32(b.fpre[g]), n = lo32(b.fpack[g]);
        if (n == 0) continue;
        const uint32_t s = find_scene_bin(b.scenes, ns, g);
        for (uint32_t c = 0; c < n && first + c < cap; ++c) b.fchunk_rec[first + c] = make_uint2(g, s);
    }
}
__global__ __launch_bounds__(256) void k_fchunk_map(JobBufs b) { fchunk_map(b); }

// k_fine_prep, the u64 scan of fpack and k_fchunk_map by one workgroup of
// 256 threads, for jobs of at most FINE_PREP_ONE_MAX coarse bins (one launch
// instead of three; the workgroup barriers order each step's global writes
// before the next step's reads). A frame's bin buffers hold at least 1280
// bins; with many more the three parallel launches win: a 4096-bin batch
// lost 0.03 ms in one workgroup. 256 threads, not 1024: beside a resident
// raster service, whose waves hold part of every SIMD's registers, a
// 1024-thread workgroup of this kernel found no CU with room for its 16
// waves until the service exited.
constexpr uint32_t FINE_PREP_ONE_MAX = 2048;
__global__ __launch_bounds__(256) void k_fine_prep_one(JobBufs b) {
    __shared__ unsigned long long lds[16];
    fine_prep(b);
    __syncthreads();
    (void)scan_one<unsigned long long>(b.fpack, scan_len(b.bins_cap, &b.head->nbins), b.fpre,
                                       reinterpret_cast<unsigned long long*>(&b.cnt->fine_chunks), lds);
    __syncthreads();
    fchunk_map(b);
}

template <bool SCATTER>
__global__ __launch_bounds__(256) void k_fine_bin(JobBufs b, const uint2* __restrict__ fchunk_rec,
                                                  const SceneCfg* __restrict__ scenes,
                                                  const unsigned long long* __restrict__ fpre,
                                                  const uint2* __restrict__ bin_rng, const uint4* __restrict__ list,
                                                  const uint32_t* __restrict__ fine_off) {
    // one workgroup per fine-binning chunk of CHUNK list entries of one coarse
    // bin, thread t holding entry t; the same LDS bit matrix as k_bin, with
    // the bin's local tiles in place of the scene's bins
    __shared__ uint32_t bits[BIN_WORDS * MAX_BINS];
    __shared__ uint32_t wbase[SCATTER ? BIN_WORDS * MAX_BINS : 1];
    if (uniform(b.cnt->overflow) != 0) return;
    const uint32_t NCH = uniform(min(b.cnt->fine_chunks, b.fchunk_cap));
    if (uniform(b.cnt->fmat_len) > b.fmat_cap) {
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(&b.cnt->overflow, OVF_FINE);
        return;
    }
    const uint32_t t = threadIdx.x, w = t >> 5, below = (1u << (t & 31)) - 1u;
    bool ovf = false;
    for (ChunkBand cb_(NCH); cb_.more(); cb_.j += cb_.step) {
        const uint32_t gc = cb_.chunk();
        if (gc >= NCH) continue;  // past the end inside this XCD's last run
        const uint2 fr = fchunk_rec[gc];
        const uint32_t g = uniform(fr.x), s = uniform(fr.y);
        const SceneCfg& sc = scenes[s];
        const uint32_t fts = sc.fts, cb = sc.cb, TS = cb >> fts, T = TS * TS;
        if (T > MAX_BINS) continue;  // the host never configures more tiles per coarse bin
        const unsigned long long fp = fpre[g];
        const uint2 rng = bin_rng[g];
        const uint32_t c = gc - lo32(fp);
        const uint32_t nchb = (rng.y - rng.x + CHUNK - 1) / CHUNK;
        const uint32_t bi = g - sc.bin_base;
        const uint32_t binx0 = (bi % sc.nbx) * cb, biny0 = (bi / sc.nbx) * cb;
        const size_t mt = (size_t)hi32(fp) + c + (size_t)t * nchb;
        const uint32_t e = rng.x + c * CHUNK + t;
        const bool valid = e < rng.y;
        // both loads unconditional (clamped to the chunk's first entry and
        // first tile, in bounds whenever the chunk exists) so that they are in
        // flight together; every use of en below is guarded by valid
        const uint4 en = list[valid ? e : rng.x];
        uint32_t run = 0;
        if (SCATTER) run = fine_off[t < T ? mt : (size_t)hi32(fp) + c];
        if (t < T)
            for (uint32_t k = 0; k < BIN_WORDS; ++k) bits[k * MAX_BINS + t] = 0u;
        const uint32_t x0 = en.y & 0xFFFFu, x1 = en.y >> 16, y0 = en.z & 0xFFFFu, y1 = en.z >> 16;
        // local tile range inside this coarse bin (the entry overlaps it)
        const uint32_t lx0 = (max(x0, binx0) - binx0) >> fts;
        const uint32_t lx1 = valid ? (min(x1, binx0 + cb) - 1 - binx0) >> fts : 0u;
        const uint32_t ly0 = (max(y0, biny0) - biny0) >> fts;
        const uint32_t ly1 = valid ? (min(y1, biny0 + cb) - 1 - biny0) >> fts : 0u;
        uint32_t* row = bits + w * MAX_BINS;
        __syncthreads();
        if (valid)
            for (uint32_t ty = ly0; ty <= ly1; ++ty)
                for (uint32_t tx = lx0; tx <= lx1; ++tx) atomicOr(&row[ty * TS + tx], 1u << (t & 31));
        __syncthreads();
        if (t < T) {
            for (uint32_t k = 0; k < BIN_WORDS; ++k) {
                const uint32_t n = (uint32_t)__popc(bits[k * MAX_BINS + t]);
                if (SCATTER) wbase[k * MAX_BINS + t] = run;
                run += n;
            }
            if (!SCATTER) b.fine_counts[mt] = run;
        }
        if (SCATTER) {
            __syncthreads();
            const uint32_t* wb = wbase + w * MAX_BINS;
            const uint32_t key = zlo_key(en.w, quad_bits_of(fts));
            if (valid)
                for (uint32_t ty = ly0; ty <= ly1; ++ty)
                    for (uint32_t tx = lx0; tx <= lx1; ++tx) {
                        const uint32_t lt = ty * TS + tx;
                        const uint32_t pos = wb[lt] + (uint32_t)__popc(row[lt] & below);
                        const uint32_t qm =
                            fine_quad_bits(fts, x0, x1, y0, y1, binx0 + (tx << fts), biny0 + (ty << fts));
                        if (pos < b.fine_cap) b.flist[pos] = make_uint2(en.x, key | qm);
                        else ovf = true;
                    }
        }
        __syncthreads();  // bits and wbase are reused by the next chunk
    }
    if (SCATTER && __ballot(ovf) && lane_id() == 0) atomicOr(&b.cnt->overflow, OVF_FINE);
}

// ---------------------------------------------------------------------------
// Rank of raster tile (tx, ty) of a scene of ntx x nty tiles in blocked
// order: 4x4-tile blocks in row-major order, tiles row-major inside a block;
// the blocks of the last column and row are narrower or shorter. A bijection
// of [0, ntx * nty): block row by holds rows * ntx tiles, every block left of
// column bx is 4 wide.
__device__ __forceinline__ uint32_t blocked_rank(uint32_t tx, uint32_t ty, uint32_t ntx, uint32_t nty) {
    const uint32_t bx = tx / 4u, by = ty / 4u;
    const uint32_t rows = min(4u, nty - by * 4u), cols = min(4u, ntx - bx * 4u);
    return by * 4u * ntx + bx * 4u * rows + (ty % 4u) * cols + tx % 4u;
}

// k_tile_prep: one thread per raster tile and 8x8 quadrant: the tile's scene
// and the range of the fine list that quadrant's wave walks -- the list of
// the fine tile holding the quadrant (the tile itself for 16x16 fine tiles,
// the quadrant for 8x8 ones, the enclosing 32x32 tile for 32x32 ones): the
// fine matrix entry of (its coarse bin, its local tile, chunk 0) and of the
// next tile, so the raster starts with one load instead of a chain of
// searches and lookups. Ranges are clamped to the list buffer: a bad range
// can never read past it. Records are stored in raster order: the tile of
// scene rank r (blocked_rank) at raster_base + r, each quadrant's record
// {scene, list begin, list end, first pixel x | y << 16}, so that
// consecutive raster slots are 4x4 squares of the screen. The host
// gives the scenes of a batch their slot ranges densest first (triangles per
// pixel), so the longest tiles start early rather than trailing at the end
// of the launch.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_tile_prep(JobBufs b) {
    const uint32_t nt = min(b.head->ntiles, b.tile_cap), ns = b.head->nscenes;
    if (b.cnt->overflow != 0) return;
    const uint32_t fmat_len = b.cnt->fmat_len, ftot = b.cnt->fine_total;
    for (uint32_t gq = blockIdx.x * blockDim.x + threadIdx.x; gq < 4u * nt; gq += gridDim.x * blockDim.x) {
        const uint32_t gt = gq >> 2, q = gq & 3u;
        const uint32_t s = find_scene_tile(b.scenes, ns, gt);
        const SceneCfg& sc = b.scenes[s];
        const uint32_t ntx = (sc.W + 15) / 16, nty = (sc.H + 15) / 16, tile = gt - sc.tile_base;
        const uint32_t tx = tile % ntx, ty = tile / ntx;
        const uint32_t px = tx * 16 + (q & 1u) * 8, py = ty * 16 + (q >> 1) * 8;
        const uint32_t cbs = sc.cb, fts = sc.fts, TS = cbs >> fts;
        const uint32_t cbx = px / cbs, cby = py / cbs;
        const uint32_t g = sc.bin_base + cby * sc.nbx + cbx;
        const uint32_t lt = ((py - cby * cbs) >> fts) * TS + ((px - cbx * cbs) >> fts);
        const uint2 rng = b.bin_rng[g];
        const uint32_t nchb = (rng.y - rng.x + CHUNK - 1) / CHUNK;
        const size_t s_idx = (size_t)hi32(b.fpre[g]) + (size_t)lt * nchb, e_idx = s_idx + nchb;
        const uint32_t lend = min(e_idx < fmat_len ? b.fine_off[e_idx] : ftot, b.fine_cap);
        const uint32_t lbeg = min(s_idx < fmat_len ? b.fine_off[s_idx] : ftot, lend);
        const uint32_t slot = sc.raster_base + blocked_rank(tx, ty, ntx, nty);
        if (slot < b.tile_cap) b.tile_rec[4u * slot + q] = make_uint4(s, lbeg, lend, px | py << 16);
    }
}

// ---------------------------------------------------------------------------
// k_raster: one workgroup per 16x16 screen tile at a time, one wave per 8x8
// quadrant, one lane per pixel. The pixel's 4 MSAA colours and depths live in
// the lane's registers for the whole frame; the wave walks its tile's fine
// list in submission order and applies mod.rs:443-473 to each rasterizer
// triangle that can touch the quadrant, then resolves (buffers.rs:111-125)
// and writes the ARGB word once.
//
// Tiles of all scenes form one index space of raster slots (each scene's
// tiles in 4x4 blocks, k_tile_prep). With a grid of one single-wave
// workgroup per quadrant (k_raster_q) the order is static; with a smaller
// (persistent) grid of four-wave workgroups (k_raster) the workgroups pull
// tiles from 8 queues, one band of tiles per XCD, and move on to the other
// bands when theirs is empty.
//
// Work that provably cannot change a sample is skipped (exact culls):
//  - list entries whose bbox misses the quadrant;
//  - hierarchical depth: the quadrant keeps qzmax >= max of its stored
//    depths (refreshed after each batch of list entries), each pixel its own
//    maximum; a prim whose depth lower bound is >= that maximum fails the
//    strict LESS test on every sample there;
//  - edge reject: if an edge function is negative, with a margin that covers
//    the rounding of the reference's evaluation, at the quadrant corner where
//    it is largest, no sample of the quadrant is inside that edge;
//  - depth plane: a lower bound of the prim's depth over the quadrant (and
//    per pixel) that is not below the stored depths.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float sample_x(int s) {  // mod.rs:108-114
    return s == 0 ? 0.625f : (s == 1 ? 0.875f : (s == 2 ? 0.375f : 0.125f));
}
__device__ __forceinline__ float sample_y(int s) {
    return s == 0 ? 0.125f : (s == 1 ? 0.625f : (s == 2 ? 0.875f : 0.375f));
}

// texture.rs:47-63: linear texel index x + y*w; past the buffer -> last texel + flag
__device__ __forceinline__ uint32_t read_texel(const uint32_t* __restrict__ tex, uint32_t tw, uint32_t th,
                                               uint32_t x, uint32_t y, bool& oob) {
    uint64_t n = (uint64_t)tw * th;
    uint64_t idx;
    if (y >= th || (uint64_t)x >= n - (uint64_t)y * tw) {
        oob = true;
        idx = n - 1;
    } else {
        idx = (uint64_t)x + (uint64_t)y * tw;
    }
    return tex[idx];
}

struct RGBA {
    float r, g, b, a;
};

// color.rs:24-27: channel byte b as b / 255.0, correctly rounded (the
// compiler folds the divisions with IEEE round-to-nearest, as the device's
// division rounds). The raster copies it into LDS at start.
struct alignas(16) ByteToFloat {
    float v[256];
};
constexpr ByteToFloat make_byte_to_float() {
    ByteToFloat t{};
    for (int i = 0; i < 256; ++i) t.v[i] = (float)i / 255.0f;
    return t;
}
__device__ constexpr ByteToFloat kByteToFloat = make_byte_to_float();

// texture.rs:65-83 (channels via color.rs:22-29, byte / 255.0 from a table)
// checked: bounds-checked texel reads (texture.rs:47-63); otherwise the prim's
// coordinates are proven in bounds and the index x + y*tw fits 32 bits.
__device__ __forceinline__ RGBA texture_sample(const uint32_t* __restrict__ tex, uint32_t tw, uint32_t th,
                                               float u, float v, const float* lut, bool checked, bool& oob) {
    float x = u * (float)(tw - 1);
    float y = v * (float)(th - 1);
    uint32_t xf0 = sat_u32(floorf(x)), xc = sat_u32(ceilf(x));
    uint32_t yf0 = sat_u32(floorf(y)), yc = sat_u32(ceilf(y));
    uint32_t tl, tr, bl, br;
    if (checked) {
        tl = read_texel(tex, tw, th, xf0, yf0, oob);
        tr = read_texel(tex, tw, th, xc, yf0, oob);
        bl = read_texel(tex, tw, th, xf0, yc, oob);
        br = read_texel(tex, tw, th, xc, yc, oob);
    } else {
        const uint32_t r0 = __umul24(yf0, tw), r1 = __umul24(yc, tw);
        tl = tex[r0 + xf0];
        tr = tex[r0 + xc];
        bl = tex[r1 + xf0];
        br = tex[r1 + xc];
    }
    float x_f = x - truncf(x);
    float y_f = y - truncf(y);
    float ix = 1.0f - x_f, iy = 1.0f - y_f;
    RGBA o;
#define RZ_BILERP(field, sh)                                                       \
    {                                                                              \
        float a_ = lut[(tl >> sh) & 255u], b_ = lut[(tr >> sh) & 255u];            \
        float c_ = lut[(bl >> sh) & 255u], d_ = lut[(br >> sh) & 255u];            \
        float y0_ = a_ * ix + b_ * x_f;                                            \
        float y1_ = c_ * ix + d_ * x_f;                                            \
        o.field = y0_ * iy + y1_ * y_f;                                            \
    }
    RZ_BILERP(r, 0)
    RZ_BILERP(g, 8)
    RZ_BILERP(b, 16)
    RZ_BILERP(a, 24)
#undef RZ_BILERP
    return o;
}

// color.rs:15-20
__device__ __forceinline__ uint32_t to_argb(float r, float g, float b, float a) {
    return (sat_u32(a * 255.0f) << 24) | (sat_u32(r * 255.0f) << 16) | (sat_u32(g * 255.0f) << 8) |
           sat_u32(b * 255.0f);
}

// vector.rs:17-23 on (n, p - a) computes (0.0 + nx*(x-ax)) + ny*(y-ay). The
// leading 0.0 + only turns a -0.0 product into +0.0, which can flip the sign
// of a zero edge value; no output depends on that sign: the inside test
// treats +-0 alike, a zero barycentric of either sign gives the same depth
// (a sum that ends in +0 or a positive term) and the same attributes, and
// to_argb / texel addressing map +-0 to the same value. So it is left out.
__device__ __forceinline__ float edge(float nx, float ny, float ax, float ay, float x, float y) {
    return nx * (x - ax) + ny * (y - ay);
}

// false only if every sample position in [xl, xh] x [yl, yh] evaluates the
// edge (reference formula and rounding) strictly below zero
__device__ __forceinline__ bool edge_may_cover(float nx, float ny, float ax, float ay, float xl, float xh,
                                               float yl, float yh) {
    float xs = nx > 0.0f ? xh : xl;
    float ys = ny > 0.0f ? yh : yl;
    float dx = xs - ax, dy = ys - ay;
    float e = nx * dx + ny * dy;
    float mag = fabsf(nx) * (fabsf(dx) + 16.0f) + fabsf(ny) * (fabsf(dy) + 16.0f);
    return !(e < -(mag * 0x1p-16f + 0x1p-126f));
}

__device__ __forceinline__ float wave_max(float v) {
    for (int o = 32; o >= 1; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

// The largest value over a full wave's 64 lanes, as a wave-uniform value:
// DPP steps (no LDS round trips) -- pairs and quads within each row of 16
// lanes, the mirrored half-rows and rows, then row 0's result broadcast into
// row 1 and rows 0-1's into rows 2-3 -- leave the maximum in lane 63. Every
// lane must be active.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_max_step(float x) {
    const int t = __builtin_amdgcn_update_dpp(__float_as_int(x), __float_as_int(x), CTRL, ROW_MASK, 0xF, false);
    return fmaxf(x, __int_as_float(t));
}
__device__ __forceinline__ float wave_max_uniform(float v) {
    v = dpp_max_step<0xB1, 0xF>(v);   // quad_perm [1, 0, 3, 2]
    v = dpp_max_step<0x4E, 0xF>(v);   // quad_perm [2, 3, 0, 1]
    v = dpp_max_step<0x141, 0xF>(v);  // row_half_mirror
    v = dpp_max_step<0x140, 0xF>(v);  // row_mirror
    v = dpp_max_step<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
    v = dpp_max_step<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

// raster slots are dealt to the XCDs in runs of this many consecutive
// slots (one 4x4 block of tiles; k_raster_q deals their quadrants)
constexpr uint32_t RASTER_RUN = 16;

// band x of the persistent queue: tiles [band_lo(x), band_lo(x + 1))
__device__ __forceinline__ uint32_t band_lo(uint32_t nt, uint32_t x) {
    const uint32_t q = nt / 8u, r = nt % 8u;
    return x * q + min(x, r);
}

// mod.rs:70-99: perspective-correct barycentrics of a fragment, at the
// pixel centre when all four samples passed the depth test, else at the
// first passing sample
__device__ __forceinline__ void frag_bary(const PrimGeom& g, float fx, float fy, uint32_t pass, float& u, float& v,
                                          float& w) {
    float ix = fx + 0.5f, iy = fy + 0.5f;
    if (pass != 0xFu) {
        const int k = __builtin_ctz(pass);
        ix = fx + sample_x(k);
        iy = fy + sample_y(k);
    }
    const float ef0 = edge(g.nx0, g.ny0, g.ax0, g.ay0, ix, iy);
    const float ef1 = edge(g.nx1, g.ny1, g.ax1, g.ay1, ix, iy);
    const float ef2 = edge(g.nx2, g.ny2, g.ax2, g.ay2, ix, iy);
    const float f_u = ef1 / g.w0;
    const float f_v = ef2 / g.w1;
    const float f_w = ef0 / g.w2;
    const float sum = f_u + f_v + f_w;
    u = clamp_bary(f_u / sum);
    v = clamp_bary(f_v / sum);
    w = clamp_bary(1.0f - u - v);
}

// texture.rs:65-83 coordinates of (tu, tv): true if one of the four texel
// reads would fall past the texture (the read_texel condition)
__device__ __forceinline__ bool texel_oob(uint32_t tw, uint32_t th, float tu, float tv) {
    const float x = tu * (float)(tw - 1);
    const float y = tv * (float)(th - 1);
    const uint32_t xf0 = sat_u32(floorf(x)), xc = sat_u32(ceilf(x));
    const uint32_t yf0 = sat_u32(floorf(y)), yc = sat_u32(ceilf(y));
    const uint64_t n = (uint64_t)tw * th;
    auto past = [&](uint32_t X, uint32_t Y) { return Y >= th || (uint64_t)X >= n - (uint64_t)Y * tw; };
    return past(xf0, yf0) || past(xc, yf0) || past(xf0, yc) || past(xc, yc);
}

// FragCoords.depths[0] (mod.rs:225-262): the prim's depth at sample 0 when
// the prim covers that sample, else 0.0
__device__ __forceinline__ float sample0_depth(const PrimGeom& g, float fx, float fy) {
    const float sx = fx + sample_x(0), sy = fy + sample_y(0);
    const float e0 = edge(g.nx0, g.ny0, g.ax0, g.ay0, sx, sy);
    const float e1 = edge(g.nx1, g.ny1, g.ax1, g.ay1, sx, sy);
    const float e2 = edge(g.nx2, g.ny2, g.ax2, g.ay2, sx, sy);
    const bool t0 = g.flags & 1u, t1 = g.flags & 2u, t2 = g.flags & 4u;
    const bool in = ((e0 > 0.0f) | (t0 & !(e0 < 0.0f))) & ((e1 > 0.0f) | (t1 & !(e1 < 0.0f))) &
                    ((e2 > 0.0f) | (t2 & !(e2 < 0.0f)));
    const float b0 = clamp_bary(e1 * g.inv_area);
    const float b1 = clamp_bary(e2 * g.inv_area);
    const float b2 = clamp_bary(1.0f - b0 - b1);
    const float z = b0 * g.z0 + b1 * g.z1 + b2 * g.z2;
    return in ? z : 0.0f;
}

// main.rs:67-77 fragment shaders, then mod.rs:380-397 (to ARGB): the colour
// one fragment writes to its passing samples
__device__ __forceinline__ uint32_t shade_fragment(const PrimGeom& g, const float* __restrict__ at,
                                                   const float* __restrict__ uv, const DrawDev* __restrict__ draws,
                                                   float fx, float fy, uint32_t pass, const float* lut, bool& oob) {
    float u, v, w;
    frag_bary(g, fx, fy, pass, u, v, w);
    const uint32_t fs = (g.flags >> 8) & 3u;
    float cr, cg, cbl, ca;
    if (fs == RZ_FS_COLOR) {
        cr = at[0] * u + at[4] * v + at[8] * w;
        cg = at[1] * u + at[5] * v + at[9] * w;
        cbl = at[2] * u + at[6] * v + at[10] * w;
        ca = at[3] * u + at[7] * v + at[11] * w;
    } else if (fs == RZ_FS_TEXTURE) {
        const DrawDev& dd = draws[g.flags >> 16];
        const float tu = uv[0] * u + uv[2] * v + uv[4] * w;
        const float tv = uv[1] * u + uv[3] * v + uv[5] * w;
        const RGBA c = texture_sample(dd.tex, dd.tw, dd.th, tu, tv, lut, (g.flags & PRIM_UV_CHECK) != 0, oob);
        cr = c.r; cg = c.g; cbl = c.b; ca = c.a;
    } else {  // Debug: Color::grayscale(depths[0]) (color.rs:65-72)
        const float d = sample0_depth(g, fx, fy);
        cr = d; cg = d; cbl = d; ca = 1.0f;
    }
    return to_argb(cr, cg, cbl, ca);
}

// Read-only inputs of the raster, passed as separate __restrict__ kernel
// arguments: the no-alias guarantee is what lets the compiler fetch prim
// records with scalar loads while the kernel writes its outputs.
struct RasterIn {
    const PrimGeom* __restrict__ geom;
    const float* __restrict__ attrs;
    const float* __restrict__ uvat;
    const uint2* __restrict__ flist;
    const uint4* __restrict__ tile_rec;
    const SceneCfg* __restrict__ scenes;
    const DrawDev* __restrict__ draws;
};

// The words of a prim record that the list walk reads: the leading 17 of
// PrimGeom (flags included), loaded with one 16-word and one 1-word scalar
// load and held in scalar registers while the next record loads.
struct WalkRec {
    float ax0, ay0, ax1, ay1, ax2, ay2;
    float nx0, ny0, nx1, ny1, nx2, ny2;
    float inv_area;
    float z0, z1, z2;
    uint32_t flags;
};
static_assert(offsetof(PrimGeom, flags) == 16 * 4, "PrimGeom: the walk's words lead the record");

__device__ __forceinline__ WalkRec walk_rec(const PrimGeom* __restrict__ g) {
    return WalkRec{g->ax0, g->ay0, g->ax1, g->ay1, g->ax2, g->ay2, g->nx0,   g->ny0, g->nx1,
                   g->ny1, g->nx2, g->ny2, g->inv_area, g->z0, g->z1, g->z2, g->flags};
}

// LDS of the deferred shading stage: per wave (R waves), its shading tasks
// {prim, owner lane | pass mask << 6}, their colours, and its task count (a
// wave's 64 pixels hold at most 256 distinct prims)
template <uint32_t R>
struct ShadeLds {
    uint2 task[R][256];
    uint32_t argb[R][256];
    uint32_t count[R];
};

// R = 4: the workgroup's four waves take the tile's four quadrants and
// share the shading tasks; R = 1: the workgroup is one wave and takes
// quadrant quad of the tile alone.
template <bool STATS, uint32_t R>
__device__ __forceinline__ void raster_tile(SceneRun* runs, const RasterIn& in, uint32_t slot, uint32_t quad,
                                            const float* lut, ShadeLds<R>& sh,
                                            uint32_t& st_entries, uint32_t& st_cand, uint32_t& st_hits,
                                            uint32_t& st_cov, uint32_t& st_shaded) {
    // the tile's scene and this wave's list range (k_tile_prep)
    // wave: the quadrant; row: this wave's rows of the shading LDS
    const uint32_t wave = R == 1 ? quad : uniform(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const uint32_t row = R == 1 ? 0u : wave;
    const uint4 tr = in.tile_rec[4u * slot + wave];
    const uint32_t s = uniform(tr.x);
    const SceneCfg& sc = in.scenes[s];
    const uint32_t W = sc.W, H = sc.H;
    const float clear_depth = sc.clear_depth;
    // the quadrant's first pixel (k_tile_prep), and its tile's
    const uint32_t qx0 = uniform(tr.w) & 0xFFFFu, qy0 = uniform(tr.w) >> 16;
    const uint32_t tx = qx0 >> 4, ty = qy0 >> 4;
    const uint32_t px = qx0 + (lane & 7), py = qy0 + (lane >> 3);
    const float fx = (float)px, fy = (float)py;
    const bool onscreen = px < W && py < H;
    const float qxl = (float)qx0 + 0.125f, qxh = (float)qx0 + 7.875f;
    const float qyl = (float)qy0 + 0.125f, qyh = (float)qy0 + 7.875f;

    // per sample: stored depth (registers), and in LDS the prim that wrote
    // it last (NO_PRIM: cleared) with that prim's passing samples when it
    // wrote, which fix where it is shaded. The winners are written during the
    // list walk and read only after it, so they live in this wave's
    // shading-task row (sample k of the lane at k * 64 + lane) rather than
    // in registers, which keeps the walk's register count down.
    float zb0 = clear_depth, zb1 = clear_depth, zb2 = clear_depth, zb3 = clear_depth;
    uint2* const win = sh.task[row];
#pragma unroll
    for (int k = 0; k < 4; ++k) win[k * 64 + lane] = make_uint2(NO_PRIM, 0u);
    bool oob = false;
    // qzmax: largest depth stored in the quadrant's on-screen samples (all
    // cleared to the same depth): the maximum of clear_depth over the
    // on-screen lanes and -FLT_MAX over the others
    const uint64_t on = __ballot(onscreen);
    float qzmax = on == 0 ? -FLT_MAX_F : (on == ~0ull ? clear_depth : fmaxf(clear_depth, -FLT_MAX_F));
    // lane_zbmax: largest depth stored in this pixel's 4 samples
    float lane_zbmax = onscreen ? clear_depth : -FLT_MAX_F;

    const uint32_t lbeg = uniform(tr.y), lend = uniform(tr.z);

    const PrimGeom* __restrict__ geom = in.geom;
    // The list is scanned 64 entries at a time with the entries' own bits
    // (the quadrant's bit, the depth key against qzmax); the entries that
    // pass queue up in list order (in this wave's shading-colour row, which
    // the shading stage fills only after the walk), and the record-based
    // culls and the walk run on batches of queued candidates (the rest at
    // the end of the list), so a long list with sparse candidates pays for
    // them once per batch of candidates rather than once per 64 entries.
    // A batch runs once QUEUE_RUN candidates are queued (up to 64 at once):
    // waiting for fewer than 64 lets each walk tighten qzmax for the next
    // entries sooner, which pays on short lists (batched scenes 1% faster at
    // 32 than at 64); long lists, with sparse candidates, do better with
    // fuller batches (the 1080p soup's raster 2.5% faster at 48 than at 32).
    const uint32_t QUEUE_RUN = uniform(tr.z - tr.y) > 1024u ? 48u : 32u;
    uint32_t* const cq = sh.argb[row];
    uint32_t qn = 0;  // queued candidates (at most QUEUE_RUN + 63)
    // this quadrant's bit in its fine tile's entries (its position in the
    // 32x32 fine tile, or in the 16x16 raster tile), and the key's bits
    const uint32_t qbit = sc.fts == 5u ? ((((qy0 >> 3) & 3u) << 2) | ((qx0 >> 3) & 3u)) : wave;
    const uint32_t kmask = sc.fts == 5u ? ~0xFFFFu : ~0xFu;
    // each batch's entries are loaded one batch ahead
    uint2 en_next = make_uint2(0u, 0u);
    if (lbeg + lane < lend) en_next = in.flist[lbeg + lane];
    for (uint32_t base = lbeg;; base += 64) {
        const bool scan = base < lend;
        if (scan) {
            const uint32_t i = base + lane;
            const uint2 en = en_next;
            if (i + 64 < lend) en_next = in.flist[i + 64];
            if (STATS) st_entries += min(64u, lend - base);
            const bool c0 = i < lend && ((en.y >> qbit) & 1u) && __uint_as_float(en.y & kmask) < qzmax;
            const uint64_t mc = __ballot(c0);
            if (c0) cq[qn + mbcnt(mc)] = en.x;
            qn += (uint32_t)__popcll(mc);
            if (qn < QUEUE_RUN) continue;
        } else if (qn == 0) {
            break;
        }
        // the next batch of queued candidates; the rest (< 64) move to the front
        const uint32_t take = min(qn, 64u), left = qn - take;
        __builtin_amdgcn_wave_barrier();
        const uint32_t myprim = lane < take ? cq[lane] : 0u;
        if (left != 0) {
            const uint32_t r = lane < left ? cq[take + lane] : 0u;
            __builtin_amdgcn_wave_barrier();
            if (lane < left) cq[lane] = r;
        }
        qn = left;
        float lane_zlb = 0.0f;
        uint64_t lane_box = 0;  // quadrant pixels (bit 8 * y + x) inside the prim's bbox
        bool cand = lane < take;
        if (cand) {
            const float4* gp = reinterpret_cast<const float4*>(geom + myprim);
            const float4 g0 = gp[0], g1 = gp[1], g2 = gp[2];  // a0 a1 | a2 n0 | n1 n2
            const float4 g5 = gp[5], g6 = gp[6];              // bbx bby zlo pz0 | pzx pzy
            // depth lower bound over the quadrant's samples inside the bbox
            const uint32_t bbx = __float_as_uint(g5.x), bby = __float_as_uint(g5.y);
            const float fbx0 = (float)(bbx & 0xFFFFu), fby0 = (float)(bby & 0xFFFFu);
            const float xl = fmaxf(qxl, fbx0 + 0.125f), xh = fminf(qxh, (float)(bbx >> 16) - 0.125f);
            const float yl = fmaxf(qyl, fby0 + 0.125f), yh = fminf(qyh, (float)(bby >> 16) - 0.125f);
            const float zq = g5.w + g6.x * ((g6.x >= 0.0f ? xl : xh) - fbx0) + g6.y * ((g6.y >= 0.0f ? yl : yh) - fby0);
            lane_zlb = fmaxf(zq, g5.z);
            {
                const uint32_t cx0 = min(max(bbx & 0xFFFFu, qx0) - qx0, 8u), cx1 = min(max(bbx >> 16, qx0) - qx0, 8u);
                const uint32_t cy0 = min(max(bby & 0xFFFFu, qy0) - qy0, 8u), cy1 = min(max(bby >> 16, qy0) - qy0, 8u);
                const uint64_t row = ((1ull << cx1) - (1ull << cx0)) * 0x0101010101010101ull;
                const uint64_t rows = (cy1 >= 8u ? ~0ull : (1ull << (8u * cy1)) - 1u) &
                                      (cy0 >= 8u ? 0ull : ~((1ull << (8u * cy0)) - 1u));
                lane_box = row & rows;
            }
            cand = edge_may_cover(g1.z, g1.w, g0.x, g0.y, qxl, qxh, qyl, qyh) &&
                   edge_may_cover(g2.x, g2.y, g0.z, g0.w, qxl, qxh, qyl, qyh) &&
                   edge_may_cover(g2.z, g2.w, g1.x, g1.y, qxl, qxh, qyl, qyh) && !(lane_zlb >= qzmax);
        }
        uint64_t m = __ballot(cand);
        if (STATS) st_cand += (uint32_t)__popcll(m);
        if (m == 0) continue;
        // the next candidate (list order) that some lane can still take --
        // pixel inside its bbox and a stored depth above its bound -- is
        // found from the candidates' registers; the others are skipped
        // before their record is fetched
        uint64_t nbox = 0;
        float nzlb = 0.0f;
        WalkRec nxt;
        uint32_t nxt_prim = 0;
        auto next_live = [&]() -> bool {
            while (m != 0) {
                const uint32_t k = (uint32_t)__ffsll((unsigned long long)m) - 1;
                m &= m - 1;
                nbox = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(lane_box >> 32), (int)k) << 32) |
                       (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)lane_box, (int)k);
                nzlb = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(lane_zlb), (int)k));
                if ((__ballot(!(nzlb >= lane_zbmax)) & nbox) != 0) {
                    const uint32_t nprim = (uint32_t)__builtin_amdgcn_readlane((int)myprim, (int)k);
                    nxt = walk_rec(geom + nprim);  // start loading its record now
                    nxt_prim = nprim;
                    return true;
                }
            }
            return false;
        };
        if (!next_live()) continue;
        bool dirty = false;
        for (;;) {
            const WalkRec gm = nxt;
            const uint32_t prim = nxt_prim;
            const float zlb = nzlb;  // the prim's depth lower bound over the quadrant
            const uint64_t box = nbox;
            // this record is waited for before the next one is asked for:
            // scalar loads return out of order, so a wait for this record
            // inside the walk would also wait for the next one
            asm volatile("" ::"s"(gm.ax0), "s"(gm.flags) : "memory");
            // pick the next live candidate now, so its record loads while this
            // prim is processed (chosen with the stored depths before this
            // prim's writes: it may turn out to have no live lane)
            const bool more = next_live();
            bool wrote = false;
            // live lanes: pixel inside the bbox, and some stored depth of the
            // pixel above the bound (else no covered sample can pass LESS)
            const bool live = ((box >> lane) & 1u) && !(zlb >= lane_zbmax);
            if (__ballot(live) != 0) {
                // mod.rs:134-146 + 148-170: coverage of the 4 samples
                // the tie-rule test (e > 0) | (t & !(e < 0)) is a class test of
                // e: +denormal, +normal, +inf without the rule; also NaN and
                // both zeros with it (class bits: 0 sNaN, 1 qNaN, 5 -0, 6 +0,
                // 7 +denormal, 8 +normal, 9 +inf)
                const int cl0 = (gm.flags & 1u) ? 0x3E3 : 0x380, cl1 = (gm.flags & 2u) ? 0x3E3 : 0x380,
                          cl2 = (gm.flags & 4u) ? 0x3E3 : 0x380;
                float e1s[4], e2s[4];
                uint32_t cov = 0;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const float sx = fx + sample_x(k), sy = fy + sample_y(k);
                    const float e0 = edge(gm.nx0, gm.ny0, gm.ax0, gm.ay0, sx, sy);
                    const float e1 = edge(gm.nx1, gm.ny1, gm.ax1, gm.ay1, sx, sy);
                    const float e2 = edge(gm.nx2, gm.ny2, gm.ax2, gm.ay2, sx, sy);
                    const bool in = __builtin_amdgcn_classf(e0, cl0) & __builtin_amdgcn_classf(e1, cl1) &
                                    __builtin_amdgcn_classf(e2, cl2);
                    cov |= (in ? 1u : 0u) << k;
                    e1s[k] = e1;
                    e2s[k] = e2;
                }
                if (!live) cov = 0;
                if (STATS) {
                    st_hits += 1;
                    st_cov += (uint32_t)__popc(cov);  // this lane's covered samples
                }
                if (__ballot(cov != 0) != 0) {
                    // mod.rs:225-253: linear depth at covered samples; mod.rs:363-378: strict LESS
                    float sd[4];
                    uint32_t pass = 0;
                    const float zbuf[4] = {zb0, zb1, zb2, zb3};
                    if (gm.flags & PRIM_FINITE_BARY) {
                        // covered samples of this prim have finite barycentric
                        // arguments, where med3(x, 0, 1) is f32::clamp (a zero's
                        // sign may differ, which no depth or colour depends on)
#pragma unroll
                        for (int k = 0; k < 4; ++k) {
                            const float b0 = __builtin_amdgcn_fmed3f(e1s[k] * gm.inv_area, 0.0f, 1.0f);
                            const float b1 = __builtin_amdgcn_fmed3f(e2s[k] * gm.inv_area, 0.0f, 1.0f);
                            const float b2 = __builtin_amdgcn_fmed3f(1.0f - b0 - b1, 0.0f, 1.0f);
                            const float z = b0 * gm.z0 + b1 * gm.z1 + b2 * gm.z2;
                            sd[k] = z;  // written only where the sample passes
                            pass |= (((cov >> k) & 1u) && z < zbuf[k]) ? (1u << k) : 0u;
                        }
                    } else {
#pragma unroll
                        for (int k = 0; k < 4; ++k) {
                            // opaque copies: keeps the compiler from sharing these
                            // products with the other branch, where each folds into
                            // its clamp
                            float e1 = e1s[k], e2 = e2s[k];
                            asm volatile("" : "+v"(e1), "+v"(e2));
                            const float b0 = clamp_bary(e1 * gm.inv_area);
                            const float b1 = clamp_bary(e2 * gm.inv_area);
                            const float b2 = clamp_bary(1.0f - b0 - b1);
                            const float z = b0 * gm.z0 + b1 * gm.z1 + b2 * gm.z2;
                            sd[k] = z;  // written only where the sample passes
                            pass |= (((cov >> k) & 1u) && z < zbuf[k]) ? (1u << k) : 0u;
                        }
                    }
                    if (pass != 0) {
                        if (gm.flags & PRIM_UV_CHECK) {
                            // texture coordinates not proven in bounds: check this
                            // fragment's texel reads exactly (the flag covers every
                            // fragment, also ones overwritten later)
                            float u, v, w;
                            frag_bary(geom[prim], fx, fy, pass, u, v, w);
                            const float* at = in.uvat + (size_t)prim * UV_STRIDE;
                            const DrawDev& dd = in.draws[gm.flags >> 16];
                            oob |= texel_oob(dd.tw, dd.th, at[0] * u + at[2] * v + at[4] * w,
                                             at[1] * u + at[3] * v + at[5] * w);
                        }
                        // mod.rs:380-397 writes the fragment's colour and depth to its
                        // passing samples; the colour is computed after the list, once
                        // per surviving (prim, pixel), from the same inputs
                        const uint2 wv = make_uint2(prim, pass);
                        if (pass & 1u) { win[lane] = wv; zb0 = sd[0]; }
                        if (pass & 2u) { win[64 + lane] = wv; zb1 = sd[1]; }
                        if (pass & 4u) { win[128 + lane] = wv; zb2 = sd[2]; }
                        if (pass & 8u) { win[192 + lane] = wv; zb3 = sd[3]; }
                        lane_zbmax = fmaxf(fmaxf(zb0, zb1), fmaxf(zb2, zb3));
                        wrote = true;
                    }
                }
            }
            dirty |= wrote;
            if (!more) break;
        }
        // stored depths only fall, so a stale qzmax is still an upper bound;
        // it is refreshed once per batch of entries
        if (__ballot(dirty)) {
            qzmax = wave_max_uniform(lane_zbmax);
        }
    }
    {
        // the pixel's coordinates are recomputed here and below rather than
        // held across the walk and the shading (registers are the raster's
        // occupancy limit)
        uint32_t l2 = lane;
        asm volatile("" : "+v"(l2));
        const uint32_t px2 = qx0 + (l2 & 7), py2 = qy0 + (l2 >> 3);
        if (px2 < W && py2 < H && sc.msaa_depth) {
            const float4 d4 = make_float4(zb0, zb1, zb2, zb3);
            *reinterpret_cast<float4*>(sc.msaa_depth + ((size_t)py2 * W + px2) * 4) = d4;
        }
    }
    // deferred shading: one task per distinct surviving prim of a pixel (its
    // first sample heads it); the workgroup's tasks are shaded 64 at a time
    // by all four waves, then each pixel gathers its samples' colours. The
    // winners are read back before the task row they share is rewritten.
    const uint2 w0 = win[lane], w1 = win[64 + lane], w2 = win[128 + lane], w3 = win[192 + lane];
    const uint32_t pr0 = w0.x, pr1 = w1.x, pr2 = w2.x, pr3 = w3.x;
    const uint32_t pm = w0.y | w1.y << 4 | w2.y << 8 | w3.y << 12;
    const bool h1 = pr1 != pr0, h2 = (pr2 != pr0) & (pr2 != pr1), h3 = (pr3 != pr0) & (pr3 != pr1) & (pr3 != pr2);
    const bool v0 = pr0 != NO_PRIM, v1 = pr1 != NO_PRIM, v2 = pr2 != NO_PRIM, v3 = pr3 != NO_PRIM;
    const bool hd0 = v0, hd1 = v1 & h1, hd2 = v2 & h2, hd3 = v3 & h3;
    const uint32_t ntask = (uint32_t)hd0 + (uint32_t)hd1 + (uint32_t)hd2 + (uint32_t)hd3;
    // exclusive prefix of ntask (<= 4) over the wave: three bit planes
    const uint64_t np0 = __ballot(ntask & 1u), np1 = __ballot(ntask & 2u), np2 = __ballot(ntask & 4u);
    const uint32_t s0 = mbcnt(np0) + 2u * mbcnt(np1) + 4u * mbcnt(np2);
    const uint32_t s1 = s0 + (uint32_t)hd0, s2 = s1 + (uint32_t)hd1, s3 = s2 + (uint32_t)hd2;
    uint2* const wtask = sh.task[row];
    if (hd0) wtask[s0] = make_uint2(pr0, lane | (pm & 0xFu) << 6);
    if (hd1) wtask[s1] = make_uint2(pr1, lane | ((pm >> 4) & 0xFu) << 6);
    if (hd2) wtask[s2] = make_uint2(pr2, lane | ((pm >> 8) & 0xFu) << 6);
    if (hd3) wtask[s3] = make_uint2(pr3, lane | ((pm >> 12) & 0xFu) << 6);
    const uint32_t own = uniform((uint32_t)__popcll(np0) + 2u * (uint32_t)__popcll(np1) + 4u * (uint32_t)__popcll(np2));
    if (R > 1 && lane == 0) sh.count[row] = own;
    // each sample's task: that of its prim's first sample (8-bit slots)
    uint32_t slots = s0 | (h1 ? s1 : s0) << 8 | (h2 ? s2 : (pr2 == pr0 ? s0 : s1)) << 16 |
                           (h3 ? s3 : (pr3 == pr0 ? s0 : (pr3 == pr1 ? s1 : s2))) << 24;
    uint32_t vmask = (uint32_t)v0 | (uint32_t)v1 << 1 | (uint32_t)v2 << 2 | (uint32_t)v3 << 3;
    // formed here, so that their inputs are dead during the shading loop
    asm volatile("" : "+v"(slots), "+v"(vmask));
    if (R == 1) {
        // one wave: its own tasks, 64 at a time (LDS accesses of one wave
        // stay in order; the fences keep the compiler from moving them)
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        for (uint32_t g = lane; g < own; g += 64u) {
            const uint2 t = sh.task[0][g];
            const uint32_t P = t.x, ol = t.y & 63u, pass = t.y >> 6;
            const float ofx = (float)(qx0 + (ol & 7u)), ofy = (float)(qy0 + (ol >> 3));
            const PrimGeom gp = geom[P];
            sh.argb[0][g] = shade_fragment(gp, in.attrs + (size_t)P * ATTR_STRIDE, in.uvat + (size_t)P * UV_STRIDE,
                                           in.draws, ofx, ofy, pass, lut, oob);
            if (STATS) st_shaded += 1;  // shading tasks run by this lane
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    } else {
        __syncthreads();
        const uint32_t p1 = uniform(sh.count[0]), p2 = p1 + uniform(sh.count[1]);
        const uint32_t p3 = p2 + uniform(sh.count[2]), tot = p3 + uniform(sh.count[3]);
        for (uint32_t gb = wave * 64u; gb < tot; gb += 256u) {
            const uint32_t g = gb + lane;
            if (g < tot) {
                // task g of the workgroup: wave r's task g - (its first index)
                const uint32_t r = (uint32_t)(g >= p1) + (uint32_t)(g >= p2) + (uint32_t)(g >= p3);
                const uint32_t i = g - (r == 0 ? 0u : (r == 1 ? p1 : (r == 2 ? p2 : p3)));
                const uint2 t = sh.task[r][i];
                const uint32_t P = t.x, ol = t.y & 63u, pass = t.y >> 6;
                // the owner's pixel
                const float ofx = (float)(tx * 16 + (r & 1u) * 8 + (ol & 7u));
                const float ofy = (float)(ty * 16 + (r >> 1) * 8 + (ol >> 3));
                const PrimGeom gp = geom[P];
                sh.argb[r][i] =
                    shade_fragment(gp, in.attrs + (size_t)P * ATTR_STRIDE, in.uvat + (size_t)P * UV_STRIDE, in.draws,
                                   ofx, ofy, pass, lut, oob);
                if (STATS) st_shaded += 1;  // shading tasks run by this lane
            }
        }
        __syncthreads();
    }
    const uint32_t* const wargb = sh.argb[row];
    const uint32_t cb0 = (vmask & 1u) ? wargb[slots & 0xFFu] : CLEAR_COLOR;
    const uint32_t cb1 = (vmask & 2u) ? wargb[(slots >> 8) & 0xFFu] : CLEAR_COLOR;
    const uint32_t cb2 = (vmask & 4u) ? wargb[(slots >> 16) & 0xFFu] : CLEAR_COLOR;
    const uint32_t cb3 = (vmask & 8u) ? wargb[slots >> 24] : CLEAR_COLOR;
    if (__ballot(oob) && lane == 0) atomicOr(&runs[s].flags, RZ_FLAG_TEXEL_OOB);
    uint32_t l3 = lane;
    asm volatile("" : "+v"(l3));
    const uint32_t px3 = qx0 + (l3 & 7), py3 = qy0 + (l3 >> 3);
    if (px3 < W && py3 < H) {
        // buffers.rs:111-125
        uint32_t rs = ((cb0 & 0x00FF0000u) >> 16) + ((cb1 & 0x00FF0000u) >> 16) + ((cb2 & 0x00FF0000u) >> 16) +
                      ((cb3 & 0x00FF0000u) >> 16);
        uint32_t gs = ((cb0 & 0x0000FF00u) >> 8) + ((cb1 & 0x0000FF00u) >> 8) + ((cb2 & 0x0000FF00u) >> 8) +
                      ((cb3 & 0x0000FF00u) >> 8);
        uint32_t bs = (cb0 & 0xFFu) + (cb1 & 0xFFu) + (cb2 & 0xFFu) + (cb3 & 0xFFu);
        size_t o = (size_t)py3 * W + px3;
        sc.out[o] = (0xFFu << 24) | (rs / 4u) << 16 | (gs / 4u) << 8 | (bs / 4u);
        if (sc.msaa_color) {
            uint4 c4 = make_uint4(cb0, cb1, cb2, cb3);
            *reinterpret_cast<uint4*>(sc.msaa_color + o * 4) = c4;
        }
    }
}

// Persistent raster (raster_per_cu > 0): raster_per_cu workgroups of four
// waves per CU, each wave one quadrant of the tile its workgroup holds;
// contiguous bands of tiles become 8 queues, drained XCD-local first. Held
// to 64 VGPRs (8 waves per SIMD), it keeps 40 bytes of its loop state per
// lane in scratch, and at 8 workgroups per CU runs a quarter faster than
// unconstrained (6 waves).
template <bool STATS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(STATS ? 1 : 8, 8))) void k_raster(
    JobBufs b, const PrimGeom* __restrict__ geom, const float* __restrict__ attrs, const float* __restrict__ uvat,
    const uint2* __restrict__ flist,
    const uint4* __restrict__ tile_rec, const SceneCfg* __restrict__ scenes, const DrawDev* __restrict__ draws) {
    const RasterIn in{geom, attrs, uvat, flist, tile_rec, scenes, draws};
    __shared__ float lut[256];
    __shared__ uint32_t next_tile;
    __shared__ ShadeLds<4> sh;
    lut[threadIdx.x] = kByteToFloat.v[threadIdx.x];
    __syncthreads();
    if (uniform(b.cnt->overflow) != 0) return;
    uint32_t st_entries = 0, st_cand = 0, st_hits = 0, st_cov = 0, st_shaded = 0;
    const uint32_t nt = b.head->ntiles;
    const uint32_t xcd = blockIdx.x & 7u;
    for (;;) {
        if (threadIdx.x == 0) {
            uint32_t got = NO_TILE;
            for (uint32_t k = 0; k < 8 && got == NO_TILE; ++k) {
                const uint32_t x = (xcd + k) & 7u;
                const uint32_t lo = band_lo(nt, x), n = band_lo(nt, x + 1) - lo;
                if (n == 0) continue;
                const uint32_t i = atomicAdd(&b.cnt->q[x], 1u);
                if (i < n) got = lo + i;
            }
            next_tile = got;
        }
        __syncthreads();
        const uint32_t tile = next_tile;
        __syncthreads();
        if (tile == NO_TILE) break;
        raster_tile<STATS, 4>(b.runs, in, tile, 0u, lut, sh, st_entries, st_cand, st_hits, st_cov, st_shaded);
    }
    if (STATS) {  // the last two counts are per lane
        for (int o = 32; o >= 1; o >>= 1) {
            st_cov += __shfl_xor(st_cov, o);
            st_shaded += __shfl_xor(st_shaded, o);
        }
    }
    if (STATS) {
        // the workgroup's sums, one plain store each (k_rstat_reduce adds them
        // up; an atomic per wave on five shared counters serializes)
        __shared__ unsigned long long wsum[4][5];
        const uint32_t w = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0) {
            wsum[w][0] = st_entries;
            wsum[w][1] = st_cand;
            wsum[w][2] = st_hits;
            wsum[w][3] = st_cov;
            wsum[w][4] = st_shaded;
        }
        __syncthreads();
        if (threadIdx.x < 5) {
            const unsigned long long v =
                wsum[0][threadIdx.x] + wsum[1][threadIdx.x] + wsum[2][threadIdx.x] + wsum[3][threadIdx.x];
            if (blockIdx.x < b.rstat_cap) b.rstat_parts[(size_t)blockIdx.x * 5 + threadIdx.x] = v;
            else atomicAdd(&b.cnt->rstat[threadIdx.x], v);
        }
    }
}

// The raster (raster_per_cu = 0): one workgroup of one wave per 8x8
// quadrant. The quadrants of a tile are independent (each walks its own list
// range and shades its own pixels), so a wave that finishes early frees its
// slot at once instead of waiting at a barrier for the tile's slowest
// quadrant. Quadrant slots (4 per raster slot, in the same order) are dealt
// to the XCDs in runs of 4 * RASTER_RUN, a 4x4 square of tiles: workgroups
// go round-robin to the 8 XCDs (workgroup 8k + x takes the k-th quadrant of
// XCD x's runs), so each XCD renders neighbouring tiles (the prims they
// share stay in its L2) while every XCD gets work from all over the job at
// the same time -- the dispatcher hands out workgroups in order, so an XCD
// that lags behind (a band of dense tiles next to bands of empty sky) would
// hold up all of them. Any grid of at least raster_grid_q(nt) workgroups
// covers every quadrant; extra ones exit. Held to 64 VGPRs (8 waves per
// SIMD): the extra waves hide more of the walk's load latency than anything
// the limit costs.
template <bool STATS>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(STATS ? 1 : 8, 8))) void k_raster_q(
    JobBufs b, const PrimGeom* __restrict__ geom, const float* __restrict__ attrs, const float* __restrict__ uvat,
    const uint2* __restrict__ flist,
    const uint4* __restrict__ tile_rec, const SceneCfg* __restrict__ scenes, const DrawDev* __restrict__ draws) {
    const RasterIn in{geom, attrs, uvat, flist, tile_rec, scenes, draws};
    __shared__ __align__(16) float lut[256];
    __shared__ ShadeLds<1> sh;
    reinterpret_cast<float4*>(lut)[threadIdx.x] = reinterpret_cast<const float4*>(kByteToFloat.v)[threadIdx.x];
    __syncthreads();
    if (uniform(b.cnt->overflow) != 0) return;
    constexpr uint32_t QRUN = 4u * RASTER_RUN;
    const uint32_t nq = 4u * b.head->ntiles;
    const uint32_t k = blockIdx.x >> 3, xcd = blockIdx.x & 7u;
    const uint32_t qs = ((k / QRUN) * 8u + xcd) * QRUN + k % QRUN;
    uint32_t st_entries = 0, st_cand = 0, st_hits = 0, st_cov = 0, st_shaded = 0;
    if (qs < nq)
        raster_tile<STATS, 1>(b.runs, in, qs >> 2, qs & 3u, lut, sh, st_entries, st_cand, st_hits, st_cov, st_shaded);
    if (STATS) {  // every workgroup writes its counts, zeros past the last quadrant
        for (int o = 32; o >= 1; o >>= 1) {  // the last two counts are per lane
            st_cov += __shfl_xor(st_cov, o);
            st_shaded += __shfl_xor(st_shaded, o);
        }
        if (threadIdx.x < 5) {
            const unsigned long long v = threadIdx.x == 0 ? st_entries
                                         : threadIdx.x == 1 ? st_cand
                                         : threadIdx.x == 2 ? st_hits
                                         : threadIdx.x == 3 ? st_cov
                                                            : st_shaded;
            if (blockIdx.x < b.rstat_cap) b.rstat_parts[(size_t)blockIdx.x * 5 + threadIdx.x] = v;
            else atomicAdd(&b.cnt->rstat[threadIdx.x], v);
        }
    }
}

static uint32_t raster_grid_q(uint32_t nt) {
    const uint32_t runs = (4u * nt + 4u * RASTER_RUN - 1) / (4u * RASTER_RUN);
    return std::max<uint32_t>(1u, (runs + 7) / 8 * 8 * 4u * RASTER_RUN);
}

// stats variant: the raster workgroups' counts (nwg workgroups) into the counters
__global__ __launch_bounds__(256) void k_rstat_reduce(JobBufs b, uint32_t nwg) {
    __shared__ unsigned long long part[4][5];
    if (b.cnt->overflow != 0) return;
    const uint32_t n = min(nwg, b.rstat_cap);
    unsigned long long s[5] = {0, 0, 0, 0, 0};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        for (int k = 0; k < 5; ++k) s[k] += b.rstat_parts[(size_t)i * 5 + k];
    for (int k = 0; k < 5; ++k)
        for (int o = 32; o >= 1; o >>= 1) s[k] += __shfl_xor(s[k], o);
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 5; ++k) part[threadIdx.x >> 6][k] = s[k];
    __syncthreads();
    if (threadIdx.x < 5)
        atomicAdd(&b.cnt->rstat[threadIdx.x], part[0][threadIdx.x] + part[1][threadIdx.x] + part[2][threadIdx.x] +
                                                   part[3][threadIdx.x]);
}

// largest index of a device index buffer (device-resident mesh upload check)
__global__ __launch_bounds__(256) void k_max_index(const uint32_t* __restrict__ idx, uint64_t n,
                                                   uint32_t* __restrict__ out) {
    uint32_t m = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        m = max(m, idx[i]);
    for (int o = 32; o >= 1; o >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, o));
    if (lane_id() == 0) atomicMax(out, m);
}

// copies a block of words passed by value (a job table built on the host)
// into device memory, so a job can be set up without a host-to-device copy
__global__ __launch_bounds__(256) void k_write_words(WordBlock blk, uint32_t n, uint32_t* __restrict__ dst) {
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) dst[i] = blk.w[i];
}

__global__ __launch_bounds__(64) void k_fill_u32(uint32_t* __restrict__ p, uint32_t n, uint32_t v) {
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = v;
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
void launch_max_index(hipStream_t s, const uint32_t* idx, uint64_t n, uint32_t* out) {
    if (n == 0) return;
    uint64_t blocks = (n + 255) / 256;
    k_max_index<<<(uint32_t)(blocks < 1024 ? blocks : 1024), 256, 0, s>>>(idx, n, out);
}

void launch_write_words(hipStream_t s, const WordBlock& blk, uint32_t n, uint32_t* dst) {
    if (n == 0) return;
    k_write_words<<<1, 256, 0, s>>>(blk, n < WORD_BLOCK ? n : WORD_BLOCK, dst);
}

void launch_fill_u32(hipStream_t s, uint32_t* p, uint32_t n, uint32_t v) {
    if (n == 0) return;
    k_fill_u32<<<1, 64, 0, s>>>(p, n, v);
}

static uint32_t at_least1(uint32_t v) { return v ? v : 1u; }

// ---------------------------------------------------------------------------
// Raster service (see SvcEntry). k_svc_post ends a posting context's front
// end: unless the job overflowed a buffer (the host then grows and replays
// it), it stores the job's buffers in the context's entry and then, after a
// release fence at device scope, the 8 band claim words that make the job
// visible. A job without tiles is finished at once.
// ---------------------------------------------------------------------------
__global__ void k_svc_post(JobBufs b, SvcEntry* e, SvcHost* host, SvcCtl* ctl, uint32_t index) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (b.cnt->overflow != 0) return;
    const uint32_t nt = b.head->ntiles;
    const uint32_t gen = b.head->svc_gen & 0xFFFFu;
    if (nt == 0) {
        for (uint32_t s = 0; s < b.head->nscenes; ++s) b.svc_flags[s] = b.runs[s].flags;
        __threadfence_system();
        __hip_atomic_store(&host->done[index], gen, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        return;
    }
    e->job = SvcJob{nt, b.geom, b.attrs, b.uvat, b.flist, b.tile_rec, b.scenes, b.draws, b.runs, b.svc_flags,
                    b.head->nscenes};
    e->finished = 0;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    for (uint32_t x = 0; x < 8; ++x)
        __hip_atomic_store(&e->band[x].claim,
                           (unsigned long long)gen << 48 |
                               (unsigned long long)(band_lo(nt, x + 1) - band_lo(nt, x)) << SVC_TILE_BITS,
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_fetch_add(&ctl->posted, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// a pointer of a posted job: an atomic (vector) load, marked as global
// memory so that the loads through it are global rather than flat loads
template <class T, class P>
__device__ __forceinline__ const T* svc_ptr(P* field) {
    const unsigned long long v =
        __hip_atomic_load(reinterpret_cast<unsigned long long*>(field), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return (const T*)(const __attribute__((address_space(1))) T*)v;
}

// a pointer field of a posted job as a plain (generic) pointer
template <class T, class P>
__device__ __forceinline__ T* svc_ptr_mut(P* field) {
    return (T*)__hip_atomic_load(reinterpret_cast<unsigned long long*>(field), __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT);
}

// claim the next tile of band x of entry e: true with (index in the band,
// job number) when the add landed inside the band
__device__ __forceinline__ bool svc_claim(SvcEntry* ents, uint32_t e, uint32_t x, uint32_t& c, uint32_t& g) {
    const unsigned long long w =
        __hip_atomic_fetch_add(&ents[e].band[x].claim, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    c = (uint32_t)w & SVC_TILE_MASK;
    g = (uint32_t)(w >> 48);
    return c < ((uint32_t)(w >> SVC_TILE_BITS) & SVC_TILE_MASK);
}

// tiles left in band x of entry e (0 when claimed out)
__device__ __forceinline__ uint32_t svc_left(const SvcEntry* ents, uint32_t e, uint32_t x) {
    const unsigned long long w =
        __hip_atomic_load(&ents[e].band[x].claim, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t n = (uint32_t)(w >> SVC_TILE_BITS) & SVC_TILE_MASK, c = (uint32_t)w & SVC_TILE_MASK;
    return c < n ? n - c : 0u;
}

// The service loop. Thread 0 of a workgroup claims tiles (one atomic add on
// a band's claim word each): from the band it claimed from last while no
// new job has been posted, else from its own XCD's band of the posted job
// with the fewest unclaimed tiles there (short jobs are not held up behind
// long ones), else from another XCD's band (stealing balances the end of
// a job). The
// workgroup renders each tile with the raster's own tile routine.
// Synchronisation is paid once per workgroup and job, not per tile: the
// first claim of a job is followed by an acquire fence at device scope (the
// job's buffers and tables, written by the front end's kernels before the
// post's release, are visible after it), and the tiles a workgroup renders
// of one job are released together -- every wave fences at device scope,
// then one add to the job's finished count -- when its next claim is of
// another job or finds nothing. The workgroup whose add completes a job
// writes the scenes' flags and then the job number to host memory at
// system scope. The job's data changes between jobs while this kernel runs,
// so every read of it is a vector load through the coherent caches: the
// descriptor is read with atomic loads, and the tile routine's reads follow
// stores of the loop's earlier tiles that may alias them, so the compiler
// cannot turn them into scalar loads (whose cache would keep an earlier
// job's lines).
// Exit: when the host's stop flag is seen (workgroup 0 reads host memory
// and relays it through ctl) and no work is found, when no workgroup has
// claimed a tile for idle_ticks (all workgroups go idle together, so the
// service does not dwindle to a few), or, before claiming anything, once
// life_ticks have passed since the launch -- every path of the loop reaches
// one of these, releasing its finished tiles first, and the host relaunches
// the service when it finds it gone with work posted.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_raster_service(
    SvcEntry* ents, uint32_t nent, SvcCtl* ctl, SvcHost* host, unsigned long long idle_ticks,
    unsigned long long life_ticks) {
    __shared__ float lut[256];
    __shared__ ShadeLds<4> sh;
    __shared__ uint32_t pick[6];
    __shared__ uint32_t acq_gen[SVC_MAX_ENTRIES];  // thread 0: the job of each entry it has acquired
    __shared__ uint32_t acq_nt[SVC_MAX_ENTRIES];   // and that job's tile count
    constexpr uint32_t NONE = 0xFFFFFFFFu, EXIT = 0xFFFFFFFEu, RETRY = 0xFFFFFFFDu;
    lut[threadIdx.x] = kByteToFloat.v[threadIdx.x];
    if (threadIdx.x < SVC_MAX_ENTRIES) acq_gen[threadIdx.x] = NONE;
    const uint32_t xcd = blockIdx.x & 7u;  // workgroups are dealt to the 8 XCDs in turn
    const unsigned long long t0 = (unsigned long long)wall_clock64();
    // thread 0's claim state
    unsigned long long noted = t0;  // the last claim time this workgroup stored
    uint32_t polls = 0, cur_e = NONE, cur_x = 0, seen = 0;
    bool quiet = false;  // the last scan found no work; nothing new until `posted` moves past `seen`
    // finished tiles not yet released (uniform: every thread keeps them)
    uint32_t pend_e = NONE, pend_g = 0, pend_nt = 0, pend_n = 0;
    uint32_t st0 = 0, st1 = 0, st2 = 0, st3 = 0, st4 = 0;
    __syncthreads();
    for (;;) {
        if (threadIdx.x == 0) {
            uint32_t e_out = NONE, t_out = 0, g_out = 0, n_out = 0;
            const unsigned long long now = (unsigned long long)wall_clock64();
            if (now - t0 > life_ticks) {
                e_out = EXIT;
            } else {
                const uint32_t posted = __hip_atomic_load(&ctl->posted, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                uint32_t c = 0, g = 0, x = 0;
                bool got = false;
                if (cur_e != NONE && posted == seen) {
                    got = svc_claim(ents, cur_e, cur_x, c, g);
                    if (got) {
                        e_out = cur_e;
                        x = cur_x;
                    } else {
                        cur_e = NONE;
                    }
                }
                if (!got && (!quiet || posted != seen)) {
                    // own band first (shortest there), then the other bands
                    uint32_t best = NONE, bx = 0, left = 0xFFFFFFFFu;
                    for (uint32_t k = 0; k < 8 && best == NONE; ++k) {
                        const uint32_t xb = (xcd + k) & 7u;
                        for (uint32_t e = 0; e < nent; ++e) {
                            const uint32_t l = svc_left(ents, e, xb);
                            if (l != 0 && l < left) {
                                left = l;
                                best = e;
                                bx = xb;
                            }
                        }
                    }
                    seen = posted;
                    quiet = best == NONE;
                    if (best != NONE) {
                        got = svc_claim(ents, best, bx, c, g);
                        cur_e = got ? best : NONE;
                        cur_x = bx;
                        if (got) {
                            e_out = best;
                            x = bx;
                        } else {
                            e_out = RETRY;  // others took the tiles meanwhile
                        }
                    }
                }
                if (got) {
                    if (acq_gen[e_out] != g) {
                        // first tile of this job here: see what the post published
                        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                        acq_gen[e_out] = g;
                        acq_nt[e_out] = __hip_atomic_load(&ents[e_out].job.ntiles, __ATOMIC_RELAXED,
                                                          __HIP_MEMORY_SCOPE_AGENT);
                    }
                    n_out = acq_nt[e_out];
                    t_out = band_lo(n_out, x) + c;
                    g_out = g;
                    // idleness is the whole service's: claims refresh a shared
                    // time stamp (a store every idle/16 ticks at most)
                    if (now - noted > idle_ticks / 16) {
                        __hip_atomic_store(&ctl->last_claim, now, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        noted = now;
                    }
                } else if (e_out == NONE) {
                    if (blockIdx.x == 0 && (++polls & 15u) == 0 &&
                        __hip_atomic_load(&host->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != 0)
                        __hip_atomic_store(&ctl->stop, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    unsigned long long last =
                        __hip_atomic_load(&ctl->last_claim, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    last = last > t0 ? last : t0;
                    if (__hip_atomic_load(&ctl->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ||
                        (now > last && now - last > idle_ticks))
                        e_out = EXIT;
                }
            }
            pick[0] = e_out;
            pick[1] = t_out;
            pick[2] = g_out;
            pick[3] = n_out;
        }
        __syncthreads();
        const uint32_t e = pick[0], t = pick[1], g = pick[2], nt = pick[3];
        __syncthreads();
        const bool tile = e < SVC_MAX_ENTRIES;
        if (pend_n != 0 && !(tile && e == pend_e && g == pend_g)) {
            // release this workgroup's finished tiles of its last job
            __threadfence();
            __syncthreads();
            if (threadIdx.x == 0) {
                const uint32_t f =
                    __hip_atomic_fetch_add(&ents[pend_e].finished, pend_n, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
                pick[4] = f + pend_n == pend_nt ? 1u : 0u;
            }
            __syncthreads();
            if (pick[4] != 0) {
                // the job's last tiles: the scenes' flags (raised by every
                // stage and tile) to host memory, then the job number
                SvcJob* pj = &ents[pend_e].job;
                uint32_t* hf = svc_ptr_mut<uint32_t>(&pj->host_flags);
                SceneRun* pr = svc_ptr_mut<SceneRun>(&pj->runs);
                const uint32_t ns = __hip_atomic_load(&pj->nscenes, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                for (uint32_t s = threadIdx.x; s < ns; s += blockDim.x)
                    __hip_atomic_store(&hf[s],
                                       __hip_atomic_load(&pr[s].flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                __threadfence_system();
                __syncthreads();
                if (threadIdx.x == 0)
                    __hip_atomic_store(&host->done[pend_e], pend_g, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
            }
            __syncthreads();
            pend_n = 0;
        }
        if (e == EXIT) break;
        if (!tile) {
            if (e == NONE) __builtin_amdgcn_s_sleep(32);
            continue;
        }
        SvcJob* j = &ents[e].job;
        const RasterIn in{svc_ptr<PrimGeom>(&j->geom), svc_ptr<float>(&j->attrs), svc_ptr<float>(&j->uvat),
                          svc_ptr<uint2>(&j->flist),   svc_ptr<uint4>(&j->tile_rec), svc_ptr<SceneCfg>(&j->scenes),
                          svc_ptr<DrawDev>(&j->draws)};
        SceneRun* runs = svc_ptr_mut<SceneRun>(&j->runs);
        raster_tile<false, 4>(runs, in, t, 0u, lut, sh, st0, st1, st2, st3, st4);
        pend_e = e;
        pend_g = g;
        pend_nt = nt;
        pend_n++;
    }
}

void launch_raster_service(hipStream_t s, uint32_t grid, SvcEntry* ents, uint32_t nent, SvcCtl* ctl,
                           SvcHost* host, unsigned long long idle_ticks, unsigned long long life_ticks) {
    k_raster_service<<<grid, 256, 0, s>>>(ents, nent, ctl, host, idle_ticks, life_ticks);
}

void launch_job(hipStream_t s, const JobBufs& b, const JobGrid& g, bool raster_stats, hipEvent_t ev_bin,
                hipEvent_t ev_raster, SvcEntry* svc, SvcHost* svc_host, SvcCtl* svc_ctl, uint32_t svc_index) {
    const uint32_t scene_blocks = (g.scenes + 255) / 256;
    const uint32_t tchunk_blocks = std::min<uint32_t>((g.tris + 255) / 256 + 1, 1024u);
    k_job_init<<<at_least1(std::max({scene_blocks, g.draws, tchunk_blocks})), 256, 0, s>>>(b);
    k_classify<<<at_least1(std::min<uint32_t>(g.tris, 2048u)), 256, 0, s>>>(b, b.tchunk_draw, b.draws);
    k_clip<<<at_least1((g.clip + CLIP_GROUPS - 1) / CLIP_GROUPS), 64, 0, s>>>(b);
    if (b.tri_cap <= SCAN_ONE_MAX) {
        k_scene_prep_one<<<1, 1024, 0, s>>>(b);
    } else {
        launch_scan<uint32_t>(s, b.nsub, b.suboff, b.tri_cap, &b.head->ntri, b.parts, &b.cnt->nprims_raw);
        k_scene_prep<<<1, 1024, 0, s>>>(b);
        k_pchunk_map<<<at_least1(std::min<uint32_t>(g.bins / 256 + 1, 1024u)), 256, 0, s>>>(b);
    }
    k_setup<<<at_least1(g.tris), 256, 0, s>>>(b, b.draws, b.scenes, b.runs, b.tchunk_draw, b.suboff, b.nsub, b.cls,
                                              b.clip_out);
    if (ev_bin) (void)hipEventRecord(ev_bin, s);
    k_bin<false><<<at_least1(g.bins), 256, 0, s>>>(b, b.pchunk_scene, b.scenes, b.runs, b.pbin, b.bin_offsets);
    launch_scan<uint32_t>(s, b.bin_counts, b.bin_offsets, b.mat_cap, &b.cnt->mat_len, b.parts,
                          &b.cnt->pairs_binned);
    k_bin<true><<<at_least1(g.bins), 256, 0, s>>>(b, b.pchunk_scene, b.scenes, b.runs, b.pbin, b.bin_offsets);
    // fine_chunks and fmat_len are adjacent: the u64 total fills both
    if (b.bins_cap <= FINE_PREP_ONE_MAX) {
        k_fine_prep_one<<<1, 256, 0, s>>>(b);
    } else {
        k_fine_prep<<<at_least1(std::min<uint32_t>((b.bins_cap + 255) / 256, 1024u)), 256, 0, s>>>(b);
        launch_scan<unsigned long long>(s, b.fpack, b.fpre, b.bins_cap, &b.head->nbins, b.parts64,
                                        reinterpret_cast<unsigned long long*>(&b.cnt->fine_chunks));
        k_fchunk_map<<<at_least1(std::min<uint32_t>((b.bins_cap + 255) / 256, 1024u)), 256, 0, s>>>(b);
    }
    k_fine_bin<false><<<at_least1(g.fine), 256, 0, s>>>(b, b.fchunk_rec, b.scenes, b.fpre, b.bin_rng, b.list,
                                                 b.fine_off);
    launch_scan<uint32_t>(s, b.fine_counts, b.fine_off, b.fmat_cap, &b.cnt->fmat_len, b.parts, &b.cnt->fine_total);
    k_fine_bin<true><<<at_least1(g.fine), 256, 0, s>>>(b, b.fchunk_rec, b.scenes, b.fpre, b.bin_rng, b.list,
                                                b.fine_off);
    k_tile_prep<<<at_least1(std::min<uint32_t>((4u * g.ntiles + 255) / 256, 4096u)), 256, 0, s>>>(b);
    if (ev_raster) (void)hipEventRecord(ev_raster, s);
    if (svc) {
        k_svc_post<<<1, 64, 0, s>>>(b, svc, svc_host, svc_ctl, svc_index);
        return;
    }
    // persistent: g.raster workgroups of four waves; else one one-wave
    // workgroup per quadrant
    const uint32_t rgrid = g.raster ? g.raster : raster_grid_q(g.ntiles);
#define RZ_RASTER(ST)                                                                                              \
    if (g.raster)                                                                                                  \
        k_raster<ST><<<rgrid, 256, 0, s>>>(b, b.geom, b.attrs, b.uvat, b.flist, b.tile_rec, b.scenes, b.draws);    \
    else                                                                                                           \
        k_raster_q<ST><<<rgrid, 64, 0, s>>>(b, b.geom, b.attrs, b.uvat, b.flist, b.tile_rec, b.scenes, b.draws)
    if (raster_stats) {
        RZ_RASTER(true);
        k_rstat_reduce<<<64, 256, 0, s>>>(b, rgrid);
    } else {
        RZ_RASTER(false);
    }
#undef RZ_RASTER
}

}  // namespace rzk
