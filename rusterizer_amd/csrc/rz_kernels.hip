// rz_kernels.hip -- gfx950 kernels of the rasterizer (device side).
//
// Arithmetic contract: this file is compiled with -ffp-contract=off, IEEE
// f32 division and denormals kept, so each expression below produces the
// same bits as the reference's Rust expression it cites (no FMA fusion,
// same association, Rust's saturating float->int casts).
//
// Pipeline of one frame (all on the context stream):
//   per draw   k_mvp -> k_classify -> k_clip -> scan(nsub) -> k_setup
//   at resolve k_bin<count> -> scan(bins) -> k_bin<scatter> -> k_raster
// Rasterizer triangles ("prims") are stored in submission order
// (input triangle, then clip-fan index), so every later stage can keep the
// reference's order per pixel without sorting.

#include "rz_kernels.h"

#include <hip/hip_runtime.h>

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

// main.rs:147-152: projection * view * world, evaluated left to right
// (matrix.rs:56-79). One thread; written into the draw descriptor.
__global__ void k_mvp(DrawDev* __restrict__ draws, uint32_t d, Mat3 m) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
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
    for (int i = 0; i < 16; ++i) draws[d].M[i] = out[i];
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
// Frame bookkeeping kernels (single thread; no memset in the stream)
// ---------------------------------------------------------------------------
__global__ void k_frame_reset(FrameCounters* __restrict__ cnt) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        cnt->nprims = 0;
        cnt->nclip = 0;
        cnt->clip_need = 0;
        cnt->overflow = 0;
        cnt->flags = 0;
        cnt->nclipped = 0;
        cnt->npairs = 0;
        cnt->pairs_binned = 0;
        cnt->nch = 0;
        cnt->fine_total = 0;
        for (int i = 0; i < 5; ++i) cnt->rstat[i] = 0;
    }
}
__global__ void k_draw_reset(FrameCounters* __restrict__ cnt) {
    if (threadIdx.x == 0 && blockIdx.x == 0) cnt->nclip = 0;
}

// ---------------------------------------------------------------------------
// k_classify: one thread per input triangle.
// render.rs:104-108 (VS), render.rs:75-96 (assembly), clipping.rs:62-112
// (degenerate cull, trivial reject / accept). Triangles that need the
// Sutherland-Hodgman clipper get a slot in the clip list.
// cls[t]: 0 = Outside, 1 = Inside, 2 + slot = clipped.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_classify(const DrawDev* __restrict__ draws, uint32_t d,
                                                  FrameCounters* __restrict__ cnt,
                                                  uint32_t* __restrict__ cls, uint32_t* __restrict__ nsub,
                                                  uint32_t* __restrict__ clip_tri, uint32_t clip_cap) {
    const DrawDev& dd = draws[d];
    uint32_t ntri = dd.ntri;
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    bool live = t < ntri;
    uint32_t c = 0;
    if (live) {
        float4 v[3];
        for (int k = 0; k < 3; ++k) {
            uint32_t vi = dd.idx[(size_t)t * 3 + k];
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
    // one clip-list reservation per workgroup: wave counts -> LDS -> one atomic
    __shared__ uint32_t wcount[4], wbase[4];
    const bool need = live && c == 2;
    const uint64_t m = __ballot(need);
    const uint32_t wave = threadIdx.x >> 6;
    if (lane_id() == 0) wcount[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = wcount[0] + wcount[1] + wcount[2] + wcount[3];
        uint32_t b = tot ? atomicAdd(&cnt->nclip, tot) : 0u;
        for (int w = 0; w < 4; ++w) {
            wbase[w] = b;
            b += wcount[w];
        }
    }
    __syncthreads();
    const uint32_t slot = wbase[wave] + mbcnt(m);
    if (live) {
        if (need) {
            cls[t] = 2u + slot;
            nsub[t] = 0;  // written by k_clip
            if (slot < clip_cap)
                clip_tri[slot] = t;
            else
                atomicOr(&cnt->overflow, OVF_CLIP);
        } else {
            cls[t] = c;
            nsub[t] = c;  // Inside -> 1 rasterizer triangle, Outside -> 0
        }
    }
}

// ---------------------------------------------------------------------------
// k_clip: one wave per clipped triangle, one lane per polygon vertex.
// Sutherland-Hodgman over the planes L, R, B, T, N, F (clipping.rs:114-177):
// lane i owns the edge (prev = i-1 mod n, cur = i) and emits 0, 1 or 2
// vertices; a wave prefix count places them, so the output keeps the
// reference's vertex order exactly. The final polygon (<= MAXV vertices)
// goes to its clip slot; the fan (clipping.rs:183-190) is formed in k_setup.
// ---------------------------------------------------------------------------
struct ClipV {
    float x, y, z, w;
    float a[6];
};

__global__ __launch_bounds__(64) void k_clip(const DrawDev* __restrict__ draws, uint32_t d,
                                             FrameCounters* __restrict__ cnt,
                                             const uint32_t* __restrict__ clip_tri, uint32_t clip_cap,
                                             uint32_t* __restrict__ nsub, float* __restrict__ clip_out,
                                             uint32_t* __restrict__ clip_n) {
    __shared__ float sh[CLIP_LDS_V * 10];
    const DrawDev& dd = draws[d];
    const uint32_t lane = threadIdx.x;
    uint32_t nslots = min(uniform(cnt->nclip), clip_cap);
    if (blockIdx.x == 0 && lane == 0) {  // per-draw bookkeeping of the clip list
        uint32_t used = cnt->nclip;
        cnt->clip_need = max(cnt->clip_need, used);
        cnt->nclipped += used;
    }
    if (uniform(cnt->overflow) != 0) return;
    for (uint32_t s = blockIdx.x; s < nslots; s += gridDim.x) {
        uint32_t t = clip_tri[s];
        ClipV v;
        if (lane < 3) {
            uint32_t vi = dd.idx[(size_t)t * 3 + lane];
            float4 p;
            load_vertex(dd, vi, p, v.a);
            v.x = p.x; v.y = p.y; v.z = p.z; v.w = p.w;
        } else {
            v.x = v.y = v.z = v.w = 0.0f;
            for (int j = 0; j < 6; ++j) v.a[j] = 0.0f;
        }
        uint32_t n = 3;
        bool poly_overflow = false;
        for (int plane = 0; plane < 6; ++plane) {
            bool act = lane < n;
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
            int prev = (lane == 0) ? (int)n - 1 : (int)lane - 1;
            ClipV pv;
            pv.x = __shfl(v.x, prev);
            pv.y = __shfl(v.y, prev);
            pv.z = __shfl(v.z, prev);
            pv.w = __shfl(v.w, prev);
            for (int j = 0; j < 6; ++j) pv.a[j] = __shfl(v.a[j], prev);
            float dprev = __shfl(dcur, prev);
            bool pin = dprev >= 0.0f, cin = dcur >= 0.0f;
            uint32_t ncnt = act ? ((pin && cin) ? 1u : (pin ? 1u : (cin ? 2u : 0u))) : 0u;
            uint64_t m1 = __ballot(ncnt >= 1);
            uint64_t m2 = __ballot(ncnt == 2);
            uint32_t pos = mbcnt(m1) + mbcnt(m2);
            uint32_t newn = (uint32_t)__popcll(m1) + (uint32_t)__popcll(m2);
            ClipV o0 = v;
            if (act && (pin != cin)) {
                // clipping.rs:43-51 with p0 = prev, p1 = cur
                float alpha = dprev / (dprev - dcur);
                float s1 = 1.0f - alpha;
                o0.x = pv.x * s1 + v.x * alpha;
                o0.y = pv.y * s1 + v.y * alpha;
                o0.z = pv.z * s1 + v.z * alpha;
                o0.w = pv.w * s1 + v.w * alpha;
                // clipping.rs:149,161: (cur - prev) * alpha + prev
                for (int j = 0; j < 6; ++j) o0.a[j] = (v.a[j] - pv.a[j]) * alpha + pv.a[j];
            }
            __syncthreads();
            if (ncnt >= 1 && pos < CLIP_LDS_V) {
                float* q = sh + pos * 10;
                q[0] = o0.x; q[1] = o0.y; q[2] = o0.z; q[3] = o0.w;
                for (int j = 0; j < 6; ++j) q[4 + j] = o0.a[j];
            }
            if (ncnt == 2 && pos + 1 < CLIP_LDS_V) {
                float* q = sh + (pos + 1) * 10;
                q[0] = v.x; q[1] = v.y; q[2] = v.z; q[3] = v.w;
                for (int j = 0; j < 6; ++j) q[4 + j] = v.a[j];
            }
            __syncthreads();
            n = newn;
            if (n > 64) {  // more vertices than lanes: cannot continue
                poly_overflow = true;
                break;
            }
            if (n == 0) break;
            if (lane < n) {
                const float* q = sh + lane * 10;
                v.x = q[0]; v.y = q[1]; v.z = q[2]; v.w = q[3];
                for (int j = 0; j < 6; ++j) v.a[j] = q[4 + j];
            }
        }
        uint32_t ntris = 0;
        if (poly_overflow || n > MAXV) {
            if (lane == 0) atomicOr(&cnt->overflow, OVF_POLY);
        } else if (n >= 3) {
            ntris = n - 2;
            if (lane < n) {
                float* q = clip_out + ((size_t)s * MAXV + lane) * 10;
                q[0] = v.x; q[1] = v.y; q[2] = v.z; q[3] = v.w;
                for (int j = 0; j < 6; ++j) q[4 + j] = v.a[j];
            }
        } else if (n > 0) {
            if (lane == 0) atomicOr(&cnt->flags, RZ_FLAG_CLIP_FEWER);
        }
        if (lane == 0) {
            nsub[t] = ntris;
            clip_n[s] = n;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Exclusive scan of u32 (3 launches: tile sums, scan of sums, tile scan).
// Element count is n_host, or (when n_dev is set) bins * ceil(*n_dev / CHUNK).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t scan_len(uint32_t n_host, const uint32_t* n_dev, uint32_t mul, uint32_t div) {
    if (n_dev == nullptr) return n_host;
    uint32_t s = uniform(*n_dev);
    uint64_t n = (uint64_t)mul * ((s + div - 1) / div);
    return n < n_host ? (uint32_t)n : n_host;  // never past the buffers the host sized
}

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* lds, uint32_t* total) {
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o);
        if (lane >= (uint32_t)o) x += y;
    }
    if (lane == 63) lds[wave] = x;
    __syncthreads();
    uint32_t nw = blockDim.x >> 6, pre = 0, tot = 0;
    for (uint32_t w = 0; w < nw; ++w) {
        uint32_t s = lds[w];
        if (w < wave) pre += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return pre + x - v;
}

__global__ __launch_bounds__(256) void k_scan_reduce(const uint32_t* __restrict__ in, uint32_t n_host,
                                                     const uint32_t* __restrict__ n_dev, uint32_t mul, uint32_t div,
                                                     uint32_t* __restrict__ parts) {
    __shared__ uint32_t lds[8];
    uint32_t n = scan_len(n_host, n_dev, mul, div);
    uint32_t base = blockIdx.x * SCAN_TILE;
    if (base >= n) return;
    uint32_t s = 0;
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        uint32_t k = base + threadIdx.x * SCAN_ITEMS + i;
        if (k < n) s += in[k];
    }
    uint32_t tot;
    block_exclusive_scan(s, lds, &tot);
    if (threadIdx.x == 0) parts[blockIdx.x] = tot;
}

// one block: exclusive scan of parts, plus base (*base_total, if given);
// writes the grand total back to *base_total and to *total2 (if given).
__global__ __launch_bounds__(1024) void k_scan_top(uint32_t* __restrict__ parts, uint32_t n_host,
                                                   const uint32_t* __restrict__ n_dev, uint32_t mul, uint32_t div,
                                                   uint32_t* __restrict__ base_total,
                                                   uint32_t* __restrict__ total2) {
    __shared__ uint32_t lds[16];
    uint32_t n = scan_len(n_host, n_dev, mul, div);
    uint32_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    uint32_t run = base_total ? *base_total : 0u;
    for (uint32_t off = 0; off < nb; off += blockDim.x) {
        uint32_t k = off + threadIdx.x;
        uint32_t v = k < nb ? parts[k] : 0u;
        uint32_t tot;
        uint32_t ex = block_exclusive_scan(v, lds, &tot);
        if (k < nb) parts[k] = run + ex;
        run += tot;
    }
    if (threadIdx.x == 0) {
        if (base_total) *base_total = run;
        if (total2) *total2 = run;
    }
}

__global__ __launch_bounds__(256) void k_scan_down(const uint32_t* __restrict__ in, uint32_t n_host,
                                                   const uint32_t* __restrict__ n_dev, uint32_t mul, uint32_t div,
                                                   const uint32_t* __restrict__ parts,
                                                   uint32_t* __restrict__ out) {
    __shared__ uint32_t lds[8];
    uint32_t n = scan_len(n_host, n_dev, mul, div);
    uint32_t base = blockIdx.x * SCAN_TILE;
    if (base >= n) return;
    uint32_t v[SCAN_ITEMS];
    uint32_t s = 0;
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        uint32_t k = base + threadIdx.x * SCAN_ITEMS + i;
        v[i] = k < n ? in[k] : 0u;
        s += v[i];
    }
    uint32_t tot;
    uint32_t ex = block_exclusive_scan(s, lds, &tot) + parts[blockIdx.x];
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        uint32_t k = base + threadIdx.x * SCAN_ITEMS + i;
        if (k < n) out[k] = ex;
        ex += v[i];
    }
}

// ---------------------------------------------------------------------------
// k_setup: one thread per input triangle; writes its rasterizer triangles
// at sub_off[t] + k in fan order.
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

__device__ __forceinline__ void setup_prim(const ClipV* tv, uint32_t W, uint32_t H, uint32_t draw,
                                           uint32_t fs, PrimGeom* __restrict__ g, float* __restrict__ attr,
                                           uint4* __restrict__ pbin, BinGeom bg, uint32_t& pairs) {
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
    if (x0 < x1 && y0 < y1) {
        uint32_t cx0 = x0 / bg.cb, cx1 = (x1 - 1) / bg.cb, cy0 = y0 / bg.cb, cy1 = (y1 - 1) / bg.cb;
        pairs += (cx1 - cx0 + 1) * (cy1 - cy0 + 1);
    }
    if (fs == RZ_FS_COLOR) {
        for (int k = 0; k < 3; ++k)
            for (int j = 0; j < 4; ++j) attr[k * 4 + j] = tv[k].a[j];
    } else if (fs == RZ_FS_TEXTURE) {
        for (int k = 0; k < 3; ++k) {
            attr[k * 2] = tv[k].a[4];
            attr[k * 2 + 1] = tv[k].a[5];
        }
    }
}

__global__ __launch_bounds__(256) void k_setup(const DrawDev* __restrict__ draws, uint32_t d,
                                               FrameCounters* __restrict__ cnt,
                                               const uint32_t* __restrict__ cls, const uint32_t* __restrict__ nsub,
                                               const uint32_t* __restrict__ sub_off,
                                               const float* __restrict__ clip_out, uint32_t clip_cap,
                                               PrimGeom* __restrict__ geom, float* __restrict__ attrs,
                                               uint4* __restrict__ pbin, uint32_t prim_cap,
                                               uint32_t W, uint32_t H, BinGeom bg) {
    const DrawDev& dd = draws[d];
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t pairs = 0;
    bool ovf = false;
    if (t < dd.ntri) {
        uint32_t c = cls[t];
        uint32_t base = sub_off[t];
        uint32_t fs = dd.fs;
        if (c == 1) {
            ClipV tv[3];
            for (int k = 0; k < 3; ++k) {
                uint32_t vi = dd.idx[(size_t)t * 3 + k];
                float4 p;
                load_vertex(dd, vi, p, tv[k].a);
                tv[k].x = p.x; tv[k].y = p.y; tv[k].z = p.z; tv[k].w = p.w;
            }
            if (base < prim_cap)
                setup_prim(tv, W, H, d, fs, geom + base, attrs + (size_t)base * ATTR_STRIDE, pbin + base, bg, pairs);
            else
                ovf = true;
        } else if (c >= 2) {
            uint32_t slot = c - 2;
            uint32_t n = nsub[t];
            if (slot < clip_cap) {
                const float* poly = clip_out + (size_t)slot * MAXV * 10;
                for (uint32_t k = 0; k < n; ++k) {
                    // clipping.rs:185-190: (v0, v_{k+1}, v_{k+2})
                    ClipV tv[3];
                    const uint32_t vid[3] = {0u, k + 1, k + 2};
                    for (int q = 0; q < 3; ++q) {
                        const float* p = poly + vid[q] * 10;
                        tv[q].x = p[0]; tv[q].y = p[1]; tv[q].z = p[2]; tv[q].w = p[3];
                        for (int j = 0; j < 6; ++j) tv[q].a[j] = p[4 + j];
                    }
                    if (base + k < prim_cap)
                        setup_prim(tv, W, H, d, fs, geom + base + k, attrs + (size_t)(base + k) * ATTR_STRIDE,
                                   pbin + base + k, bg, pairs);
                    else
                        ovf = true;
                }
            }
        }
    }
    if (__ballot(ovf) && lane_id() == 0) atomicOr(&cnt->overflow, OVF_PRIMS);
    // wave sum of coarse-bin entries
    uint32_t s = pairs;
    for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o);
    if (lane_id() == 0 && s) atomicAdd(&cnt->npairs, (unsigned long long)s);
}

// ---------------------------------------------------------------------------
// k_bin: ordered coarse binning, one wave per chunk of CHUNK prims.
// COUNT:   counts[b * NC + chunk] = prims of the chunk overlapping bin b.
// SCATTER: writes {prim, bbox, zlo} to list position offsets[b * NC + chunk]
//          + rank, rank = number of earlier prims of the chunk in bin b.
// Both loop bins in the same order and rank lanes with ballots, so each bin's
// list holds its prims in submission order.
// ---------------------------------------------------------------------------
template <bool SCATTER>
__global__ __launch_bounds__(256) void k_bin(const FrameCounters* __restrict__ cnt,
                                             const uint4* __restrict__ pbin, BinGeom bg,
                                             uint32_t* __restrict__ counts, const uint32_t* __restrict__ offsets,
                                             uint4* __restrict__ list, uint32_t pair_cap,
                                             FrameCounters* __restrict__ cnt_w) {
    if (uniform(cnt->overflow) != 0) return;
    const uint32_t S = uniform(cnt->nprims);
    const uint32_t NC = (S + CHUNK - 1) / CHUNK;
    const uint32_t lane = lane_id();
    const uint32_t nb = bg.nbx * bg.nby;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    bool ovf = false;
    for (uint32_t chunk = wave; chunk < NC; chunk += nwaves) {
        // per-bin running rank, held by lane (b & 63) in register (b >> 6)
        uint32_t run0 = 0, run1 = 0, run2 = 0, run3 = 0;
        if (SCATTER) {
            if (lane < nb) run0 = offsets[(size_t)lane * NC + chunk];
            if (lane + 64 < nb) run1 = offsets[(size_t)(lane + 64) * NC + chunk];
            if (lane + 128 < nb) run2 = offsets[(size_t)(lane + 128) * NC + chunk];
            if (lane + 192 < nb) run3 = offsets[(size_t)(lane + 192) * NC + chunk];
        }
        for (uint32_t sub = 0; sub < CHUNK / 64; ++sub) {
            uint32_t p = chunk * CHUNK + sub * 64 + lane;
            uint4 pb = make_uint4(0u, 0u, 0u, 0u);
            if (p < S) pb = pbin[p];
            uint32_t x0 = pb.y & 0xFFFFu, x1 = pb.y >> 16, y0 = pb.z & 0xFFFFu, y1 = pb.z >> 16;
            bool ne = (x0 < x1) && (y0 < y1);
            uint32_t cx0 = x0 / bg.cb, cx1 = ne ? (x1 - 1) / bg.cb : 0u;
            uint32_t cy0 = y0 / bg.cb, cy1 = ne ? (y1 - 1) / bg.cb : 0u;
            for (uint32_t by = 0; by < bg.nby; ++by) {
                bool rowhit = ne && by >= cy0 && by <= cy1;
                if (__ballot(rowhit) == 0) continue;
                for (uint32_t bx = 0; bx < bg.nbx; ++bx) {
                    bool hit = rowhit && bx >= cx0 && bx <= cx1;
                    uint64_t m = __ballot(hit);
                    if (m == 0) continue;
                    uint32_t b = by * bg.nbx + bx;
                    uint32_t j = b >> 6, bl = b & 63;
                    uint32_t c = (uint32_t)__popcll(m);
                    if (SCATTER) {
                        uint32_t rv = j == 0 ? run0 : (j == 1 ? run1 : (j == 2 ? run2 : run3));
                        uint32_t r0 = (uint32_t)__builtin_amdgcn_readlane((int)rv, (int)bl);
                        if (hit) {
                            uint32_t pos = r0 + mbcnt(m);
                            if (pos < pair_cap)
                                list[pos] = make_uint4(p, pb.y, pb.z, pb.w);
                            else
                                ovf = true;
                        }
                    }
                    if (lane == bl) {
                        if (j == 0) run0 += c;
                        else if (j == 1) run1 += c;
                        else if (j == 2) run2 += c;
                        else run3 += c;
                    }
                }
            }
        }
        if (!SCATTER) {
            if (lane < nb) counts[(size_t)lane * NC + chunk] = run0;
            if (lane + 64 < nb) counts[(size_t)(lane + 64) * NC + chunk] = run1;
            if (lane + 128 < nb) counts[(size_t)(lane + 128) * NC + chunk] = run2;
            if (lane + 192 < nb) counts[(size_t)(lane + 192) * NC + chunk] = run3;
        }
    }
    if (SCATTER && __ballot(ovf) && lane == 0) atomicOr(&cnt_w->overflow, OVF_PAIRS);
}

// ---------------------------------------------------------------------------
// Fine binning: each coarse bin's list is split again into its 16x16 screen
// tiles (the raster workgroups), keeping submission order.
// k_fine_prep numbers the coarse lists' chunks of CHUNK entries:
//   chunk_base[b] = chunks of all bins before b, chunk_base[nb] = total (nch).
// The fine count matrix holds, for coarse bin b, local tile lt and chunk c,
// the entry (T * chunk_base[b] + lt * nch_b + c); its exclusive scan gives
// every tile one contiguous, ordered list.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(512) void k_fine_prep(FrameCounters* __restrict__ cnt,
                                                   const uint32_t* __restrict__ coff, BinGeom bg,
                                                   uint32_t* __restrict__ chunk_base) {
    __shared__ uint32_t lds[8];
    if (uniform(cnt->overflow) != 0) return;
    const uint32_t S = uniform(cnt->nprims);
    const uint32_t NC = (S + CHUNK - 1) / CHUNK;
    const uint32_t nb = bg.nbx * bg.nby;
    const uint32_t b = threadIdx.x;
    uint32_t len = 0;
    if (b < nb && NC > 0) {
        uint32_t beg = coff[(size_t)b * NC];
        uint32_t end = (b + 1 < nb) ? coff[(size_t)(b + 1) * NC] : cnt->pairs_binned;
        len = end - beg;
    }
    uint32_t tot;
    uint32_t ex = block_exclusive_scan((len + CHUNK - 1) / CHUNK, lds, &tot);
    if (b <= nb) chunk_base[b] = ex;
    if (threadIdx.x == 0) cnt->nch = tot;
}

template <bool SCATTER>
__global__ __launch_bounds__(256) void k_fine_bin(const FrameCounters* __restrict__ cnt,
                                                  const uint32_t* __restrict__ coff, const uint4* __restrict__ clist,
                                                  const uint32_t* __restrict__ chunk_base, BinGeom bg,
                                                  uint32_t* __restrict__ fcounts, const uint32_t* __restrict__ foff,
                                                  uint4* __restrict__ flist, uint32_t fcap,
                                                  FrameCounters* __restrict__ cnt_w) {
    __shared__ uint32_t cbase[MAX_BINS + 1];
    if (uniform(cnt->overflow) != 0) return;
    const uint32_t nb = bg.nbx * bg.nby;
    for (uint32_t i = threadIdx.x; i <= nb; i += blockDim.x) cbase[i] = chunk_base[i];
    __syncthreads();
    const uint32_t NCH = cbase[nb];
    const uint32_t S = uniform(cnt->nprims);
    const uint32_t NC = (S + CHUNK - 1) / CHUNK;
    const uint32_t TS = bg.cb / 16, T = TS * TS;
    if (T > 256) return;  // the host never configures more tiles per coarse bin
    const uint32_t lane = lane_id();
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    bool ovf = false;
    for (uint32_t g = wave; g < NCH; g += nwaves) {
        uint32_t lo = 0, hi = nb;  // largest b < nb with cbase[b] <= g
        while (hi - lo > 1) {
            uint32_t mid = (lo + hi) >> 1;
            if (cbase[mid] <= g) lo = mid;
            else hi = mid;
        }
        const uint32_t b = lo;
        const uint32_t c = g - cbase[b], nch_b = cbase[b + 1] - cbase[b];
        const uint32_t cbeg = uniform(coff[(size_t)b * NC]);
        const uint32_t cend = (b + 1 < nb) ? uniform(coff[(size_t)(b + 1) * NC]) : uniform(cnt->pairs_binned);
        const uint32_t binx0 = (b % bg.nbx) * bg.cb, biny0 = (b / bg.nbx) * bg.cb;
        const size_t mbase = (size_t)T * cbase[b] + c;
        uint32_t run0 = 0, run1 = 0, run2 = 0, run3 = 0;
        if (SCATTER) {
            if (lane < T) run0 = foff[mbase + (size_t)lane * nch_b];
            if (lane + 64 < T) run1 = foff[mbase + (size_t)(lane + 64) * nch_b];
            if (lane + 128 < T) run2 = foff[mbase + (size_t)(lane + 128) * nch_b];
            if (lane + 192 < T) run3 = foff[mbase + (size_t)(lane + 192) * nch_b];
        }
        for (uint32_t sub = 0; sub < CHUNK / 64; ++sub) {
            const uint32_t e = cbeg + c * CHUNK + sub * 64 + lane;
            const bool valid = e < cend;
            uint4 en = make_uint4(0u, 0u, 0u, 0u);
            if (valid) en = clist[e];
            const uint32_t x0 = en.y & 0xFFFFu, x1 = en.y >> 16, y0 = en.z & 0xFFFFu, y1 = en.z >> 16;
            // local tile range inside this coarse bin (the entry overlaps it)
            const uint32_t lx0 = (max(x0, binx0) - binx0) >> 4;
            const uint32_t lx1 = valid ? (min(x1, binx0 + bg.cb) - 1 - binx0) >> 4 : 0u;
            const uint32_t ly0 = (max(y0, biny0) - biny0) >> 4;
            const uint32_t ly1 = valid ? (min(y1, biny0 + bg.cb) - 1 - biny0) >> 4 : 0u;
            for (uint32_t ty = 0; ty < TS; ++ty) {
                const bool rowhit = valid & (ty >= ly0) & (ty <= ly1);
                if (__ballot(rowhit) == 0) continue;
                for (uint32_t tx = 0; tx < TS; ++tx) {
                    const bool hit = rowhit & (tx >= lx0) & (tx <= lx1);
                    const uint64_t m = __ballot(hit);
                    if (m == 0) continue;
                    const uint32_t lt = ty * TS + tx;
                    const uint32_t j = lt >> 6, bl = lt & 63;
                    const uint32_t cnt_m = (uint32_t)__popcll(m);
                    if (SCATTER) {
                        const uint32_t rv = j == 0 ? run0 : (j == 1 ? run1 : (j == 2 ? run2 : run3));
                        const uint32_t r0 = (uint32_t)__builtin_amdgcn_readlane((int)rv, (int)bl);
                        if (hit) {
                            const uint32_t pos = r0 + mbcnt(m);
                            if (pos < fcap) flist[pos] = en;
                            else ovf = true;
                        }
                    }
                    if (lane == bl) {
                        if (j == 0) run0 += cnt_m;
                        else if (j == 1) run1 += cnt_m;
                        else if (j == 2) run2 += cnt_m;
                        else run3 += cnt_m;
                    }
                }
            }
        }
        if (!SCATTER) {
            if (lane < T) fcounts[mbase + (size_t)lane * nch_b] = run0;
            if (lane + 64 < T) fcounts[mbase + (size_t)(lane + 64) * nch_b] = run1;
            if (lane + 128 < T) fcounts[mbase + (size_t)(lane + 128) * nch_b] = run2;
            if (lane + 192 < T) fcounts[mbase + (size_t)(lane + 192) * nch_b] = run3;
        }
    }
    if (SCATTER && __ballot(ovf) && lane == 0) atomicOr(&cnt_w->overflow, OVF_FINE);
}

// ---------------------------------------------------------------------------
// k_raster: one workgroup per 16x16 screen tile, one wave per 8x8 quadrant,
// one lane per pixel. The pixel's 4 MSAA colours and depths live in the
// lane's registers for the whole frame; the wave walks its coarse bin's list
// in submission order and applies mod.rs:443-473 to each rasterizer
// triangle that can touch the quadrant, then resolves (buffers.rs:111-125)
// and writes the ARGB word once.
//
// Work that provably cannot change a sample is skipped (exact culls):
//  - list entries whose bbox misses the quadrant;
//  - hierarchical depth: the quadrant keeps qzmax = max of its stored
//    depths; a prim whose depth lower bound zlo is >= qzmax fails the strict
//    LESS test on every sample of the quadrant;
//  - edge reject: if an edge function is negative, with a margin that covers
//    the rounding of the reference's evaluation, at the quadrant corner where
//    it is largest, no sample of the quadrant is inside that edge.
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

// texture.rs:65-83 (channels via color.rs:22-29, byte / 255.0 from a table)
__device__ __forceinline__ RGBA texture_sample(const uint32_t* __restrict__ tex, uint32_t tw, uint32_t th,
                                               float u, float v, const float* lut, bool& oob) {
    float x = u * (float)(tw - 1);
    float y = v * (float)(th - 1);
    uint32_t xf0 = sat_u32(floorf(x)), xc = sat_u32(ceilf(x));
    uint32_t yf0 = sat_u32(floorf(y)), yc = sat_u32(ceilf(y));
    uint32_t tl = read_texel(tex, tw, th, xf0, yf0, oob);
    uint32_t tr = read_texel(tex, tw, th, xc, yf0, oob);
    uint32_t bl = read_texel(tex, tw, th, xf0, yc, oob);
    uint32_t br = read_texel(tex, tw, th, xc, yc, oob);
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

template <bool STATS>
__global__ __launch_bounds__(256) void k_raster(const FrameCounters* __restrict__ cnt,
                                                const PrimGeom* __restrict__ geom, const float* __restrict__ attrs,
                                                const DrawDev* __restrict__ draws,
                                                const uint32_t* __restrict__ chunk_base,
                                                const uint32_t* __restrict__ foff,
                                                const uint4* __restrict__ list, uint32_t fcap, BinGeom bg,
                                                uint32_t W, uint32_t H, float clear_depth,
                                                uint32_t* __restrict__ out, uint32_t* __restrict__ msaa_color,
                                                float* __restrict__ msaa_depth, FrameCounters* __restrict__ cnt_w) {
    __shared__ float lut[256];
    lut[threadIdx.x] = (float)threadIdx.x / 255.0f;  // color.rs:24-27
    __syncthreads();
    if (uniform(cnt->overflow) != 0) return;

    // XCD-aware tile order: workgroups go round-robin to the 8 XCDs, so give
    // each XCD a contiguous band of tiles (shared coarse lists stay in its L2).
    const uint32_t ntx = (W + 15) / 16, nty = (H + 15) / 16, nt = ntx * nty;
    const uint32_t b = blockIdx.x, xcd = b & 7u, k = b >> 3;
    const uint32_t q = nt / 8u, r = nt % 8u;
    const uint32_t tile = xcd * q + min(xcd, r) + k;
    const uint32_t tx = tile % ntx, ty = tile / ntx;
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t qx0 = tx * 16 + (wave & 1) * 8, qy0 = ty * 16 + (wave >> 1) * 8;
    const uint32_t px = qx0 + (lane & 7), py = qy0 + (lane >> 3);
    const float fx = (float)px, fy = (float)py;
    const bool onscreen = px < W && py < H;
    const float qxl = (float)qx0 + 0.125f, qxh = (float)qx0 + 7.875f;
    const float qyl = (float)qy0 + 0.125f, qyh = (float)qy0 + 7.875f;

    float zb0 = clear_depth, zb1 = clear_depth, zb2 = clear_depth, zb3 = clear_depth;
    uint32_t cb0 = CLEAR_COLOR, cb1 = CLEAR_COLOR, cb2 = CLEAR_COLOR, cb3 = CLEAR_COLOR;
    bool oob = false;
    // qzmax: largest depth stored in the quadrant's on-screen samples
    float qzmax = __int_as_float(
        __builtin_amdgcn_readfirstlane(__float_as_int(wave_max(onscreen ? clear_depth : -FLT_MAX_F))));
    // lane_zbmax: largest depth stored in this pixel's 4 samples
    float lane_zbmax = onscreen ? clear_depth : -FLT_MAX_F;

    // this tile's fine list (see k_fine_bin for the layout)
    const uint32_t nb = bg.nbx * bg.nby;
    const uint32_t TS = bg.cb / 16, T = TS * TS;
    const uint32_t cbx = (tx * 16) / bg.cb, cby = (ty * 16) / bg.cb;
    const uint32_t cbin = cby * bg.nbx + cbx;
    const uint32_t lt = ((ty * 16 - cby * bg.cb) >> 4) * TS + ((tx * 16 - cbx * bg.cb) >> 4);
    const uint32_t chb = uniform(chunk_base[cbin]), nch_b = uniform(chunk_base[cbin + 1]) - chb;
    const size_t nfine = (size_t)T * uniform(chunk_base[nb]);
    const size_t s_idx = (size_t)T * chb + (size_t)lt * nch_b, e_idx = s_idx + nch_b;
    const uint32_t ftot = uniform(cnt->fine_total);
    // (clamped to the list buffer: a bad range can never read past it)
    const uint32_t lend = min(e_idx < nfine ? uniform(foff[e_idx]) : ftot, fcap);
    const uint32_t lbeg = min(s_idx < nfine ? uniform(foff[s_idx]) : ftot, lend);

    uint32_t st_entries = 0, st_cand = 0, st_hits = 0, st_cov = 0, st_shaded = 0;
    for (uint32_t base = lbeg; base < lend; base += 64) {
        const uint32_t i = base + lane;
        uint32_t myprim = 0;
        bool cand = false;
        if (STATS) st_entries += min(64u, lend - base);
        if (i < lend) {
            const uint4 en = list[i];
            myprim = en.x;
            const uint32_t x0 = en.y & 0xFFFFu, x1 = en.y >> 16, y0 = en.z & 0xFFFFu, y1 = en.z >> 16;
            cand = (x0 < qx0 + 8) & (x1 > qx0) & (y0 < qy0 + 8) & (y1 > qy0) & (__uint_as_float(en.w) < qzmax);
        }
        if (cand) {
            const float4* gp = reinterpret_cast<const float4*>(geom + myprim);
            const float4 g0 = gp[0], g1 = gp[1], g2 = gp[2];  // a0 a1 | a2 n0 | n1 n2
            cand = edge_may_cover(g1.z, g1.w, g0.x, g0.y, qxl, qxh, qyl, qyh) &&
                   edge_may_cover(g2.x, g2.y, g0.z, g0.w, qxl, qxh, qyl, qyh) &&
                   edge_may_cover(g2.z, g2.w, g1.x, g1.y, qxl, qxh, qyl, qyh);
        }
        uint64_t m = __ballot(cand);
        if (STATS) st_cand += (uint32_t)__popcll(m);
        if (m == 0) continue;
        uint32_t nprim = (uint32_t)__builtin_amdgcn_readlane((int)myprim, (int)(__ffsll((unsigned long long)m) - 1));
        PrimGeom nxt = geom[nprim];
        for (;;) {
            m &= m - 1;
            const PrimGeom g = nxt;
            const uint32_t prim = nprim;
            if (m) {  // start loading the next hit's record now
                nprim = (uint32_t)__builtin_amdgcn_readlane((int)myprim, (int)(__ffsll((unsigned long long)m) - 1));
                nxt = geom[nprim];
            }
            bool wrote = false;
            const uint32_t bx0 = g.bbx & 0xFFFFu, bx1 = g.bbx >> 16, by0 = g.bby & 0xFFFFu, by1 = g.bby >> 16;
            // depth lower bound over the quadrant's samples inside the bbox
            const float fbx0 = (float)bx0, fby0 = (float)by0;
            const float xl = fmaxf(qxl, fbx0 + 0.125f), xh = fminf(qxh, (float)bx1 - 0.125f);
            const float yl = fmaxf(qyl, fby0 + 0.125f), yh = fminf(qyh, (float)by1 - 0.125f);
            const float zq = g.pz0 + g.pzx * ((g.pzx >= 0.0f ? xl : xh) - fbx0) +
                             g.pzy * ((g.pzy >= 0.0f ? yl : yh) - fby0);
            const float zlb = fmaxf(zq, g.zlo);
            if (!(zq >= qzmax) && !(g.zlo >= qzmax)) {
                const bool inb = (px >= bx0) & (px < bx1) & (py >= by0) & (py < by1);
                // mod.rs:134-146 + 148-170: coverage of the 4 samples
                const bool t0 = g.flags & 1u, t1 = g.flags & 2u, t2 = g.flags & 4u;
                float e1s[4], e2s[4];
                uint32_t cov = 0;
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const float sx = fx + sample_x(s), sy = fy + sample_y(s);
                    const float e0 = edge(g.nx0, g.ny0, g.ax0, g.ay0, sx, sy);
                    const float e1 = edge(g.nx1, g.ny1, g.ax1, g.ay1, sx, sy);
                    const float e2 = edge(g.nx2, g.ny2, g.ax2, g.ay2, sx, sy);
                    // bitwise on purpose: straight-line compares, no branches
                    const bool in = ((e0 > 0.0f) | (t0 & !(e0 < 0.0f))) & ((e1 > 0.0f) | (t1 & !(e1 < 0.0f))) &
                                    ((e2 > 0.0f) | (t2 & !(e2 < 0.0f)));
                    cov |= (in ? 1u : 0u) << s;
                    e1s[s] = e1;
                    e2s[s] = e2;
                }
                if (!inb) cov = 0;
                // a covered lane can pass LESS only if some stored depth exceeds the
                // bound; when no lane can, the wave skips the depth math
                if (zlb >= lane_zbmax) cov = 0;
                if (STATS) {
                    st_hits += 1;
                    st_cov += __ballot(cov != 0) ? 1u : 0u;
                }
                if (__ballot(cov != 0) != 0) {
                    // mod.rs:225-253: linear depth at covered samples; mod.rs:363-378: strict LESS
                    float sd[4];
                    uint32_t pass = 0;
                    const float zbuf[4] = {zb0, zb1, zb2, zb3};
#pragma unroll
                    for (int s = 0; s < 4; ++s) {
                        const float b0 = clamp_bary(e1s[s] * g.inv_area);
                        const float b1 = clamp_bary(e2s[s] * g.inv_area);
                        const float b2 = clamp_bary(1.0f - b0 - b1);
                        const float z = b0 * g.z0 + b1 * g.z1 + b2 * g.z2;
                        const bool c = (cov >> s) & 1u;
                        sd[s] = c ? z : 0.0f;
                        pass |= (c & (z < zbuf[s])) ? (1u << s) : 0u;
                    }
                    if (pass != 0) {
                        // mod.rs:70-99: perspective-correct barycentrics at the pixel
                        // centre or at the first depth-passing sample
                        float ix = fx + 0.5f, iy = fy + 0.5f;
                        if (pass != 0xFu) {
                            const int s = __builtin_ctz(pass);
                            ix = fx + sample_x(s);
                            iy = fy + sample_y(s);
                        }
                        const float ef0 = edge(g.nx0, g.ny0, g.ax0, g.ay0, ix, iy);
                        const float ef1 = edge(g.nx1, g.ny1, g.ax1, g.ay1, ix, iy);
                        const float ef2 = edge(g.nx2, g.ny2, g.ax2, g.ay2, ix, iy);
                        const float f_u = ef1 / g.w0;
                        const float f_v = ef2 / g.w1;
                        const float f_w = ef0 / g.w2;
                        const float sum = f_u + f_v + f_w;
                        const float u = clamp_bary(f_u / sum);
                        const float v = clamp_bary(f_v / sum);
                        const float w = clamp_bary(1.0f - u - v);
                        // main.rs:67-77 fragment shaders; Debug is Color::grayscale(depths[0])
                        const uint32_t fs = (g.flags >> 8) & 3u;
                        const float* at = attrs + (size_t)prim * ATTR_STRIDE;
                        float cr = sd[0], cg = sd[0], cbl = sd[0], ca = 1.0f;
                        if (fs == RZ_FS_COLOR) {
                            cr = at[0] * u + at[4] * v + at[8] * w;
                            cg = at[1] * u + at[5] * v + at[9] * w;
                            cbl = at[2] * u + at[6] * v + at[10] * w;
                            ca = at[3] * u + at[7] * v + at[11] * w;
                        } else if (fs == RZ_FS_TEXTURE) {
                            const DrawDev& dd = draws[g.flags >> 16];
                            const float tu = at[0] * u + at[2] * v + at[4] * w;
                            const float tv = at[1] * u + at[3] * v + at[5] * w;
                            const RGBA c = texture_sample(dd.tex, dd.tw, dd.th, tu, tv, lut, oob);
                            cr = c.r; cg = c.g; cbl = c.b; ca = c.a;
                        }
                        // mod.rs:380-397
                        const uint32_t argb = to_argb(cr, cg, cbl, ca);
                        if (pass & 1u) { cb0 = argb; zb0 = sd[0]; }
                        if (pass & 2u) { cb1 = argb; zb1 = sd[1]; }
                        if (pass & 4u) { cb2 = argb; zb2 = sd[2]; }
                        if (pass & 8u) { cb3 = argb; zb3 = sd[3]; }
                        lane_zbmax = fmaxf(fmaxf(zb0, zb1), fmaxf(zb2, zb3));
                        wrote = true;
                    }
                }
            }
            if (STATS) st_shaded += __ballot(wrote) ? 1u : 0u;
            if (__ballot(wrote)) {
                qzmax = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(wave_max(lane_zbmax))));
            }
            if (m == 0) break;
        }
    }
    if (__ballot(oob) && lane == 0) atomicOr(&cnt_w->flags, RZ_FLAG_TEXEL_OOB);
    if (STATS && lane == 0) {
        atomicAdd(&cnt_w->rstat[0], (unsigned long long)st_entries);
        atomicAdd(&cnt_w->rstat[1], (unsigned long long)st_cand);
        atomicAdd(&cnt_w->rstat[2], (unsigned long long)st_hits);
        atomicAdd(&cnt_w->rstat[3], (unsigned long long)st_cov);
        atomicAdd(&cnt_w->rstat[4], (unsigned long long)st_shaded);
    }
    if (onscreen) {
        // buffers.rs:111-125
        uint32_t rs = ((cb0 & 0x00FF0000u) >> 16) + ((cb1 & 0x00FF0000u) >> 16) + ((cb2 & 0x00FF0000u) >> 16) +
                      ((cb3 & 0x00FF0000u) >> 16);
        uint32_t gs = ((cb0 & 0x0000FF00u) >> 8) + ((cb1 & 0x0000FF00u) >> 8) + ((cb2 & 0x0000FF00u) >> 8) +
                      ((cb3 & 0x0000FF00u) >> 8);
        uint32_t bs = (cb0 & 0xFFu) + (cb1 & 0xFFu) + (cb2 & 0xFFu) + (cb3 & 0xFFu);
        size_t o = (size_t)py * W + px;
        out[o] = (0xFFu << 24) | (rs / 4u) << 16 | (gs / 4u) << 8 | (bs / 4u);
        if (msaa_color) {
            uint4 c4 = make_uint4(cb0, cb1, cb2, cb3);
            *reinterpret_cast<uint4*>(msaa_color + o * 4) = c4;
        }
        if (msaa_depth) {
            float4 d4 = make_float4(zb0, zb1, zb2, zb3);
            *reinterpret_cast<float4*>(msaa_depth + o * 4) = d4;
        }
    }
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

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
void launch_max_index(hipStream_t s, const uint32_t* idx, uint64_t n, uint32_t* out) {
    if (n == 0) return;
    uint64_t blocks = (n + 255) / 256;
    k_max_index<<<(uint32_t)(blocks < 1024 ? blocks : 1024), 256, 0, s>>>(idx, n, out);
}
void launch_frame_reset(hipStream_t s, FrameCounters* cnt) { k_frame_reset<<<1, 64, 0, s>>>(cnt); }
void launch_draw_reset(hipStream_t s, FrameCounters* cnt) { k_draw_reset<<<1, 64, 0, s>>>(cnt); }
void launch_mvp(hipStream_t s, DrawDev* draws, uint32_t d, const Mat3& m) { k_mvp<<<1, 64, 0, s>>>(draws, d, m); }

void launch_classify(hipStream_t s, const DrawDev* draws, uint32_t d, uint32_t ntri, FrameCounters* cnt,
                     uint32_t* cls, uint32_t* nsub, uint32_t* clip_tri, uint32_t clip_cap) {
    if (ntri == 0) return;
    k_classify<<<(ntri + 255) / 256, 256, 0, s>>>(draws, d, cnt, cls, nsub, clip_tri, clip_cap);
}

void launch_clip(hipStream_t s, const DrawDev* draws, uint32_t d, FrameCounters* cnt, const uint32_t* clip_tri,
                 uint32_t clip_cap, uint32_t* nsub, float* clip_out, uint32_t* clip_n) {
    if (clip_cap == 0) return;
    uint32_t grid = clip_cap < 8192u ? clip_cap : 8192u;
    k_clip<<<grid, 64, 0, s>>>(draws, d, cnt, clip_tri, clip_cap, nsub, clip_out, clip_n);
}

void launch_scan(hipStream_t s, const uint32_t* in, uint32_t* out, uint32_t n_host, const uint32_t* n_dev,
                 uint32_t mul, uint32_t div, uint32_t* parts, uint32_t* base_total, uint32_t* total2) {
    uint32_t nb = (n_host + SCAN_TILE - 1) / SCAN_TILE;
    if (nb > 0) k_scan_reduce<<<nb, 256, 0, s>>>(in, n_host, n_dev, mul, div, parts);
    k_scan_top<<<1, 1024, 0, s>>>(parts, n_host, n_dev, mul, div, base_total, total2);
    if (nb > 0) k_scan_down<<<nb, 256, 0, s>>>(in, n_host, n_dev, mul, div, parts, out);
}

void launch_setup(hipStream_t s, const DrawDev* draws, uint32_t d, uint32_t ntri, FrameCounters* cnt,
                  const uint32_t* cls, const uint32_t* nsub, const uint32_t* sub_off, const float* clip_out,
                  uint32_t clip_cap, PrimGeom* geom, float* attrs, uint4* pbin, uint32_t prim_cap, uint32_t W,
                  uint32_t H, BinGeom bg) {
    if (ntri == 0) return;
    k_setup<<<(ntri + 255) / 256, 256, 0, s>>>(draws, d, cnt, cls, nsub, sub_off, clip_out, clip_cap, geom, attrs,
                                               pbin, prim_cap, W, H, bg);
}

void launch_bin(hipStream_t s, bool scatter, uint32_t grid, FrameCounters* cnt, const uint4* pbin, BinGeom bg,
                uint32_t* counts, const uint32_t* offsets, uint4* list, uint32_t pair_cap) {
    if (grid == 0) return;
    if (scatter)
        k_bin<true><<<grid, 256, 0, s>>>(cnt, pbin, bg, counts, offsets, list, pair_cap, cnt);
    else
        k_bin<false><<<grid, 256, 0, s>>>(cnt, pbin, bg, counts, offsets, list, pair_cap, cnt);
}

void launch_fine(hipStream_t s, FrameCounters* cnt, const uint32_t* coff, const uint4* clist, uint32_t* chunk_base,
                 BinGeom bg, uint32_t* fcounts, uint32_t* foff, uint4* flist, uint32_t fcap, uint32_t nch_max,
                 uint32_t* parts) {
    k_fine_prep<<<1, 512, 0, s>>>(cnt, coff, bg, chunk_base);
    const uint32_t grid = nch_max / 4 + 1 < 4096 ? nch_max / 4 + 1 : 4096;
    k_fine_bin<false><<<grid, 256, 0, s>>>(cnt, coff, clist, chunk_base, bg, fcounts, nullptr, nullptr, 0, cnt);
    const uint32_t T = (bg.cb / 16) * (bg.cb / 16);
    launch_scan(s, fcounts, foff, T * nch_max, &cnt->nch, T, 1, parts, nullptr, &cnt->fine_total);
    k_fine_bin<true><<<grid, 256, 0, s>>>(cnt, coff, clist, chunk_base, bg, nullptr, foff, flist, fcap, cnt);
}

void launch_raster(hipStream_t s, FrameCounters* cnt, const PrimGeom* geom, const float* attrs,
                   const DrawDev* draws, const uint32_t* chunk_base, const uint32_t* foff, const uint4* list,
                   uint32_t fcap, BinGeom bg, uint32_t W, uint32_t H, float clear_depth, uint32_t* out,
                   uint32_t* msaa_color, float* msaa_depth, bool stats) {
    uint32_t nt = ((W + 15) / 16) * ((H + 15) / 16);
    if (stats)
        k_raster<true><<<nt, 256, 0, s>>>(cnt, geom, attrs, draws, chunk_base, foff, list, fcap, bg, W, H,
                                          clear_depth, out, msaa_color, msaa_depth, cnt);
    else
        k_raster<false><<<nt, 256, 0, s>>>(cnt, geom, attrs, draws, chunk_base, foff, list, fcap, bg, W, H,
                                           clear_depth, out, msaa_color, msaa_depth, cnt);
}

}  // namespace rzk
