// rz_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A line-faithful C++ restatement of rusterizer's CPU draw path (clip-space
// vertices -> resolved ARGB framebuffer). It exists to check the HIP
// product library bit for bit; the product never links, loads or calls it.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
//
// Arithmetic contract (held by compiling with -O2 -ffp-contract=off and no
// fast-math): every + and * below keeps the association of the Rust source
// it cites; f32 division is IEEE; float->int casts saturate like Rust's
// `as` (NaN -> 0); f32::min/max are fminf/fmaxf; f32::clamp lets NaN through.
//
// Reference files are cited as path:line relative to the rusterizer crate.

#include "rz_oracle.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace {

// ---------------------------------------------------------------------------
// Rust scalar semantics
// ---------------------------------------------------------------------------

// `f as usize` (x86-64): saturating, NaN -> 0.
inline uint64_t sat_usize(float f) {
    if (!(f > 0.0f)) return 0;
    if (f >= 18446744073709551616.0f) return UINT64_MAX;
    return (uint64_t)f;
}
// `f as u32`: saturating, NaN -> 0.
inline uint32_t sat_u32(float f) {
    if (!(f > 0.0f)) return 0;
    if (f >= 4294967296.0f) return UINT32_MAX;
    return (uint32_t)f;
}
// f32::clamp (core::f32): NaN passes through.
inline float clampf(float x, float lo, float hi) {
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    return x;
}
// mod.rs:102-106 (debug_assert dropped: release semantics)
inline float clamp_bary(float x) { return clampf(x, 0.0f, 1.0f); }
// f32::fract = self - self.trunc()
inline float fract(float x) { return x - truncf(x); }

// ---------------------------------------------------------------------------
// math/vector.rs, math/matrix.rs
// ---------------------------------------------------------------------------

// vector.rs:17-23: sum starts at 0.0 and accumulates left to right.
inline float dot(const float *a, const float *b, int n) {
    float sum = 0.0f;
    for (int i = 0; i < n; ++i) sum += a[i] * b[i];
    return sum;
}

// matrix.rs:56-79: result[i][j] = row_i(a) . col_j(b)
void mat_mul(const float *a, const float *b, float *out) {
    float tmp[16];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            float col[4] = {b[0 * 4 + j], b[1 * 4 + j], b[2 * 4 + j], b[3 * 4 + j]};
            tmp[i * 4 + j] = dot(a + i * 4, col, 4);
        }
    memcpy(out, tmp, sizeof(tmp));
}

// vector.rs:219-240: result[i] = row_i(m) . v
void mat_vec(const float *m, const float *v, float *out) {
    float tmp[4];
    for (int i = 0; i < 4; ++i) tmp[i] = dot(m + i * 4, v, 4);
    memcpy(out, tmp, sizeof(tmp));
}

// ---------------------------------------------------------------------------
// color.rs, graphics_primitives.rs
// ---------------------------------------------------------------------------

struct Color {
    float r, g, b, a;
};

// color.rs:15-20 (u32 shifts drop high bits; `as u32` saturates)
inline uint32_t to_argb(Color c) {
    return (sat_u32(c.a * 255.0f) << 24) | (sat_u32(c.r * 255.0f) << 16) |
           (sat_u32(c.g * 255.0f) << 8) | sat_u32(c.b * 255.0f);
}

// color.rs:22-29
inline Color from_rgba(const uint8_t *rgba) {
    return Color{(float)rgba[0] / 255.0f, (float)rgba[1] / 255.0f, (float)rgba[2] / 255.0f,
                 (float)rgba[3] / 255.0f};
}

inline Color cmul(Color c, float s) { return Color{c.r * s, c.g * s, c.b * s, c.a * s}; }  // color.rs:75-85
inline Color cadd(Color x, Color y) {  // color.rs:101-111
    return Color{x.r + y.r, x.g + y.g, x.b + y.b, x.a + y.a};
}

// graphics_primitives.rs:9-13
struct VertexAttribute {
    Color color;
    float uv[2];
};

// graphics_primitives.rs:21-30
inline VertexAttribute amul(VertexAttribute v, float s) {
    return VertexAttribute{cmul(v.color, s), {v.uv[0] * s, v.uv[1] * s}};
}
// graphics_primitives.rs:43-51
inline VertexAttribute aadd(VertexAttribute x, VertexAttribute y) {
    return VertexAttribute{cadd(x.color, y.color), {x.uv[0] + y.uv[0], x.uv[1] + y.uv[1]}};
}
// graphics_primitives.rs:53-61
inline VertexAttribute asub(VertexAttribute x, VertexAttribute y) {
    return VertexAttribute{Color{x.color.r - y.color.r, x.color.g - y.color.g, x.color.b - y.color.b,
                                 x.color.a - y.color.a},
                           {x.uv[0] - y.uv[0], x.uv[1] - y.uv[1]}};
}

inline VertexAttribute attr_from(const float *a18, int i) {
    const float *p = a18 + i * 6;
    return VertexAttribute{Color{p[0], p[1], p[2], p[3]}, {p[4], p[5]}};
}
inline void attr_to(VertexAttribute v, float *p) {
    p[0] = v.color.r; p[1] = v.color.g; p[2] = v.color.b; p[3] = v.color.a;
    p[4] = v.uv[0]; p[5] = v.uv[1];
}

struct P4 {
    float x, y, z, w;
};

// graphics_primitives.rs:63-71
struct Triangle {
    P4 v[3];
    VertexAttribute a[3];
};

// ---------------------------------------------------------------------------
// texture.rs
// ---------------------------------------------------------------------------

struct Texture {
    std::vector<uint8_t> buf;
    uint64_t width = 0, height = 0;
};

// texture.rs:47-63. The reference indexes buf[x*4 + y*4*w ..]; an index past
// the buffer panics there. Here it is flagged and clamped to the last texel.
inline Color read_texel(const uint8_t *buf, uint64_t w, uint64_t h, uint64_t x, uint64_t y,
                        uint32_t *flags) {
    uint64_t n = w * h;
    uint64_t idx;
    if (y >= h || x >= n - y * w) {
        *flags |= RZO_FLAG_TEXEL_OOB;
        idx = n - 1;
    } else {
        idx = x + y * w;
    }
    uint8_t rgba[4] = {buf[idx * 4], buf[idx * 4 + 1], buf[idx * 4 + 2], 255};
    rgba[3] = buf[idx * 4 + 3];
    return from_rgba(rgba);
}

// texture.rs:65-83
Color texture_sample(const uint8_t *buf, uint64_t w, uint64_t h, float u, float v,
                     uint32_t *flags) {
    float x = u * (float)(w - 1);
    float y = v * (float)(h - 1);

    Color topleft = read_texel(buf, w, h, sat_usize(floorf(x)), sat_usize(floorf(y)), flags);
    Color topright = read_texel(buf, w, h, sat_usize(ceilf(x)), sat_usize(floorf(y)), flags);
    Color botleft = read_texel(buf, w, h, sat_usize(floorf(x)), sat_usize(ceilf(y)), flags);
    Color botright = read_texel(buf, w, h, sat_usize(ceilf(x)), sat_usize(ceilf(y)), flags);

    float x_f = fract(x);
    float y_f = fract(y);

    Color y0 = cadd(cmul(topleft, 1.0f - x_f), cmul(topright, x_f));
    Color y1 = cadd(cmul(botleft, 1.0f - x_f), cmul(botright, x_f));

    return cadd(cmul(y0, 1.0f - y_f), cmul(y1, y_f));
}

// ---------------------------------------------------------------------------
// rasterizer/mod.rs
// ---------------------------------------------------------------------------

const int N_MSAA_SAMPLES = 4;  // mod.rs:23

// mod.rs:108-114
const float RGSS[4][2] = {
    {5.0f / 8.0f, 1.0f / 8.0f},
    {7.0f / 8.0f, 5.0f / 8.0f},
    {3.0f / 8.0f, 7.0f / 8.0f},
    {1.0f / 8.0f, 3.0f / 8.0f},
};

// mod.rs:25-60
inline bool mask_get(uint32_t m, int i) { return ((1u << i) & m) != 0; }
inline uint32_t mask_set(uint32_t m, int i, bool v) {
    uint32_t b = v ? 1u : 0u;
    return (m & ~(1u << i)) | (b << i);
}

// mod.rs:15-21 (xy of any point type)
inline float triangle_2x_area(float x0, float y0, float x1, float y1, float x2, float y2) {
    float v10x = x1 - x0, v10y = y1 - y0;
    float v20x = x2 - x0, v20y = y2 - y0;
    return v10x * v20y - v20x * v10y;  // vector.rs:183-185
}

// mod.rs:116-175 + 177-184
struct RasterizerTriangle {
    float pts[6];
    float nrm[6];
    float cov_eval[4][3];  // coverage_evaluated
    uint32_t mask = 0;  // coverage_mask
    float depths_cs[3];
    float depths[3];
    VertexAttribute attrs[3];
    float inv_2x_area;

    // mod.rs:125-132
    void eval_single(float x, float y, float e[3]) const {
        for (int k = 0; k < 3; ++k) {
            float p[2] = {x - pts[2 * k], y - pts[2 * k + 1]};  // point.rs:86-95
            e[k] = dot(nrm + 2 * k, p, 2);
        }
    }

    // mod.rs:148-170
    static bool inside(const float *n, const float *e) {
        for (int k = 0; k < 3; ++k) {
            float val = e[k];
            float nx = n[2 * k], ny = n[2 * k + 1];
            bool ok;
            if (val > 0.0f) ok = true;
            else if (val < 0.0f) ok = false;
            else if (nx > 0.0f) ok = true;
            else if (nx < 0.0f) ok = false;
            else if (ny < 0.0f) ok = true;
            else ok = false;
            if (!ok) return false;
        }
        return true;
    }

    // mod.rs:134-146
    void eval(uint64_t x, uint64_t y) {
        for (int i = 0; i < N_MSAA_SAMPLES; ++i) {
            float xs = (float)x + RGSS[i][0];
            float ys = (float)y + RGSS[i][1];
            eval_single(xs, ys, cov_eval[i]);
            mask = mask_set(mask, i, inside(nrm, cov_eval[i]));
        }
    }
};

// mod.rs:187-222
RasterizerTriangle make_rasterizer_triangle(const float scr[9], const float dcs[3],
                                            const VertexAttribute attrs[3]) {
    RasterizerTriangle t;
    for (int k = 0; k < 3; ++k) {
        const float *p = scr + 3 * k;
        const float *q = scr + 3 * ((k + 1) % 3);
        float vx = q[0] - p[0];
        float vy = q[1] - p[1];
        t.nrm[2 * k] = -vy;
        t.nrm[2 * k + 1] = vx;
        t.pts[2 * k] = p[0];
        t.pts[2 * k + 1] = p[1];
        t.depths[k] = p[2];
        t.depths_cs[k] = dcs[k];
        t.attrs[k] = attrs[k];
    }
    t.inv_2x_area = 1.0f / triangle_2x_area(scr[0], scr[1], scr[3], scr[4], scr[6], scr[7]);
    memset(t.cov_eval, 0, sizeof(t.cov_eval));
    t.mask = 0;
    return t;
}

// mod.rs:225-253: sampled depths for raster-covered samples, 0.0 elsewhere
void fragment_depths(const RasterizerTriangle &t, float out[4]) {
    for (int i = 0; i < N_MSAA_SAMPLES; ++i) {
        out[i] = 0.0f;
        if (mask_get(t.mask, i)) {
            const float *ef = t.cov_eval[i];
            float b0 = clamp_bary(ef[1] * t.inv_2x_area);
            float b1 = clamp_bary(ef[2] * t.inv_2x_area);
            float b2 = clamp_bary(1.0f - b0 - b1);
            out[i] = b0 * t.depths[0] + b1 * t.depths[1] + b2 * t.depths[2];
        }
    }
}

// mod.rs:70-99
VertexAttribute interpolate(const RasterizerTriangle &t, uint64_t x, uint64_t y, uint32_t cov) {
    float xs = (float)x + 0.5f;
    float ys = (float)y + 0.5f;
    if (cov != 0b1111) {
        for (int i = 0; i < N_MSAA_SAMPLES; ++i) {
            if (mask_get(cov, i)) {
                xs = (float)x + RGSS[i][0];
                ys = (float)y + RGSS[i][1];
                break;
            }
        }
    }
    float efs[3];
    t.eval_single(xs, ys, efs);
    float f_u = efs[1] / t.depths_cs[0];
    float f_v = efs[2] / t.depths_cs[1];
    float f_w = efs[0] / t.depths_cs[2];
    float sum = f_u + f_v + f_w;
    float u = clamp_bary(f_u / sum);
    float v = clamp_bary(f_v / sum);
    float w = clamp_bary(1.0f - u - v);
    return aadd(aadd(amul(t.attrs[0], u), amul(t.attrs[1], v)), amul(t.attrs[2], w));
}

// mod.rs:284-313
Triangle perspective_divide(const Triangle &t) {
    Triangle o = t;
    for (int k = 0; k < 3; ++k) {
        const P4 &p = t.v[k];
        o.v[k] = P4{p.x / p.w, p.y / p.w, p.z / p.w, p.w};
    }
    return o;
}

// mod.rs:315-345
RasterizerTriangle viewport_transform(uint64_t width, uint64_t height, const Triangle &tri) {
    const float zmin = 0.0f;
    const float zmax = 1.0f;
    float scr[9], dcs[3];
    for (int k = 0; k < 3; ++k) {
        const P4 &v = tri.v[k];
        float x = (float)width * (v.x + 1.0f) / 2.0f;
        float y = (float)height * (1.0f - (v.y + 1.0f) / 2.0f);
        float z = (v.z + 1.0f) * 0.5f * (zmax - zmin) + zmin;
        scr[3 * k] = x;
        scr[3 * k + 1] = y;
        scr[3 * k + 2] = z;
        dcs[k] = v.w;
    }
    return make_rasterizer_triangle(scr, dcs, tri.a);
}

struct PixelBoundingBox {
    uint64_t min_x, max_x, min_y, max_y;
};

// bounding_box.rs:13-42
PixelBoundingBox bbox_from(const float *pts) {
    float a0 = FLT_MAX, a1 = -FLT_MAX, a2 = FLT_MAX, a3 = -FLT_MAX;
    for (int k = 0; k < 3; ++k) {
        float px = pts[2 * k], py = pts[2 * k + 1];
        a0 = fminf(a0, px);
        a1 = fmaxf(a1, px);
        a2 = fminf(a2, py);
        a3 = fmaxf(a3, py);
    }
    return PixelBoundingBox{sat_usize(floorf(a0)), sat_usize(ceilf(a1)), sat_usize(floorf(a2)),
                            sat_usize(ceilf(a3))};
}

// ---------------------------------------------------------------------------
// rasterizer/clipping.rs
// ---------------------------------------------------------------------------

enum ClipResult { OUTSIDE = 0, INSIDE = 1, CLIPPED = 2 };
const float CULL_DEGENERATE_TRIANGLE_AREA_EPS = 0.000001f;  // clipping.rs:12

// clipping.rs:29-38; plane order L, R, B, T, N, F (clipping.rs:53-60)
inline float distance_measure(int plane, const P4 &p) {
    switch (plane) {
        case 0: return p.w + p.x;
        case 1: return p.w - p.x;
        case 2: return p.w + p.y;
        case 3: return p.w - p.y;
        case 4: return p.w + p.z;
        default: return p.w - p.z;
    }
}

// clipping.rs:43-51: (1-a)*p0 + a*p1 via point.rs:99-106 and vector.rs:124-136
inline P4 compute_intersection(const P4 &p0, float d0, const P4 &p1, float d1, float *alpha) {
    float a = d0 / (d0 - d1);
    *alpha = a;
    float s = 1.0f - a;
    return P4{p0.x * s + p1.x * a, p0.y * s + p1.y * a, p0.z * s + p1.z * a, p0.w * s + p1.w * a};
}

// clipping.rs:62-195
ClipResult try_clip(const Triangle &tri, std::vector<Triangle> &out, uint32_t *flags) {
    out.clear();
    if (fabsf(triangle_2x_area(tri.v[0].x, tri.v[0].y, tri.v[1].x, tri.v[1].y, tri.v[2].x,
                               tri.v[2].y)) < CULL_DEGENERATE_TRIANGLE_AREA_EPS)
        return OUTSIDE;

    bool inside[3][2] = {{true, true}, {true, true}, {true, true}};
    bool outside[3][2] = {{true, true}, {true, true}, {true, true}};
    for (int k = 0; k < 3; ++k) {
        const P4 &v = tri.v[k];
        inside[0][0] &= v.x >= -v.w;
        inside[1][0] &= v.y >= -v.w;
        inside[2][0] &= v.z >= -v.w;
        inside[0][1] &= v.x <= v.w;
        inside[1][1] &= v.y <= v.w;
        inside[2][1] &= v.z <= v.w;
        outside[0][0] &= v.x < -v.w;
        outside[1][0] &= v.y < -v.w;
        outside[2][0] &= v.z < -v.w;
        outside[0][1] &= v.x > v.w;
        outside[1][1] &= v.y > v.w;
        outside[2][1] &= v.z > v.w;
    }
    bool any_out = false, all_in = true;
    for (int a = 0; a < 3; ++a)
        for (int s = 0; s < 2; ++s) {
            any_out |= outside[a][s];
            all_in &= inside[a][s];
        }
    if (any_out) return OUTSIDE;
    if (all_in) return INSIDE;

    std::vector<P4> out_v(tri.v, tri.v + 3);
    std::vector<VertexAttribute> out_a(tri.a, tri.a + 3);
    std::vector<P4> in_v;
    std::vector<VertexAttribute> in_a;
    for (int plane = 0; plane < 6; ++plane) {
        in_v = out_v;
        in_a = out_a;
        out_v.clear();
        out_a.clear();
        size_t n = in_v.size();
        for (size_t i = 0; i < n; ++i) {
            size_t prev_i = (i + n - 1) % n;
            P4 prev_vert = in_v[prev_i];
            VertexAttribute prev_attr = in_a[prev_i];
            P4 cur_vert = in_v[i];
            VertexAttribute cur_attr = in_a[i];
            float prev_d = distance_measure(plane, prev_vert);
            float cur_d = distance_measure(plane, cur_vert);
            bool pin = prev_d >= 0.0f, cin = cur_d >= 0.0f;
            if (pin && cin) {
                out_v.push_back(cur_vert);
                out_a.push_back(cur_attr);
            } else if (pin && !cin) {
                float alpha;
                P4 p = compute_intersection(prev_vert, prev_d, cur_vert, cur_d, &alpha);
                out_v.push_back(p);
                out_a.push_back(aadd(amul(asub(cur_attr, prev_attr), alpha), prev_attr));
            } else if (!pin && cin) {
                float alpha;
                P4 p = compute_intersection(prev_vert, prev_d, cur_vert, cur_d, &alpha);
                out_v.push_back(p);
                out_a.push_back(aadd(amul(asub(cur_attr, prev_attr), alpha), prev_attr));
                out_v.push_back(cur_vert);
                out_a.push_back(cur_attr);
            }
        }
    }
    if (out_v.empty()) return OUTSIDE;
    // clipping.rs:183 computes len-2 in usize; SH over a polygon never yields
    // 1 or 2 vertices from >= 3, but flag it rather than wrap.
    if (out_v.size() < 3) {
        *flags |= RZO_FLAG_CLIP_FEWER;
        return CLIPPED;
    }
    for (size_t i = 0; i + 2 < out_v.size(); ++i) {
        Triangle t;
        t.v[0] = out_v[0];
        t.v[1] = out_v[i + 1];
        t.v[2] = out_v[i + 2];
        t.a[0] = out_a[0];
        t.a[1] = out_a[i + 1];
        t.a[2] = out_a[i + 2];
        out.push_back(t);
    }
    return CLIPPED;
}

// ---------------------------------------------------------------------------
// rasterizer/buffers.rs
// ---------------------------------------------------------------------------

const uint32_t CLEAR_COLOR = 0xFF191919u;  // buffers.rs:5
const uint64_t TILE_SIZE = 64;             // buffers.rs:8

// buffers.rs:111-125
inline uint32_t box_filter_color(const uint32_t *c) {
    uint32_t red = 0, green = 0, blue = 0;
    for (int i = 0; i < 4; ++i) {
        red += (c[i] & 0x00FF0000u) >> 16;
        green += (c[i] & 0x0000FF00u) >> 8;
        blue += c[i] & 0x000000FFu;
    }
    return (0xFFu << 24) | (red / 4u) << 16 | (green / 4u) << 8 | (blue / 4u);
}

}  // namespace

// buffers.rs:10-80 (masks kept as bytes so tile-parallel workers can mark
// disjoint tiles without sharing words)
struct rzo_tiles {
    std::vector<PixelBoundingBox> tiles;
    std::vector<uint8_t> masks[2];
    size_t mask_idx = 0;
    uint64_t n_horizontal = 0;

    rzo_tiles(uint64_t width, uint64_t height) {  // buffers.rs:19-47
        uint64_t nh = width / TILE_SIZE + (width % TILE_SIZE == 0 ? 0 : 1);
        uint64_t nv = height / TILE_SIZE + (height % TILE_SIZE == 0 ? 0 : 1);
        for (uint64_t j = 0; j < nv; ++j)
            for (uint64_t i = 0; i < nh; ++i)
                tiles.push_back(PixelBoundingBox{i * TILE_SIZE, std::min((i + 1) * TILE_SIZE, width),
                                                 j * TILE_SIZE, std::min((j + 1) * TILE_SIZE, height)});
        masks[0].assign(nh * nv, 0);
        masks[1].assign(nh * nv, 0);
        n_horizontal = nh;
    }
    uint64_t tile_idx(uint64_t row, uint64_t col) const {  // buffers.rs:49-51
        return (row / TILE_SIZE) * n_horizontal + (col / TILE_SIZE);
    }
    void mark(uint64_t row, uint64_t col) { masks[mask_idx][tile_idx(row, col)] = 1; }  // :53-56
    void next() {  // buffers.rs:58-63
        mask_idx = (mask_idx + 1) % 2;
        std::fill(masks[mask_idx].begin(), masks[mask_idx].end(), 0);
    }
};

// Rasterizer (mod.rs:264-281) + Uniforms (uniform.rs:12-46)
struct rzo_ctx {
    uint64_t width, height;
    std::vector<uint32_t> color;  // ColorBuffer.buffer, [u32;4] per pixel
    std::vector<float> depth;     // DepthBuffer.buffer, [f32;4] per pixel
    std::vector<uint32_t> resolve;
    rzo_tiles tiles;
    float clear_depth = FLT_MAX;  // buffers.rs:6 CLEAR_DEPTH = f32::MAX
    float world[16], view[16], proj[16];
    std::vector<Texture> textures;
    uint32_t flags = 0;
    uint64_t stats[5] = {0, 0, 0, 0, 0};

    rzo_ctx(uint64_t w, uint64_t h)
        : width(w), height(h), color(w * h * 4, CLEAR_COLOR), depth(w * h * 4, FLT_MAX),
          resolve(w * h, CLEAR_COLOR), tiles(w, h) {
        for (int i = 0; i < 16; ++i) world[i] = view[i] = proj[i] = (i % 5 == 0) ? 1.0f : 0.0f;
    }
};

namespace {

struct DrawState {
    int fs;
    const Texture *tex;  // uniforms.get_texture(0) for the Texture FS
};

// main.rs:67-77 fragment shaders
inline Color run_fs(const DrawState &ds, const float sampled[4], VertexAttribute attr,
                    uint32_t *flags) {
    switch (ds.fs) {
        case RZO_FS_TEXTURE:
            return texture_sample(ds.tex->buf.data(), ds.tex->width, ds.tex->height, attr.uv[0],
                                  attr.uv[1], flags);
        case RZO_FS_COLOR:
            return attr.color;
        default:  // Debug: Color::grayscale(frag_coords.depths[0]) (color.rs:65-72)
            return Color{sampled[0], sampled[0], sampled[0], 1.0f};
    }
}

// One pixel of the loop at mod.rs:443-473. Returns true if it wrote.
inline bool shade_pixel(rzo_ctx *c, RasterizerTriangle &t, uint64_t x, uint64_t y,
                        const DrawState &ds, uint32_t *flags, rzo_tiles *tiles) {
    t.eval(x, y);
    if (t.mask == 0) return false;
    float sampled[4];
    fragment_depths(t, sampled);
    size_t idx = y * c->width + x;
    const float *cur = &c->depth[idx * 4];
    uint32_t cov = 0;  // depth_coverage, mod.rs:363-378 (strict LESS)
    for (int i = 0; i < 4; ++i)
        if (mask_get(t.mask, i)) cov = mask_set(cov, i, sampled[i] < cur[i]);
    if (cov == 0) return false;
    VertexAttribute attr = interpolate(t, x, y, cov);
    Color col = run_fs(ds, sampled, attr, flags);
    // write_pixel, mod.rs:380-397
    tiles->mark(y, x);
    uint32_t argb = to_argb(col);
    for (int i = 0; i < 4; ++i)
        if (mask_get(cov, i)) {
            c->color[idx * 4 + i] = argb;
            c->depth[idx * 4 + i] = sampled[i];
        }
    return true;
}

// Rasterizer::bounding_box, mod.rs:347-361
inline PixelBoundingBox clamped_bbox(const rzo_ctx *c, const RasterizerTriangle &t) {
    PixelBoundingBox b = bbox_from(t.pts);
    return PixelBoundingBox{std::max<uint64_t>(b.min_x, 0), std::min(b.max_x, c->width),
                            std::max<uint64_t>(b.min_y, 0), std::min(b.max_y, c->height)};
}

// render.rs:98-114 front half: VS map (main.rs:147-152) + primitive_assembly
// (render.rs:75-96), then the clip / divide / viewport stages of mod.rs:425-441.
// Produces the ordered list of rasterizer triangles.
int build_setups(rzo_ctx *c, const float *pos, const float *color, const float *uv, uint32_t nv,
                 const uint32_t *idx, uint64_t ni, int vs, int fs,
                 std::vector<RasterizerTriangle> &setups, DrawState &ds) {
    if (vs != RZO_VS_MVP) return -2;
    if (fs < RZO_FS_TEXTURE || fs > RZO_FS_DEBUG) return -2;
    if (ni % 3 != 0) return -2;
    for (uint64_t i = 0; i < ni; ++i)
        if (idx[i] >= nv) return -2;
    ds.fs = fs;
    ds.tex = nullptr;
    if (fs == RZO_FS_TEXTURE) {
        if (c->textures.empty()) {
            c->flags |= RZO_FLAG_NO_TEXTURE;
            return -3;
        }
        ds.tex = &c->textures[0];
    }
    float m[16];
    rzo_mvp(c->world, c->view, c->proj, m);
    std::vector<float> clip((size_t)nv * 4);
    for (uint32_t i = 0; i < nv; ++i) {
        float v4[4] = {pos[i * 3], pos[i * 3 + 1], pos[i * 3 + 2], 1.0f};  // point.rs:81-83
        mat_vec(m, v4, &clip[(size_t)i * 4]);
    }
    for (int k = 0; k < 5; ++k) c->stats[k] = 0;
    std::vector<Triangle> clipped;
    setups.clear();
    for (uint64_t t = 0; t < ni / 3; ++t) {
        Triangle tri;
        for (int k = 0; k < 3; ++k) {
            uint32_t vi = idx[t * 3 + k];
            const float *p = &clip[(size_t)vi * 4];
            tri.v[k] = P4{p[0], p[1], p[2], p[3]};
            const float *cc = color ? color + (size_t)vi * 4 : nullptr;
            const float *tc = uv ? uv + (size_t)vi * 2 : nullptr;
            tri.a[k] = VertexAttribute{cc ? Color{cc[0], cc[1], cc[2], cc[3]} : Color{0, 0, 0, 0},
                                       {tc ? tc[0] : 0.0f, tc ? tc[1] : 0.0f}};
        }
        c->stats[0]++;
        ClipResult r = try_clip(tri, clipped, &c->flags);
        if (r == OUTSIDE) {
            c->stats[1]++;
            continue;
        }
        if (r == INSIDE) {
            c->stats[2]++;
            clipped.assign(1, tri);
        } else {
            c->stats[3]++;
        }
        for (const Triangle &sub : clipped) {
            setups.push_back(viewport_transform(c->width, c->height, perspective_divide(sub)));
            c->stats[4]++;
        }
    }
    return 0;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================

extern "C" {

float rzo_dot4(const float a[4], const float b[4]) { return dot(a, b, 4); }
void rzo_mat_mul(const float a[16], const float b[16], float out[16]) { mat_mul(a, b, out); }
void rzo_mat_vec(const float m[16], const float v[4], float out[4]) { mat_vec(m, v, out); }

// main.rs:147-152: projection * view * world (left to right), then * v
void rzo_mvp(const float world[16], const float view[16], const float proj[16], float out[16]) {
    float pv[16];
    mat_mul(proj, view, pv);
    mat_mul(pv, world, out);
}

void rzo_perspective_divide(const float clip[12], float ndc[12]) {
    Triangle t;
    for (int k = 0; k < 3; ++k) t.v[k] = P4{clip[4 * k], clip[4 * k + 1], clip[4 * k + 2], clip[4 * k + 3]};
    Triangle o = perspective_divide(t);
    for (int k = 0; k < 3; ++k) {
        ndc[4 * k] = o.v[k].x;
        ndc[4 * k + 1] = o.v[k].y;
        ndc[4 * k + 2] = o.v[k].z;
        ndc[4 * k + 3] = o.v[k].w;
    }
}

static void export_setup(const RasterizerTriangle &t, rzo_setup *out) {
    memcpy(out->pts, t.pts, sizeof(out->pts));
    memcpy(out->normals, t.nrm, sizeof(out->normals));
    memcpy(out->depths, t.depths, sizeof(out->depths));
    memcpy(out->depths_cs, t.depths_cs, sizeof(out->depths_cs));
    out->inv_2x_area = t.inv_2x_area;
}
static RasterizerTriangle import_setup(const rzo_setup *s) {
    RasterizerTriangle t;
    memcpy(t.pts, s->pts, sizeof(t.pts));
    memcpy(t.nrm, s->normals, sizeof(t.nrm));
    memcpy(t.depths, s->depths, sizeof(t.depths));
    memcpy(t.depths_cs, s->depths_cs, sizeof(t.depths_cs));
    t.inv_2x_area = s->inv_2x_area;
    memset(t.cov_eval, 0, sizeof(t.cov_eval));
    t.mask = 0;
    for (int k = 0; k < 3; ++k) t.attrs[k] = VertexAttribute{Color{0, 0, 0, 0}, {0, 0}};
    return t;
}

void rzo_viewport_setup(uint32_t w, uint32_t h, const float ndc[12], rzo_setup *out) {
    Triangle t;
    for (int k = 0; k < 3; ++k) {
        t.v[k] = P4{ndc[4 * k], ndc[4 * k + 1], ndc[4 * k + 2], ndc[4 * k + 3]};
        t.a[k] = VertexAttribute{Color{0, 0, 0, 0}, {0, 0}};
    }
    export_setup(viewport_transform(w, h, t), out);
}

void rzo_setup_from_screen(const float scr[9], const float w3[3], rzo_setup *out) {
    VertexAttribute a[3] = {};
    export_setup(make_rasterizer_triangle(scr, w3, a), out);
}

void rzo_eval_single(const rzo_setup *s, float x, float y, float e[3]) {
    import_setup(s).eval_single(x, y, e);
}

int rzo_inside(const rzo_setup *s, const float e[3]) {
    return RasterizerTriangle::inside(s->normals, e) ? 1 : 0;
}

uint32_t rzo_coverage(const rzo_setup *s, uint32_t x, uint32_t y, float evals[12]) {
    RasterizerTriangle t = import_setup(s);
    t.eval(x, y);
    if (evals) memcpy(evals, t.cov_eval, sizeof(t.cov_eval));
    return t.mask;
}

void rzo_fragment_depths(const rzo_setup *s, uint32_t mask, const float evals[12], float depths[4]) {
    RasterizerTriangle t = import_setup(s);
    memcpy(t.cov_eval, evals, sizeof(t.cov_eval));
    t.mask = mask;
    fragment_depths(t, depths);
}

void rzo_interpolate(const rzo_setup *s, const float attrs[18], uint32_t x, uint32_t y,
                     uint32_t cov, float out[6]) {
    RasterizerTriangle t = import_setup(s);
    for (int k = 0; k < 3; ++k) t.attrs[k] = attr_from(attrs, k);
    attr_to(interpolate(t, x, y, cov), out);
}

uint32_t rzo_mask_set(uint32_t mask, uint32_t i, int v) { return mask_set(mask, (int)i, v != 0); }
int rzo_mask_get(uint32_t mask, uint32_t i) { return mask_get(mask, (int)i) ? 1 : 0; }

int rzo_try_clip(const float verts[12], const float attrs[18], float *out_verts, float *out_attrs,
                 int max_tris, int *n_tris) {
    Triangle t;
    for (int k = 0; k < 3; ++k) {
        t.v[k] = P4{verts[4 * k], verts[4 * k + 1], verts[4 * k + 2], verts[4 * k + 3]};
        t.a[k] = attr_from(attrs, k);
    }
    std::vector<Triangle> out;
    uint32_t flags = 0;
    ClipResult r = try_clip(t, out, &flags);
    int n = (int)out.size();
    if (n_tris) *n_tris = n;
    for (int i = 0; i < n && i < max_tris; ++i) {
        for (int k = 0; k < 3; ++k) {
            out_verts[i * 12 + 4 * k] = out[i].v[k].x;
            out_verts[i * 12 + 4 * k + 1] = out[i].v[k].y;
            out_verts[i * 12 + 4 * k + 2] = out[i].v[k].z;
            out_verts[i * 12 + 4 * k + 3] = out[i].v[k].w;
            attr_to(out[i].a[k], out_attrs + i * 18 + 6 * k);
        }
    }
    return (int)r;
}

void rzo_bbox(const float pts[6], uint64_t out[4]) {
    PixelBoundingBox b = bbox_from(pts);
    out[0] = b.min_x;
    out[1] = b.max_x;
    out[2] = b.min_y;
    out[3] = b.max_y;
}

uint32_t rzo_box_filter(const uint32_t c[4]) { return box_filter_color(c); }

rzo_tiles *rzo_tiles_new(uint32_t w, uint32_t h) { return new rzo_tiles(w, h); }
void rzo_tiles_free(rzo_tiles *t) { delete t; }
uint32_t rzo_tiles_len(const rzo_tiles *t) { return (uint32_t)t->tiles.size(); }
void rzo_tiles_get(const rzo_tiles *t, uint32_t i, uint64_t box[4]) {
    const PixelBoundingBox &b = t->tiles[i];
    box[0] = b.min_x;
    box[1] = b.max_x;
    box[2] = b.min_y;
    box[3] = b.max_y;
}
void rzo_tiles_mark(rzo_tiles *t, uint64_t row, uint64_t col) { t->mark(row, col); }
void rzo_tiles_next(rzo_tiles *t) { t->next(); }
uint32_t rzo_tiles_marked(const rzo_tiles *t) {  // buffers.rs:72-79
    uint32_t n = 0;
    for (uint8_t m : t->masks[t->mask_idx]) n += m ? 1 : 0;
    return n;
}
uint32_t rzo_tiles_prev_marked(const rzo_tiles *t) {  // buffers.rs:65-70
    uint32_t n = 0;
    for (uint8_t m : t->masks[(t->mask_idx + 1) % 2]) n += m ? 1 : 0;
    return n;
}
int rzo_tiles_mask(const rzo_tiles *t, uint32_t which, uint32_t idx) { return t->masks[which][idx]; }

uint32_t rzo_to_argb(const float rgba[4]) { return to_argb(Color{rgba[0], rgba[1], rgba[2], rgba[3]}); }
void rzo_from_rgba(const uint8_t rgba[4], float out[4]) {
    Color c = from_rgba(rgba);
    out[0] = c.r;
    out[1] = c.g;
    out[2] = c.b;
    out[3] = c.a;
}
void rzo_texture_sample(const uint8_t *rgba, uint32_t tw, uint32_t th, float u, float v,
                        float out[4], uint32_t *flags) {
    uint32_t f = 0;
    Color c = texture_sample(rgba, tw, th, u, v, &f);
    out[0] = c.r;
    out[1] = c.g;
    out[2] = c.b;
    out[3] = c.a;
    if (flags) *flags = f;
}

rzo_ctx *rzo_create(uint32_t width, uint32_t height) {
    if (width == 0 || height == 0) return nullptr;
    return new rzo_ctx(width, height);
}
void rzo_destroy(rzo_ctx *c) { delete c; }

void rzo_set_clear_depth(rzo_ctx *c, float depth) {
    // Applies to the current MSAA depth contents too: only valid between frames.
    c->clear_depth = depth;
    std::fill(c->depth.begin(), c->depth.end(), depth);
}

void rzo_set_uniforms(rzo_ctx *c, const float world[16], const float view[16], const float proj[16]) {
    memcpy(c->world, world, sizeof(c->world));
    memcpy(c->view, view, sizeof(c->view));
    memcpy(c->proj, proj, sizeof(c->proj));
}

// uniform.rs:29-33: index must equal the number bound so far
int rzo_bind_texture(rzo_ctx *c, uint32_t index, const uint8_t *rgba, uint32_t tw, uint32_t th) {
    if (index != c->textures.size()) return -2;
    if (tw == 0 || th == 0) return -2;
    Texture t;
    t.buf.assign(rgba, rgba + (size_t)tw * th * 4);
    t.width = tw;
    t.height = th;
    c->textures.push_back(std::move(t));
    return 0;
}

// render.rs:98-114 -> rasterizer/mod.rs:399-476
int rzo_render(rzo_ctx *c, const float *pos, const float *color, const float *uv, uint32_t nv,
               const uint32_t *idx, uint64_t ni, int vs, int fs) {
    std::vector<RasterizerTriangle> setups;
    DrawState ds;
    int rc = build_setups(c, pos, color, uv, nv, idx, ni, vs, fs, setups, ds);
    if (rc) return rc;
    for (RasterizerTriangle &t : setups) {
        PixelBoundingBox b = clamped_bbox(c, t);
        for (uint64_t i = b.min_y; i < b.max_y; ++i)
            for (uint64_t j = b.min_x; j < b.max_x; ++j) shade_pixel(c, t, j, i, ds, &c->flags, &c->tiles);
    }
    return 0;
}

// Same per-pixel sequence as rzo_render; rows of 64-pixel tiles are dealt
// round-robin to workers, so every pixel still sees the triangles in order.
int rzo_render_mt(rzo_ctx *c, const float *pos, const float *color, const float *uv, uint32_t nv,
                  const uint32_t *idx, uint64_t ni, int vs, int fs, int nthreads) {
    std::vector<RasterizerTriangle> setups;
    DrawState ds;
    int rc = build_setups(c, pos, color, uv, nv, idx, ni, vs, fs, setups, ds);
    if (rc) return rc;
    if (nthreads < 1) nthreads = 1;
    std::vector<uint32_t> wflags(nthreads, 0);
    auto worker = [&](int tid) {
        for (const RasterizerTriangle &src : setups) {
            PixelBoundingBox b = clamped_bbox(c, src);
            if (b.min_y >= b.max_y || b.min_x >= b.max_x) continue;
            RasterizerTriangle t = src;
            uint64_t band0 = b.min_y / TILE_SIZE, band1 = (b.max_y - 1) / TILE_SIZE;
            for (uint64_t band = band0; band <= band1; ++band) {
                if ((int)(band % nthreads) != tid) continue;
                uint64_t y0 = std::max(b.min_y, band * TILE_SIZE);
                uint64_t y1 = std::min(b.max_y, (band + 1) * TILE_SIZE);
                for (uint64_t i = y0; i < y1; ++i)
                    for (uint64_t j = b.min_x; j < b.max_x; ++j)
                        shade_pixel(c, t, j, i, ds, &wflags[tid], &c->tiles);
            }
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nthreads; ++t) pool.emplace_back(worker, t);
    worker(0);
    for (auto &th : pool) th.join();
    for (uint32_t f : wflags) c->flags |= f;
    return 0;
}

void rzo_read_msaa(const rzo_ctx *c, uint32_t *color4, float *depth4) {
    if (color4) memcpy(color4, c->color.data(), c->color.size() * sizeof(uint32_t));
    if (depth4) memcpy(depth4, c->depth.data(), c->depth.size() * sizeof(float));
}

// mod.rs:478-522
void rzo_framebuffer(rzo_ctx *c, uint32_t *out) {
    rzo_tiles &tl = c->tiles;
    const std::vector<uint8_t> &prev = tl.masks[(tl.mask_idx + 1) % 2];
    const std::vector<uint8_t> &cur = tl.masks[tl.mask_idx];
    for (size_t t = 0; t < tl.tiles.size(); ++t) {
        if (!prev[t]) continue;
        const PixelBoundingBox &b = tl.tiles[t];
        for (uint64_t y = b.min_y; y < b.max_y; ++y)
            for (uint64_t x = b.min_x; x < b.max_x; ++x) c->resolve[y * c->width + x] = CLEAR_COLOR;
    }
    for (size_t t = 0; t < tl.tiles.size(); ++t) {
        if (!cur[t]) continue;
        const PixelBoundingBox &b = tl.tiles[t];
        for (uint64_t y = b.min_y; y < b.max_y; ++y)
            for (uint64_t x = b.min_x; x < b.max_x; ++x) {
                size_t idx = y * c->width + x;
                c->resolve[idx] = box_filter_color(&c->color[idx * 4]);
                for (int s = 0; s < 4; ++s) {
                    c->color[idx * 4 + s] = CLEAR_COLOR;
                    c->depth[idx * 4 + s] = c->clear_depth;
                }
            }
    }
    tl.next();
    if (out) memcpy(out, c->resolve.data(), c->resolve.size() * sizeof(uint32_t));
}

uint32_t rzo_flags(const rzo_ctx *c) { return c->flags; }
void rzo_stats(const rzo_ctx *c, uint64_t out[5]) { memcpy(out, c->stats, sizeof(c->stats)); }

}  // extern "C"
