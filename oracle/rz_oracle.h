/*
 * rz_oracle.h -- CPU restatement of rusterizer's draw path, used ONLY as a
 * test oracle (pytest, the smoke check and the bench's CPU-baseline leg).
 * Nothing in the product library links or calls this code.
 *
 * Every function follows the reference operation-for-operation (same
 * association of every + and *, IEEE f32, no FMA contraction, Rust's
 * saturating float->int casts), citing the Rust source it restates.
 * Built with g++ -O2 -ffp-contract=off (oracle/Makefile).
 */
#ifndef RZ_ORACLE_H
#define RZ_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* shader selectors: src/main.rs:67-77 (FS) and :147-152 (the one VS) */
enum { RZO_VS_MVP = 0 };
enum { RZO_FS_TEXTURE = 0, RZO_FS_COLOR = 1, RZO_FS_DEBUG = 2 };

/* conditions where the reference panics; the oracle flags them instead */
#define RZO_FLAG_TEXEL_OOB 1u  /* read_texel past the buffer (texture.rs:51-61) */
#define RZO_FLAG_NO_TEXTURE 2u /* Texture FS with no texture 0 (uniform.rs:35-37) */
#define RZO_FLAG_CLIP_FEWER 4u /* SH output of 1..2 vertices (clipping.rs:183) */

/* ---- src/math/vector.rs, src/math/matrix.rs ---------------------------- */
float rzo_dot4(const float a[4], const float b[4]);
void rzo_mat_mul(const float a[16], const float b[16], float out[16]);
void rzo_mat_vec(const float m[16], const float v[4], float out[4]);
/* ((P*V)*W) as the VS closure computes it (main.rs:147-152) */
void rzo_mvp(const float world[16], const float view[16], const float proj[16], float out[16]);

/* ---- src/rasterizer/mod.rs pieces ------------------------------------- */
typedef struct {
    float pts[6];        /* edge_functions.points (x0,y0,x1,y1,x2,y2)  */
    float normals[6];    /* edge_functions.normals                      */
    float depths[3];     /* screen-space z                              */
    float depths_cs[3];  /* clip-space w ("depths_camera_space")        */
    float inv_2x_area;
} rzo_setup;

void rzo_perspective_divide(const float clip[12], float ndc[12]);
void rzo_viewport_setup(uint32_t w, uint32_t h, const float ndc[12], rzo_setup *out);
void rzo_setup_from_screen(const float scr[9], const float w3[3], rzo_setup *out);
void rzo_eval_single(const rzo_setup *s, float x, float y, float e[3]);
int rzo_inside(const rzo_setup *s, const float e[3]);
uint32_t rzo_coverage(const rzo_setup *s, uint32_t x, uint32_t y, float evals[12]);
void rzo_fragment_depths(const rzo_setup *s, uint32_t mask, const float evals[12], float depths[4]);
void rzo_interpolate(const rzo_setup *s, const float attrs[18], uint32_t x, uint32_t y,
                     uint32_t cov, float out[6]);
uint32_t rzo_mask_set(uint32_t mask, uint32_t i, int v);
int rzo_mask_get(uint32_t mask, uint32_t i);

/* src/rasterizer/clipping.rs: returns 0 Outside, 1 Inside, 2 Clipped.
 * For Clipped, writes up to max_tris triangles (12 floats + 18 attrs each). */
int rzo_try_clip(const float verts[12], const float attrs[18], float *out_verts,
                 float *out_attrs, int max_tris, int *n_tris);

/* src/rasterizer/bounding_box.rs (unclamped, saturating casts) */
void rzo_bbox(const float pts[6], uint64_t out[4]);

/* src/rasterizer/buffers.rs */
uint32_t rzo_box_filter(const uint32_t c[4]);
typedef struct rzo_tiles rzo_tiles;
rzo_tiles *rzo_tiles_new(uint32_t w, uint32_t h);
void rzo_tiles_free(rzo_tiles *t);
uint32_t rzo_tiles_len(const rzo_tiles *t);
void rzo_tiles_get(const rzo_tiles *t, uint32_t i, uint64_t box[4]);
void rzo_tiles_mark(rzo_tiles *t, uint64_t row, uint64_t col);
void rzo_tiles_next(rzo_tiles *t);
uint32_t rzo_tiles_marked(const rzo_tiles *t);
uint32_t rzo_tiles_prev_marked(const rzo_tiles *t);
int rzo_tiles_mask(const rzo_tiles *t, uint32_t which, uint32_t idx);

/* src/color.rs, src/texture.rs */
uint32_t rzo_to_argb(const float rgba[4]);
void rzo_from_rgba(const uint8_t rgba[4], float out[4]);
void rzo_texture_sample(const uint8_t *rgba, uint32_t tw, uint32_t th, float u, float v,
                        float out[4], uint32_t *flags);

/* ---- whole draw path: Renderer::render + Rasterizer::framebuffer ------- */
typedef struct rzo_ctx rzo_ctx;
rzo_ctx *rzo_create(uint32_t width, uint32_t height);
void rzo_destroy(rzo_ctx *c);
void rzo_set_clear_depth(rzo_ctx *c, float depth);
void rzo_set_uniforms(rzo_ctx *c, const float world[16], const float view[16], const float proj[16]);
int rzo_bind_texture(rzo_ctx *c, uint32_t index, const uint8_t *rgba, uint32_t tw, uint32_t th);
/* pos: nv*3, color: nv*4, uv: nv*2, idx: ni (multiple of 3). Returns 0 or <0. */
int rzo_render(rzo_ctx *c, const float *pos, const float *color, const float *uv, uint32_t nv,
               const uint32_t *idx, uint64_t ni, int vs, int fs);
/* pre-resolve MSAA dump: color4/depth4 are W*H*4 */
void rzo_read_msaa(const rzo_ctx *c, uint32_t *color4, float *depth4);
/* resolve + clear (mod.rs:478-522); out is W*H ARGB */
void rzo_framebuffer(rzo_ctx *c, uint32_t *out);
uint32_t rzo_flags(const rzo_ctx *c);
/* counters of the last render: [input tris, culled/outside, inside, clipped, sub-tris] */
void rzo_stats(const rzo_ctx *c, uint64_t out[5]);
/* tile-parallel variant of rzo_render: same output bits, nthreads workers
 * each owning whole rows of 64x64 screen tiles (submission order kept per pixel). */
int rzo_render_mt(rzo_ctx *c, const float *pos, const float *color, const float *uv, uint32_t nv,
                  const uint32_t *idx, uint64_t ni, int vs, int fs, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
