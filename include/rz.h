/*
 * rz.h -- C ABI of the MI355X (gfx950) rasterizer: clip-space vertices to the
 * resolved framebuffer, 4x RGSS MSAA, strict-LESS depth, perspective-correct
 * attributes, bilinear RGBA8 textures, box-filter resolve.
 *
 * It is the drop-in for rusterizer's draw API. Each entry point names the
 * reference item it replaces (paths relative to the rusterizer crate).
 *
 * Conventions
 *  - Every call returns 0 on success or a negative RZ_ERR_* code; nothing
 *    aborts. rz_last_error() gives the message of the last failure.
 *  - One context per host thread; contexts are independent. All work of a
 *    context runs in order on its own HIP stream (rz_stream()). Contexts of
 *    one group (rz_create_shared) share its meshes, textures and raster
 *    service: calls on them are made from one host thread, or serialized.
 *  - Host arrays passed in are copied during the call; meshes and textures
 *    are device-resident and owned by the context until destroyed.
 *  - Draws accumulate into the frame until rz_resolve*, which resolves the
 *    4x MSAA buffers to ARGB and clears them, as Rasterizer::framebuffer does.
 */
#ifndef RZ_H
#define RZ_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rz_ctx rz_ctx;
typedef struct rz_mesh rz_mesh;

/* Vertex shaders (src/main.rs:147-152: projection * view * world * v). */
enum rz_vertex_shader { RZ_VS_MVP = 0 };

/* Fragment shaders (src/main.rs:67-77). */
enum rz_fragment_shader {
    RZ_FS_TEXTURE = 0, /* uniforms.get_texture(0).sample(uv)        */
    RZ_FS_COLOR = 1,   /* interpolated vertex colour                 */
    RZ_FS_DEBUG = 2    /* Color::grayscale(frag_coords.depths[0])    */
};

enum rz_status {
    RZ_OK = 0,
    RZ_ERR_ARG = -1,      /* bad argument (null, size, enum, index)     */
    RZ_ERR_HIP = -2,      /* HIP runtime error                           */
    RZ_ERR_STATE = -3,    /* call not valid now (e.g. missing texture)   */
    RZ_ERR_NOMEM = -4,    /* device allocation failed                    */
    RZ_ERR_LIMIT = -5     /* a documented limit was exceeded             */
};

/* Bits reported by rz_frame_flags() for the last resolved frame: inputs on
 * which the reference would panic; the device applies a defined behaviour. */
#define RZ_FLAG_TEXEL_OOB 1u   /* texel index past the texture: last texel used (texture.rs:51-61) */
#define RZ_FLAG_CLIP_FEWER 4u  /* clipper returned 1-2 vertices: no triangle (clipping.rs:183)   */

/* Pipeline stages timed by rz_profile_read(). */
enum rz_stage { RZ_STAGE_GEOMETRY = 0, RZ_STAGE_BINNING = 1, RZ_STAGE_RASTER = 2, RZ_NSTAGES = 3 };

/* ------------------------------------------------------------------ context */

/* Renderer::new / Rasterizer::new (src/render.rs:48-69, src/rasterizer/mod.rs:273-281),
 * headless: no window. width, height in [1, 16384] and the screen must fit 256
 * bins of 256x256 pixels (e.g. up to 4096x4096 or 3840x2160), else RZ_ERR_LIMIT.
 * Uses the current HIP device. */
int rz_create(uint32_t width, uint32_t height, rz_ctx **out);
/* Environment read by rz_create / rz_create_shared (debugging aids):
 *   RZ_RASTER_STATS=1  count raster work per frame (rz_frame_stats; slower)
 *   RZ_CLIP_WIDE=1     clip each polygon with a full 64-lane wave instead of
 *                      8 lanes (same results; exercises the wide clipper)
 *   RZ_EXACT_PATHS=1   use none of the per-triangle proofs that select the
 *                      raster's fast paths (finite barycentric clamps, texel
 *                      reads without bounds checks): same results, slower;
 *                      exercises the exact paths */
/* A further context on the same device sharing ctx's meshes and textures
 * (no reference counterpart: lets several jobs -- frames or batches -- be in
 * flight at once on their own streams over the same resident data). The
 * shared data lives until the last context of the group is destroyed. */
int rz_create_shared(rz_ctx *ctx, uint32_t width, uint32_t height, rz_ctx **out);
/* Drop for Renderer. Accepts NULL. Waits for the context's work. */
int rz_destroy(rz_ctx *ctx);
/* Message of the last failed call on ctx ("" if none); static text for NULL ctx. */
const char *rz_last_error(const rz_ctx *ctx);
/* The HIP stream (hipStream_t) all work of ctx runs on. */
void *rz_stream(const rz_ctx *ctx);
/* Restricts the CUs ctx's work may run on (no reference counterpart: lets a
 * server keep some CUs free of long jobs, so that short jobs on another
 * context do not wait behind their workgroups). mask holds `words` 32-bit
 * words, bit i of word w selecting CU 32*w + i, as for
 * hipExtStreamCreateWithCUMask; words = 0 lifts the restriction. ctx gets a
 * new stream (rz_stream changes), so it must be idle: no frame or batch
 * recorded or in flight (RZ_ERR_STATE). RZ_ERR_ARG if the mask selects no CU
 * of the device. Results do not depend on the mask. */
int rz_set_cu_mask(rz_ctx *ctx, const uint32_t *mask, uint32_t words);

/* Depth the MSAA buffers are cleared to (CLEAR_DEPTH, src/rasterizer/buffers.rs:6;
 * default FLT_MAX). Takes effect from the next frame. */
int rz_set_clear_depth(rz_ctx *ctx, float depth);

/* Uniforms::write_block (src/uniform.rs:43-45): world, view, projection as
 * row-major f32[16] (Mat4.array, src/math/matrix.rs:8-14). Applies to later draws. */
int rz_set_uniforms(rz_ctx *ctx, const float world[16], const float view[16], const float proj[16]);

/* Uniforms::bind_texture (src/uniform.rs:29-33): index must equal the number
 * of textures bound so far. rgba8 is th rows of tw texels, (0,0) top-left
 * (src/texture.rs:47-63); copied to the device. */
int rz_bind_texture(rz_ctx *ctx, uint32_t index, const uint8_t *rgba8, uint32_t tw, uint32_t th);
/* Number of textures bound in ctx's group (the next valid bind index). */
uint32_t rz_texture_count(const rz_ctx *ctx);

/* --------------------------------------------------------------------- mesh */

/* Mesh<WorldSpace> (src/mesh.rs:5-12) uploaded once and kept resident:
 * pos nv*3 f32, color nv*4 f32 (or NULL: zeros), uv nv*2 f32 (or NULL: zeros),
 * idx ni u32 (ni a multiple of 3, every index < nv). Host arrays are copied. */
int rz_mesh_create(rz_ctx *ctx, const float *pos, const float *color, const float *uv, uint32_t nv,
                   const uint32_t *idx, uint64_t ni, rz_mesh **out);
/* Same, from device pointers already in HBM (copied device-to-device). */
int rz_mesh_create_device(rz_ctx *ctx, const float *pos, const float *color, const float *uv,
                          uint32_t nv, const uint32_t *idx, uint64_t ni, rz_mesh **out);
/* Fails with RZ_ERR_STATE while a draw of the mesh waits for rz_resolve* or a
 * batch that uses it is unfinished. */
int rz_mesh_destroy(rz_ctx *ctx, rz_mesh *mesh);

/* -------------------------------------------------------------------- draws */

/* Renderer::render(mesh, vertex_shader, fragment_shader) (src/render.rs:98-114):
 * records the draw with the current uniforms; its triangles are processed in
 * index order, after earlier draws of the frame, when the frame is resolved. */
int rz_draw(rz_ctx *ctx, const rz_mesh *mesh, int vertex_shader, int fragment_shader);

/* Rasterizer::framebuffer (src/rasterizer/mod.rs:478-522): rasterizes all
 * queued draws in order, resolves the 4x MSAA buffers into width*height
 * 0xAARRGGBB words, clears them for the next frame. Blocks until done. */
int rz_resolve(rz_ctx *ctx, uint32_t *argb_out);
/* Same, writing to device memory (width*height u32). Returns after the frame
 * finished on the device. */
int rz_resolve_device(rz_ctx *ctx, uint32_t *argb_dev);
/* Debug variant: also copies the pre-resolve MSAA buffers to host memory
 * (width*height*4 colour words / depths, sample-major per pixel;
 * ColorBuffer/DepthBuffer of src/rasterizer/buffers.rs:82-157). Any pointer may be NULL. */
int rz_resolve_msaa(rz_ctx *ctx, uint32_t *argb_out, uint32_t *color4_out, float *depth4_out);
/* RZ_FLAG_* bits raised while rendering the last resolved frame (or, after a
 * batch, by any of its scenes). */
uint32_t rz_frame_flags(const rz_ctx *ctx);
/* Counts of the last resolved frame or batch: [input triangles, triangles clipped,
 * rasterizer triangles, coarse bin entries, frame replays, then -- only when
 * the context was created with RZ_RASTER_STATS=1 in the environment, else 0 --
 * raster work: tile-list entries scanned and candidates left by the list
 * culls (both summed over the 4 quadrant waves of a tile), prims processed
 * per quadrant, MSAA samples they covered, and shading tasks (distinct
 * surviving prims of a pixel, each shaded once), then the buffer-overflow
 * reasons that caused replays
 * (bits: 1 triangles, 2 clip slots, 4 coarse entries, 16 fine entries)]. */
int rz_frame_stats(const rz_ctx *ctx, uint64_t out[11]);

/* ------------------------------------------------------------------ batches
 * Many independent scenes rendered by one job (no reference counterpart: the
 * reference renders one frame per Renderer). Each scene is a render target
 * with its own size, clear depth and draws; a scene's draws are rendered in
 * the order they were recorded, exactly as rz_draw/rz_resolve would render
 * them in a context of that size, so every scene's output is identical to
 * rendering it alone. The batch runs as one launch sequence (captured once
 * as a HIP graph and replayed while the buffers fit). */

/* Start recording a batch (no frame may be pending, no batch in flight). */
int rz_batch_begin(rz_ctx *ctx);
/* Add a scene: width x height (same limits as rz_create), clear depth, and
 * device memory for its width*height resolved 0xAARRGGBB words. Returns its
 * index in *scene_out (may be NULL). */
int rz_batch_scene(rz_ctx *ctx, uint32_t width, uint32_t height, float clear_depth, uint32_t *argb_dev,
                   uint32_t *scene_out);
/* Debug: also write the scene's pre-resolve MSAA colours / depths
 * (width*height*4 each, sample-major per pixel, as rz_resolve_msaa) to
 * device memory. Either pointer may be NULL. */
int rz_batch_scene_msaa(rz_ctx *ctx, uint32_t scene, uint32_t *color4_dev, float *depth4_dev);
/* Renderer::render into a scene with explicit uniforms (row-major world,
 * view, projection) and texture index (for RZ_FS_TEXTURE; see
 * rz_bind_texture). */
int rz_batch_draw(rz_ctx *ctx, uint32_t scene, const rz_mesh *mesh, const float world[16], const float view[16],
                  const float proj[16], uint32_t texture, int vertex_shader, int fragment_shader);
/* One scene with one draw, for rz_batch_add. */
typedef struct rz_batch_item {
    uint32_t width, height;   /* scene size (same limits as rz_create)          */
    float clear_depth;
    uint32_t texture;         /* texture index for RZ_FS_TEXTURE                */
    int32_t vertex_shader, fragment_shader;
    uint32_t *argb_dev;       /* width*height resolved 0xAARRGGBB words, device */
    const rz_mesh *mesh;
    float world[16], view[16], proj[16]; /* row-major uniforms                 */
} rz_batch_item;

/* Append n scenes, each drawn once, in one call (rz_batch_scene +
 * rz_batch_draw per item). Either all items are added or none (on error).
 * *first_scene (may be NULL) receives the index of the first one. */
int rz_batch_add(rz_ctx *ctx, const rz_batch_item *items, uint32_t n, uint32_t *first_scene);
/* Queue the batch on the context stream and return. */
int rz_batch_submit(rz_ctx *ctx);
/* 1 if the submitted batch has finished (or none is in flight), 0 if it is
 * still running, < 0 on error. Never blocks except to grow buffers. */
int rz_batch_query(rz_ctx *ctx);
/* Block until the submitted batch has finished. */
int rz_batch_wait(rz_ctx *ctx);
/* RZ_FLAG_* bits of one scene of the last finished batch. */
int rz_batch_scene_flags(const rz_ctx *ctx, uint32_t scene, uint32_t *flags);
/* Batch tuning: bin_target = most coarse bins per scene when choosing the
 * bin side automatically (1..256); raster_per_cu = persistent raster
 * workgroups per CU (0: one single-wave workgroup per 8x8 quadrant);
 * use_graph = replay the captured HIP graph (1) or launch the kernels
 * directly (0), for batches and for frames of at most 65536 triangles
 * (while profiling is on, frames launch directly). Between jobs. */
int rz_batch_tuning(rz_ctx *ctx, uint32_t bin_target, uint32_t raster_per_cu, int use_graph);
/* The current tuning: [bin_target, raster_per_cu, use_graph]. */
int rz_batch_get_tuning(const rz_ctx *ctx, uint32_t out[3]);

/* ------------------------------------------------------------ raster service
 * No counterpart in the reference (it renders one frame at a time on the
 * calling thread). For servers that keep several contexts of one group busy
 * at once: a persistent raster kernel stays resident on a stream of its own
 * and renders the tiles of every batch that an attached context submits.
 * The context's own stream runs the batch up to binning and then posts the
 * job to the service through device memory; the service's workgroups take
 * tiles from the posted job with the fewest tiles left, so a small job is
 * not held up behind a long one, and jobs of several contexts render at
 * the same time in one kernel. Results are identical to the context's own
 * raster. rz_batch_query / rz_batch_wait report a batch done once the
 * service has finished it.
 *
 * rz_service_start launches the service of ctx's group with wg_per_cu
 * (1..4) workgroups per CU -- the rest of each CU stays free for the
 * contexts' front ends -- that exit after idle_us (1 us .. 5 s) in which no
 * tile was claimed, and after 4 s in any case; the library relaunches the
 * service when work is posted and it has exited. RZ_ERR_STATE if it is
 * running. rz_service_stop asks it to stop and waits (RZ_ERR_STATE while an
 * attached context has a batch in flight). rz_service_attach(ctx, 1) routes
 * ctx's batches (not its plain frames) through the running service (at most
 * 16 contexts per group; RZ_ERR_LIMIT beyond; not with RZ_RASTER_STATS);
 * rz_service_attach(ctx, 0) detaches it. Both between jobs. A batch of more
 * than 16,769,023 16x16 tiles cannot be posted (RZ_ERR_LIMIT).
 * rz_service_stats: out[0] started, out[1] a launch of it is running now,
 * out[2] launches so far, out[3] workgroups per launch.
 * The service runs on a CU-masked stream (a hardware queue of its own),
 * which HIP creates as a blocking stream: work on the legacy null stream and
 * device-wide synchronisation (hipDeviceSynchronize) wait until the service
 * exits, up to idle_us after its last tile. Keep other work on non-blocking
 * streams while it runs; the library's own copies from host memory and
 * to host memory use the context's (non-blocking) stream, while
 * rz_mesh_create_device copies on the null stream, after the caller's work
 * there. */
int rz_service_start(rz_ctx *ctx, uint32_t wg_per_cu, uint32_t idle_us);
int rz_service_stop(rz_ctx *ctx);
int rz_service_attach(rz_ctx *ctx, int on);
int rz_service_stats(const rz_ctx *ctx, uint64_t out[4]);

/* ------------------------------------------------------- allocation-free frame
 * No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): one frame of one draw, rendered exactly as
 * rz_set_uniforms + rz_draw + rz_resolve_device would render it in a fresh
 * context of that size, with every working buffer taken from a fixed
 * 128 MiB device array that the library holds per GPU until the process
 * ends. The call makes no device allocation and no host-to-device copy.
 *
 * Inputs are device pointers: pos nv*3 f32, color nv*4 f32 (or NULL:
 * zeros), uv nv*2 f32 (or NULL: zeros), idx ni u32 (ni a multiple of 3, at
 * most 3 * 262144, every index < nv, checked on the device), the texture as
 * th rows of tw RGBA8 texels (for RZ_FS_TEXTURE; else may be NULL); world,
 * view, proj are host row-major f32[16]; argb_dev receives width*height
 * 0xAARRGGBB words. Work runs on `stream` (a hipStream_t; NULL = the null
 * stream); the call returns after the frame finished there. *flags_out (may
 * be NULL) receives the frame's RZ_FLAG_* bits; *message (may be NULL) a
 * static description of a failure ("" on success).
 *
 * Errors: RZ_ERR_ARG / RZ_ERR_STATE as rz_draw and rz_resolve_device;
 * RZ_ERR_LIMIT when the screen or the triangle count is beyond the limits
 * above, or when the frame's primitives or bin lists outgrow the array
 * (rz_resolve_device grows its buffers instead); the output is then not
 * defined. Calls on one GPU share its array: calls that overlap in time on
 * the same GPU (two host threads, or two streams) race and must not be
 * made. */
int rz_render_noalloc(void *stream, uint32_t width, uint32_t height, float clear_depth, const float world[16],
                      const float view[16], const float proj[16], const float *pos_dev, const float *color_dev,
                      const float *uv_dev, uint32_t nv, const uint32_t *idx_dev, uint64_t ni,
                      const uint32_t *texture_dev, uint32_t tw, uint32_t th, int vertex_shader,
                      int fragment_shader, uint32_t *argb_dev, uint32_t *flags_out, const char **message);

/* ------------------------------------------------------------ tuning/timing */

/* Coarse bin side in pixels: a multiple of 16 up to 256 with at most 256 bins
 * on the screen, or 0 = automatic. Not inside a frame. */
int rz_set_coarse_bin(rz_ctx *ctx, uint32_t pixels);
/* Device memory in bytes (no reference counterpart: lets a server admit
 * work against what the device really holds):
 *   out[0]  held by ctx now: its pipeline buffers, job tables and outputs
 *           (buffers grow to fit the jobs run so far and are kept);
 *   out[1]  the meshes and textures resident in ctx's group;
 *   out[2]  what ctx will hold once the frame or batch being recorded has
 *           run, by the same estimate the library sizes its buffers with
 *           (a job that outgrows the estimate grows them further when it
 *           runs, which out[0] then shows); out[0] when nothing is recorded. */
int rz_memory(const rz_ctx *ctx, uint64_t out[3]);
/* What ctx would hold (as out[2] above) after running a batch of exactly
 * these n items with its current tuning, without recording anything: lets
 * a server size a batch to its memory budget before it records it. */
int rz_batch_need(const rz_ctx *ctx, const rz_batch_item *items, uint32_t n, uint64_t *bytes);
/* Frees ctx's pipeline buffers (they grow again with the next job, which
 * then re-captures its graph): lets a server give memory back under
 * pressure. Between jobs; meshes, textures and tables stay. */
int rz_trim(rz_ctx *ctx);
/* Side in pixels (8, 16 or 32; 0 = 32) of the screen tiles whose primitive
 * lists binning builds and the raster walks (no reference counterpart: the
 * reference walks each triangle's bounding box; it tunes work per list
 * against list count). The raster still works on 16x16 tiles, one wave per
 * 8x8 quadrant: with 8 each wave walks its own quadrant's list, with 32 four
 * raster tiles share one list. A scene whose coarse bins cannot hold tiles of
 * that side (more than 256 tiles per bin, or a side that does not divide the
 * bin) uses 16. Results do not depend on it. Between jobs; frames and
 * batches. */
int rz_set_fine_tile(rz_ctx *ctx, uint32_t side);
/* The side last set (32 by default). */
uint32_t rz_get_fine_tile(const rz_ctx *ctx);
/* Enable (1) / disable (0) per-stage HIP event timing; enabling resets sums. */
int rz_profile(rz_ctx *ctx, int enable);
/* Summed milliseconds per stage (RZ_NSTAGES floats) and the number of frames
 * or batches. A batch replayed from its graph is timed as a whole, under
 * RZ_STAGE_RASTER. */
int rz_profile_read(const rz_ctx *ctx, float ms[RZ_NSTAGES], uint32_t *frames);

#ifdef __cplusplus
}
#endif
#endif
