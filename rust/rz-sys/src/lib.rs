//! Raw bindings of `include/rz.h`, the C ABI of the gfx950 rasterizer
//! (`librz.so`). Every function returns 0 or a negative `RZ_ERR_*` status;
//! `rz_last_error` describes the last failure of a context. See `rz.h` for
//! the contract of each entry point and the reference item it replaces.
#![allow(non_camel_case_types)]

use std::os::raw::{c_char, c_int, c_void};

/// Opaque rendering context (one HIP stream and its pipeline buffers).
#[repr(C)]
pub struct rz_ctx {
    _opaque: [u8; 0],
}

/// Opaque device-resident mesh.
#[repr(C)]
pub struct rz_mesh {
    _opaque: [u8; 0],
}

pub const RZ_VS_MVP: c_int = 0;

pub const RZ_FS_TEXTURE: c_int = 0;
pub const RZ_FS_COLOR: c_int = 1;
pub const RZ_FS_DEBUG: c_int = 2;

pub const RZ_OK: c_int = 0;
pub const RZ_ERR_ARG: c_int = -1;
pub const RZ_ERR_HIP: c_int = -2;
pub const RZ_ERR_STATE: c_int = -3;
pub const RZ_ERR_NOMEM: c_int = -4;
pub const RZ_ERR_LIMIT: c_int = -5;

pub const RZ_FLAG_TEXEL_OOB: u32 = 1;
pub const RZ_FLAG_CLIP_FEWER: u32 = 4;

pub const RZ_STAGE_GEOMETRY: usize = 0;
pub const RZ_STAGE_BINNING: usize = 1;
pub const RZ_STAGE_RASTER: usize = 2;
pub const RZ_NSTAGES: usize = 3;

/// One scene with one draw, for `rz_batch_add` (232 bytes, the C layout).
#[repr(C)]
#[derive(Clone, Copy)]
pub struct rz_batch_item {
    pub width: u32,
    pub height: u32,
    pub clear_depth: f32,
    pub texture: u32,
    pub vertex_shader: i32,
    pub fragment_shader: i32,
    pub argb_dev: *mut u32,
    pub mesh: *const rz_mesh,
    pub world: [f32; 16],
    pub view: [f32; 16],
    pub proj: [f32; 16],
}

const _: () = assert!(std::mem::size_of::<rz_batch_item>() == 232);

extern "C" {
    // context
    pub fn rz_create(width: u32, height: u32, out: *mut *mut rz_ctx) -> c_int;
    pub fn rz_create_shared(ctx: *mut rz_ctx, width: u32, height: u32, out: *mut *mut rz_ctx) -> c_int;
    pub fn rz_destroy(ctx: *mut rz_ctx) -> c_int;
    pub fn rz_last_error(ctx: *const rz_ctx) -> *const c_char;
    pub fn rz_stream(ctx: *const rz_ctx) -> *mut c_void;
    pub fn rz_set_cu_mask(ctx: *mut rz_ctx, mask: *const u32, words: u32) -> c_int;
    pub fn rz_set_clear_depth(ctx: *mut rz_ctx, depth: f32) -> c_int;
    pub fn rz_set_uniforms(ctx: *mut rz_ctx, world: *const f32, view: *const f32, proj: *const f32) -> c_int;
    pub fn rz_bind_texture(ctx: *mut rz_ctx, index: u32, rgba8: *const u8, tw: u32, th: u32) -> c_int;
    pub fn rz_texture_count(ctx: *const rz_ctx) -> u32;

    // meshes
    pub fn rz_mesh_create(
        ctx: *mut rz_ctx,
        pos: *const f32,
        color: *const f32,
        uv: *const f32,
        nv: u32,
        idx: *const u32,
        ni: u64,
        out: *mut *mut rz_mesh,
    ) -> c_int;
    pub fn rz_mesh_create_device(
        ctx: *mut rz_ctx,
        pos: *const f32,
        color: *const f32,
        uv: *const f32,
        nv: u32,
        idx: *const u32,
        ni: u64,
        out: *mut *mut rz_mesh,
    ) -> c_int;
    pub fn rz_mesh_destroy(ctx: *mut rz_ctx, mesh: *mut rz_mesh) -> c_int;

    // draws and frames
    pub fn rz_draw(ctx: *mut rz_ctx, mesh: *const rz_mesh, vertex_shader: c_int, fragment_shader: c_int) -> c_int;
    pub fn rz_resolve(ctx: *mut rz_ctx, argb_out: *mut u32) -> c_int;
    pub fn rz_resolve_device(ctx: *mut rz_ctx, argb_dev: *mut u32) -> c_int;
    pub fn rz_resolve_msaa(ctx: *mut rz_ctx, argb_out: *mut u32, color4_out: *mut u32, depth4_out: *mut f32)
        -> c_int;
    pub fn rz_frame_flags(ctx: *const rz_ctx) -> u32;
    pub fn rz_frame_stats(ctx: *const rz_ctx, out: *mut u64) -> c_int;

    // batches
    pub fn rz_batch_begin(ctx: *mut rz_ctx) -> c_int;
    pub fn rz_batch_scene(
        ctx: *mut rz_ctx,
        width: u32,
        height: u32,
        clear_depth: f32,
        argb_dev: *mut u32,
        scene_out: *mut u32,
    ) -> c_int;
    pub fn rz_batch_scene_msaa(ctx: *mut rz_ctx, scene: u32, color4_dev: *mut u32, depth4_dev: *mut f32) -> c_int;
    pub fn rz_batch_draw(
        ctx: *mut rz_ctx,
        scene: u32,
        mesh: *const rz_mesh,
        world: *const f32,
        view: *const f32,
        proj: *const f32,
        texture: u32,
        vertex_shader: c_int,
        fragment_shader: c_int,
    ) -> c_int;
    pub fn rz_batch_add(ctx: *mut rz_ctx, items: *const rz_batch_item, n: u32, first_scene: *mut u32) -> c_int;
    pub fn rz_batch_submit(ctx: *mut rz_ctx) -> c_int;
    pub fn rz_batch_query(ctx: *mut rz_ctx) -> c_int;
    pub fn rz_batch_wait(ctx: *mut rz_ctx) -> c_int;
    pub fn rz_batch_scene_flags(ctx: *const rz_ctx, scene: u32, flags: *mut u32) -> c_int;
    pub fn rz_batch_tuning(ctx: *mut rz_ctx, bin_target: u32, raster_per_cu: u32, use_graph: c_int) -> c_int;
    pub fn rz_batch_get_tuning(ctx: *const rz_ctx, out: *mut u32) -> c_int;

    // persistent raster service (no reference counterpart; for servers
    // that keep several contexts of one group busy)
    pub fn rz_service_start(ctx: *mut rz_ctx, wg_per_cu: u32, idle_us: u32) -> c_int;
    pub fn rz_service_stop(ctx: *mut rz_ctx) -> c_int;
    pub fn rz_service_attach(ctx: *mut rz_ctx, on: c_int) -> c_int;
    pub fn rz_service_stats(ctx: *const rz_ctx, out: *mut u64) -> c_int;

    // allocation-free frame (no reference counterpart; for callers that
    // cannot allocate device memory)
    pub fn rz_render_noalloc(
        stream: *mut c_void,
        width: u32,
        height: u32,
        clear_depth: f32,
        world: *const f32,
        view: *const f32,
        proj: *const f32,
        pos_dev: *const f32,
        color_dev: *const f32,
        uv_dev: *const f32,
        nv: u32,
        idx_dev: *const u32,
        ni: u64,
        texture_dev: *const u32,
        tw: u32,
        th: u32,
        vertex_shader: c_int,
        fragment_shader: c_int,
        argb_dev: *mut u32,
        flags_out: *mut u32,
        message: *mut *const c_char,
    ) -> c_int;

    // tuning and timing
    pub fn rz_set_coarse_bin(ctx: *mut rz_ctx, pixels: u32) -> c_int;
    pub fn rz_memory(ctx: *const rz_ctx, out: *mut u64) -> c_int;
    pub fn rz_batch_need(ctx: *const rz_ctx, items: *const rz_batch_item, n: u32, bytes: *mut u64) -> c_int;
    pub fn rz_trim(ctx: *mut rz_ctx) -> c_int;
    pub fn rz_set_fine_tile(ctx: *mut rz_ctx, side: u32) -> c_int;
    pub fn rz_get_fine_tile(ctx: *const rz_ctx) -> u32;
    pub fn rz_profile(ctx: *mut rz_ctx, enable: c_int) -> c_int;
    pub fn rz_profile_read(ctx: *const rz_ctx, ms: *mut f32, frames: *mut u32) -> c_int;
}
