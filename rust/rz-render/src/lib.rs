//! The draw API of rusterizer's `src/render.rs` and `src/uniform.rs`
//! (`Renderer::new`, `uniforms()`, `render(mesh, vertex_shader,
//! fragment_shader)`, `framebuffer()`; `Uniforms::bind_texture`,
//! `write_block`) running on the gfx950 rasterizer through `rz-sys`.
//!
//! Shaders are Rust fn pointers in the reference (`render.rs:33-36`); they
//! cannot run on the device, so they are enums naming the demo shaders of
//! `main.rs:67-77` and `main.rs:147-152`, compiled into the kernels. As in
//! the reference, invalid input panics (with the library's error message).
//! There is no window: `framebuffer()` returns the resolved 0xAARRGGBB image
//! that `display()` hands to minifb.

use std::ffi::CStr;
use std::ptr;

use rz_sys as sys;

/// Row-major 4x4 matrix, as `Mat4.array` (`src/math/matrix.rs:8-14`).
pub type Mat4 = [[f32; 4]; 4];

pub const IDENTITY: Mat4 = [[1.0, 0.0, 0.0, 0.0], [0.0, 1.0, 0.0, 0.0], [0.0, 0.0, 1.0, 0.0], [0.0, 0.0, 0.0, 1.0]];

/// `projection * view * world * v` (`main.rs:147-152`).
#[derive(Clone, Copy, Debug, PartialEq, Eq)]
pub enum VertexShader {
    Mvp,
}

/// The demo fragment shaders (`main.rs:67-77`).
#[derive(Clone, Copy, Debug, PartialEq, Eq)]
pub enum FragmentShader {
    /// `uniforms.get_texture(0).sample(uv)`
    Texture,
    /// the interpolated vertex colour
    Color,
    /// `Color::grayscale(frag_coords.depths[0])`
    Debug,
}

/// Per-vertex attributes (`graphics_primitives.rs:9-13`): RGBA colour and uv.
#[derive(Clone, Copy, Debug, Default)]
pub struct VertexAttribute {
    pub color: [f32; 4],
    pub uvs: [f32; 2],
}

/// A world-space mesh (`mesh.rs:5-12`).
#[derive(Clone, Debug, Default)]
pub struct Mesh {
    pub vertices: Vec<[f32; 3]>,
    pub indices: Vec<usize>,
    pub attributes: Vec<VertexAttribute>,
}

/// RGBA8 texels, row-major from the upper-left texel (`texture.rs:7-13`).
#[derive(Clone, Debug)]
pub struct Texture {
    pub width: usize,
    pub height: usize,
    pub rgba8: Vec<u8>,
}

/// `uniform.rs:5-9`.
#[derive(Clone, Debug)]
pub struct UniformBlock {
    pub world: Mat4,
    pub view: Mat4,
    pub projection: Mat4,
}

/// `uniform.rs:12-46`. Textures are uploaded to the device by the renderer
/// the next time it draws.
#[derive(Clone, Debug)]
pub struct Uniforms {
    textures: Vec<Texture>,
    uniform_block: UniformBlock,
}

impl Uniforms {
    fn new() -> Self {
        Uniforms {
            textures: Vec::new(),
            uniform_block: UniformBlock { world: IDENTITY, view: IDENTITY, projection: IDENTITY },
        }
    }

    /// Takes ownership; `index` must equal the number of bound textures.
    pub fn bind_texture(&mut self, index: usize, tex: Texture) {
        assert!(self.textures.len() == index, "texture index must equal the number of bound textures");
        assert!(tex.rgba8.len() == tex.width * tex.height * 4, "texture size does not match its texels");
        self.textures.push(tex);
    }

    pub fn get_texture(&self, index: usize) -> &Texture {
        &self.textures[index]
    }

    pub fn read_block(&self) -> &UniformBlock {
        &self.uniform_block
    }

    pub fn write_block(&mut self) -> &mut UniformBlock {
        &mut self.uniform_block
    }
}

pub struct Renderer {
    ctx: *mut sys::rz_ctx,
    uniforms: Uniforms,
    // meshes uploaded for the frame being drawn; freed after framebuffer()
    frame_meshes: Vec<*mut sys::rz_mesh>,
    fb: Vec<u32>,
    width: usize,
    height: usize,
}

fn check(rc: i32, ctx: *const sys::rz_ctx, what: &str) {
    if rc != sys::RZ_OK {
        let msg = unsafe { CStr::from_ptr(sys::rz_last_error(ctx)) }.to_string_lossy().into_owned();
        panic!("{what} failed ({rc}): {msg}");
    }
}

fn flat(m: &Mat4) -> [f32; 16] {
    let mut out = [0.0f32; 16];
    for (i, row) in m.iter().enumerate() {
        out[4 * i..4 * i + 4].copy_from_slice(row);
    }
    out
}

impl Renderer {
    /// Headless `Renderer::new` (`render.rs:48-69`) on the current HIP device.
    pub fn new(width: usize, height: usize) -> Renderer {
        let mut ctx = ptr::null_mut();
        check(unsafe { sys::rz_create(width as u32, height as u32, &mut ctx) }, ctx, "rz_create");
        Renderer {
            ctx,
            uniforms: Uniforms::new(),
            frame_meshes: Vec::new(),
            fb: vec![0; width * height],
            width,
            height,
        }
    }

    pub fn uniforms(&mut self) -> &mut Uniforms {
        &mut self.uniforms
    }

    pub fn width(&self) -> usize {
        self.width
    }

    pub fn height(&self) -> usize {
        self.height
    }

    /// `render.rs:98-114`: the draw is recorded with the current uniforms
    /// and rendered, in order after earlier draws, at `framebuffer()`.
    pub fn render(&mut self, mesh: &Mesh, vertex_shader: VertexShader, fragment_shader: FragmentShader) {
        self.upload_textures();
        let b = &self.uniforms.uniform_block;
        let (w, v, p) = (flat(&b.world), flat(&b.view), flat(&b.projection));
        check(unsafe { sys::rz_set_uniforms(self.ctx, w.as_ptr(), v.as_ptr(), p.as_ptr()) }, self.ctx, "rz_set_uniforms");
        let m = self.upload(mesh);
        self.frame_meshes.push(m);
        let vs = match vertex_shader {
            VertexShader::Mvp => sys::RZ_VS_MVP,
        };
        let fs = match fragment_shader {
            FragmentShader::Texture => sys::RZ_FS_TEXTURE,
            FragmentShader::Color => sys::RZ_FS_COLOR,
            FragmentShader::Debug => sys::RZ_FS_DEBUG,
        };
        check(unsafe { sys::rz_draw(self.ctx, m, vs, fs) }, self.ctx, "rz_draw");
    }

    /// `Rasterizer::framebuffer` (`mod.rs:478-522`): resolves the frame and
    /// clears the MSAA buffers; width*height 0xAARRGGBB words.
    pub fn framebuffer(&mut self) -> &[u32] {
        check(unsafe { sys::rz_resolve(self.ctx, self.fb.as_mut_ptr()) }, self.ctx, "rz_resolve");
        for m in self.frame_meshes.drain(..) {
            check(unsafe { sys::rz_mesh_destroy(self.ctx, m) }, self.ctx, "rz_mesh_destroy");
        }
        &self.fb
    }

    fn upload_textures(&mut self) {
        let bound = unsafe { sys::rz_texture_count(self.ctx) } as usize;
        for (i, t) in self.uniforms.textures.iter().enumerate().skip(bound) {
            let rc = unsafe { sys::rz_bind_texture(self.ctx, i as u32, t.rgba8.as_ptr(), t.width as u32, t.height as u32) };
            check(rc, self.ctx, "rz_bind_texture");
        }
    }

    fn upload(&mut self, mesh: &Mesh) -> *mut sys::rz_mesh {
        // the reference indexes the vertex buffer with usize (render.rs:78-83);
        // indices past u32 cannot address a vertex of a mesh this size
        let idx: Vec<u32> = mesh.indices.iter().map(|&i| u32::try_from(i).expect("index exceeds u32")).collect();
        let pos: Vec<f32> = mesh.vertices.iter().flat_map(|v| v.iter().copied()).collect();
        let col: Vec<f32> = mesh.attributes.iter().flat_map(|a| a.color.iter().copied()).collect();
        let uv: Vec<f32> = mesh.attributes.iter().flat_map(|a| a.uvs.iter().copied()).collect();
        assert!(mesh.attributes.len() == mesh.vertices.len(), "one attribute per vertex");
        let mut out = ptr::null_mut();
        let rc = unsafe {
            sys::rz_mesh_create(
                self.ctx,
                pos.as_ptr(),
                col.as_ptr(),
                uv.as_ptr(),
                mesh.vertices.len() as u32,
                idx.as_ptr(),
                idx.len() as u64,
                &mut out,
            )
        };
        check(rc, self.ctx, "rz_mesh_create");
        out
    }
}

impl Drop for Renderer {
    fn drop(&mut self) {
        unsafe {
            if !self.frame_meshes.is_empty() {
                // a recorded draw holds its mesh until the frame is resolved
                let _ = sys::rz_resolve(self.ctx, self.fb.as_mut_ptr());
            }
            for m in self.frame_meshes.drain(..) {
                let _ = sys::rz_mesh_destroy(self.ctx, m);
            }
            sys::rz_destroy(self.ctx);
        }
    }
}
