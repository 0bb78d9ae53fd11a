//! `AptRenderer`: Appearance's `NodeRenderer` on libapt, a drop-in for
//! `DistributedRenderer` (crates/appearance-distributed-renderer/src/lib.rs:60-82);
//! `apps/render-node` constructs it instead.  Built with the `node-renderer`
//! feature from a workspace that has the reference's crates.
#![cfg(feature = "node-renderer")]

use std::ffi::{c_void, CStr};

use appearance_render_loop::node::NodeRenderer;
use appearance_world::visible_world_action::VisibleWorldActionType;
use glam::UVec2;

use crate::*;

fn last_error() -> String {
    unsafe { CStr::from_ptr(apt_last_error()) }.to_string_lossy().into_owned()
}

/// The render-node renderer: one libapt renderer with the reference's
/// default configuration (PathTracerGpuConfig::default, src/lib.rs:230-245).
pub struct AptRenderer {
    raw: *mut apt_renderer,
}

impl AptRenderer {
    pub fn new() -> Self {
        let mut cfg = apt_config::default();
        let mut raw = std::ptr::null_mut();
        unsafe {
            apt_config_default(&mut cfg);
            let st = apt_create(&cfg, &mut raw);
            assert_eq!(st, APT_OK, "apt_create: {}", last_error());
        }
        Self { raw }
    }

    pub fn with_config(cfg: &apt_config) -> Result<Self, String> {
        let mut raw = std::ptr::null_mut();
        match unsafe { apt_create(cfg, &mut raw) } {
            APT_OK => Ok(Self { raw }),
            _ => Err(last_error()),
        }
    }

    /// Node::start_render without the transport: the host's StartRenderData
    /// bytes in, one call per 64x64 RGBA8 block (row, column block,
    /// frame_idx, pixels) out, in the order the node sends them.
    pub fn start_render<F: FnMut(u32, u32, u32, &[u8])>(&mut self, start_render_data: &[u8], mut block: F) {
        extern "C" fn tramp<F: FnMut(u32, u32, u32, &[u8])>(row: u32, col: u32, frame: u32, p: *const u8,
                                                            n: usize, user: *mut c_void) {
            let f = unsafe { &mut *(user as *mut F) };
            f(row, col, frame, unsafe { std::slice::from_raw_parts(p, n) });
        }
        let st = unsafe {
            apt_start_render(self.raw, start_render_data.as_ptr().cast(), start_render_data.len(), tramp::<F>,
                             (&mut block as *mut F).cast())
        };
        assert_eq!(st, APT_OK, "apt_start_render: {}", last_error());
    }

    /// Node::start_render with the node's compression: as start_render, but
    /// each block arrives as its JPEG file (`turbojpeg::compress(image,
    /// quality, Sub2x2)`, node.rs:65-79, compressed on the device), ready to
    /// be RenderPartialFinishedData::compressed_pixel_bytes.
    pub fn start_render_jpeg<F: FnMut(u32, u32, u32, &[u8])>(&mut self, start_render_data: &[u8], quality: i32,
                                                             mut block: F) {
        extern "C" fn tramp<F: FnMut(u32, u32, u32, &[u8])>(row: u32, col: u32, frame: u32, p: *const u8,
                                                            n: usize, user: *mut c_void) {
            let f = unsafe { &mut *(user as *mut F) };
            f(row, col, frame, unsafe { std::slice::from_raw_parts(p, n) });
        }
        let st = unsafe {
            apt_start_render_jpeg(self.raw, start_render_data.as_ptr().cast(), start_render_data.len(), quality,
                                  tramp::<F>, (&mut block as *mut F).cast())
        };
        assert_eq!(st, APT_OK, "apt_start_render_jpeg: {}", last_error());
    }
}

impl Default for AptRenderer {
    fn default() -> Self {
        Self::new()
    }
}

impl NodeRenderer for AptRenderer {
    fn visible_world_action(&mut self, action: &VisibleWorldActionType) {
        let ty: u32 = (*action).into();
        let bytes = action.as_bytes();
        let st = unsafe { apt_world_action(self.raw, ty, bytes.as_ptr().cast(), bytes.len()) };
        if st != APT_OK {
            log::warn!("apt_world_action({ty}): {}", last_error());
        }
    }

    fn render<F: FnMut(&[u8])>(&mut self, resolution: UVec2, start_row: u32, end_row: u32, mut result_callback: F) {
        extern "C" fn tramp<F: FnMut(&[u8])>(p: *const u8, n: usize, user: *mut c_void) {
            let f = unsafe { &mut *(user as *mut F) };
            f(unsafe { std::slice::from_raw_parts(p, n) });
        }
        let st = unsafe {
            apt_render_cb(self.raw, resolution.x, resolution.y, start_row, end_row, tramp::<F>,
                          (&mut result_callback as *mut F).cast())
        };
        // The reference panics on a failed render (e.g. crates/appearance-wgpu/src/lib.rs:305-307).
        assert_eq!(st, APT_OK, "apt_render_cb: {}", last_error());
    }
}

impl Drop for AptRenderer {
    fn drop(&mut self) {
        unsafe { apt_destroy(self.raw) }
    }
}
