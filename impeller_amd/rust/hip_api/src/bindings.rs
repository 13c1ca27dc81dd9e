//! The `impeller_rt_*` declarations of include/impeller_hip.h.
#![allow(non_camel_case_types)]
use std::ffi::{c_char, c_int, c_void};

pub type hipError_t = c_int;
pub type hipStream_t = *mut c_void;
pub type hipEvent_t = *mut c_void;
pub const HIP_SUCCESS: hipError_t = 0;

#[repr(C)]
pub struct ImpRtEvent {
    _opaque: [u8; 0],
}
#[repr(C)]
pub struct ImpRtStatic {
    _opaque: [u8; 0],
}
#[repr(C)]
pub struct ImpRtPages {
    _opaque: [u8; 0],
}

pub const MEMCPY_H2D: c_int = 1;
pub const MEMCPY_D2H: c_int = 2;
pub const MEMCPY_D2D: c_int = 3;

extern "C" {
    pub fn impeller_rt_stream_create(device: c_int, out: *mut hipStream_t) -> hipError_t;
    pub fn impeller_rt_stream_destroy(s: hipStream_t) -> hipError_t;
    pub fn impeller_rt_stream_sync(s: hipStream_t) -> hipError_t;
    pub fn impeller_rt_stream_alloc(s: hipStream_t, bytes: usize, out: *mut *mut c_void) -> hipError_t;
    pub fn impeller_rt_stream_free(s: hipStream_t, p: *mut c_void) -> hipError_t;
    pub fn impeller_rt_memcpy_async(device: c_int, dst: *mut c_void, src: *const c_void, bytes: usize, kind: c_int, s: hipStream_t) -> hipError_t;

    pub fn impeller_rt_event_create(device: c_int, host_ordered: c_int, out: *mut *mut ImpRtEvent) -> hipError_t;
    pub fn impeller_rt_event_record(e: *mut ImpRtEvent, s: hipStream_t) -> hipError_t;
    pub fn impeller_rt_event_sync(e: *mut ImpRtEvent) -> hipError_t;
    pub fn impeller_rt_stream_wait_event(s: hipStream_t, e: *mut ImpRtEvent) -> hipError_t;
    pub fn impeller_rt_event_reset(e: *mut ImpRtEvent) -> hipError_t;
    pub fn impeller_rt_event_elapsed(a: *mut ImpRtEvent, b: *mut ImpRtEvent, ms: *mut f32) -> hipError_t;
    pub fn impeller_rt_event_raw(e: *mut ImpRtEvent) -> hipEvent_t;
    pub fn impeller_rt_event_destroy(e: *mut ImpRtEvent) -> hipError_t;

    pub fn impeller_rt_alloc_pinned(bytes: usize, out: *mut *mut c_void) -> hipError_t;
    pub fn impeller_rt_free_pinned(p: *mut c_void) -> hipError_t;

    pub fn impeller_rt_static_create(device: c_int, max_size: usize, check_overlap: c_int, out: *mut *mut ImpRtStatic) -> hipError_t;
    pub fn impeller_rt_static_allocate(a: *mut ImpRtStatic, offset: usize, bytes: usize, out: *mut *mut c_void) -> hipError_t;
    pub fn impeller_rt_static_free(a: *mut ImpRtStatic, p: *mut c_void) -> hipError_t;
    pub fn impeller_rt_static_base(a: *mut ImpRtStatic) -> *mut c_void;
    pub fn impeller_rt_static_destroy(a: *mut ImpRtStatic) -> hipError_t;

    pub fn impeller_rt_page_granularity(device: c_int, out: *mut usize) -> hipError_t;
    pub fn impeller_rt_page_create(device: c_int, page_size: usize, page_num: usize, memory_check: c_int, out: *mut *mut ImpRtPages) -> hipError_t;
    pub fn impeller_rt_page_allocate(a: *mut ImpRtPages, va_size: usize, page_ids: *const usize, n_pages: usize, vaid: usize, out: *mut *mut c_void) -> hipError_t;
    pub fn impeller_rt_page_deallocate(a: *mut ImpRtPages, p: *mut c_void, vaid: usize) -> hipError_t;
    pub fn impeller_rt_page_size(a: *mut ImpRtPages) -> usize;
    pub fn impeller_rt_page_destroy(a: *mut ImpRtPages) -> hipError_t;

    pub fn impeller_rt_num_cus(device: c_int, out: *mut c_int) -> hipError_t;
    pub fn impeller_rt_arch_name(device: c_int, buf: *mut c_char, len: usize) -> hipError_t;

    pub fn impeller_rt_file_read_to_device(fd: c_int, file_offset: u64, bytes: usize, dst: *mut c_void, staging: *mut c_void, staging_bytes: usize, s: hipStream_t) -> hipError_t;
    pub fn impeller_rt_file_write_from_device(fd: c_int, file_offset: u64, bytes: usize, src: *const c_void, staging: *mut c_void, staging_bytes: usize, s: hipStream_t) -> hipError_t;
}
