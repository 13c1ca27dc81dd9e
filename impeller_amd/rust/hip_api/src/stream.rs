//! cuda_api/src/stream.rs: HipEvent (host-ordered: a wait blocks until the
//! event has been recorded), HipEventRaw and HipStream.
use std::ffi::c_void;

use crate::bindings::*;
use crate::hip_check;

pub struct HipEvent {
    e: *mut ImpRtEvent,
    device_id: i32,
}

impl HipEvent {
    pub fn new(device_id: i32) -> Self {
        let mut e = std::ptr::null_mut();
        unsafe { hip_check!(impeller_rt_event_create(device_id, 1, &mut e)) };
        Self { e, device_id }
    }
    pub fn get_device(&self) -> i32 {
        self.device_id
    }
    pub fn reset(&mut self) {
        unsafe { hip_check!(impeller_rt_event_reset(self.e)) };
    }
    pub fn record(&self, stream: &HipStream) {
        unsafe { hip_check!(impeller_rt_event_record(self.e, stream.raw())) };
    }
    pub fn sync(&self) {
        unsafe { hip_check!(impeller_rt_event_sync(self.e)) };
    }
    pub fn as_ptr(&self) -> hipEvent_t {
        unsafe { impeller_rt_event_raw(self.e) }
    }
    pub fn elapsed(&self, other: &HipEvent) -> f32 {
        let mut ms = 0f32;
        unsafe { hip_check!(impeller_rt_event_elapsed(self.e, other.e, &mut ms)) };
        ms
    }
}

impl Drop for HipEvent {
    fn drop(&mut self) {
        unsafe { hip_check!(impeller_rt_event_destroy(self.e)) };
    }
}

pub struct HipEventRaw {
    e: *mut ImpRtEvent,
}

impl HipEventRaw {
    pub fn new(device_id: i32) -> Self {
        let mut e = std::ptr::null_mut();
        unsafe { hip_check!(impeller_rt_event_create(device_id, 0, &mut e)) };
        Self { e }
    }
    pub fn record(&self, stream: &HipStream) {
        unsafe { hip_check!(impeller_rt_event_record(self.e, stream.raw())) };
    }
    pub fn sync(&self) {
        unsafe { hip_check!(impeller_rt_event_sync(self.e)) };
    }
    pub fn as_ptr(&self) -> hipEvent_t {
        unsafe { impeller_rt_event_raw(self.e) }
    }
    pub fn elapsed(&self, other: &HipEventRaw) -> f32 {
        let mut ms = 0f32;
        unsafe { hip_check!(impeller_rt_event_elapsed(self.e, other.e, &mut ms)) };
        ms
    }
}

impl Drop for HipEventRaw {
    fn drop(&mut self) {
        unsafe { hip_check!(impeller_rt_event_destroy(self.e)) };
    }
}

#[derive(Debug, Clone)]
pub struct HipStream {
    stream: hipStream_t,
    gpu_id: i32,
}

impl HipStream {
    pub fn new(device: i32) -> Self {
        let mut stream = std::ptr::null_mut();
        unsafe { hip_check!(impeller_rt_stream_create(device, &mut stream)) };
        Self { stream, gpu_id: device }
    }
    pub fn sync(&self) {
        unsafe { hip_check!(impeller_rt_stream_sync(self.stream)) };
    }
    pub fn raw(&self) -> hipStream_t {
        self.stream
    }
    pub fn get_device(&self) -> i32 {
        self.gpu_id
    }
    pub fn allocate<T: Sized>(&self, len: usize) -> *mut T {
        let mut p: *mut c_void = std::ptr::null_mut();
        let bytes = std::mem::size_of::<T>() * len;
        unsafe { hip_check!(impeller_rt_stream_alloc(self.stream, bytes, &mut p)) };
        p as *mut T
    }
    pub fn free<T: Sized>(&self, ptr: *mut T) {
        unsafe { hip_check!(impeller_rt_stream_free(self.stream, ptr as *mut c_void)) };
    }
    pub fn record(&self, event: &HipEvent) {
        event.record(self);
    }
    pub fn wait(&self, event: &HipEvent) {
        unsafe { hip_check!(impeller_rt_stream_wait_event(self.stream, event.e)) };
    }
    pub fn wait_raw(&self, event: &HipEventRaw) {
        unsafe { hip_check!(impeller_rt_stream_wait_event(self.stream, event.e)) };
    }
    fn copy<T: Sized>(&self, dst: *mut T, src: *const T, len: usize, kind: i32) {
        let bytes = std::mem::size_of::<T>() * len;
        unsafe {
            hip_check!(impeller_rt_memcpy_async(self.gpu_id, dst as *mut c_void, src as *const c_void, bytes, kind, self.stream))
        };
    }
    pub fn memcpy_h2d<T: Sized>(&self, dst: *mut T, src: *const T, len: usize) {
        self.copy(dst, src, len, MEMCPY_H2D);
    }
    pub fn memcpy_d2h<T: Sized>(&self, dst: *mut T, src: *const T, len: usize) {
        self.copy(dst, src, len, MEMCPY_D2H);
    }
    pub fn memcpy_d2d<T: Sized>(&self, dst: *mut T, src: *const T, len: usize) {
        self.copy(dst, src, len, MEMCPY_D2D);
    }
    pub fn destroy(&self) {
        unsafe { hip_check!(impeller_rt_stream_destroy(self.stream)) };
    }
}

unsafe impl Send for HipStream {}
unsafe impl Sync for HipStream {}
unsafe impl Send for HipEvent {}
unsafe impl Sync for HipEvent {}
unsafe impl Send for HipEventRaw {}
unsafe impl Sync for HipEventRaw {}
