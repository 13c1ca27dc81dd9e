//! cuda_api/src/mem.rs, mem/static_allocator.rs, mem/page_allocator.rs:
//! pinned host memory, the caller-planned static allocator and the page
//! allocator over HIP virtual memory (device pages).
use std::ffi::c_void;

use crate::bindings::*;
use crate::hip_check;

pub unsafe fn alloc_pinned<T: Sized>(len: usize) -> *mut T {
    let mut p: *mut c_void = std::ptr::null_mut();
    hip_check!(impeller_rt_alloc_pinned(len * std::mem::size_of::<T>(), &mut p));
    p as *mut T
}

pub unsafe fn free_pinned<T: Sized>(ptr: *mut T, _len: usize) {
    hip_check!(impeller_rt_free_pinned(ptr as *mut c_void));
}

pub struct StaticAllocator {
    a: *mut ImpRtStatic,
}

unsafe impl Send for StaticAllocator {}

impl StaticAllocator {
    pub fn new(device_id: i32, max_size: usize, check_overlap: bool) -> Self {
        let mut a = std::ptr::null_mut();
        unsafe { hip_check!(impeller_rt_static_create(device_id, max_size, check_overlap as i32, &mut a)) };
        Self { a }
    }
    /// panics on an overlapping or out-of-range request, as the reference does
    pub fn allocate<F: Sized>(&mut self, offset: usize, len: usize) -> *mut F {
        let mut p: *mut c_void = std::ptr::null_mut();
        let bytes = len * std::mem::size_of::<F>();
        unsafe { hip_check!(impeller_rt_static_allocate(self.a, offset, bytes, &mut p)) };
        p as *mut F
    }
    pub fn free<F: Sized>(&mut self, ptr: *mut F) {
        unsafe { hip_check!(impeller_rt_static_free(self.a, ptr as *mut c_void)) };
    }
}

impl Drop for StaticAllocator {
    fn drop(&mut self) {
        unsafe { hip_check!(impeller_rt_static_destroy(self.a)) };
    }
}

pub struct PageAllocator {
    a: *mut ImpRtPages,
}

unsafe impl Send for PageAllocator {}

impl PageAllocator {
    /// device pages of page_size bytes (a multiple of the granularity)
    pub fn new(device_id: i32, page_size: usize, page_num: usize, memory_check: bool) -> Self {
        let mut a = std::ptr::null_mut();
        unsafe { hip_check!(impeller_rt_page_create(device_id, page_size, page_num, memory_check as i32, &mut a)) };
        Self { a }
    }
    pub fn allocate<T: Sized>(&mut self, va_size: usize, page_ids: Vec<usize>, vaid: usize) -> *mut T {
        let mut p: *mut c_void = std::ptr::null_mut();
        unsafe {
            hip_check!(impeller_rt_page_allocate(self.a, va_size, page_ids.as_ptr(), page_ids.len(), vaid, &mut p))
        };
        p as *mut T
    }
    pub fn deallocate<T: Sized>(&mut self, ptr: *mut T, vaid: usize) {
        unsafe { hip_check!(impeller_rt_page_deallocate(self.a, ptr as *mut c_void, vaid)) };
    }
    pub fn page_size(&self) -> usize {
        unsafe { impeller_rt_page_size(self.a) }
    }
}

impl Drop for PageAllocator {
    fn drop(&mut self) {
        unsafe { hip_check!(impeller_rt_page_destroy(self.a)) };
    }
}

pub fn page_granularity(device_id: i32) -> usize {
    let mut g = 0usize;
    unsafe { hip_check!(impeller_rt_page_granularity(device_id, &mut g)) };
    g
}
