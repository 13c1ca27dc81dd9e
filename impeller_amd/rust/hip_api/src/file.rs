//! cuda_api/src/file.rs registers a file with cuFile; here a file is its
//! descriptor, and the transfers go through a pinned staging buffer.
use std::ffi::c_void;
use std::fs::File;
use std::os::fd::AsRawFd;

use crate::bindings::*;
use crate::hip_check;
use crate::stream::HipStream;

pub fn read_to_device(file: &File, offset: u64, bytes: usize, dst: *mut c_void, staging: *mut c_void, staging_bytes: usize, stream: &HipStream) {
    unsafe {
        hip_check!(impeller_rt_file_read_to_device(file.as_raw_fd(), offset, bytes, dst, staging, staging_bytes, stream.raw()))
    };
}

pub fn write_from_device(file: &File, offset: u64, bytes: usize, src: *const c_void, staging: *mut c_void, staging_bytes: usize, stream: &HipStream) {
    unsafe {
        hip_check!(impeller_rt_file_write_from_device(file.as_raw_fd(), offset, bytes, src, staging, staging_bytes, stream.raw()))
    };
}
