//! cuda_api/src/device_info.rs: compute units and the gfx target.
use crate::bindings::*;
use crate::hip_check;

pub fn get_num_sms(device: i32) -> i32 {
    let mut n = 0;
    unsafe { hip_check!(impeller_rt_num_cus(device, &mut n)) };
    n
}

pub fn sm_version(device: i32) -> String {
    let mut buf = [0 as std::ffi::c_char; 64];
    unsafe {
        hip_check!(impeller_rt_arch_name(device, buf.as_mut_ptr(), buf.len()));
        std::ffi::CStr::from_ptr(buf.as_ptr()).to_string_lossy().into_owned()
    }
}
