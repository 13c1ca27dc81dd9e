//! The HIP counterpart of the reference's `zkpoly_cuda_api` crate
//! (cuda_api/src/lib.rs): the same modules and type shapes, bound to the
//! `impeller_rt_*` C ABI of libimpeller_hip.so (include/impeller_hip.h)
//! instead of bindgen output over the CUDA headers.
pub mod bindings;
pub mod device_info;
pub mod error;
pub mod file;
pub mod mem;
pub mod stream;
