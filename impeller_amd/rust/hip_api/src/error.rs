//! cuda_api/src/error.rs: a failed call panics with the call site, as the
//! reference's `cuda_check!` does (the library prints the HIP error text).
#[macro_export]
macro_rules! hip_check {
    ($x:expr) => {{
        let err = $x;
        if err != $crate::bindings::HIP_SUCCESS {
            eprintln!("HIP Error [{}:{}]: code {}", file!(), line!(), err);
            panic!("HIP Error");
        }
    }};
}
