//! Library errors.  The reference's API panics where the library returns an error code (a point
//! that coincides with a vertex panics in `locate`, src/delaunay_tree.rs:47-54); the
//! reference-named methods of this crate panic the same way, and a `try_` method beside each
//! returns the error instead.
use std::ffi::CStr;
use std::fmt;
use std::os::raw::c_int;

use crate::ffi;

pub const VO_OK: i32 = 0;
/// bad argument (dimension, null pointer, NaN / inf coordinate)
pub const VO_ERR_INVALID: i32 = -1;
/// device or host allocation failed
pub const VO_ERR_OOM: i32 = -2;
/// a HIP runtime call failed
pub const VO_ERR_HIP: i32 = -3;
/// some inserted points coincide with existing vertices (they are skipped)
pub const VO_ERR_DUPLICATE: i32 = -4;
/// a cavity exceeded the per-point buffers, or an output buffer was too small
pub const VO_ERR_OVERFLOW: i32 = -5;
/// an exact evaluation left the f64 exponent range
pub const VO_ERR_PRECISION: i32 = -6;
/// a walk or chain exceeded its bound
pub const VO_ERR_INTERNAL: i32 = -7;

/// An error code of libvoronoids_hip with the library's message for it.
#[derive(Debug, Clone, PartialEq, Eq)]
pub struct VoError {
    pub code: i32,
    pub message: String,
}

impl VoError {
    pub(crate) fn invalid(message: String) -> Self {
        VoError { code: VO_ERR_INVALID, message }
    }
}

impl fmt::Display for VoError {
    fn fmt(&self, f: &mut fmt::Formatter<'_>) -> fmt::Result {
        write!(f, "{} (libvoronoids_hip error {})", self.message, self.code)
    }
}

impl std::error::Error for VoError {}

/// 0 -> Ok; a negative code -> the error with `vo_last_error`'s message.
pub(crate) fn check(rc: c_int) -> Result<(), VoError> {
    if rc == 0 {
        return Ok(());
    }
    // SAFETY: vo_last_error returns null or a NUL-terminated thread-local string that stays valid
    // until the next failing call on this thread; it is copied before anything else is called.
    let message = unsafe {
        let p = ffi::vo_last_error();
        if p.is_null() {
            String::new()
        } else {
            CStr::from_ptr(p).to_string_lossy().into_owned()
        }
    };
    Err(VoError { code: rc as i32, message })
}
