// fft_plan.hpp -- host construction of device FFT descriptors (cached per
// length and device).
#pragma once
#include "fft_device.hpp"

namespace so3g {

// Returns a descriptor whose tables live in device memory of the current
// device for the lifetime of the process.  Thread-safe.
FftDesc fft_plan(int n);

// LDS stride (in double2) needed per sequence to run FftDesc d.
inline int fft_ld(const FftDesc &d) { return d.len + (d.len % 32 == 0 ? 1 : 0); }

}  // namespace so3g
