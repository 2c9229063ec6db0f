// fft_plan.hpp -- host construction of device FFT descriptors (cached per
// length and device).
#pragma once
#include <complex>
#include <vector>

#include "fft_device.hpp"

namespace so3g {

// Returns a descriptor whose tables live in device memory of the current
// device for the lifetime of the process.  Thread-safe.
FftDesc fft_plan(int n);

// The plan for the workgroup (LDS) engine: as fft_plan, except that a smooth
// n in [257, 512] runs directly instead of through Bluestein with M = 1024
// (which only the register-resident FFTs run faster).
FftDesc fft_plan_lds(int n);

// The plan of the 2048- and 4096-point register FFTs (fft_reg2k.hpp,
// fft_reg4k.hpp): Bluestein with M = 2048 for every n in [513, 1024] and
// M = 4096 for every n in [1025, 2048].
FftDesc fft_plan_reg(int n);

// The plan of the 256-point register FFT (fft_reg16.hpp): Bluestein with
// M = 256 for every n in [33, 128].
FftDesc fft_plan_r256(int n);

// Largest S * len one workgroup can run through every pass of d (each
// thread holds fft_nb(R) butterflies of radix R).
int fft_capacity(const FftDesc &d);

// In-place DFT in long double on the host (plan tables only), length 2^a or
// 3 * 2^a; sign -1 forward (exp(-2 pi i jk/n)), +1 backward.
void host_dft_ld(std::vector<std::complex<long double>> &x, int sign);

// LDS stride (in double2) needed per sequence to run FftDesc d.
inline int fft_ld(const FftDesc &d) { return d.len + (d.len % 32 == 0 ? 1 : 0); }

}  // namespace so3g
