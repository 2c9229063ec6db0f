// dft_api.hip -- so3x_dft: the library's batched DFT kernels on contiguous
// sequences (the alpha-stage launcher), exported so that their numerics can
// be tested in isolation.
#include <algorithm>
#include <string>

#include "fft_plan.hpp"
#include "kernels.hpp"
#include "so3/so3x.h"

extern "C" int so3x_dft(void *x, int n, int howmany, int dir, void *stream) {
  using namespace so3g;
  try {
    if (n < 1 || howmany < 0 || (dir != 1 && dir != -1)) throw HipError("bad so3x_dft arguments");
    launch_rows((double2 *)x, howmany, n, n, dir, 1.0, (hipStream_t)stream);
    return 0;
  } catch (const std::exception &e) {
    fprintf(stderr, "so3x_dft: %s\n", e.what());
    return -1;
  }
}
