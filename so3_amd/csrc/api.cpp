// api.cpp -- the C ABI: drop-in replacements of the reference's so3_core_*,
// so3_adjoint_* and so3_conv_* functions, plus the so3x_* extensions.
//
// Drop-in functions keep the reference contract (so3_core.h:15-45): host
// pointers, synchronous completion, invalid parameters print "ERROR: ..."
// and exit(1) (so3_error.h:19-33).  Device pointers are detected and used
// in place.  All arithmetic runs on the GPU.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>

#include "fft_plan.hpp"
#include "so3/so3_core.h"
#include "so3/so3_sampling.h"
#include "so3/so3x.h"
#include "transform.hpp"

using so3g::HipError;
using so3g::Method;

namespace {

thread_local std::string t_err;

[[noreturn]] void fatal(const char *func, const std::string &msg) {
  printf("ERROR: %s.\n", msg.c_str());
  printf("ERROR: Occurred in function <%s> of file %s.\n", func, __FILE__);
  fflush(stdout);
  exit(1);
}

bool is_device_ptr(const void *p) {
  if (!p) return false;
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice;
}

// staging buffers for host-pointer calls (one pair per device)
struct Staging {
  void *a = nullptr, *b = nullptr;
  size_t ca = 0, cb = 0;
};
std::mutex g_stage_mu;
Staging g_stage[64];

void *grow(void *&p, size_t &cap, size_t bytes) {
  if (cap < bytes) {
    if (p) SO3_HIP(hipFree(p));
    p = nullptr;
    cap = 0;
    SO3_HIP(hipMalloc(&p, bytes));
    cap = bytes;
  }
  return p;
}

size_t f_bytes(const so3_parameters_t *p, Method m) {
  so3_parameters_t q = *p;
  if (m == so3g::kDirect) q.sampling_scheme = SO3_SAMPLING_MW;  // direct is MW only
  return (size_t)so3_sampling_f_size(&q) * (p->reality ? sizeof(double) : 2 * sizeof(double));
}
size_t flmn_bytes(const so3_parameters_t *p) {
  return (size_t)so3_sampling_flmn_size(p) * 2 * sizeof(double);
}

void host_inverse(const char *fn, void *f, const void *flmn, const so3_parameters_t *p, Method m) {
  try {
    so3g::validate(p, m);
    if (is_device_ptr(f) && is_device_ptr(flmn)) {
      so3g::run_inverse(p, m, flmn, f, nullptr);
      SO3_HIP(hipDeviceSynchronize());
      return;
    }
    int dev = 0;
    SO3_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_stage_mu);
    Staging &s = g_stage[dev & 63];
    const size_t nf = f_bytes(p, m), nl = flmn_bytes(p);
    void *dl = grow(s.a, s.ca, nl);
    void *df = grow(s.b, s.cb, nf);
    SO3_HIP(hipMemcpy(dl, flmn, nl, hipMemcpyHostToDevice));
    so3g::run_inverse(p, m, dl, df, nullptr);
    SO3_HIP(hipMemcpy(f, df, nf, hipMemcpyDeviceToHost));
  } catch (const std::exception &e) {
    fatal(fn, e.what());
  }
}

void host_forward(const char *fn, void *flmn, const void *f, const so3_parameters_t *p, Method m) {
  try {
    so3g::validate(p, m);
    if (is_device_ptr(f) && is_device_ptr(flmn)) {
      so3g::run_forward(p, m, f, flmn, nullptr);
      SO3_HIP(hipDeviceSynchronize());
      return;
    }
    int dev = 0;
    SO3_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_stage_mu);
    Staging &s = g_stage[dev & 63];
    const size_t nf = f_bytes(p, m), nl = flmn_bytes(p);
    void *dl = grow(s.a, s.ca, nl);
    void *df = grow(s.b, s.cb, nf);
    SO3_HIP(hipMemcpy(df, f, nf, hipMemcpyHostToDevice));
    // blocks the transform does not write keep the caller's values
    if (!so3g::forward_writes_all(p, m)) SO3_HIP(hipMemcpy(dl, flmn, nl, hipMemcpyHostToDevice));
    so3g::run_forward(p, m, df, dl, nullptr);
    SO3_HIP(hipMemcpy(flmn, dl, nl, hipMemcpyDeviceToHost));
  } catch (const std::exception &e) {
    fatal(fn, e.what());
  }
}

int status(const std::exception &e) {
  t_err = e.what();
  return -1;
}

}  // namespace

extern "C" {

void so3_core_inverse_via_ssht(std::complex<double> *f, const std::complex<double> *flmn,
                               const so3_parameters_t *parameters) {
  host_inverse(__func__, f, flmn, parameters, so3g::kViaSsht);
}
void so3_core_inverse_via_ssht_real(double *f, const std::complex<double> *flmn,
                                    const so3_parameters_t *parameters) {
  so3_parameters_t q = *parameters;
  q.reality = 1;
  host_inverse(__func__, f, flmn, &q, so3g::kViaSsht);
}
void so3_core_forward_via_ssht(std::complex<double> *flmn, const std::complex<double> *f,
                               const so3_parameters_t *parameters) {
  so3_parameters_t q = *parameters;
  q.reality = 0;
  host_forward(__func__, flmn, f, &q, so3g::kViaSsht);
}
void so3_core_forward_via_ssht_real(std::complex<double> *flmn, const double *f,
                                    const so3_parameters_t *parameters) {
  so3_parameters_t q = *parameters;
  q.reality = 1;
  host_forward(__func__, flmn, f, &q, so3g::kViaSsht);
}
void so3_core_inverse_direct(std::complex<double> *f, const std::complex<double> *flmn,
                             const so3_parameters_t *parameters) {
  so3_parameters_t q = *parameters;
  q.reality = 0;
  host_inverse(__func__, f, flmn, &q, so3g::kDirect);
}
void so3_core_inverse_direct_real(double *f, const std::complex<double> *flmn,
                                  const so3_parameters_t *parameters) {
  so3_parameters_t q = *parameters;
  q.reality = 1;
  host_inverse(__func__, f, flmn, &q, so3g::kDirect);
}
void so3_core_forward_direct(std::complex<double> *flmn, const std::complex<double> *f,
                             const so3_parameters_t *parameters) {
  so3_parameters_t q = *parameters;
  q.reality = 0;
  host_forward(__func__, flmn, f, &q, so3g::kDirect);
}
void so3_core_forward_direct_real(std::complex<double> *flmn, const double *f,
                                  const so3_parameters_t *parameters) {
  so3_parameters_t q = *parameters;
  q.reality = 1;
  host_forward(__func__, flmn, f, &q, so3g::kDirect);
}

const char *so3x_last_error(void) { return t_err.c_str(); }

int so3x_inverse(void *f, const void *flmn, const so3_parameters_t *parameters, int method,
                 void *stream) {
  try {
    so3g::run_inverse(parameters, method == SO3X_DIRECT ? so3g::kDirect : so3g::kViaSsht, flmn, f,
                      (hipStream_t)stream);
    return 0;
  } catch (const std::exception &e) {
    return status(e);
  }
}

int so3x_forward(void *flmn, const void *f, const so3_parameters_t *parameters, int method,
                 void *stream) {
  try {
    so3g::run_forward(parameters, method == SO3X_DIRECT ? so3g::kDirect : so3g::kViaSsht, f, flmn,
                      (hipStream_t)stream);
    return 0;
  } catch (const std::exception &e) {
    return status(e);
  }
}

int so3x_prepare(const so3_parameters_t *parameters) {
  try {
    so3g::prepare(parameters, so3g::kViaSsht);
    return 0;
  } catch (const std::exception &e) {
    return status(e);
  }
}

long long so3x_workspace_bytes(const so3_parameters_t *parameters) {
  try {
    return (long long)so3g::workspace_bytes(parameters, so3g::kViaSsht);
  } catch (const std::exception &e) {
    return status(e);
  }
}

int so3x_release_all(void) {
  try {
    so3g::release_all();
    return 0;
  } catch (const std::exception &e) {
    return status(e);
  }
}

int so3x_stage_timing(int on) {
  so3g::set_stage_timing(on != 0);
  return 0;
}

int so3x_stage_times(const char **names, double *ms, int *count, int cap) {
  try {
    return so3g::stage_times(names, ms, count, cap);
  } catch (const std::exception &e) {
    return status(e);
  }
}

int so3x_stage_reset(void) {
  try {
    so3g::reset_stage_times();
    return 0;
  } catch (const std::exception &e) {
    return status(e);
  }
}

int so3x_device_info(char *buf, int len) {
  try {
    int dev = 0;
    SO3_HIP(hipGetDevice(&dev));
    hipDeviceProp_t pr;
    SO3_HIP(hipGetDeviceProperties(&pr, dev));
    snprintf(buf, (size_t)len, "%s, %d CUs, %.0f MHz, LDS/block %zu B, HBM %.1f GiB",
             pr.gcnArchName, pr.multiProcessorCount, pr.clockRate / 1000.0, (size_t)pr.sharedMemPerBlock,
             pr.totalGlobalMem / 1073741824.0);
    return 0;
  } catch (const std::exception &e) {
    return status(e);
  }
}

}  // extern "C"
