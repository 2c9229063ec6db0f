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
#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "batch.hpp"
#include "conv.hpp"
#include "fft_plan.hpp"
#include "so3/so3_adjoint.h"
#include "so3/so3_conv.h"
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

// ---- host <-> device copies of the host-pointer entry points
//
// A host array that is already pinned (hipHostMalloc) is copied by DMA
// directly.  A pageable one goes through two pinned 64 MiB staging buffers
// per device, chunk by chunk: host to device, the CPU copy of chunk i into
// one buffer (several threads) overlaps the DMA of chunk i - 1 from the
// other; device to host, the DMA of chunk i overlaps the CPU copy of chunk
// i - 1 out of the other buffer.  The copies run on a stream of their own
// and end synchronised, as the drop-in calls are synchronous.
constexpr size_t kChunk = (size_t)64 << 20;

struct Pinned {
  char *buf[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  bool busy[2] = {false, false};
  hipStream_t cs = nullptr;
};
std::mutex g_pin_mu;
Pinned g_pin[64];

bool is_pinned_host(const void *p) {
  if (!p) return false;
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeHost;
}

void par_memcpy(void *dst, const void *src, size_t n) {
  const unsigned hw = std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
  const size_t parts = std::min<size_t>(hw, std::max<size_t>(1, n >> 22));  // >= 4 MiB per thread
  if (parts <= 1) {
    std::memcpy(dst, src, n);
    return;
  }
  std::vector<std::thread> th;
  const size_t step = (n + parts - 1) / parts;
  for (size_t k = 0; k < parts; ++k) {
    const size_t o = k * step;
    if (o >= n) break;
    th.emplace_back([=] { std::memcpy((char *)dst + o, (const char *)src + o, std::min(step, n - o)); });
  }
  for (auto &t : th) t.join();
}

Pinned &pinned_for_device() {
  int dev = 0;
  SO3_HIP(hipGetDevice(&dev));
  Pinned &P = g_pin[dev & 63];
  if (!P.cs) {
    SO3_HIP(hipStreamCreateWithFlags(&P.cs, hipStreamNonBlocking));
    for (int b = 0; b < 2; ++b) {
      SO3_HIP(hipHostMalloc((void **)&P.buf[b], kChunk, hipHostMallocDefault));
      SO3_HIP(hipEventCreateWithFlags(&P.ev[b], hipEventDisableTiming));
    }
  }
  return P;
}

void copy_h2d(void *dst, const void *src, size_t n) {
  if (!n) return;
  std::lock_guard<std::mutex> lk(g_pin_mu);
  Pinned &P = pinned_for_device();
  if (is_pinned_host(src)) {
    SO3_HIP(hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, P.cs));
    SO3_HIP(hipStreamSynchronize(P.cs));
    return;
  }
  for (size_t o = 0, i = 0; o < n; o += kChunk, ++i) {
    const int b = (int)(i & 1);
    const size_t m = std::min(kChunk, n - o);
    if (P.busy[b]) SO3_HIP(hipEventSynchronize(P.ev[b]));  // its previous DMA is done
    par_memcpy(P.buf[b], (const char *)src + o, m);
    SO3_HIP(hipMemcpyAsync((char *)dst + o, P.buf[b], m, hipMemcpyHostToDevice, P.cs));
    SO3_HIP(hipEventRecord(P.ev[b], P.cs));
    P.busy[b] = true;
  }
  SO3_HIP(hipStreamSynchronize(P.cs));
  P.busy[0] = P.busy[1] = false;
}

void copy_d2h(void *dst, const void *src, size_t n) {
  if (!n) return;
  std::lock_guard<std::mutex> lk(g_pin_mu);
  Pinned &P = pinned_for_device();
  if (is_pinned_host(dst)) {
    SO3_HIP(hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, P.cs));
    SO3_HIP(hipStreamSynchronize(P.cs));
    return;
  }
  const size_t nch = (n + kChunk - 1) / kChunk;
  auto issue = [&](size_t i) {
    const int b = (int)(i & 1);
    const size_t o = i * kChunk, m = std::min(kChunk, n - o);
    SO3_HIP(hipMemcpyAsync(P.buf[b], (const char *)src + o, m, hipMemcpyDeviceToHost, P.cs));
    SO3_HIP(hipEventRecord(P.ev[b], P.cs));
  };
  issue(0);
  for (size_t i = 0; i < nch; ++i) {
    const int b = (int)(i & 1);
    // the DMA of chunk i + 1 goes into the other buffer, whose copy-out
    // (chunk i - 1) finished in the previous iteration
    if (i + 1 < nch) issue(i + 1);
    SO3_HIP(hipEventSynchronize(P.ev[b]));
    const size_t o = i * kChunk;
    par_memcpy((char *)dst + o, P.buf[b], std::min(kChunk, n - o));
  }
}

void release_pinned() {
  std::lock_guard<std::mutex> lk(g_pin_mu);
  for (Pinned &P : g_pin) {
    if (!P.cs) continue;
    (void)hipStreamSynchronize(P.cs);
    for (int b = 0; b < 2; ++b) {
      (void)hipHostFree(P.buf[b]);
      (void)hipEventDestroy(P.ev[b]);
    }
    (void)hipStreamDestroy(P.cs);
    P = Pinned{};
  }
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
  if (m == so3g::kDirect) {
    q.sampling_scheme = SO3_SAMPLING_MW;  // direct is MW only
    q.steerable = 0;                      // and writes all 2N - 1 gamma samples
  }
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
    copy_h2d(dl, flmn, nl);
    so3g::run_inverse(p, m, dl, df, nullptr);
    SO3_HIP(hipDeviceSynchronize());
    copy_d2h(f, df, nf);
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
    copy_h2d(df, f, nf);
    // blocks the transform does not write keep the caller's values
    if (!so3g::forward_writes_all(p, m)) copy_h2d(dl, flmn, nl);
    so3g::run_forward(p, m, df, dl, nullptr);
    SO3_HIP(hipDeviceSynchronize());
    copy_d2h(flmn, dl, nl);
  } catch (const std::exception &e) {
    fatal(fn, e.what());
  }
}

// a device buffer that is either the caller's device pointer or a staged copy
struct DevBuf {
  void *d = nullptr;
  bool owned = false;
  DevBuf(const void *host, size_t bytes, bool copy_in) {
    if (is_device_ptr(host)) {
      d = const_cast<void *>(host);
      return;
    }
    SO3_HIP(hipMalloc(&d, bytes ? bytes : 16));
    owned = true;
    if (copy_in && bytes) copy_h2d(d, host, bytes);
  }
  void copy_out(void *host, size_t bytes) {
    if (owned && bytes) {
      SO3_HIP(hipDeviceSynchronize());
      copy_d2h(host, d, bytes);
    }
  }
  ~DevBuf() {
    if (owned) (void)hipFree(d);
  }
};

void check_conv_params(const so3_parameters_t *p) {
  if (p->storage != SO3_STORAGE_PADDED && p->storage != SO3_STORAGE_COMPACT)
    throw HipError("Invalid storage method.");
  if (p->n_order != SO3_N_ORDER_ZERO_FIRST && p->n_order != SO3_N_ORDER_NEGATIVE_FIRST)
    throw HipError("Invalid n-order.");
  if ((int)p->n_mode < 0 || (int)p->n_mode >= (int)SO3_N_MODE_SIZE) throw HipError("Invalid n-mode.");
}

int status(const std::exception &e) {
  t_err = e.what();
  return -1;
}

}  // namespace

extern "C" {

void so3_core_inverse_via_ssht(std::complex<double> *f, const std::complex<double> *flmn,
                               const so3_parameters_t *parameters) {
  // the complex transform ignores `reality`, as the reference's does
  so3_parameters_t q = *parameters;
  q.reality = 0;
  host_inverse(__func__, f, flmn, &q, so3g::kViaSsht);
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

// ---------------------------------------------------------------- adjoints
static void host_adjoint(const char *fn, void *out, const void *in, const so3_parameters_t *p,
                         bool of_inverse) {
  try {
    so3g::validate(p, so3g::kDirect);
    const size_t nf = f_bytes(p, so3g::kDirect), nl = flmn_bytes(p);
    if (of_inverse) {  // f -> flmn
      DevBuf df(in, nf, true), dl(out, nl, true);
      so3g::run_adjoint_inverse(p, df.d, dl.d, nullptr);
      SO3_HIP(hipDeviceSynchronize());
      dl.copy_out(out, nl);
    } else {  // flmn -> f
      DevBuf dl(in, nl, true), df(out, nf, false);
      so3g::run_adjoint_forward(p, dl.d, df.d, nullptr);
      SO3_HIP(hipDeviceSynchronize());
      df.copy_out(out, nf);
    }
  } catch (const std::exception &e) {
    fatal(fn, e.what());
  }
}

void so3_adjoint_inverse_direct(std::complex<double> *flmn, const std::complex<double> *f,
                                const so3_parameters_t *parameters) {
  so3_parameters_t q = *parameters;
  q.reality = 0;
  host_adjoint(__func__, flmn, f, &q, true);
}
void so3_adjoint_forward_direct(std::complex<double> *f, const std::complex<double> *flmn,
                                const so3_parameters_t *parameters) {
  so3_parameters_t q = *parameters;
  q.reality = 0;
  host_adjoint(__func__, f, flmn, &q, false);
}
void so3_adjoint_inverse_direct_real(std::complex<double> *flmn, const double *f,
                                     const so3_parameters_t *parameters) {
  so3_parameters_t q = *parameters;
  q.reality = 1;
  host_adjoint(__func__, flmn, f, &q, true);
}
void so3_adjoint_forward_direct_real(double *f, const std::complex<double> *flmn,
                                     const so3_parameters_t *parameters) {
  so3_parameters_t q = *parameters;
  q.reality = 1;
  host_adjoint(__func__, f, flmn, &q, false);
}

// ---------------------------------------------------------------- conv
void so3_conv_s2toso3_harmonic_convolution(std::complex<double> *hlmn, const so3_parameters_t *hp,
                                           const std::complex<double> *flm,
                                           const std::complex<double> *glm) {
  try {
    check_conv_params(hp);
    const size_t nl = (size_t)hp->L * hp->L * 16, nh = flmn_bytes(hp);
    DevBuf f(flm, nl, true), g(glm, nl, true), h(hlmn, nh, false);
    so3g::s2toso3(hp, f.d, g.d, h.d, nullptr);
    h.copy_out(hlmn, nh);
  } catch (const std::exception &e) {
    fatal(__func__, e.what());
  }
}

void so3_conv_harmonic_convolution(std::complex<double> *hlmn, const so3_parameters_t *hp,
                                   const std::complex<double> *flmn, const so3_parameters_t *fp,
                                   const std::complex<double> *glmn, const so3_parameters_t *gp) {
  try {
    check_conv_params(hp);
    check_conv_params(fp);
    check_conv_params(gp);
    const size_t nh = flmn_bytes(hp), nf = flmn_bytes(fp), ng = flmn_bytes(gp);
    DevBuf f(flmn, nf, true), g(glmn, ng, true), h(hlmn, nh, true);
    so3g::harmonic_convolution(hp, fp, gp, f.d, g.d, h.d, nullptr);
    h.copy_out(hlmn, nh);
  } catch (const std::exception &e) {
    fatal(__func__, e.what());
  }
}

so3_parameters_t so3_conv_get_parameters_of_convolved_lmn(const so3_parameters_t *fp,
                                                          const so3_parameters_t *gp) {
  // so3_conv.c:104-125, including h.N = h.L
  if (fp->sampling_scheme != gp->sampling_scheme) fatal(__func__, "f and g must have the same sampling_scheme");
  if (fp->n_order != gp->n_order) fatal(__func__, "f and g must have the same n_order");
  if (fp->storage != gp->storage) fatal(__func__, "f and g must have the same storage");
  if (fp->n_mode != gp->n_mode) fatal(__func__, "f and g must have the same n_mode");
  if (fp->reality != gp->reality) fatal(__func__, "f and g must have the same reality");
  so3_parameters_t h = *fp;
  h.L = std::min(fp->L, gp->L);
  h.N = h.L;
  h.L0 = std::max(fp->L0, gp->L0);
  return h;
}

void so3_conv_get_parameters_of_convolved_lmn_void(so3_parameters_t *hp, const so3_parameters_t *fp,
                                                   const so3_parameters_t *gp) {
  *hp = so3_conv_get_parameters_of_convolved_lmn(fp, gp);
}

void so3_conv_convolution(std::complex<double> *h, const so3_parameters_t *hp,
                          const std::complex<double> *f, const so3_parameters_t *fp,
                          const std::complex<double> *g, const so3_parameters_t *gp) {
  // so3_conv.c:154-184: forward_direct(f), forward_direct(g), harmonic
  // product, inverse_direct(h), all on the GPU (temporaries are freed).
  // f, g and h are complex signals (the reference calls the complex direct
  // transforms), so every step uses the complex parameter blocks and the
  // complex coefficient layout whatever `reality` says.
  try {
    so3_parameters_t fq = *fp, gq = *gp, hq = *hp;
    fq.reality = gq.reality = hq.reality = 0;
    so3g::validate(&fq, so3g::kDirect);
    so3g::validate(&gq, so3g::kDirect);
    so3g::validate(&hq, so3g::kDirect);
    const size_t ff = f_bytes(&fq, so3g::kDirect), fg = f_bytes(&gq, so3g::kDirect),
                 fh = f_bytes(&hq, so3g::kDirect);
    const size_t lf = flmn_bytes(&fq), lg = flmn_bytes(&gq), lh = flmn_bytes(&hq);
    DevBuf df(f, ff, true), dg(g, fg, true), dh(h, fh, false);
    DevBuf fl(nullptr, lf, false), gl(nullptr, lg, false), hl(nullptr, lh, false);
    SO3_HIP(hipMemset(fl.d, 0, lf));
    SO3_HIP(hipMemset(gl.d, 0, lg));
    SO3_HIP(hipMemset(hl.d, 0, lh));
    so3g::run_forward(&fq, so3g::kDirect, df.d, fl.d, nullptr);
    so3g::run_forward(&gq, so3g::kDirect, dg.d, gl.d, nullptr);
    so3g::harmonic_convolution(&hq, &fq, &gq, fl.d, gl.d, hl.d, nullptr);
    so3g::run_inverse(&hq, so3g::kDirect, hl.d, dh.d, nullptr);
    SO3_HIP(hipDeviceSynchronize());
    dh.copy_out(h, fh);
  } catch (const std::exception &e) {
    fatal(__func__, e.what());
  }
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

int so3x_inverse_noalloc(void *f, const void *flmn, const so3_parameters_t *parameters, int method,
                         void *stream) {
  try {
    so3g::run_inverse_noalloc(parameters, method == SO3X_DIRECT ? so3g::kDirect : so3g::kViaSsht, flmn, f,
                              (hipStream_t)stream);
    return 0;
  } catch (const std::exception &e) {
    return status(e);
  }
}

int so3x_tune(const so3_parameters_t *parameters, int method, void *stream, int *choice, double *ms) {
  try {
    so3g::tune(parameters, method == SO3X_DIRECT ? so3g::kDirect : so3g::kViaSsht, (hipStream_t)stream, choice, ms);
    return 0;
  } catch (const std::exception &e) {
    return status(e);
  }
}

int so3x_tune_get(const so3_parameters_t *parameters, int method, int *choice) {
  try {
    if (!choice) throw HipError("choice pointer is NULL");
    so3g::tune_get(parameters, method == SO3X_DIRECT ? so3g::kDirect : so3g::kViaSsht, choice);
    return 0;
  } catch (const std::exception &e) {
    return status(e);
  }
}

int so3x_tune_save(const char *path) {
  try {
    if (!path) throw HipError("path is NULL");
    so3g::tune_save(path);
    return 0;
  } catch (const std::exception &e) {
    return status(e);
  }
}

int so3x_tune_clear(void) {
  so3g::tune_clear();
  return 0;
}

int so3x_tune_load(const char *path) {
  try {
    if (!path) throw HipError("path is NULL");
    return so3g::tune_load(path);
  } catch (const std::exception &e) {
    return status(e);
  }
}

int so3x_to_mex_layout(double *re, double *im, const void *f, const so3_parameters_t *parameters, void *stream) {
  try {
    so3g::launch_mex_layout(true, const_cast<void *>(f), re, im, parameters, (hipStream_t)stream);
    return 0;
  } catch (const std::exception &e) {
    return status(e);
  }
}

int so3x_from_mex_layout(void *f, const double *re, const double *im, const so3_parameters_t *parameters,
                         void *stream) {
  try {
    so3g::launch_mex_layout(false, f, const_cast<double *>(re), const_cast<double *>(im), parameters,
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
    so3g::release_batch();
    release_pinned();
    so3g::release_all();
    so3g::release_conv_caches();
    return 0;
  } catch (const std::exception &e) {
    return status(e);
  }
}

int so3x_cache_limit(long long bytes) {
  try {
    if (bytes < 0) throw HipError("negative cache limit");
    so3g::dl_tables_limit((size_t)bytes);
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

// ---------------------------------------------------------------- batches
int so3x_batch(const so3x_item_t *items, int count, int lanes, void *stream) {
  try {
    so3g::batch_enqueue(items, count, lanes, SO3X_BATCH_AUTO, (hipStream_t)stream);
    return 0;
  } catch (const std::exception &e) {
    return status(e);
  }
}

int so3x_batch_persistent(const so3x_item_t *items, int count, int lanes, void *stream) {
  try {
    so3g::batch_enqueue(items, count, lanes, SO3X_BATCH_PERSISTENT, (hipStream_t)stream);
    return 0;
  } catch (const std::exception &e) {
    return status(e);
  }
}

int so3x_batch_choice(const so3x_item_t *items, int count, int *mode, int *lanes) {
  try {
    so3g::batch_choice(items, count, mode, lanes);
    return 0;
  } catch (const std::exception &e) {
    return status(e);
  }
}

int so3x_batch_tune(const so3x_item_t *items, int count, void *stream) {
  try {
    return so3g::batch_tune(items, count, (hipStream_t)stream);
  } catch (const std::exception &e) {
    return status(e);
  }
}

int so3x_graph_create(const so3x_item_t *items, int count, int lanes, long long *graph) {
  try {
    if (!graph) throw HipError("graph pointer is NULL");
    *graph = so3g::graph_create(items, count, lanes, SO3X_BATCH_AUTO);
    return 0;
  } catch (const std::exception &e) {
    return status(e);
  }
}

int so3x_graph_create_mode(const so3x_item_t *items, int count, int lanes, int mode, long long *graph) {
  try {
    if (!graph) throw HipError("graph pointer is NULL");
    if (mode != SO3X_BATCH_AUTO && mode != SO3X_BATCH_LANES && mode != SO3X_BATCH_PERSISTENT)
      throw HipError("invalid batch mode");
    *graph = so3g::graph_create(items, count, lanes, mode);
    return 0;
  } catch (const std::exception &e) {
    return status(e);
  }
}

int so3x_graph_launch(long long graph, void *stream) {
  try {
    so3g::graph_launch(graph, (hipStream_t)stream);
    return 0;
  } catch (const std::exception &e) {
    return status(e);
  }
}

int so3x_graph_destroy(long long graph) {
  try {
    so3g::graph_destroy(graph);
    return 0;
  } catch (const std::exception &e) {
    return status(e);
  }
}

int so3x_ticket(void *stream, long long *ticket) {
  try {
    if (!ticket) throw HipError("ticket pointer is NULL");
    *ticket = so3g::ticket_new((hipStream_t)stream);
    return 0;
  } catch (const std::exception &e) {
    return status(e);
  }
}

int so3x_ticket_done(long long ticket) {
  try {
    return so3g::ticket_done(ticket);
  } catch (const std::exception &e) {
    return status(e);
  }
}

int so3x_ticket_wait(long long ticket) {
  try {
    so3g::ticket_wait(ticket);
    return 0;
  } catch (const std::exception &e) {
    return status(e);
  }
}

}  // extern "C"
