// fft_plan.cpp -- factorisation, twiddle and Bluestein tables for the
// device FFT (fft_device.hpp).  Tables are computed in long double on the
// host and rounded once to FP64.
#include "fft_plan.hpp"

#include <algorithm>
#include <cmath>
#include <complex>
#include <map>
#include <tuple>
#include <mutex>
#include <vector>

namespace so3g {

namespace {

using cld = std::complex<long double>;
const long double kTwoPi = 6.283185307179586476925286766559005768L;

// exp(-2 pi i k / n) with the angle reduced to [0, 2pi) exactly in integers
cld root(long long k, long long n) {
  k %= n;
  if (k < 0) k += n;
  const long double a = kTwoPi * (long double)k / (long double)n;
  return cld(cosl(a), -sinl(a));
}

// in-place radix-2 FFT in long double (host precomputation only)
void host_fft_pow2(std::vector<cld> &x, int sign) {
  const size_t n = x.size();
  for (size_t i = 1, j = 0; i < n; ++i) {
    size_t bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(x[i], x[j]);
  }
  for (size_t len = 2; len <= n; len <<= 1) {
    for (size_t i = 0; i < n; i += len)
      for (size_t k = 0; k < len / 2; ++k) {
        cld w = root((long long)k, (long long)len);
        if (sign > 0) w = std::conj(w);
        cld u = x[i + k], v = x[i + k + len / 2] * w;
        x[i + k] = u + v;
        x[i + k + len / 2] = u - v;
      }
  }
}

// in-place FFT of length 2^k or 3 * 2^k in long double: three interleaved
// power-of-two transforms and one radix-3 combination for the latter
void host_fft(std::vector<cld> &x, int sign) {
  const size_t n = x.size();
  if ((n & (n - 1)) == 0) {
    host_fft_pow2(x, sign);
    return;
  }
  const size_t m = n / 3;
  std::vector<cld> part[3];
  for (int r = 0; r < 3; ++r) {
    part[r].resize(m);
    for (size_t j = 0; j < m; ++j) part[r][j] = x[3 * j + r];
    host_fft_pow2(part[r], sign);
  }
  for (size_t k = 0; k < n; ++k) {
    cld acc(0, 0);
    for (int r = 0; r < 3; ++r) {
      cld w = root((long long)(r * k), (long long)n);
      if (sign > 0) w = std::conj(w);
      acc += w * part[r][k % m];
    }
    x[k] = acc;
  }
}

bool factor(int n, std::vector<int> &rad) {
  rad.clear();
  int m = n;
  while (m % 16 == 0) { rad.push_back(16); m /= 16; }
  while (m % 8 == 0) { rad.push_back(8); m /= 8; }
  while (m % 4 == 0) { rad.push_back(4); m /= 4; }
  while (m % 2 == 0) { rad.push_back(2); m /= 2; }
  for (int p : {3, 5, 7})
    while (m % p == 0) { rad.push_back(p); m /= p; }
  return m == 1 && (int)rad.size() <= kMaxPass;
}

double2 *upload(const std::vector<cld> &v) {
  std::vector<double2> h(v.size());
  for (size_t i = 0; i < v.size(); ++i) h[i] = make_double2((double)v[i].real(), (double)v[i].imag());
  double2 *d = nullptr;
  SO3_HIP(hipMalloc(&d, sizeof(double2) * (h.size() ? h.size() : 1)));
  SO3_HIP(hipMemcpy(d, h.data(), sizeof(double2) * h.size(), hipMemcpyHostToDevice));
  return d;
}

// mode 0: the general rule (the workgroup engine); 1: lengths in [257, 512]
// take Bluestein with M = 1024 (the register-resident FFTs); 2: as 1, and
// lengths in [513, 2048] take Bluestein with M = 2048 or 4096; 3: lengths in
// [33, 128] take Bluestein with M = 256
FftDesc build(int n, int mode) {
  const bool reg1024 = mode >= 1;
  FftDesc d{};
  d.n = n;
  std::vector<int> rad;
  if (n <= 1) {
    d.len = 1;
    d.npass = 0;
    d.tw = upload(std::vector<cld>{cld(1, 0)});
    return d;
  }
  // a smooth length runs directly when one workgroup holds it in registers
  // (radix 5 and 7 passes hold fewer points per thread), else via Bluestein;
  // with reg1024 every n in [257, 512] takes Bluestein with M = 1024, which
  // runs on the register-resident FFTs (fft_reg.hpp, fft_reg32.hpp) faster
  // than the workgroup engine runs even a smooth n directly
  const bool r1024 = reg1024 && n >= 257 && n <= 512;
  // mode 2: every n in [513, 2048] takes Bluestein with M = 2048 or 4096,
  // which the 2048- and 4096-point register FFTs (fft_reg2k.hpp,
  // fft_reg4k.hpp) run
  const bool r2048 = mode == 2 && n >= 513 && n <= 2048;
  // mode 3: every n in [33, 128] takes Bluestein with M = 256, which the
  // 256-point register FFT (fft_reg16.hpp) runs
  const bool r256 = mode == 3 && n >= 33 && n <= 128;
  bool direct = factor(n, rad) && !r1024 && !r2048 && !r256;
  if (direct) {
    int cap = kFftThreads * kFftElems * 4;
    for (int R : rad) cap = std::min(cap, kFftThreads * fft_nb(R) * R);
    direct = cap >= n;
  }
  if (direct) {
    d.blue = 0;
    d.len = n;
  } else {
    int M = 1;
    while (M < 2 * n - 1) M *= 2;
    // 3 * 2^k when that suffices and one workgroup holds it (fewer points
    // than the next power of two); n in [257, 512] keeps M = 1024
    if (!r1024 && !r2048 && !r256 && M >= 4 && 3 * (M / 4) >= 2 * n - 1 && 3 * (M / 4) <= 1536) M = 3 * (M / 4);
    if (r256) M = 256;
    d.blue = 1;
    d.len = M;
    factor(M, rad);
    std::vector<cld> chirp(n), bf(M, cld(0, 0)), bb(M, cld(0, 0));
    for (long long j = 0; j < n; ++j) {
      // exp(-pi i j^2 / n) = root(j^2 mod 2n, 2n)
      chirp[j] = root((j * j) % (2LL * n), 2LL * n);
    }
    for (int j = 0; j < n; ++j) {
      bf[j] = std::conj(chirp[j]);
      bb[j] = chirp[j];
      if (j) {
        bf[M - j] = std::conj(chirp[j]);
        bb[M - j] = chirp[j];
      }
    }
    host_fft(bf, -1);
    host_fft(bb, -1);
    d.chirp = upload(chirp);
    d.bhat_f = upload(bf);
    d.bhat_b = upload(bb);
  }
  d.npass = (int)rad.size();
  d.radcode = 0;
  for (int i = 0; i < d.npass; ++i) {
    int code = -1;
    for (int c = 0; c < kNumRadix; ++c)
      if (kRadixOf[c] == rad[i]) code = c;
    if (code < 0) throw HipError("unsupported radix");
    d.radcode |= (unsigned long long)code << (4 * i);
  }
  std::vector<cld> tw(d.len);
  for (int k = 0; k < d.len; ++k) tw[k] = root(k, d.len);
  d.tw = upload(tw);
  if (d.len == 1024) {
    // tws[k1 * 64 + l] = w1024^{l k1} (k1 < 16, l < 64), then
    // tws[1024 + c * 4 + a] = w64^{a c} (c < 16, a < 4)
    std::vector<cld> tws(1024 + 64);
    for (int k1 = 0; k1 < 16; ++k1)
      for (int l = 0; l < 64; ++l) tws[k1 * 64 + l] = root((long long)l * k1, 1024);
    for (int c = 0; c < 16; ++c)
      for (int a = 0; a < 4; ++a) tws[1024 + c * 4 + a] = root((long long)a * c, 64);
    d.tws = upload(tws);
  }
  return d;
}

}  // namespace

void host_dft_ld(std::vector<std::complex<long double>> &x, int sign) {
  const size_t n = x.size();
  const bool pow2 = (n & (n - 1)) == 0, three = n % 3 == 0 && ((n / 3) & (n / 3 - 1)) == 0;
  if (!pow2 && !three) throw HipError("host_dft_ld: length not 2^a or 3 * 2^a");
  host_fft(x, sign);
}

int fft_capacity(const FftDesc &d) {
  int cap = kFftThreads * kFftElems * 4;
  unsigned long long code = d.radcode;
  for (int p = 0; p < d.npass; ++p, code >>= 4) {
    const int R = kRadixOf[code & 15];
    cap = std::min(cap, kFftThreads * fft_nb(R) * R);
  }
  return cap;
}

static FftDesc plan_cached(int n, int mode) {
  static std::mutex mu;
  static std::map<std::tuple<int, int, int>, FftDesc> cache;
  int dev = 0;
  SO3_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> g(mu);
  auto key = std::make_tuple(dev, n, mode);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  FftDesc d = build(n, mode);
  cache[key] = d;
  return d;
}

FftDesc fft_plan(int n) { return plan_cached(n, 1); }

FftDesc fft_plan_lds(int n) { return n >= 257 && n <= 512 ? plan_cached(n, 0) : plan_cached(n, 1); }

FftDesc fft_plan_r256(int n) {
  if (n < 33 || n > 128) throw HipError("fft_plan_r256: length outside [33, 128]");
  return plan_cached(n, 3);
}

FftDesc fft_plan_reg(int n) {
  if (n < 513 || n > 2048) throw HipError("fft_plan_reg: length outside [513, 2048]");
  return plan_cached(n, 2);
}

}  // namespace so3g
