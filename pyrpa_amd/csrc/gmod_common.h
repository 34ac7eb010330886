// gmod: MI355X-native FFT-convolved many-body engine (chi0/RPA, FLEX, Eliashberg).
// Shared definitions: complex arithmetic, error channel, device context.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

namespace gmod {

// ----------------------------------------------------------------- complex
template <typename R>
struct alignas(2 * sizeof(R)) cpx {
  R re, im;
  __host__ __device__ cpx() : re(0), im(0) {}
  __host__ __device__ cpx(R r, R i = 0) : re(r), im(i) {}
  template <typename S>
  __host__ __device__ explicit cpx(const cpx<S>& o) : re((R)o.re), im((R)o.im) {}
};
using zd = cpx<double>;
using zf = cpx<float>;

template <typename R> __host__ __device__ inline cpx<R> operator+(cpx<R> a, cpx<R> b) { return {a.re + b.re, a.im + b.im}; }
template <typename R> __host__ __device__ inline cpx<R> operator-(cpx<R> a, cpx<R> b) { return {a.re - b.re, a.im - b.im}; }
template <typename R> __host__ __device__ inline cpx<R> operator-(cpx<R> a) { return {-a.re, -a.im}; }
template <typename R> __host__ __device__ inline cpx<R> operator*(cpx<R> a, cpx<R> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename R> __host__ __device__ inline cpx<R> operator*(cpx<R> a, R s) { return {a.re * s, a.im * s}; }
template <typename R> __host__ __device__ inline cpx<R> operator*(R s, cpx<R> a) { return {a.re * s, a.im * s}; }
template <typename R> __host__ __device__ inline cpx<R>& operator+=(cpx<R>& a, cpx<R> b) { a.re += b.re; a.im += b.im; return a; }
template <typename R> __host__ __device__ inline cpx<R>& operator-=(cpx<R>& a, cpx<R> b) { a.re -= b.re; a.im -= b.im; return a; }
template <typename R> __host__ __device__ inline cpx<R> conj(cpx<R> a) { return {a.re, -a.im}; }
template <typename R> __host__ __device__ inline R abs1(cpx<R> a) { return fabs(a.re) + fabs(a.im); }  // LAPACK cabs1
template <typename R> __host__ __device__ inline R absz(cpx<R> a) { return hypot(a.re, a.im); }
// complex division (Smith's algorithm)
template <typename R> __host__ __device__ inline cpx<R> operator/(cpx<R> a, cpx<R> b) {
  if (fabs(b.re) >= fabs(b.im)) {
    R r = b.im / b.re, d = b.re + b.im * r;
    return {(a.re + a.im * r) / d, (a.im - a.re * r) / d};
  } else {
    R r = b.re / b.im, d = b.im + b.re * r;
    return {(a.re * r + a.im) / d, (a.im * r - a.re) / d};
  }
}

// ----------------------------------------------------------------- errors
// The reference library prints and STOPs (kills the host process) on LAPACK
// failures; this library keeps the void ABI, records a sticky error readable
// through gmod_last_error(), and leaves the outputs untouched/partial.
struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};
void set_error(const std::string& msg);
void clear_error();

#define GMOD_HIP(call)                                                                   \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      throw ::gmod::Error(std::string("HIP error ") + hipGetErrorString(e_) + " at " +   \
                          __FILE__ + ":" + std::to_string(__LINE__));                    \
  } while (0)

#define GMOD_CHECK_LAUNCH() GMOD_HIP(hipGetLastError())

// ----------------------------------------------------------------- device buffers
// Simple size-bucketed caching allocator: FLEX iterations reuse the same
// shapes, so after the first iteration no hipMalloc happens inside the loop.
class Pool {
 public:
  void* get(size_t bytes);
  void put(void* p);
  void trim();
  ~Pool();

 private:
  struct Blk { void* p; size_t bytes; bool used; };
  std::vector<Blk> blocks_;
  std::mutex mu_;
};
Pool& pool();

template <typename T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  explicit DBuf(size_t count) { alloc(count); }
  void alloc(size_t count) {
    release();
    n = count;
    if (count) p = static_cast<T*>(pool().get(count * sizeof(T)));
  }
  void release() {
    if (p) pool().put(p);
    p = nullptr;
    n = 0;
  }
  ~DBuf() { release(); }
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DBuf& operator=(DBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  size_t bytes() const { return n * sizeof(T); }
  void upload(const T* h, size_t count, hipStream_t s) {
    GMOD_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void download(T* h, size_t count, hipStream_t s) const {
    GMOD_HIP(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
  }
  void zero(hipStream_t s) { if (n) GMOD_HIP(hipMemsetAsync(p, 0, bytes(), s)); }
};

// ----------------------------------------------------------------- context
struct Context {
  hipStream_t stream = nullptr;
  int device = 0;
  int precision = 0;  // 0: FP64 everywhere; 1: FP32 convolutions (FP64 LU/Dyson/mu/lambda)
  int verbose = 0;
  std::mutex lock;
};
Context& ctx();

// ----------------------------------------------------------------- profiling
// hipEvent pairs around kernel launches on the library stream, enabled by
// gmod_profile(1); per class: elapsed ms, launches, algorithmic bytes.
enum ProfClass {
  PROF_FFT = 0, PROF_VERTEX, PROF_DYSON, PROF_SCATTER, PROF_PRODUCT, PROF_GATHER, PROF_FFT32, PROF_NCLASS
};
bool prof_enabled();
struct ProfScope {
  int cls;
  double bytes;
  hipStream_t s;
  hipEvent_t e0 = nullptr;
  ProfScope(int cls, double bytes, hipStream_t s);
  ~ProfScope();
};
void prof_read(double* out);  // 3*PROF_NCLASS values
void prof_reset();

}  // namespace gmod
