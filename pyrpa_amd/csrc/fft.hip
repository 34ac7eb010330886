// Hand-written batched FFT for gfx950: plan construction (pass lists for the
// forward x,y,z,t and backward t,z,y,x orders), twiddle tables, plan cache.
// The pass kernel and the fused-exec template live in fft.h.
#include <map>
#include <tuple>

#include "fft.h"

namespace gmod {

constexpr int MAX_SINGLE = 256;

template <typename R> struct LdsBudget;
template <> struct LdsBudget<double> { static constexpr int elems = 4096; };
template <> struct LdsBudget<float> { static constexpr int elems = 8192; };

template <typename R>
FFTPlan<R>::FFTPlan(int nx, int ny, int nz, int m, int axes_mask) {
  dims_[0] = nx; dims_[1] = ny; dims_[2] = nz; dims_[3] = m;
  mask_ = axes_mask;
  S_ = (unsigned)((long)nx * ny * nz);
  const long total = (long)nx * ny * nz * m;
  if (total >= (1L << 31)) throw Error("FFT grid larger than 2^31 points");
  std::vector<int> tables;  // lengths needing a twiddle table
  auto add_table = [&](int len) -> long {
    long off = 0;
    for (size_t i = 0; i < tables.size(); ++i) {
      if (tables[i] == len) return off;
      off += tables[i];
    }
    tables.push_back(len);
    return off;
  };
  long strides[4];
  {
    long st = 1;
    for (int ax = 0; ax < 4; ++ax) { strides[ax] = st; st *= dims_[ax]; }
  }
  auto axis_passes = [&](int ax, bool fwd) {
    std::vector<FFTPassDesc> out;
    const int n = dims_[ax];
    const long stride = strides[ax];
    const long outer = total / (stride * n);
    if (!((mask_ >> ax) & 1) || n <= 1) return out;
    if (n <= MAX_SINGLE) {
      FFTPassDesc p;
      p.n = n;
      p.n_inner = stride; p.n_mid = 1; p.n_outer = outer;
      p.in_mid = p.out_mid = 0;
      p.in_outer = p.out_outer = stride * n;
      p.in_elem = p.out_elem = stride;
      p.tw_off = add_table(n);
      p.axis = ax;
      out.push_back(p);
      return out;
    }
    int n1 = 0;
    if (fft_is_pow2(n)) {  // balanced power-of-two factors
      int c = 1 << ((fft_ilog2(n) + 1) / 2);
      if (c <= MAX_SINGLE && n / c <= MAX_SINGLE) n1 = c;
    } else {
      for (int c = MAX_SINGLE; c >= 2; --c)
        if (n % c == 0 && n / c <= MAX_SINGLE) { n1 = c; break; }
    }
    if (!n1) throw Error("FFT length " + std::to_string(n) + " has no factorisation into two factors <= 256");
    const int n2 = n / n1;
    FFTPassDesc A;  // length n1 over t1, stride n2*s, lines (inner, t2, outer)
    A.n = n1;
    A.n_inner = stride; A.n_mid = n2; A.n_outer = outer;
    A.in_mid = A.out_mid = stride;
    A.in_outer = A.out_outer = stride * n;
    A.in_elem = A.out_elem = stride * n2;
    A.twn = n;
    A.src = 0; A.dst = 1;
    A.tw_off = add_table(n1);
    A.tw4_off = add_table(n);
    A.axis = ax;
    FFTPassDesc B;  // length n2 over t2 -> transposed store k1 + n1*k2
    B.n = n2;
    B.n_inner = stride; B.n_mid = n1; B.n_outer = outer;
    B.in_mid = stride * n2; B.out_mid = stride;
    B.in_outer = B.out_outer = stride * n;
    B.in_elem = stride; B.out_elem = stride * n1;
    B.src = 1; B.dst = 0;
    B.tw_off = add_table(n2);
    B.axis = ax;
    out.push_back(A);
    out.push_back(B);
    (void)fwd;
    return out;
  };
  for (int ax = 0; ax < 4; ++ax)
    for (auto& p : axis_passes(ax, true)) fwd_.push_back(p);
  for (int ax = 3; ax >= 0; --ax)
    for (auto& p : axis_passes(ax, false)) bwd_.push_back(p);
  for (auto* L : {&fwd_, &bwd_})
    for (auto& p : *L) {
      int budget = LdsBudget<R>::elems;
      int lines = std::max(1, std::min(budget / p.n, FFT_MAX_LINES));
      long nl = p.n_inner * p.n_mid * p.n_outer;
      if (lines > nl) lines = (int)nl;
      p.lines = lines;
    }
  // twiddle tables: tw[k] = exp(sign * 2 pi i k / len), computed in double
  long len = 0;
  for (int t : tables) len += t;
  tw_len_ = len;
  if (len) {
    for (int sg = 0; sg < 2; ++sg) {
      std::vector<cpx<R>> h(len);
      long off = 0;
      const double sign = sg == 0 ? -1.0 : 1.0;
      for (int t : tables) {
        for (int k = 0; k < t; ++k) {
          // argument reduced to the first half period for accuracy
          double c, s;
          long kk = k;
          double ang = 2.0 * M_PI * (double)kk / (double)t;
          if (2 * kk == t) { c = -1.0; s = 0.0; }
          else if (4 * kk == t) { c = 0.0; s = 1.0; }
          else if (4 * kk == 3 * t) { c = 0.0; s = -1.0; }
          else { c = std::cos(ang); s = std::sin(ang); }
          h[off + k] = cpx<R>((R)c, (R)(sign * s));
        }
        off += t;
      }
      GMOD_HIP(hipMalloc(&tw_[sg], len * sizeof(cpx<R>)));
      GMOD_HIP(hipMemcpy(tw_[sg], h.data(), len * sizeof(cpx<R>), hipMemcpyHostToDevice));
    }
  }
}

template <typename R>
FFTPlan<R>::~FFTPlan() {
  for (auto p : tw_)
    if (p) (void)hipFree(p);
}

template <typename R>
long FFTPlan<R>::transformed_points() const {
  long t = 1;
  for (int a = 0; a < 4; ++a)
    if (mask_ >> a & 1) t *= dims_[a];
  return t;
}

template <typename R>
const FFTPlan<R>& fft_plan(int nx, int ny, int nz, int m, int axes_mask) {
  static std::map<std::tuple<int, int, int, int, int>, FFTPlan<R>*> cache;
  static std::mutex mu;
  std::lock_guard<std::mutex> g(mu);
  auto key = std::make_tuple(nx, ny, nz, m, axes_mask);
  auto it = cache.find(key);
  if (it != cache.end()) return *it->second;
  auto* p = new FFTPlan<R>(nx, ny, nz, m, axes_mask);
  cache[key] = p;
  return *p;
}

template class FFTPlan<double>;
template class FFTPlan<float>;
template const FFTPlan<double>& fft_plan<double>(int, int, int, int, int);
template const FFTPlan<float>& fft_plan<float>(int, int, int, int, int);

}  // namespace gmod
