// Hand-written batched FFT for gfx950.
//
// Each pass transforms one axis of the 4D grid.  A work-group loads `lines`
// lines of length n into LDS (coalesced: the lines of one group are adjacent
// in memory for strided axes, the elements are adjacent for the contiguous x
// axis), runs an in-place radix-2^2 decimation-in-time FFT in LDS (bit-reversed
// load, natural-order result), applies the optional four-step twiddle and the
// output scale, and stores.  Axes longer than MAX_SINGLE are split n = n1*n2
// into two passes (four-step): pass A = length-n1 transforms with the
// w_n^(k1*t2) twiddle, pass B = length-n2 transforms with a transposing store.
// Non-power-of-two lengths use a direct DFT from LDS.
#include <map>
#include <tuple>

#include "fft.h"

namespace gmod {

constexpr int FFT_THREADS = 256;
constexpr int MAX_SINGLE = 256;

template <typename R> struct LdsBudget;
template <> struct LdsBudget<double> { static constexpr int elems = 4096; };
template <> struct LdsBudget<float> { static constexpr int elems = 8192; };

struct PassArgs {
  int n, logn, lines, pitch;
  long n_inner, n_mid, total_lines;
  long in_mid, in_outer, in_elem;
  long out_mid, out_outer, out_elem;
  int twn;
};

__device__ inline void line_addr(const PassArgs& a, long g, long& in_base, long& out_base, long& mid) {
  long inner = g % a.n_inner;
  long rest = g / a.n_inner;
  mid = rest % a.n_mid;
  long outer = rest / a.n_mid;
  in_base = inner + mid * a.in_mid + outer * a.in_outer;
  out_base = inner + mid * a.out_mid + outer * a.out_outer;
}

template <typename R, bool POW2>
__global__ void __launch_bounds__(FFT_THREADS)
fft_pass_kernel(const cpx<R>* __restrict__ in, cpx<R>* __restrict__ out, PassArgs a,
                const cpx<R>* __restrict__ tw, const cpx<R>* __restrict__ tw4, R scale) {
  extern __shared__ unsigned char smem_raw[];
  cpx<R>* lds = reinterpret_cast<cpx<R>*>(smem_raw);
  const int n = a.n;
  const int lines = a.lines;
  const long g0 = (long)blockIdx.x * lines;
  const int tid = threadIdx.x;
  const int total = lines * n;
  // ---- load
  if (a.in_elem == 1) {
    for (int idx = tid; idx < total; idx += FFT_THREADS) {
      int e = idx % n, L = idx / n;
      long g = g0 + L;
      if (g < a.total_lines) {
        long ib, ob, mid;
        line_addr(a, g, ib, ob, mid);
        int pe = POW2 ? (int)(__brev((unsigned)e) >> (32 - a.logn)) : e;
        if (POW2 && a.logn == 0) pe = 0;
        lds[L * a.pitch + pe] = in[ib + e];
      }
    }
  } else {
    for (int idx = tid; idx < total; idx += FFT_THREADS) {
      int L = idx % lines, e = idx / lines;
      long g = g0 + L;
      if (g < a.total_lines) {
        long ib, ob, mid;
        line_addr(a, g, ib, ob, mid);
        int pe = POW2 ? (int)(__brev((unsigned)e) >> (32 - a.logn)) : e;
        if (POW2 && a.logn == 0) pe = 0;
        lds[L * a.pitch + pe] = in[ib + (long)e * a.in_elem];
      }
    }
  }
  __syncthreads();
  if (POW2) {
    int h = 1;
    if (a.logn & 1) {  // one radix-2 stage (h = 1, twiddle 1)
      const int units = lines * (n >> 1);
      for (int u = tid; u < units; u += FFT_THREADS) {
        int L = u / (n >> 1), p = u % (n >> 1);
        cpx<R>* b = lds + L * a.pitch + 2 * p;
        cpx<R> x0 = b[0], x1 = b[1];
        b[0] = x0 + x1;
        b[1] = x0 - x1;
      }
      h = 2;
      __syncthreads();
    }
    for (; h < n; h *= 4) {  // radix-2^2 stages (h, 2h)
      const int quarter = n >> 2;
      const int units = lines * quarter;
      const int s2 = n / (2 * h), s4 = n / (4 * h);
      for (int u = tid; u < units; u += FFT_THREADS) {
        int L = u / quarter, q = u % quarter;
        int blk = q / h, j = q % h;
        cpx<R>* b = lds + L * a.pitch + blk * 4 * h + j;
        cpx<R> w1 = tw[j * s2], w2 = tw[j * s4], w3 = tw[(j + h) * s4];
        cpx<R> x0 = b[0], x1 = b[h] * w1, x2 = b[2 * h], x3 = b[3 * h] * w1;
        cpx<R> y0 = x0 + x1, y1 = x0 - x1, y2 = x2 + x3, y3 = x2 - x3;
        y2 = y2 * w2;
        y3 = y3 * w3;
        b[0] = y0 + y2;
        b[2 * h] = y0 - y2;
        b[h] = y1 + y3;
        b[3 * h] = y1 - y3;
      }
      __syncthreads();
    }
  }
  // ---- store (direct DFT for non powers of two)
  auto value = [&](int L, int k) -> cpx<R> {
    if (POW2) return lds[L * a.pitch + k];
    cpx<R> acc(0, 0);
    const cpx<R>* row = lds + L * a.pitch;
    int idx = 0;
    for (int e = 0; e < n; ++e) {
      acc += row[e] * tw[idx];
      idx += k;
      if (idx >= n) idx -= n;
    }
    return acc;
  };
  if (a.out_elem == 1) {
    for (int idx = tid; idx < total; idx += FFT_THREADS) {
      int k = idx % n, L = idx / n;
      long g = g0 + L;
      if (g < a.total_lines) {
        long ib, ob, mid;
        line_addr(a, g, ib, ob, mid);
        cpx<R> v = value(L, k);
        if (a.twn) v = v * tw4[(long)k * mid % a.twn];
        out[ob + k] = v * scale;
      }
    }
  } else {
    for (int idx = tid; idx < total; idx += FFT_THREADS) {
      int L = idx % lines, k = idx / lines;
      long g = g0 + L;
      if (g < a.total_lines) {
        long ib, ob, mid;
        line_addr(a, g, ib, ob, mid);
        cpx<R> v = value(L, k);
        if (a.twn) v = v * tw4[(long)k * mid % a.twn];
        out[ob + (long)k * a.out_elem] = v * scale;
      }
    }
  }
}

static bool is_pow2(long n) { return n > 0 && (n & (n - 1)) == 0; }
static int ilog2(long n) { int l = 0; while ((1L << l) < n) ++l; return l; }

template <typename R>
FFTPlan<R>::FFTPlan(int nx, int ny, int nz, int m, int axes_mask) {
  dims_[0] = nx; dims_[1] = ny; dims_[2] = nz; dims_[3] = m;
  mask_ = axes_mask;
  const long total = (long)nx * ny * nz * m;
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
  long stride = 1;
  for (int ax = 0; ax < 4; ++ax) {
    const int n = dims_[ax];
    const long outer = total / (stride * n);
    if ((mask_ >> ax & 1) && n > 1) {
      if (n <= MAX_SINGLE) {
        FFTPassDesc p;
        p.n = n;
        p.n_inner = stride; p.n_mid = 1; p.n_outer = outer;
        p.in_mid = p.out_mid = 0;
        p.in_outer = p.out_outer = stride * n;
        p.in_elem = p.out_elem = stride;
        p.tw_off = add_table(n);
        passes_.push_back(p);
      } else {
        // four-step split n = n1 * n2 with both <= MAX_SINGLE
        int n1 = 0;
        if (is_pow2(n)) {  // balanced power-of-two factors
          int c = 1 << ((ilog2(n) + 1) / 2);
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
        FFTPassDesc B;  // length n2 over t2 -> transposed store k1 + n1*k2
        B.n = n2;
        B.n_inner = stride; B.n_mid = n1; B.n_outer = outer;
        B.in_mid = stride * n2; B.out_mid = stride;
        B.in_outer = B.out_outer = stride * n;
        B.in_elem = stride; B.out_elem = stride * n1;
        B.src = 1; B.dst = 0;
        B.tw_off = add_table(n2);
        passes_.push_back(A);
        passes_.push_back(B);
      }
    }
    stride *= n;
  }
  for (auto& p : passes_) {
    int budget = LdsBudget<R>::elems;
    int lines = std::max(1, budget / p.n);
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
void FFTPlan<R>::exec(cpx<R>* buf, cpx<R>* tmp, int sign, R scale, hipStream_t s) const {
  const cpx<R>* tw = tw_[sign < 0 ? 0 : 1];
  const int np = (int)passes_.size();
  if (np == 0) {
    if (scale != (R)1) throw Error("FFT with nothing to transform");
    return;
  }
  for (int i = 0; i < np; ++i) {
    const FFTPassDesc& p = passes_[i];
    PassArgs a;
    a.n = p.n;
    a.logn = is_pow2(p.n) ? ilog2(p.n) : 0;
    a.lines = p.lines;
    a.pitch = p.n + 1;
    a.n_inner = p.n_inner;
    a.n_mid = p.n_mid;
    a.total_lines = p.n_inner * p.n_mid * p.n_outer;
    a.in_mid = p.in_mid; a.in_outer = p.in_outer; a.in_elem = p.in_elem;
    a.out_mid = p.out_mid; a.out_outer = p.out_outer; a.out_elem = p.out_elem;
    a.twn = p.twn;
    const cpx<R>* src = p.src ? tmp : buf;
    cpx<R>* dst = p.dst ? tmp : buf;
    const R sc = (i == np - 1) ? scale : (R)1;
    const long blocks = (a.total_lines + p.lines - 1) / p.lines;
    const size_t shm = (size_t)p.lines * a.pitch * sizeof(cpx<R>);
    ProfScope ps(PROF_FFT, 2.0 * (double)size() * sizeof(cpx<R>), s);
    if (is_pow2(p.n))
      hipLaunchKernelGGL((fft_pass_kernel<R, true>), dim3((unsigned)blocks), dim3(FFT_THREADS), shm, s,
                         src, dst, a, tw + p.tw_off, tw + p.tw4_off, sc);
    else
      hipLaunchKernelGGL((fft_pass_kernel<R, false>), dim3((unsigned)blocks), dim3(FFT_THREADS), shm, s,
                         src, dst, a, tw + p.tw_off, tw + p.tw4_off, sc);
    GMOD_CHECK_LAUNCH();
  }
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
