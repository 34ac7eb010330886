// Batched multi-dimensional complex FFT over the (k-mesh x imaginary-time) grid.
//
// Grid layout: element (x,y,z,t) at x + Nx*(y + Ny*(z + Nz*t)), x fastest, the
// same as the reference's complex(0:Nx-1,0:Ny-1,0:Nz-1,2*Nw) work arrays
// (fchi.f90:529).  Conventions of the reference's FFT wrapper
// (fmod.f90:67-160): sign -1 unnormalised forward, sign +1 backward divided by
// the number of transformed points.
//
// Each pass transforms one axis.  A work-group loads `lines` lines of length n
// into LDS (coalesced: the lines of one group are adjacent in memory for
// strided axes, the elements are adjacent for the contiguous x axis), runs an
// in-place radix-2^2 decimation-in-time FFT in LDS (bit-reversed load,
// natural-order result), applies the optional four-step twiddle and the
// output scale, and stores.  Axes longer than 256 are split n = n1*n2 into two
// passes (four-step): pass A = length-n1 transforms with the w_n^(k1*t2)
// twiddle, pass B = length-n2 transforms with a transposing store.
// Non-power-of-two lengths use a direct DFT from LDS.
//
// The first pass reads through a Load functor and the last pass writes through
// a Store functor, so the irreducible -> full-grid scatter, the pointwise
// products and the full -> irreducible gather are fused into the transform.
// Forward transforms run x, y, z, t; backward transforms run t, z, y, x, so
// the last backward pass works on lines of one fixed t and skips the slots a
// gather does not keep.
#pragma once
#include "fftbox.h"
#include "gmod_common.h"

namespace gmod {

struct FFTPassDesc {
  int n = 1;                // transform length
  long n_inner = 1, n_mid = 1, n_outer = 1;
  long in_mid = 0, in_outer = 0, in_elem = 1;
  long out_mid = 0, out_outer = 0, out_elem = 1;
  int twn = 0;              // four-step twiddle: out[k] *= w_twn^(k*mid)
  int src = 0, dst = 0;     // 0 = caller buffer, 1 = scratch
  int lines = 1;            // lines per work-group
  long tw_off = 0;          // offset of the length-n table in the twiddle pool
  long tw4_off = 0;         // offset of the length-twn table
  int axis = 0;
};

struct PassArgs {
  int n, logn, lines, pitch;
  unsigned n_inner, n_mid, total_lines;
  unsigned in_mid, in_outer, in_elem;
  unsigned out_mid, out_outer, out_elem;
  unsigned S;           // spatial size Nx*Ny*Nz: address = sp + S * t
  int twn;
  int loglines;         // log2(lines) if lines is a power of two, else -1
};

constexpr int FFT_THREADS = 256;
constexpr int FFT_MAX_LINES = 512;

// Functors see (spatial index, slot) instead of a flat address.
template <typename R> struct GridLoad {
  const cpx<R>* p;
  unsigned S;
  // S, t < 2^24 and S t < 2^31 (checked by the plan): a 24-bit multiply
  __device__ cpx<R> operator()(unsigned sp, unsigned t) const { return p[sp + __umul24(S, t)]; }
  __host__ double algo_bytes(double N) const { return N * sizeof(cpx<R>); }
};
template <typename R> struct GridStore {
  cpx<R>* p;
  unsigned S;
  struct Pre {};
  __device__ Pre prep(unsigned, unsigned) const { return Pre{}; }
  __device__ void commit(unsigned sp, unsigned t, const Pre&, cpx<R> v) const { p[sp + __umul24(S, t)] = v; }
  __device__ void operator()(unsigned sp, unsigned t, cpx<R> v) const { p[sp + __umul24(S, t)] = v; }
  __host__ double algo_bytes(double N) const { return N * sizeof(cpx<R>); }
  __device__ bool skip(unsigned) const { return false; }
};

template <typename R, bool POW2, class LD, class ST>
__global__ void __launch_bounds__(FFT_THREADS)
fft_pass_kernel(LD ld, ST st, PassArgs a, const cpx<R>* __restrict__ tw, const cpx<R>* __restrict__ tw4, R scale) {
  extern __shared__ unsigned char smem_raw[];
  __shared__ unsigned l_spi[FFT_MAX_LINES], l_ti[FFT_MAX_LINES], l_spo[FFT_MAX_LINES], l_to[FFT_MAX_LINES],
      l_mid[FFT_MAX_LINES];
  cpx<R>* lds = reinterpret_cast<cpx<R>*>(smem_raw);
  const int n = a.n;
  const int lines = a.lines;
  const unsigned g0 = blockIdx.x * (unsigned)lines;
  const int tid = threadIdx.x;
  const int total = lines * n;
  const int nl = min(lines, (int)(a.total_lines - g0));
  // per-line decomposition (once per line, 32-bit)
  for (int L = tid; L < nl; L += FFT_THREADS) {
    unsigned g = g0 + L;
    unsigned inner = g % a.n_inner, rest = g / a.n_inner;
    unsigned mid = rest % a.n_mid, outer = rest / a.n_mid;
    unsigned ib = inner + mid * a.in_mid + outer * a.in_outer;
    unsigned ob = inner + mid * a.out_mid + outer * a.out_outer;
    l_spi[L] = ib % a.S; l_ti[L] = ib / a.S;
    l_spo[L] = ob % a.S; l_to[L] = ob / a.S;
    l_mid[L] = mid;
  }
  __syncthreads();
  // a line moves along space (stride < S) or along t (stride a multiple of S)
  const unsigned dsi = a.in_elem < a.S ? a.in_elem : 0, dti = a.in_elem < a.S ? 0 : a.in_elem / a.S;
  const unsigned dso = a.out_elem < a.S ? a.out_elem : 0, dto = a.out_elem < a.S ? 0 : a.out_elem / a.S;
  // lines of one fixed t (slots increase with the line index): skip the block
  // when the store keeps nothing of its first and last line
  if (dto == 0 && st.skip(l_to[0]) && st.skip(l_to[nl - 1])) return;
  const int shift = 32 - a.logn;
  // ---- load
  if (a.in_elem == 1) {
    for (int idx = tid; idx < total; idx += FFT_THREADS) {
      int e = POW2 ? (idx & (n - 1)) : idx % n;
      int L = POW2 ? (idx >> a.logn) : idx / n;
      if (L < nl) {
        int pe = POW2 ? (int)(__brev((unsigned)e) >> shift) : e;
        lds[L * a.pitch + pe] = ld(l_spi[L] + e * dsi, l_ti[L] + e * dti);
      }
    }
  } else {
    for (int idx = tid; idx < total; idx += FFT_THREADS) {
      int L = a.loglines >= 0 ? (idx & (lines - 1)) : idx % lines;
      int e = a.loglines >= 0 ? (idx >> a.loglines) : idx / lines;
      if (L < nl) {
        int pe = POW2 ? (int)(__brev((unsigned)e) >> shift) : e;
        lds[L * a.pitch + pe] = ld(l_spi[L] + e * dsi, l_ti[L] + e * dti);
      }
    }
  }
  __syncthreads();
  if (POW2) {
    int lh = 0;
    if (a.logn & 1) {  // one radix-2 stage (h = 1, twiddle 1)
      const int half = n >> 1;
      const int units = lines * half;
      for (int u = tid; u < units; u += FFT_THREADS) {
        int L = u >> (a.logn - 1), p = u & (half - 1);
        cpx<R>* b = lds + L * a.pitch + 2 * p;
        cpx<R> x0 = b[0], x1 = b[1];
        b[0] = x0 + x1;
        b[1] = x0 - x1;
      }
      lh = 1;
      __syncthreads();
    }
    for (; lh < a.logn; lh += 2) {  // radix-2^2 stages (h, 2h), h = 2^lh
      const int h = 1 << lh;
      const int lq = a.logn - 2;
      const int units = lines << lq;
      const int s2 = a.logn - 1 - lh, s4 = a.logn - 2 - lh;  // log2 of n/(2h), n/(4h)
      for (int u = tid; u < units; u += FFT_THREADS) {
        int L = u >> lq, q = u & ((1 << lq) - 1);
        int blk = q >> lh, j = q & (h - 1);
        cpx<R>* b = lds + L * a.pitch + (blk << (lh + 2)) + j;
        cpx<R> w1 = tw[j << s2], w2 = tw[j << s4], w3 = tw[(j + h) << s4];
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
  // ---- store
  if (a.out_elem == 1) {
    for (int idx = tid; idx < total; idx += FFT_THREADS) {
      int k = POW2 ? (idx & (n - 1)) : idx % n;
      int L = POW2 ? (idx >> a.logn) : idx / n;
      if (L < nl) {
        cpx<R> v = value(L, k);
        if (a.twn) {
          unsigned ti = (unsigned)k * l_mid[L];
          v = v * tw4[(a.twn & (a.twn - 1)) == 0 ? (ti & (unsigned)(a.twn - 1)) : ti % (unsigned)a.twn];
        }
        st(l_spo[L] + k * dso, l_to[L] + k * dto, v * scale);
      }
    }
  } else {
    for (int idx = tid; idx < total; idx += FFT_THREADS) {
      int L = a.loglines >= 0 ? (idx & (lines - 1)) : idx % lines;
      int k = a.loglines >= 0 ? (idx >> a.loglines) : idx / lines;
      if (L < nl) {
        cpx<R> v = value(L, k);
        if (a.twn) {
          unsigned ti = (unsigned)k * l_mid[L];
          v = v * tw4[(a.twn & (a.twn - 1)) == 0 ? (ti & (unsigned)(a.twn - 1)) : ti % (unsigned)a.twn];
        }
        st(l_spo[L] + k * dso, l_to[L] + k * dto, v * scale);
      }
    }
  }
}

template <typename R>
class FFTPlan {
 public:
  // axes_mask: bit a set -> transform axis a (0:x 1:y 2:z 3:t)
  FFTPlan(int nx, int ny, int nz, int m, int axes_mask = 15);
  ~FFTPlan();
  // In place on `buf` (scratch `tmp` of the same size).  sign=-1 forward,
  // sign=+1 backward; `scale` multiplies the result (1/N for the normalised
  // backward transform).
  void exec(cpx<R>* buf, cpx<R>* tmp, int sign, R scale, hipStream_t s) const {
    exec_fused(GridLoad<R>{buf, S_}, GridStore<R>{buf, S_}, buf, tmp, sign, scale, s);
  }
  // First pass reads through `ld`, last pass writes through `st`; buf/tmp hold
  // the intermediate passes.
  // spatial_first (box plans): run the passes in the forward order
  // (spatial, then t) whatever the sign.  A product load that reads an
  // operand at (-r, -t) then reads whole mirrored spatial planes, not row
  // segments that straddle cache lines.
  template <class LD, class ST>
  void exec_fused(LD ld, ST st, cpx<R>* buf, cpx<R>* tmp, int sign, R scale, hipStream_t s,
                  bool spatial_first = false) const;
  // Accumulating form (box_accum_kernel): exec_head runs every pass but the
  // last for one source (ld -> out; tmp as intermediate); exec_accum_last runs
  // the last pass of nsrc such sources (src + c stride) into the accumulator
  // functor ac.  accum_ok: the plan supports it for this sign (box passes,
  // 2 or 3 of them, a tiled last pass of exactly BOX_TILE_REGS tiles with
  // radices <= 16).
  bool accum_ok(int sign) const;
  // the accumulating (last) pass transforms the t axis alone: the sources it
  // reads hold spatial transforms only
  bool accum_last_is_t(int sign) const {
    const std::vector<BoxArgs>& bp = sign < 0 ? box_.fwd : box_.bwd;
    if (!box_.ok || bp.empty()) return false;
    const BoxArgs& a = bp.back();
    return a.nrad[0] == 0 && a.nrad[1] == 0 && a.nrad[2] == 0 && a.nrad[3] > 0;
  }
  template <class LD>
  void exec_head(LD ld, cpx<R>* out, cpx<R>* tmp, int sign, hipStream_t s) const;
  // srcmap (device, nsrc entries, optional): source c is buffer srcmap[c] >> 1,
  // read conjugated at (-r, -t) when srcmap[c] & 1, plus at t = 0 the
  // correction dhat[(srcmap[c] >> 1) S + sp] (S spatial points per buffer)
  template <class AC>
  void exec_accum_last(const cpx<R>* src, long stride, int nsrc, AC ac, int sign, R scale, hipStream_t s,
                       const int* srcmap = nullptr, const cpx<R>* dhat = nullptr) const;
  // the conjugated reversed reads of a source map are available (grid range)
  bool reversal_ok() const { return sprev_ok(dims_[0], dims_[1], dims_[2]); }
  // Boxes of the forward list's last pass (the last pass of a sign -1 or a
  // spatial-first transform) that a gather keeping the spatial points with
  // kept_sp[sp] >= 0 needs: per box index (after box_remap) 1 if the box
  // holds such a point.  Empty unless that pass's boxes span the t axis and
  // hold few spatial points each; *frac receives the share of live boxes.
  std::vector<unsigned char> fwd_last_live(const std::vector<int>& kept_sp, double* frac) const;
  long size() const { return (long)dims_[0] * dims_[1] * dims_[2] * dims_[3]; }
  unsigned spatial() const { return S_; }
  long transformed_points() const;
  int num_passes() const { return box_.ok ? (int)box_.fwd.size() : (int)fwd_.size(); }
  bool uses_box() const { return box_.ok; }

 private:
  template <class LD, class ST>
  void launch(const FFTPassDesc& p, LD ld, ST st, const cpx<R>* tw, R sc, hipStream_t s) const;
  template <class LD, class ST>
  void launch_box(int sg, int i, LD ld, ST st, const cpx<R>* tw, R sc, hipStream_t s) const;
  int dims_[4];
  int mask_;
  unsigned S_ = 1;
  std::vector<FFTPassDesc> fwd_, bwd_;
  cpx<R>* tw_[2] = {nullptr, nullptr};  // [0]: sign -1 tables, [1]: sign +1 tables
  long tw_len_ = 0;
  BoxPlan<R> box_;                        // box passes (preferred when box_.ok)
  cpx<R>* twb_[2] = {nullptr, nullptr};   // their twiddle pools
  int* tabs_ = nullptr;                   // their tile tables
  BoxArgs* argsd_[2] = {nullptr, nullptr};  // device copies of box_.fwd / box_.bwd
  long persist_ = 256;                    // work-groups of a persistent tiled pass
};

// Algorithmic bytes of one pass: what its load and store functors must move
// (algo_bytes; a functor without it counts as one grid read or write).
template <class F, class = void>
struct AlgoBytes {
  template <typename R>
  static double get(const F&, double N) { return N * sizeof(cpx<R>); }
};
template <class F>
struct AlgoBytes<F, std::void_t<decltype(std::declval<const F&>().algo_bytes(1.0))>> {
  template <typename R>
  static double get(const F& f, double N) { return f.algo_bytes(N); }
};
// Share of the input a pass reads when its store drops whole boxes
// (read_fraction(t_box); t_box: the pass's boxes span the t axis).
template <class F, class = void>
struct ReadFraction {
  static double get(const F&, bool) { return 1.0; }
};
template <class F>
struct ReadFraction<F, std::void_t<decltype(std::declval<const F&>().read_fraction(true))>> {
  static double get(const F& f, bool t_box) { return f.read_fraction(t_box); }
};
template <typename R, class LD, class ST>
double pass_bytes(const LD& ld, const ST& st, double N, bool t_box = false) {
  return AlgoBytes<LD>::template get<R>(ld, N * ReadFraction<ST>::get(st, t_box)) + AlgoBytes<ST>::template get<R>(st, N);
}

inline bool fft_is_pow2(long n) { return n > 0 && (n & (n - 1)) == 0; }
inline int fft_ilog2(long n) { int l = 0; while ((1L << l) < n) ++l; return l; }

template <typename R>
template <class LD, class ST>
void FFTPlan<R>::launch(const FFTPassDesc& p, LD ld, ST st, const cpx<R>* tw, R sc, hipStream_t s) const {
  PassArgs a;
  a.n = p.n;
  a.logn = fft_is_pow2(p.n) ? fft_ilog2(p.n) : 0;
  a.lines = p.lines;
  a.pitch = p.n + 1;
  a.n_inner = (unsigned)p.n_inner;
  a.n_mid = (unsigned)p.n_mid;
  a.total_lines = (unsigned)(p.n_inner * p.n_mid * p.n_outer);
  a.in_mid = (unsigned)p.in_mid; a.in_outer = (unsigned)p.in_outer; a.in_elem = (unsigned)p.in_elem;
  a.out_mid = (unsigned)p.out_mid; a.out_outer = (unsigned)p.out_outer; a.out_elem = (unsigned)p.out_elem;
  a.S = S_;
  a.twn = p.twn;
  a.loglines = fft_is_pow2(p.lines) ? fft_ilog2(p.lines) : -1;
  const long blocks = (a.total_lines + p.lines - 1) / p.lines;
  const size_t shm = (size_t)p.lines * a.pitch * sizeof(cpx<R>);
  ProfScope ps(sizeof(R) == 4 ? PROF_FFT32 : PROF_FFT, pass_bytes<R>(ld, st, (double)size()), s);
  if (fft_is_pow2(p.n))
    hipLaunchKernelGGL((fft_pass_kernel<R, true, LD, ST>), dim3((unsigned)blocks), dim3(FFT_THREADS), shm, s, ld, st,
                       a, tw + p.tw_off, tw + p.tw4_off, sc);
  else
    hipLaunchKernelGGL((fft_pass_kernel<R, false, LD, ST>), dim3((unsigned)blocks), dim3(FFT_THREADS), shm, s, ld,
                       st, a, tw + p.tw_off, tw + p.tw4_off, sc);
  GMOD_CHECK_LAUNCH();
}

template <typename R>
template <class LD, class ST>
void FFTPlan<R>::launch_box(int sg, int i, LD ld, ST st, const cpx<R>* tw, R sc, hipStream_t s) const {
  const BoxArgs& a = (sg == 0 ? box_.fwd : box_.bwd)[i];
  const size_t shm = (size_t)(a.lds_elems + a.ntwl) * sizeof(cpx<R>);
  static bool attr_done = false;  // one per kernel instantiation: allow the device's whole LDS
  if (!attr_done) {
    GMOD_HIP(hipFuncSetAttribute((const void*)&box_fft_kernel<R, LD, ST>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)box_.lds_limit));
    attr_done = true;
  }
  ProfScope ps(sizeof(R) == 4 ? PROF_FFT32 : PROF_FFT, pass_bytes<R>(ld, st, (double)size(), a.box[3] == dims_[3]),
               s);
  // tiled passes: persistent work-groups, one per CU, a multiple of 8
  long grid = a.blocks();
  if (a.tiled && grid > persist_) grid = persist_;
  // the pool a pass reads its tables from is the sign's; the power-of-two
  // sub-passes take the sign itself at compile time
  const int sign = tw == twb_[0] ? -1 : 1;
  hipLaunchKernelGGL((box_fft_kernel<R, LD, ST>), dim3((unsigned)grid), dim3(BOX_THREADS), shm, s, ld, st,
                     (const BoxArgs*)(argsd_[sg] + i), tw, (const int*)tabs_, sc, sign);
  GMOD_CHECK_LAUNCH();
}

template <typename R>
template <class LD, class ST>
void FFTPlan<R>::exec_fused(LD ld, ST st, cpx<R>* buf, cpx<R>* tmp, int sign, R scale, hipStream_t s,
                            bool spatial_first) const {
  if (box_.ok) {
    const cpx<R>* twb = twb_[sign < 0 ? 0 : 1];  // the passes of either list run with the sign's twiddles
    const int sg = (sign < 0 || spatial_first) ? 0 : 1;
    const std::vector<BoxArgs>& bp = sg == 0 ? box_.fwd : box_.bwd;
    const int nb = (int)bp.size();
    auto bo = [&](int which) { return which ? tmp : buf; };
    for (int i = 0; i < nb; ++i) {
      const BoxArgs& a = bp[i];
      const R sc = (i == nb - 1) ? scale : (R)1;
      const bool first = i == 0, last = i == nb - 1;
      if (first && last) launch_box(sg, i, ld, st, twb, sc, s);
      else if (first) launch_box(sg, i, ld, GridStore<R>{bo(a.dst), S_}, twb, sc, s);
      else if (last) launch_box(sg, i, GridLoad<R>{bo(a.src), S_}, st, twb, sc, s);
      else launch_box(sg, i, GridLoad<R>{bo(a.src), S_}, GridStore<R>{bo(a.dst), S_}, twb, sc, s);
    }
    return;
  }
  const cpx<R>* tw = tw_[sign < 0 ? 0 : 1];
  const std::vector<FFTPassDesc>& ps = sign < 0 ? fwd_ : bwd_;
  const int np = (int)ps.size();
  if (np == 0) throw Error("FFT with nothing to transform");
  auto buf_of = [&](int which) { return which ? tmp : buf; };
  for (int i = 0; i < np; ++i) {
    const FFTPassDesc& p = ps[i];
    const R sc = (i == np - 1) ? scale : (R)1;
    const bool first = i == 0, last = i == np - 1;
    if (first && last) launch(p, ld, st, tw, sc, s);
    else if (first) launch(p, ld, GridStore<R>{buf_of(p.dst), S_}, tw, sc, s);
    else if (last) launch(p, GridLoad<R>{buf_of(p.src), S_}, st, tw, sc, s);
    else launch(p, GridLoad<R>{buf_of(p.src), S_}, GridStore<R>{buf_of(p.dst), S_}, tw, sc, s);
  }
}

template <typename R>
bool FFTPlan<R>::accum_ok(int sign) const {
  if (!box_.ok) return false;
  const std::vector<BoxArgs>& bp = sign < 0 ? box_.fwd : box_.bwd;
  const int nb = (int)bp.size();
  if (nb < 2 || nb > 3) return false;
  for (const BoxArgs& a : bp)
    if (a.tl_mul != a.ts_mul || a.tl_str != a.ts_str || a.tw4_n) return false;  // no slot-moving (split) pass
  const BoxArgs& l = bp[nb - 1];
  return l.tiled && l.ntiles == BOX_TILE_REGS<R> && box_max_radix(l) <= 16;
}

template <typename R>
template <class LD>
void FFTPlan<R>::exec_head(LD ld, cpx<R>* out, cpx<R>* tmp, int sign, hipStream_t s) const {
  const int sg = sign < 0 ? 0 : 1;
  const int nb = (int)(sg == 0 ? box_.fwd : box_.bwd).size();
  const cpx<R>* twb = twb_[sg];
  if (nb == 2) {
    launch_box(sg, 0, ld, GridStore<R>{out, S_}, twb, (R)1, s);
  } else {  // 3 passes: ld -> tmp -> out
    launch_box(sg, 0, ld, GridStore<R>{tmp, S_}, twb, (R)1, s);
    launch_box(sg, 1, GridLoad<R>{tmp, S_}, GridStore<R>{out, S_}, twb, (R)1, s);
  }
}

template <typename R>
template <class AC>
void FFTPlan<R>::exec_accum_last(const cpx<R>* src, long stride, int nsrc, AC ac, int sign, R scale,
                                 hipStream_t s, const int* srcmap, const cpx<R>* dhat) const {
  if (srcmap && !reversal_ok()) throw Error("exec_accum_last: grid too large for the reversed reads");
  if (srcmap && !dhat) throw Error("exec_accum_last: a source map needs the slot-0 corrections");
  const int sg = sign < 0 ? 0 : 1;
  const std::vector<BoxArgs>& bp = sg == 0 ? box_.fwd : box_.bwd;
  const int i = (int)bp.size() - 1;
  const BoxArgs& a = bp[i];
  const size_t shm = (size_t)(a.lds_elems + a.ntwl) * sizeof(cpx<R>);
  static bool attr_done = false;
  if (!attr_done) {
    GMOD_HIP(hipFuncSetAttribute((const void*)&box_accum_kernel<R, 16, AC>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)box_.lds_limit));
    attr_done = true;
  }
  const double N = (double)size();
  ProfScope ps(sizeof(R) == 4 ? PROF_FFT32 : PROF_FFT,
               nsrc * (N * sizeof(cpx<R>) + ac.src_bytes(N)) + ac.sum_bytes(N), s);
  long grid = a.blocks();
  if (grid > persist_) grid = persist_;
  hipLaunchKernelGGL((box_accum_kernel<R, 16, AC>), dim3((unsigned)grid), dim3(BOX_THREADS), shm, s, src, stride,
                     nsrc, ac, (const BoxArgs*)(argsd_[sg] + i), twb_[sg], (const int*)tabs_, scale, sg == 0 ? -1 : 1,
                     srcmap, make_sprev(dims_[0], dims_[1], dims_[2]), dhat);
  GMOD_CHECK_LAUNCH();
}

// Cached plans keyed by shape and precision.
template <typename R>
const FFTPlan<R>& fft_plan(int nx, int ny, int nz, int m, int axes_mask = 15);

}  // namespace gmod
