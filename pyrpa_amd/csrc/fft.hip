// Hand-written batched FFT for gfx950: plan construction (pass lists for the
// forward x,y,z,t and backward t,z,y,x orders), twiddle tables, plan cache.
// The pass kernel and the fused-exec template live in fft.h.
#include <algorithm>
#include <cmath>
#include <map>
#include <tuple>

#include "fft.h"

namespace gmod {

constexpr int MAX_SINGLE = 256;

template <typename R> struct LdsBudget;
template <> struct LdsBudget<double> { static constexpr int elems = 4096; };
template <> struct LdsBudget<float> { static constexpr int elems = 8192; };

// ---------------------------------------------------------------- box plans
// radices of one dimension: powers of two up to maxr (balanced), then 7, 5, 3
static bool box_factor(int n, int maxr, std::vector<int>& out) {
  out.clear();
  if (n <= 1) return true;
  int e = 0, o = n;
  while ((o & 1) == 0) { o >>= 1; ++e; }
  std::vector<int> odd;
  for (int p : {7, 5, 3})
    while (o % p == 0) { odd.push_back(p); o /= p; }
  if (o != 1) return false;
  const int lg = fft_ilog2(maxr);
  if (e > 0) {
    const int k = (e + lg - 1) / lg;
    const int base = e / k, extra = e % k;
    for (int i = 0; i < k; ++i) out.push_back(1 << (base + (i < extra ? 1 : 0)));
  }
  for (int p : odd) out.push_back(p);
  return (int)out.size() <= BOX_MAX_RAD;
}

static int largest_divisor_le(int n, long cap) {
  if (cap < 1) return 0;
  for (int d = (int)std::min<long>(n, cap); d >= 1; --d)
    if (n % d == 0) return d;
  return 1;
}

// Tiled load/store of a box (see BoxArgs): power-of-two tile extents, x
// first, whose product is BOX_THREADS; accepted only when every store
// position is additive over (tile origin, in-tile offset), which is checked
// here digit by digit for every dimension.
static void set_tiles(BoxArgs& a, std::vector<int>& tabs, int max_tiles) {
  a.tiled = 0;
  a.ntiles = 0;
  if (a.elems % BOX_THREADS != 0 || a.elems / BOX_THREADS > std::min(max_tiles, BOX_MAX_TILES)) return;
  if (a.tl_mul != a.ts_mul || a.tl_str != a.ts_str || a.tw4_n) return;
  int rem = BOX_THREADS, tile[4];
  for (int d = 0; d < 4; ++d) {
    int t = 1;
    while (t * 2 <= rem && a.box[d] % (t * 2) == 0) t *= 2;
    tile[d] = t;
    rem /= t;
    int lg = 0;
    while ((1 << lg) < t) ++lg;
    a.tlog[d] = lg;
  }
  if (rem != 1) return;
  for (int d = 0; d < 4; ++d)
    for (int o = 0; o < a.box[d]; o += tile[d])
      for (int c = 0; c < tile[d]; ++c)
        if (box_outpos(a, d, o + c) != box_outpos(a, d, o) + box_outpos(a, d, c)) return;
  const int nt[4] = {a.box[0] / tile[0], a.box[1] / tile[1], a.box[2] / tile[2], a.box[3] / tile[3]};
  const int n = a.elems / BOX_THREADS;
  a.ntiles = n;
  a.tab_off = (int)tabs.size();
  tabs.resize(tabs.size() + 4 * (size_t)n);
  int* t = tabs.data() + a.tab_off;
  for (int j = 0; j < n; ++j) {
    int r = j, o[4];
    for (int d = 0; d < 4; ++d) {
      o[d] = (r % nt[d]) * tile[d];
      r /= nt[d];
    }
    t[j] = o[0] + (int)a.Nx * (o[1] + (int)a.Ny * o[2]);
    t[n + j] = o[3] * a.tl_str;
    t[2 * n + j] = o[0] + o[1] * a.lds_stride[1] + o[2] * a.lds_stride[2] + o[3] * a.lds_stride[3];
    t[3 * n + j] = box_outpos(a, 0, o[0]) + box_outpos(a, 1, o[1]) * a.lds_stride[1] +
                   box_outpos(a, 2, o[2]) * a.lds_stride[2] + box_outpos(a, 3, o[3]) * a.lds_stride[3];
  }
  // the prefetch of the next box goes after the last sub-pass that reads
  // twiddles (stride L / radix > 1) from global memory: reads issued after
  // it would wait for the prefetch (memory returns in order)
  a.pf_at = 0;
  for (int i = 0; i < a.nsub; ++i)
    if (a.sub[i].L / a.sub[i].radix > 1) a.pf_at = i + 1;
  a.tiled = 1;
}

template <typename R>
void BoxPlan<R>::build(int nx, int ny, int nz, int m, int mask, long budget, size_t lds_bytes) {
  ok = false;
  fwd.clear();
  bwd.clear();
  tab_host.clear();
  const int n[4] = {nx, ny, nz, m};
  const int maxr = sizeof(R) == 4 ? 32 : 16;
  const long b = (long)sizeof(cpx<R>);
  std::vector<int> tables;
  auto add_table = [&](int len) -> int {
    int off = 0;
    for (int t : tables) {
      if (t == len) return off;
      off += t;
    }
    tables.push_back(len);
    return off;
  };
  max_lds_bytes = 0;
  lds_limit = lds_bytes;
  auto make_box = [&](const int ext[4], const std::vector<int>& tdims, BoxArgs& a) -> bool {
    a = BoxArgs{};
    long elems = 1;
    for (int d = 0; d < 4; ++d) {
      if (ext[d] < 1 || n[d] % ext[d] != 0) return false;
      a.box[d] = ext[d];
      a.nblk[d] = n[d] / ext[d];
      elems *= ext[d];
    }
    // x rows padded by one element against LDS bank conflicts when lines run
    // along x or the rows are long; narrow rows of a t box stay unpadded
    const bool x_t = std::find(tdims.begin(), tdims.end(), 0) != tdims.end();
    const int px = ext[0] + ((ext[0] > 1 && elems > ext[0] && (x_t || ext[0] >= 8)) ? 1 : 0);
    a.lds_stride[0] = 1;
    a.lds_stride[1] = px;
    a.lds_stride[2] = px * ext[1];
    a.lds_stride[3] = px * ext[1] * ext[2];
    const long lds_elems = (long)px * ext[1] * ext[2] * ext[3];
    if (lds_elems * b > (long)lds_bytes || elems > budget) return false;
    a.lds_elems = (int)lds_elems;
    a.elems = (int)elems;
    a.nsub = 0;
    a.tl_mul = a.ts_mul = ext[3];
    a.tl_str = a.ts_str = 1;
    a.tw4_n = a.tw4_off = 0;
    for (int d : tdims) {
      std::vector<int> rads;
      if (!box_factor(ext[d], maxr, rads)) return false;
      a.nrad[d] = (int)rads.size();
      int L = ext[d], w = ext[d];
      for (size_t i = 0; i < rads.size(); ++i) {
        const int r = rads[i];
        w /= r;
        a.rad[d][i] = r;
        a.radw[d][i] = w;
        if (a.nsub >= BOX_MAX_SUB) return false;
        BoxSub sb{};
        sb.dim = d;
        sb.radix = r;
        sb.L = L;
        sb.tw_off = L / r > 1 ? add_table(L) : 0;
        sb.wr_off = add_table(r);
        a.sub[a.nsub++] = sb;
        L /= r;
      }
    }
    // multiply-high divisors (exact while elems * divisor < 2^32)
    auto fd = [&](long d) -> FDiv {
      if (d < 1 || elems * d >= (1L << 32)) throw Error("box FFT: divisor out of the multiply-high range");
      return make_fdiv((unsigned)d);
    };
    for (int d = 0; d < 3; ++d) a.fbox[d] = fd(ext[d]);
    for (int d = 0; d < 4; ++d)
      for (int i = 0; i < BOX_MAX_RAD; ++i) a.frad[d][i] = fd(i < a.nrad[d] ? a.rad[d][i] : 1);
    for (int i = 0; i < a.nsub; ++i) {
      BoxSub& sb = a.sub[i];
      const int d = sb.dim;
      int od[3], c = 0;
      for (int e = 0; e < 4; ++e)
        if (e != d) od[c++] = e;
      for (int j = 0; j < 3; ++j) sb.os[j] = a.lds_stride[od[j]];
      sb.n = ext[d];
      sb.ld = a.lds_stride[d];
      sb.foe[0] = fd(ext[od[0]]);
      sb.foe[1] = fd(ext[od[1]]);
      sb.flines = fd(elems / ext[d]);
      sb.fs = fd(sb.L / sb.radix);
    }
    // twiddles of the sub-passes with stride > 1 from per-dimension w_n
    // tables in LDS after the box, when they fit
    a.ntwl = a.ntwc = 0;
    for (int i = 0; i < a.nsub; ++i) {
      a.sub[i].twl = -1;
      a.sub[i].tws = 1;
    }
    {
      long need = 0;
      for (int d : tdims)
        if (a.nrad[d] >= 2) need += ext[d];
      if (need > 0 && (lds_elems + need) * b <= (long)lds_bytes && a.nsub > 0) {
        int dst = (int)lds_elems;
        for (int d : tdims) {
          if (a.nrad[d] < 2) continue;
          int first = -1;
          for (int i = 0; i < a.nsub; ++i)
            if (a.sub[i].dim == d && first < 0) first = i;
          a.twc_src[a.ntwc] = a.sub[first].tw_off;  // the w_{ext[d]} table (L = ext[d] on the first sub-pass)
          a.twc_dst[a.ntwc] = dst;
          a.twc_len[a.ntwc] = ext[d];
          ++a.ntwc;
          for (int i = 0; i < a.nsub; ++i)
            if (a.sub[i].dim == d && a.sub[i].L / a.sub[i].radix > 1) {
              a.sub[i].twl = dst;
              a.sub[i].tws = ext[d] / a.sub[i].L;
            }
          dst += ext[d];
        }
        a.ntwl = dst - (int)lds_elems;
        max_lds_bytes = std::max(max_lds_bytes, (size_t)((lds_elems + a.ntwl) * b));
      }
    }
    a.Nx = (unsigned)nx;
    a.Ny = (unsigned)ny;
    a.S = (unsigned)((long)nx * ny * nz);
    a.M = (unsigned)m;
    set_tiles(a, tab_host, BOX_TILE_REGS<R>);
    // bytes contiguous in the grid per box row; boxes sharing a 128 B line
    // are kept on one XCD
    long seg = ext[0] * b;
    if (ext[0] == nx) {
      seg = (long)nx * ext[1] * b;
      if (ext[1] == ny) seg = (long)nx * ny * ext[2] * b;
    }
    a.xcd_group = (seg < 128 && 128 % seg == 0) ? (int)(128 / seg) : 1;
    a.src = a.dst = 0;
    max_lds_bytes = std::max(max_lds_bytes, (size_t)(lds_elems * b));
    return true;
  };
  std::vector<int> tx;
  for (int d = 0; d < 4; ++d)
    if (((mask >> d) & 1) && n[d] > 1) tx.push_back(d);
  if (tx.empty()) return;
  std::vector<int> sp;
  for (int d : tx)
    if (d < 3) sp.push_back(d);
  std::vector<std::vector<int>> groups;
  if (!sp.empty()) {
    long vol = 1;
    for (int d : sp) vol *= n[d];
    if (vol <= budget) {
      groups.push_back(sp);
    } else {
      std::vector<int> xy, rest;
      for (int d : sp) (d < 2 ? xy : rest).push_back(d);
      long vxy = 1;
      for (int d : xy) vxy *= n[d];
      if (xy.size() == 2 && vxy <= budget) {
        groups.push_back(xy);
        if (!rest.empty()) groups.push_back(rest);
      } else {
        for (int d : sp) groups.push_back({d});
      }
    }
  }
  // boxes are not capped below the LDS budget: halving them so that two
  // work-groups share a CU measured no faster on the 4-orbital 32x32x16 x
  // 2048 grid (FP64 unchanged, FP32 13 % slower with the 32 B rows of a
  // 4-column t box)
  const long small = budget;
  std::vector<BoxArgs> spatial;
  for (auto& gdims : groups) {
    int ext[4] = {1, 1, 1, 1};
    long vol = 1;
    for (int d : gdims) { ext[d] = n[d]; vol *= n[d]; }
    if (vol > budget) return;
    const long cap = std::max(vol, small);
    if (ext[0] == 1) {  // keep rows contiguous: adjacent x columns in the box
      ext[0] = largest_divisor_le(nx, std::min<long>(256 / b, cap / vol));
      vol *= ext[0];
    }
    ext[3] = largest_divisor_le(m, cap / vol);
    BoxArgs a;
    if (!make_box(ext, gdims, a)) return;
    spatial.push_back(a);
  }
  const bool do_t = std::find(tx.begin(), tx.end(), 3) != tx.end();
  BoxArgs tb, tA, tB;
  bool split = false;
  if (do_t) {
    if (m > budget) return;
    const long cap = std::max<long>(m, small);
    int ext[4] = {1, 1, 1, m};
    ext[0] = largest_divisor_le(nx, cap / m);
    if (ext[0] == nx) ext[1] = largest_divisor_le(ny, cap / ((long)m * nx));
    if (ext[0] == nx && ext[1] == ny) ext[2] = largest_divisor_le(nz, cap / ((long)m * nx * ny));
    if (!make_box(ext, {3}, tb)) return;
    // Optional (GMOD_FFT_TSPLIT=1): for a t box whose rows are shorter than a
    // 128 B cache line, split t = t2 + M2 t1 (four-step): pass A transforms
    // t1 for each t2 (twiddle w_M^(t2 k1)), pass B transforms t2 for each k1
    // and stores k1 + M1 k2 in natural order; both boxes hold whole x rows.
    // Measured on the 4-orbital 32x32x16 x 2048 grid, the two passes take
    // longer than the one narrow t box (FP32 2 x 296 us vs 375 us, FP64
    // 2 x 413 us vs 578 us), so it is off by default.
    const long row = (ext[0] < nx ? (long)ext[0] : (long)nx * ext[1]) * b;
    const char* ts = getenv("GMOD_FFT_TSPLIT");
    if (row < 128 && (long)nx * b >= 128 && ts && atoi(ts) != 0) {
      int M1 = 1;
      while ((long)M1 * M1 < m) M1 *= 2;
      while (M1 > 1 && m % M1 != 0) M1 /= 2;
      const int M2 = m / M1;
      std::vector<int> r1, r2;
      if (M1 > 1 && M2 > 1 && box_factor(M1, maxr, r1) && box_factor(M2, maxr, r2)) {
        // as many whole x rows per box as the budget allows: each slot row of
        // a box is then one contiguous run of the grid (2 KB and more at the
        // benchmark shapes), which DRAM reads and writes as whole pages
        int eA[4] = {nx, largest_divisor_le(ny, std::max<long>(1, budget / ((long)nx * M1))), 1, M1};
        int eB[4] = {nx, largest_divisor_le(ny, std::max<long>(1, budget / ((long)nx * M2))), 1, M2};
        if (eA[1] == ny) eA[2] = largest_divisor_le(nz, std::max<long>(1, budget / ((long)nx * ny * M1)));
        if (eB[1] == ny) eB[2] = largest_divisor_le(nz, std::max<long>(1, budget / ((long)nx * ny * M2)));
        if (make_box(eA, {3}, tA) && make_box(eB, {3}, tB)) {
          tA.tl_mul = tA.ts_mul = 1;            // bt = t2
          tA.tl_str = tA.ts_str = M2;
          tA.tw4_n = m;
          tA.tw4_off = add_table(m);
          tB.tl_mul = M2;                       // bt = k1: loads t = k1 M2 + t2
          tB.tl_str = 1;
          tB.ts_mul = 1;                        // stores t = k1 + M1 k2
          tB.ts_str = M1;
          set_tiles(tA, tab_host, BOX_TILE_REGS<R>);
          set_tiles(tB, tab_host, BOX_TILE_REGS<R>);
          split = true;
        }
      }
    }
  }
  // buffers (0: caller's, 1: scratch): a pass that moves data between slots
  // (pass B) never writes the buffer it reads
  if (!split) {
    for (auto& a : spatial) fwd.push_back(a);
    if (do_t) fwd.push_back(tb);
    if (do_t) bwd.push_back(tb);
    for (auto it = spatial.rbegin(); it != spatial.rend(); ++it) bwd.push_back(*it);
  } else {
    for (auto& a : spatial) { a.src = a.dst = 0; fwd.push_back(a); }
    tA.src = 0; tA.dst = 1;
    tB.src = 1; tB.dst = 0;
    fwd.push_back(tA);
    fwd.push_back(tB);
    // backward: A then B, then the spatial passes in place on the scratch
    BoxArgs bA = tA, bB = tB;
    if (spatial.empty()) { bA.src = 0; bA.dst = 1; bB.src = 1; bB.dst = 0; }
    else { bA.src = 0; bA.dst = 0; bB.src = 0; bB.dst = 1; }
    bwd.push_back(bA);
    bwd.push_back(bB);
    for (auto it = spatial.rbegin(); it != spatial.rend(); ++it) {
      BoxArgs a = *it;
      a.src = a.dst = 1;
      bwd.push_back(a);
    }
  }
  // twiddle pools: tw[k] = exp(sign 2 pi i k / len), computed in double
  for (int sg = 0; sg < 2; ++sg) {
    tw_host[sg].clear();
    const double sign = sg == 0 ? -1.0 : 1.0;
    for (int t : tables)
      for (int k = 0; k < t; ++k) {
        double c, si;
        if (2 * k == t) { c = -1.0; si = 0.0; }
        else if (4 * k == t) { c = 0.0; si = 1.0; }
        else if (4 * k == 3 * t) { c = 0.0; si = -1.0; }
        else { const double ang = 2.0 * M_PI * (double)k / (double)t; c = std::cos(ang); si = std::sin(ang); }
        tw_host[sg].push_back(cpx<R>((R)c, (R)(sign * si)));
      }
  }
  ok = true;
}

// box budget from the device's LDS limit: the largest power-of-two element
// count whose padded box fits
template <typename R>
static void box_budget(long& budget, size_t& lds_bytes) {
  int dev = 0, v = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
    v = 65536;
  lds_bytes = (size_t)v;
  long e = 1;
  while ((e * 2) * (long)sizeof(cpx<R>) * 33 / 32 <= (long)v) e *= 2;
  budget = e;
}

template <typename R>
FFTPlan<R>::FFTPlan(int nx, int ny, int nz, int m, int axes_mask) {
  dims_[0] = nx; dims_[1] = ny; dims_[2] = nz; dims_[3] = m;
  mask_ = axes_mask;
  S_ = (unsigned)((long)nx * ny * nz);
  const long total = (long)nx * ny * nz * m;
  if (total >= (1L << 31)) throw Error("FFT grid larger than 2^31 points");
  // the grid functors form sp + S t with 24-bit multiplies
  if ((long)nx * ny * nz >= (1L << 24) || m >= (1 << 24)) throw Error("FFT grid: k-mesh or slot count >= 2^24");
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
  // box passes (two or three HBM round trips per transform) unless disabled
  const char* eb = getenv("GMOD_FFT_BOX");
  if (!eb || atoi(eb) != 0) {
    long budget = 0;
    size_t lds = 0;
    box_budget<R>(budget, lds);
    box_.build(nx, ny, nz, m, axes_mask, budget, lds);
    if (box_.ok) {
      int dev = 0, ncu = 0;
      if (hipGetDevice(&dev) == hipSuccess &&
          hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && ncu >= 8)
        persist_ = (long)(ncu / 8) * 8;
      for (int sg = 0; sg < 2; ++sg) {
        const std::vector<BoxArgs>& v = sg == 0 ? box_.fwd : box_.bwd;
        GMOD_HIP(hipMalloc(&argsd_[sg], std::max<size_t>(v.size(), 1) * sizeof(BoxArgs)));
        if (!v.empty()) GMOD_HIP(hipMemcpy(argsd_[sg], v.data(), v.size() * sizeof(BoxArgs), hipMemcpyHostToDevice));
      }
      const size_t ntab = box_.tab_host.size();
      GMOD_HIP(hipMalloc(&tabs_, std::max<size_t>(ntab, 1) * sizeof(int)));
      if (ntab) GMOD_HIP(hipMemcpy(tabs_, box_.tab_host.data(), ntab * sizeof(int), hipMemcpyHostToDevice));
      for (int sg = 0; sg < 2; ++sg) {
        const size_t nt = box_.tw_host[sg].size();
        GMOD_HIP(hipMalloc(&twb_[sg], std::max<size_t>(nt, 1) * sizeof(cpx<R>)));
        if (nt)
          GMOD_HIP(hipMemcpy(twb_[sg], box_.tw_host[sg].data(), nt * sizeof(cpx<R>), hipMemcpyHostToDevice));
      }
    }
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
  for (auto p : twb_)
    if (p) (void)hipFree(p);
  if (tabs_) (void)hipFree(tabs_);
  for (auto p : argsd_)
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

template <typename R>
std::vector<unsigned char> FFTPlan<R>::fwd_last_live(const std::vector<int>& kept_sp, double* frac) const {
  std::vector<unsigned char> live;
  if (frac) *frac = 1.0;
  if (!box_.ok || box_.fwd.empty() || kept_sp.size() != (size_t)S_) return live;
  const BoxArgs& a = box_.fwd.back();
  const long pts = (long)a.box[0] * a.box[1] * a.box[2];
  if (a.box[3] != dims_[3] || pts > 64) return live;
  const long nb = a.blocks();
  live.assign((size_t)nb, 0);
  long n = 0;
  for (long lb = 0; lb < nb; ++lb) {
    const BoxOrigin O = box_origin(a, (unsigned)lb);
    bool any = false;
    for (int z = 0; z < a.box[2] && !any; ++z)
      for (int y = 0; y < a.box[1] && !any; ++y)
        for (int x = 0; x < a.box[0] && !any; ++x) {
          const unsigned sp = box_sp(a, O, (unsigned)x, (unsigned)y, (unsigned)z);
          if (sp >= (unsigned)S_) throw Error("fwd_last_live: spatial index out of range");
          any = kept_sp[sp] >= 0;
        }
    live[(size_t)lb] = any ? 1 : 0;
    n += any ? 1 : 0;
  }
  if (frac) *frac = (double)n / (double)nb;
  return live;
}

template struct BoxPlan<double>;
template struct BoxPlan<float>;
template class FFTPlan<double>;
template class FFTPlan<float>;
template const FFTPlan<double>& fft_plan<double>(int, int, int, int, int);
template const FFTPlan<float>& fft_plan<float>(int, int, int, int, int);


// ---------------------------------------------------------------- host test hook
// Runs the box passes of a plan on the host (the same load / sub-pass / store
// code as the kernel, one emulated thread, with bounds checks), so the CPU
// test suite can check plans and indexing against a reference FFT without a
// device.  Not used by any computation of the library.
namespace {
template <typename R>
struct HostGridLoad {
  const cpx<R>* p;
  unsigned S;
  cpx<R> operator()(unsigned sp, unsigned t) const { return p[sp + (size_t)S * t]; }
};
template <typename R>
struct HostGridStore {
  cpx<R>* p;
  unsigned S;
  struct Pre {};
  Pre prep(unsigned, unsigned) const { return Pre{}; }
  void commit(unsigned sp, unsigned t, const Pre&, cpx<R> v) const { p[sp + (size_t)S * t] = v; }
  bool skip(unsigned) const { return false; }
};

template <typename R>
int box_host_run(double* data, int nx, int ny, int nz, int m, int mask, int sign, long budget, size_t lds_bytes,
                 int* npasses) {
  // sign 2: the backward transform with the forward pass order
  // (FFTPlan::exec_fused's spatial_first)
  const bool spatial_first = sign == 2;
  if (spatial_first) sign = 1;
  BoxPlan<R> plan;
  plan.build(nx, ny, nz, m, mask, budget, lds_bytes);
  if (!plan.ok) return 1;
  if (!data) {  // plan only
    if (npasses) *npasses = (int)(sign < 0 ? plan.fwd.size() : plan.bwd.size());
    return 0;
  }
  const long N = (long)nx * ny * nz * m;
  const unsigned S = (unsigned)((long)nx * ny * nz);
  std::vector<cpx<R>> g(N), gt(N);
  for (long i = 0; i < N; ++i) g[i] = cpx<R>((R)data[2 * i], (R)data[2 * i + 1]);
  // the exec order of FFTPlan::exec_fused: pass i reads the input (i == 0)
  // or buffer src, writes buffer dst or the output (last pass); the output is
  // the caller's buffer here
  std::vector<cpx<R>>* bufs[2] = {&g, &gt};
  const std::vector<BoxArgs>& ps = (sign < 0 || spatial_first) ? plan.fwd : plan.bwd;
  const cpx<R>* tw = plan.tw_host[sign < 0 ? 0 : 1].data();
  long tpts = 1;
  const int n[4] = {nx, ny, nz, m};
  for (int d = 0; d < 4; ++d)
    if ((mask >> d) & 1) tpts *= n[d];
  for (size_t i = 0; i < ps.size(); ++i) {
    const BoxArgs& a = ps[i];
    const R scale = (sign > 0 && i + 1 == ps.size()) ? (R)(1.0 / (double)tpts) : (R)1;
    std::vector<cpx<R>> lds(a.lds_elems + a.ntwl);
    const long nb = a.blocks();
    std::vector<char> seen(nb, 0);
    const bool first = i == 0, last = i + 1 == ps.size();
    cpx<R>* in = first ? g.data() : bufs[a.src]->data();
    cpx<R>* out = last ? g.data() : bufs[a.dst]->data();
    if (!first && !last && in != out && a.src == a.dst) throw Error("inconsistent pass buffers");
    // a pass that moves data between slots must not read the buffer it writes
    if (a.ts_mul != a.tl_mul || a.ts_str != a.tl_str) {
      if (in == out) throw Error("slot-moving pass in place");
    }
    std::vector<cpx<R>> snapshot;
    for (long blk = 0; blk < nb; ++blk) {
      const unsigned lb = box_remap((unsigned)blk, (unsigned)nb, a.xcd_group);
      if (lb >= (unsigned)nb || seen[lb]) throw Error("box remap is not a permutation");
      seen[lb] = 1;
      const BoxOrigin O = box_origin(a, lb);
      box_twiddles_to_lds<R>(lds.data(), a, tw, 0, 1);
      box_load<R>(lds.data(), HostGridLoad<R>{in, S}, a, O, 0, 1, plan.tab_host.data());
      for (int j = 0; j < a.nsub; ++j) box_dispatch<R>(lds.data(), a, a.sub[j], tw, 0, 1, sign);
      box_store<R>(lds.data(), HostGridStore<R>{out, S}, a, O, scale, tw, 0, 1, plan.tab_host.data());
    }
  }
  for (long i = 0; i < N; ++i) { data[2 * i] = (double)g[i].re; data[2 * i + 1] = (double)g[i].im; }
  if (npasses) *npasses = (int)ps.size();
  return 0;
}
}  // namespace

}  // namespace gmod

// Host emulation of the box-pass FFT of an (nx,ny,nz,m) grid (x fastest),
// interleaved complex in place; precision 0: FP64, 1: FP32 arithmetic.
// lds_bytes sets the box budget as the device's LDS limit would.  Returns 0,
// 1 when the box engine does not cover the shape, 2 on an indexing error.
extern "C" int gmod_box_fft_host(double* data, int nx, int ny, int nz, int m, int mask, int sign, int precision,
                                 long lds_bytes, int* npasses) {
  using namespace gmod;
  try {
    if (precision == 1) {
      long e = 1;
      while ((e * 2) * 8 * 33 / 32 <= lds_bytes) e *= 2;
      return box_host_run<float>(data, nx, ny, nz, m, mask, sign, e, (size_t)lds_bytes, npasses);
    }
    long e = 1;
    while ((e * 2) * 16 * 33 / 32 <= lds_bytes) e *= 2;
    return box_host_run<double>(data, nx, ny, nz, m, mask, sign, e, (size_t)lds_bytes, npasses);
  } catch (std::exception& ex) {
    set_error(ex.what());
    return 2;
  }
}

// Device transform of an (nx,ny,nz,m) grid (x fastest, interleaved complex)
// held by the caller in device memory: `reps` transforms in place on buf with
// tmp (same size) as the intermediate buffer, on the library's stream, then a
// stream synchronisation.  sign -1 unnormalised forward, +1 backward divided
// by the number of transformed points; precision 0 FP64, 1 FP32.  Returns 0,
// or 1 on an error (message through gmod_last_error).
extern "C" int gmod_fft_device(void* buf, void* tmp, int nx, int ny, int nz, int m, int mask, int sign,
                               int precision, int reps) {
  using namespace gmod;
  try {
    if (!buf || !tmp) throw Error("fft: null buffer");
    std::lock_guard<std::mutex> g(ctx().lock);
    hipStream_t s = ctx().stream;
    for (int r = 0; r < std::max(reps, 1); ++r) {
      if (precision == 1) {
        const FFTPlan<float>& p = fft_plan<float>(nx, ny, nz, m, mask);
        const float sc = sign > 0 ? (float)(1.0 / (double)p.transformed_points()) : 1.0f;
        p.exec(static_cast<cpx<float>*>(buf), static_cast<cpx<float>*>(tmp), sign, sc, s);
      } else {
        const FFTPlan<double>& p = fft_plan<double>(nx, ny, nz, m, mask);
        const double sc = sign > 0 ? 1.0 / (double)p.transformed_points() : 1.0;
        p.exec(static_cast<cpx<double>*>(buf), static_cast<cpx<double>*>(tmp), sign, sc, s);
      }
    }
    GMOD_HIP(hipStreamSynchronize(s));
    return 0;
  } catch (std::exception& ex) {
    set_error(ex.what());
    return 1;
  }
}
