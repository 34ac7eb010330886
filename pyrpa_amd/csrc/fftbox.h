// Box passes of the batched 4D FFT: one work-group loads a whole 4D box of
// the grid into LDS, transforms every dimension of the box that spans the
// grid (a sequence of in-place mixed-radix sub-passes with the butterflies in
// registers), and stores the box.  A 4D transform is two box passes when the
// spatial volume fits in LDS (a box of all Nx*Ny*Nz points x bt slots, then a
// box of c adjacent spatial columns x all M slots), three when it does not
// (x-y planes, z columns, t columns).  Every element crosses HBM once per
// pass.  The per-axis passes of fft.h remain for shapes this engine does not
// cover.
//
// In-place mixed-radix recursion along one dimension of length n = r1 r2 ...:
// sub-pass i works on blocks of length L_i = n / (r1 ... r_{i-1}) with stride
// s_i = L_i / r_i: a radix-r_i DFT over the elements j' + s_i j of a block,
// then the twiddle w_{L_i}^(j' k), stored back in place at j' + s_i k.  The
// output index k = k1 + r1 (k2 + r2 (k3 + ...)) then sits at position
// sum_i k_i n / (r1 ... r_i), which only the store undoes.
//
// The three phases (load, sub-passes, store) are __host__ __device__ with an
// explicit (thread, thread count) so that the host test hook runs the same
// code, with bounds checks, on the CPU.
#pragma once
#include <type_traits>
#include <vector>

#include "gmod_common.h"

namespace gmod {

constexpr int BOX_THREADS = 512;
constexpr int BOX_MAX_SUB = 10;
constexpr int BOX_MAX_RAD = 4;  // sub-passes per dimension
constexpr int BOX_MAX_TILES = 32;  // tiles of BOX_THREADS elements per box (tiled load/store)
// tiles a persistent work-group holds in registers while prefetching (one element per tile and thread)
template <typename R>
constexpr int BOX_TILE_REGS = sizeof(R) == 4 ? 32 : 16;

// Division by a launch-invariant divisor as a multiply-high: with
// m = ceil(2^32 / d), n / d = (n m) >> 32 exactly whenever n (d - 1) < 2^32
// (the rounding error n (m d - 2^32) / 2^32 < 1 / d then never crosses an
// integer).  The planner checks that bound for every divisor it sets.
struct FDiv {
  unsigned d, m;
  unsigned one;  // all ones for d = 1 (m = 0), else 0: the quotient is then n
};

inline FDiv make_fdiv(unsigned d) {
  FDiv f;
  f.d = d;
  f.m = d <= 1 ? 0u : (unsigned)((0xFFFFFFFFull) / d + 1ull);
  f.one = d <= 1 ? 0xFFFFFFFFu : 0u;
  return f;
}

// branch-free: one multiply-high and one and-or
__host__ __device__ __forceinline__ unsigned fdiv(unsigned n, FDiv f) {
  return (unsigned)(((unsigned long long)n * f.m) >> 32) | (n & f.one);
}

// product of two values below 2^24 (full-rate multiply on the device)
__host__ __device__ __forceinline__ unsigned mul24(unsigned a, unsigned b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umul24(a, b);
#else
  return a * b;
#endif
}

struct BoxSub {
  int dim;      // 0:x 1:y 2:z 3:t
  int radix;
  int L;        // block length; stride s = L / radix
  int tw_off;   // w_L^m table (sign-specific pool)
  int wr_off;   // w_radix^m table
  int n;        // box extent along dim
  int ld;       // LDS stride along dim
  int twl;      // LDS element offset of the dimension's w_n table (-1: use tw_off in global memory)
  int tws;      // n / L: index step in that table
  // line decomposition: the three other dimensions of the box, x first
  int os[3];    // their LDS strides
  FDiv foe[2];  // extents of the first two of them
  FDiv flines;  // elems / box[dim] lines
  FDiv fs;      // stride L / radix
};

struct BoxArgs {
  int box[4];      // box extent per dimension
  int nblk[4];     // boxes per dimension
  int lds_stride[4];
  int lds_elems;   // LDS elements (with the x-row padding)
  int elems;       // grid elements per box
  int nsub;
  BoxSub sub[BOX_MAX_SUB];
  int nrad[4];                  // radices of a transformed dimension (0: not transformed)
  int rad[4][BOX_MAX_RAD];
  int radw[4][BOX_MAX_RAD];     // position weight n / (r1 ... r_i)
  unsigned Nx, Ny, S, M;
  FDiv fbox[3];                 // box[0..2]
  FDiv frad[4][BOX_MAX_RAD];    // rad
  int xcd_group;                // consecutive boxes that share cache lines (1: no remap)
  int src, dst;                 // intermediate buffers (0: caller buffer, 1: scratch)
  // slot of box element i along t for the box with t index bt:
  //   load t = bt tl_mul + i tl_str, store t = bt ts_mul + i ts_str
  // (bt box[3] + i for an ordinary box; strided sets for the two passes of
  // a four-step split of the t axis)
  int tl_mul, tl_str, ts_mul, ts_str;
  int tw4_n, tw4_off;           // four-step twiddle w_{tw4_n}^(bt k) on output k along t (0: none)
  // Tiled load/store (tiled = 1): the box is ntiles tiles of BOX_THREADS
  // elements with power-of-two extents 2^tlog[d] (x first), aligned with the
  // radix digits so that the store position of (tile origin + c) is the
  // position of the origin plus that of c.  A thread keeps its coordinates c
  // inside a tile; the tables give each tile's offsets from the box origin:
  // spatial index, slot, natural LDS position (load), digit-reversed LDS
  // position (store).
  // The tables live in the plan's int array (not in this kernel argument, whose
  // copy to scratch memory a run-time index would force): at tab_off, four
  // rows of ntiles entries: spatial index, slot, load and store LDS offsets.
  int tiled, ntiles;
  int tlog[4];
  int tab_off;
  int pf_at;  // persistent tiled pass: the next box is fetched before sub-pass pf_at (nsub: after the last)
  // twiddle tables copied to LDS after the box (lds_elems .. lds_elems + ntwl)
  int ntwl, ntwc;
  int twc_src[4], twc_dst[4], twc_len[4];  // copies: pool offset, LDS offset, length
  __host__ __device__ long blocks() const { return (long)nblk[0] * nblk[1] * nblk[2] * nblk[3]; }
};

#if defined(__HIP_DEVICE_COMPILE__)
#define GMOD_BOX_CHECK(cond, what) ((void)0)
#else
#define GMOD_BOX_CHECK(cond, what) \
  do {                             \
    if (!(cond)) throw ::gmod::Error(std::string("box FFT index out of range: ") + (what)); \
  } while (0)
#endif

__host__ __device__ inline unsigned box_remap(unsigned a, unsigned nb, int g) {
  // consecutive logical boxes 0..g-1 of a group land on one XCD (the
  // dispatcher deals work-groups round-robin over the 8 XCDs)
  if (g <= 1 || (nb % (8u * (unsigned)g)) != 0) return a;
  const unsigned xcd = a & 7u, slot = a >> 3;
  return (unsigned)g * (8u * (slot / (unsigned)g) + xcd) + slot % (unsigned)g;
}

template <int N>
struct ConstLog2 { static constexpr int v = 1 + ConstLog2<N / 2>::v; };
template <>
struct ConstLog2<1> { static constexpr int v = 0; };

template <int N>
__host__ __device__ constexpr int cbrev(int i) {
  int r = 0;
  for (int b = 0; b < ConstLog2<N>::v; ++b)
    if (i & (1 << b)) r |= 1 << (ConstLog2<N>::v - 1 - b);
  return r;
}

// cos(2 pi m / 32) for m = 0..8; the rest of the period by symmetry
__host__ __device__ constexpr double box_cos32_q(int m) {
  return m == 0 ? 1.0
       : m == 1 ? 0.98078528040323044912618223613424
       : m == 2 ? 0.92387953251128675612818318939679
       : m == 3 ? 0.83146961230254523707878837761791
       : m == 4 ? 0.70710678118654752440084436210485
       : m == 5 ? 0.55557023301960222474283081394853
       : m == 6 ? 0.38268343236508977172845998403040
       : m == 7 ? 0.19509032201612826784828486847702
                : 0.0;
}
__host__ __device__ constexpr double box_cos32(int m) {
  m &= 31;
  if (m > 16) m = 32 - m;                                  // cos(2 pi - x) = cos x
  return m <= 8 ? box_cos32_q(m) : -box_cos32_q(16 - m);   // cos(pi - x) = -cos x
}
__host__ __device__ constexpr double box_sin32(int m) { return box_cos32(8 - m); }

// b w^m for w = exp(S 2 pi i / 32): the twiddle of a power-of-two DFT of
// length <= 32 with compile-time values (m is a constant once the caller's
// loops are unrolled), so a factor 1, -1 or +-i costs no arithmetic and a
// factor (+-1 +- i)/sqrt(2) two multiplies
template <typename R, int S>
__host__ __device__ __forceinline__ cpx<R> box_twiddle32(cpx<R> b, int m) {
  m &= 31;
  const R c = (R)box_cos32(m), s = (R)(S * box_sin32(m));
  if (m == 0) return b;
  if (m == 16) return cpx<R>(-b.re, -b.im);
  if (m == 8 || m == 24) return cpx<R>(-s * b.im, s * b.re);                   // s = +-1
  if ((m & 7) == 4) return cpx<R>(c * (b.re - (s / c) * b.im), c * (b.im + (s / c) * b.re));  // s/c = +-1
  return cpx<R>(c * b.re - s * b.im, c * b.im + s * b.re);
}

// In-register DFT of length RAD, natural order in and out, exponent sign S
// (-1 forward, +1 backward).  Powers of two up to 32: radix-2 steps with
// compile-time twiddles; other radices: wr[m] = w_RAD^m from the plan's table.
template <typename R, int RAD, int S>
__host__ __device__ __forceinline__ void box_dft(cpx<R> (&v)[RAD], const cpx<R>* __restrict__ wr) {
  if constexpr ((RAD & (RAD - 1)) == 0) {
    static_assert(RAD <= 32, "compile-time twiddles cover lengths up to 32");
    cpx<R> u[RAD];
#pragma unroll
    for (int i = 0; i < RAD; ++i) u[cbrev<RAD>(i)] = v[i];
#pragma unroll
    for (int len = 2; len <= RAD; len <<= 1) {
#pragma unroll
      for (int i = 0; i < RAD; i += len) {
#pragma unroll
        for (int j = 0; j < len / 2; ++j) {
          const cpx<R> a = u[i + j];
          const cpx<R> b = box_twiddle32<R, S>(u[i + j + len / 2], j * (32 / len));
          u[i + j] = a + b;
          u[i + j + len / 2] = a - b;
        }
      }
    }
#pragma unroll
    for (int i = 0; i < RAD; ++i) v[i] = u[i];
  } else {
    cpx<R> u[RAD];
#pragma unroll
    for (int k = 0; k < RAD; ++k) {
      cpx<R> acc = v[0];
#pragma unroll
      for (int j = 1; j < RAD; ++j) acc += v[j] * wr[(j * k) % RAD];
      u[k] = acc;
    }
#pragma unroll
    for (int i = 0; i < RAD; ++i) v[i] = u[i];
  }
}

struct BoxOrigin {
  int o[3];  // spatial origin
  int bt;    // box index along t
};

__host__ __device__ inline BoxOrigin box_origin(const BoxArgs& a, unsigned b) {
  BoxOrigin O;
  unsigned r = b;
  for (int d = 0; d < 3; ++d) {
    O.o[d] = (int)(r % (unsigned)a.nblk[d]) * a.box[d];
    r /= (unsigned)a.nblk[d];
  }
  O.bt = (int)r;
  return O;
}

// coordinates of thread tid inside a tile (x fastest)
__host__ __device__ __forceinline__ void box_tile_coord(const BoxArgs& a, unsigned tid, unsigned (&c)[4]) {
  c[0] = tid & ((1u << a.tlog[0]) - 1u);
  const unsigned r0 = tid >> a.tlog[0];
  c[1] = r0 & ((1u << a.tlog[1]) - 1u);
  const unsigned r1 = r0 >> a.tlog[1];
  c[2] = r1 & ((1u << a.tlog[2]) - 1u);
  c[3] = r1 >> a.tlog[2];
}

// spatial index of the box origin + (c0, c1, c2)
__host__ __device__ __forceinline__ unsigned box_sp(const BoxArgs& a, const BoxOrigin& O, unsigned c0, unsigned c1,
                                                    unsigned c2) {
  return (unsigned)O.o[0] + c0 + mul24(a.Nx, (unsigned)O.o[1] + c1) + mul24(a.Nx * a.Ny, (unsigned)O.o[2] + c2);
}

template <typename R, class LD>
__host__ __device__ __forceinline__ void box_load_tiled(cpx<R>* lds, const LD& ld, const BoxArgs& a,
                                                        const BoxOrigin& O, unsigned tid,
                                                        const int* __restrict__ tabs) {
  const int* tsp = tabs + a.tab_off;
  const int* tt = tsp + a.ntiles;
  const int* tld = tt + a.ntiles;
  constexpr int U = sizeof(R) == 4 ? 16 : 8;
  unsigned c[4];
  box_tile_coord(a, tid, c);
  const unsigned sp0 = box_sp(a, O, c[0], c[1], c[2]);
  const unsigned t0 = mul24((unsigned)O.bt, (unsigned)a.tl_mul) + mul24(c[3], (unsigned)a.tl_str);
  const unsigned l0 = c[0] + mul24(c[1], (unsigned)a.lds_stride[1]) + mul24(c[2], (unsigned)a.lds_stride[2]) +
                      mul24(c[3], (unsigned)a.lds_stride[3]);
  for (int j0 = 0; j0 < a.ntiles; j0 += U) {
    cpx<R> v[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (j0 + u < a.ntiles) {
        const unsigned sp = sp0 + (unsigned)tsp[j0 + u], t = t0 + (unsigned)tt[j0 + u];
        GMOD_BOX_CHECK(sp < a.S && t < a.M, "tiled load grid");
        v[u] = ld(sp, t);
      }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (j0 + u < a.ntiles) {
        const unsigned off = l0 + (unsigned)tld[j0 + u];
        GMOD_BOX_CHECK(off < (unsigned)a.lds_elems, "tiled load lds");
        lds[off] = v[u];
      }
  }
}

// ---- phase 1: grid -> LDS (x fastest: coalesced along the grid's contiguous axis)
template <typename R, class LD>
__host__ __device__ __forceinline__ void box_load(cpx<R>* lds, const LD& ld, const BoxArgs& a, const BoxOrigin& O,
                                                  int tid, int nthr, const int* __restrict__ tabs) {
  if (a.tiled) {  // one pass on the device (nthr = BOX_THREADS); the host emulation loops over the threads
    for (int t = tid; t < BOX_THREADS; t += nthr) box_load_tiled<R>(lds, ld, a, O, (unsigned)t, tabs);
    return;
  }
  // U loads in flight per thread before the first LDS store: one work-group
  // per CU has to keep ~64 KB of HBM reads outstanding
  constexpr int U = sizeof(R) == 4 ? 16 : 8;
  const int ex = a.box[0], ey = a.box[1], ez = a.box[2];
  for (int e0 = tid; e0 < a.elems; e0 += nthr * U) {
    cpx<R> v[U];
    int off[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int e = e0 + u * nthr;
      off[u] = -1;
      if (e < a.elems) {
        const unsigned q1 = fdiv((unsigned)e, a.fbox[0]), q2 = fdiv(q1, a.fbox[1]);
        const int it = (int)fdiv(q2, a.fbox[2]);
        const int ix = e - (int)mul24(q1, ex);
        const int iy = (int)(q1 - mul24(q2, ey));
        const int iz = (int)(q2 - mul24(it, ez));
        const unsigned x = (unsigned)(O.o[0] + ix), y = (unsigned)(O.o[1] + iy), z = (unsigned)(O.o[2] + iz);
        const unsigned t = mul24((unsigned)O.bt, (unsigned)a.tl_mul) + mul24((unsigned)it, (unsigned)a.tl_str);
        const unsigned sp = x + mul24(a.Nx, y) + mul24(a.Nx * a.Ny, z);
        GMOD_BOX_CHECK(sp < a.S && t < a.M, "load grid");
        v[u] = ld(sp, t);
        off[u] = ix + (int)(mul24((unsigned)iy, (unsigned)a.lds_stride[1]) + mul24((unsigned)iz, (unsigned)a.lds_stride[2]) +
                            mul24((unsigned)it, (unsigned)a.lds_stride[3]));
        GMOD_BOX_CHECK(off[u] < a.lds_elems, "load lds");
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (off[u] >= 0) lds[off[u]] = v[u];
  }
}

// ---- phase 2: one radix-RAD sub-pass over the whole box (in place in LDS)
template <typename R, int RAD, int S>
__host__ __device__ __forceinline__ void box_subpass(cpx<R>* lds, const BoxArgs& a, const BoxSub& sp,
                                                     const cpx<R>* __restrict__ tw, int tid, int nthr) {
  const int n = sp.n;
  const int L = sp.L, s = L / RAD;
  const int lines = (int)sp.flines.d;
  const int groups = a.elems / RAD;
  const int ld = sp.ld;
  const cpx<R>* wr = tw + sp.wr_off;
  const cpx<R>* wl = tw + sp.tw_off;
  // LDS distance of consecutive elements of a group (below 2^24: one 24-bit multiply)
  const unsigned str = mul24((unsigned)s, (unsigned)ld);
  for (int g = tid; g < groups; g += nthr) {
    const unsigned gl = fdiv((unsigned)g, sp.flines);
    const unsigned line = (unsigned)g - mul24(gl, (unsigned)lines);
    const unsigned r1 = fdiv(line, sp.foe[0]), c2 = fdiv(r1, sp.foe[1]);
    const unsigned c0 = line - mul24(r1, sp.foe[0].d), c1 = r1 - mul24(c2, sp.foe[1].d);
    const unsigned base = mul24(c0, (unsigned)sp.os[0]) + mul24(c1, (unsigned)sp.os[1]) + mul24(c2, (unsigned)sp.os[2]);
    const unsigned blk = fdiv(gl, sp.fs);
    const int jp = (int)(gl - mul24(blk, (unsigned)s));
    const int p0 = (int)mul24(blk, (unsigned)L) + jp;
    GMOD_BOX_CHECK(line < (unsigned)lines && p0 + s * (RAD - 1) < n, "sub-pass position");
    GMOD_BOX_CHECK((int)base + (p0 + s * (RAD - 1)) * ld < a.lds_elems, "sub-pass lds");
    // element j of the group at a0 + j str (additions only)
    const unsigned a0 = base + mul24((unsigned)p0, (unsigned)ld);
    cpx<R> v[RAD];
    {
      unsigned ad = a0;
#pragma unroll
      for (int j = 0; j < RAD; ++j, ad += str) v[j] = lds[ad];
    }
    box_dft<R, RAD, S>(v, wr);
    if (s > 1) {
      // w_L^(jp k) = w_n^(jp k n / L): the dimension's table, in LDS when
      // the plan placed it there (sp.twl >= 0), else in global memory
      const unsigned step = mul24((unsigned)jp, (unsigned)sp.tws);
      GMOD_BOX_CHECK(jp * (RAD - 1) < L, "sub-pass twiddle");
      if (sp.twl >= 0) {
        const cpx<R>* wt = lds + sp.twl;
        unsigned ti = step;
#pragma unroll
        for (int k = 1; k < RAD; ++k, ti += step) v[k] = v[k] * wt[ti];
      } else {
        unsigned ti = (unsigned)jp;
#pragma unroll
        for (int k = 1; k < RAD; ++k, ti += (unsigned)jp) v[k] = v[k] * wl[ti];
      }
    }
    {
      unsigned ad = a0;
#pragma unroll
      for (int k = 0; k < RAD; ++k, ad += str) lds[ad] = v[k];
    }
  }
}

// a.sub[i] read with constant indices only: an index into the by-value
// kernel argument that is not a constant makes the compiler copy the whole
// argument to scratch memory
__host__ __device__ __forceinline__ BoxSub box_sub_at(const BoxArgs& a, int i) {
  static_assert(BOX_MAX_SUB == 10, "one case per sub-pass slot");
  switch (i) {
    case 0: return a.sub[0];
    case 1: return a.sub[1];
    case 2: return a.sub[2];
    case 3: return a.sub[3];
    case 4: return a.sub[4];
    case 5: return a.sub[5];
    case 6: return a.sub[6];
    case 7: return a.sub[7];
    case 8: return a.sub[8];
    default: return a.sub[9];
  }
}

// MAXR: the largest radix compiled in (a kernel that never runs radix 32
// then does not reserve its registers)
template <typename R, int MAXR, int S>
__host__ __device__ __forceinline__ void box_dispatch_pow2(cpx<R>* lds, const BoxArgs& a, const BoxSub& sp,
                                                           const cpx<R>* __restrict__ tw, int tid, int nthr) {
  switch (sp.radix) {
    case 2: box_subpass<R, 2, S>(lds, a, sp, tw, tid, nthr); break;
    case 4: box_subpass<R, 4, S>(lds, a, sp, tw, tid, nthr); break;
    case 8: box_subpass<R, 8, S>(lds, a, sp, tw, tid, nthr); break;
    case 16:
      if constexpr (MAXR >= 16) box_subpass<R, 16, S>(lds, a, sp, tw, tid, nthr);
      break;
    case 32:
      if constexpr (sizeof(R) == 4 && MAXR >= 32) box_subpass<R, 32, S>(lds, a, sp, tw, tid, nthr);
      break;
    default: break;
  }
}

// sign: the exponent sign of the transform (-1 forward, +1 backward); the
// power-of-two radices take it at compile time, the odd ones from the
// sign's twiddle table tw
template <typename R, int MAXR = 32>
__host__ __device__ __forceinline__ void box_dispatch(cpx<R>* lds, const BoxArgs& a, const BoxSub& sp,
                                                      const cpx<R>* __restrict__ tw, int tid, int nthr, int sign) {
  switch (sp.radix) {
    case 3: box_subpass<R, 3, 0>(lds, a, sp, tw, tid, nthr); break;
    case 5: box_subpass<R, 5, 0>(lds, a, sp, tw, tid, nthr); break;
    case 7: box_subpass<R, 7, 0>(lds, a, sp, tw, tid, nthr); break;
    default:
      if (sign < 0) box_dispatch_pow2<R, MAXR, -1>(lds, a, sp, tw, tid, nthr);
      else box_dispatch_pow2<R, MAXR, 1>(lds, a, sp, tw, tid, nthr);
      break;
  }
}

// largest radix of a pass's sub-passes
inline int box_max_radix(const BoxArgs& a) {
  int r = 1;
  for (int i = 0; i < a.nsub; ++i) r = a.sub[i].radix > r ? a.sub[i].radix : r;
  return r;
}

// position of output index k of a transformed dimension (undoes the in-place
// digit order of the sub-passes)
__host__ __device__ __forceinline__ int box_outpos(const BoxArgs& a, int d, int k) {
  const int nr = a.nrad[d];
  if (nr <= 1) return k;
  unsigned p = 0, q = (unsigned)k;
#pragma unroll
  for (int i = 0; i < BOX_MAX_RAD; ++i) {
    if (i < nr) {
      const unsigned qn = fdiv(q, a.frad[d][i]);
      p += mul24(q - mul24(qn, (unsigned)a.rad[d][i]), (unsigned)a.radw[d][i]);
      q = qn;
    }
  }
  return (int)p;
}

template <typename R, class ST>
__host__ __device__ __forceinline__ void box_store_tiled(const cpx<R>* lds, const ST& st, const BoxArgs& a,
                                                         const BoxOrigin& O, R scale, unsigned tid,
                                                         const int* __restrict__ tabs) {
  constexpr int U = 8;
  const int* tsp = tabs + a.tab_off;
  const int* tt = tsp + a.ntiles;
  const int* tst = tt + 2 * a.ntiles;
  unsigned c[4];
  box_tile_coord(a, tid, c);
  const unsigned sp0 = box_sp(a, O, c[0], c[1], c[2]);
  const unsigned t0 = mul24((unsigned)O.bt, (unsigned)a.ts_mul) + mul24(c[3], (unsigned)a.ts_str);
  const unsigned l0 = (unsigned)box_outpos(a, 0, (int)c[0]) +
                      mul24((unsigned)box_outpos(a, 1, (int)c[1]), (unsigned)a.lds_stride[1]) +
                      mul24((unsigned)box_outpos(a, 2, (int)c[2]), (unsigned)a.lds_stride[2]) +
                      mul24((unsigned)box_outpos(a, 3, (int)c[3]), (unsigned)a.lds_stride[3]);
  // the preparation (every load of the store functor) carries no branch: an
  // element past the last tile repeats the last tile's element, which its
  // batch commits twice with the same value; prep() is valid on slots that
  // skip() drops, and those are not committed
  const int last = a.ntiles - 1;
  for (int j0 = 0; j0 < a.ntiles; j0 += U) {
    typename ST::Pre pre[U];
    unsigned spv[U], tv[U], offv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int jj = j0 + u < last ? j0 + u : last;
      spv[u] = sp0 + (unsigned)tsp[jj];
      tv[u] = t0 + (unsigned)tt[jj];
      offv[u] = l0 + (unsigned)tst[jj];
      GMOD_BOX_CHECK(spv[u] < a.S && tv[u] < a.M, "tiled store grid");
      GMOD_BOX_CHECK(offv[u] < (unsigned)a.lds_elems, "tiled store lds");
      pre[u] = st.prep(spv[u], tv[u]);
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (!st.skip(tv[u])) st.commit(spv[u], tv[u], pre[u], lds[offv[u]] * scale);
  }
}

// ---- phase 3: LDS -> grid through the store functor
template <typename R, class ST>
__host__ __device__ __forceinline__ void box_store(const cpx<R>* lds, const ST& st, const BoxArgs& a,
                                                   const BoxOrigin& O, R scale, const cpx<R>* __restrict__ tw,
                                                   int tid, int nthr, const int* __restrict__ tabs) {
  if (a.tiled) {
    for (int t = tid; t < BOX_THREADS; t += nthr) box_store_tiled<R>(lds, st, a, O, scale, (unsigned)t, tabs);
    return;
  }
  // batches of U elements: every load a store functor needs (accumulators,
  // operands, gather indices) is issued before the batch's first store
  constexpr int U = 8;
  const int ex = a.box[0], ey = a.box[1], ez = a.box[2];
  for (int e0 = tid; e0 < a.elems; e0 += nthr * U) {
    unsigned spv[U], tv[U];
    int offv[U];
    typename ST::Pre pre[U];
    cpx<R> twv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int e = e0 + u * nthr;
      offv[u] = -1;
      if (e < a.elems) {
        const unsigned q1 = fdiv((unsigned)e, a.fbox[0]), q2 = fdiv(q1, a.fbox[1]);
        const int kt = (int)fdiv(q2, a.fbox[2]);
        const int kx = e - (int)mul24(q1, ex);
        const int ky = (int)(q1 - mul24(q2, ey));
        const int kz = (int)(q2 - mul24(kt, ez));
        const unsigned t = mul24((unsigned)O.bt, (unsigned)a.ts_mul) + mul24((unsigned)kt, (unsigned)a.ts_str);
        if (!st.skip(t)) {
          offv[u] = box_outpos(a, 0, kx) +
                    (int)(mul24((unsigned)box_outpos(a, 1, ky), (unsigned)a.lds_stride[1]) +
                          mul24((unsigned)box_outpos(a, 2, kz), (unsigned)a.lds_stride[2]) +
                          mul24((unsigned)box_outpos(a, 3, kt), (unsigned)a.lds_stride[3]));
          GMOD_BOX_CHECK(offv[u] < a.lds_elems, "store lds");
          const unsigned x = (unsigned)(O.o[0] + kx), y = (unsigned)(O.o[1] + ky), z = (unsigned)(O.o[2] + kz);
          spv[u] = x + mul24(a.Nx, y) + mul24(a.Nx * a.Ny, z);
          tv[u] = t;
          GMOD_BOX_CHECK(spv[u] < a.S && t < a.M, "store grid");
          pre[u] = st.prep(spv[u], t);
          if (a.tw4_n) {
            twv[u] = tw[a.tw4_off + (int)(((long)O.bt * kt) % a.tw4_n)];
            GMOD_BOX_CHECK(((long)O.bt * kt) % a.tw4_n < a.tw4_n, "four-step twiddle");
          }
        }
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (offv[u] >= 0) {
        cpx<R> v = lds[offv[u]] * scale;
        if (a.tw4_n) v = v * twv[u];
        st.commit(spv[u], tv[u], pre[u], v);
      }
  }
}

// the pass's twiddle tables into LDS (after the box)
template <typename R>
__host__ __device__ __forceinline__ void box_twiddles_to_lds(cpx<R>* lds, const BoxArgs& a,
                                                             const cpx<R>* __restrict__ tw, int tid, int nthr) {
  for (int c = 0; c < a.ntwc; ++c) {
    const int src = c == 0 ? a.twc_src[0] : c == 1 ? a.twc_src[1] : c == 2 ? a.twc_src[2] : a.twc_src[3];
    const int dst = c == 0 ? a.twc_dst[0] : c == 1 ? a.twc_dst[1] : c == 2 ? a.twc_dst[2] : a.twc_dst[3];
    const int len = c == 0 ? a.twc_len[0] : c == 1 ? a.twc_len[1] : c == 2 ? a.twc_len[2] : a.twc_len[3];
    for (int i = tid; i < len; i += nthr) {
      GMOD_BOX_CHECK(dst + i < a.lds_elems + a.ntwl, "twiddle copy");
      lds[dst + i] = tw[src + i];
    }
  }
}

// A store may also drop whole boxes by their index (skip_box(lb), lb the
// box index after box_remap): the gather of a pass whose boxes span the
// whole t axis keeps only the boxes that hold a kept spatial point.
template <class ST, class = void>
struct BoxSkip {
  __host__ __device__ static bool skip(const ST&, unsigned) { return false; }
};
template <class ST>
struct BoxSkip<ST, std::void_t<decltype(std::declval<const ST&>().skip_box(0u))>> {
  __device__ static bool skip(const ST& st, unsigned lb) { return st.skip_box(lb); }
};

// a box whose slots the store keeps none of (the gather keeps t < Nw; the
// store slots of a box increase with the element index), or that the store
// drops as a whole
template <class ST>
__host__ __device__ __forceinline__ bool box_skipped(const ST& st, const BoxArgs& a, unsigned lb,
                                                     const BoxOrigin& O) {
  const unsigned t_first = (unsigned)(O.bt * a.ts_mul), t_last = t_first + (unsigned)((a.box[3] - 1) * a.ts_str);
  return (st.skip(t_first) && st.skip(t_last)) || BoxSkip<ST>::skip(st, lb);
}

// Tiled passes run persistent work-groups (one per CU: a box fills the LDS)
// that loop over boxes b = blockIdx.x + k gridDim.x.  The next box is read
// into registers (TM elements per thread) once the last sub-pass that reads
// twiddles from global memory has run, so its HBM reads overlap the
// remaining sub-passes and the store of the current box.  gridDim.x is a
// multiple of 8, so a work-group keeps its XCD and box_remap's grouping holds.
template <class ST>
__device__ __forceinline__ unsigned box_next_live(const ST& st, const BoxArgs& a, unsigned b, unsigned nb) {
  while (b < nb) {
    const unsigned lb = box_remap(b, nb, a.xcd_group);
    if (!box_skipped(st, a, lb, box_origin(a, lb))) break;
    b += gridDim.x;
  }
  return b;
}

// Load functors that read an index before the value (an irreducible-k map,
// the -r map) expose key(sp) / value(key, sp, t); the fetch then issues the
// index loads of all its elements before any value load.  Others are called
// as ld(sp, t).
template <class LD, class = void>
struct LoadKey {
  struct Key {};
  __device__ static Key key(const LD&, unsigned) { return Key{}; }
  __device__ static auto value(const LD& ld, const Key&, unsigned sp, unsigned t) { return ld(sp, t); }
};
template <class LD>
struct LoadKey<LD, std::void_t<typename LD::Key>> {
  using Key = typename LD::Key;
  __device__ static Key key(const LD& ld, unsigned sp) { return ld.key(sp); }
  __device__ static auto value(const LD& ld, const Key& k, unsigned sp, unsigned t) { return ld.value(k, sp, t); }
};

// Load functors whose Raw type is the element type cpx<R> split a value into
// raw() (loads only) and finish() (arithmetic on the loaded value and the
// slot); the persistent pass keeps raw values in its prefetch registers and
// finishes them as it writes them to LDS.  Others: raw = value, no finish.
template <typename R, class LD, class = void>
struct LoadRaw {
  template <class K>
  __device__ static cpx<R> raw(const LD& ld, const K& k, unsigned sp, unsigned t) {
    return LoadKey<LD>::value(ld, k, sp, t);
  }
  __device__ static cpx<R> finish(const LD&, const cpx<R>& v, unsigned) { return v; }
};
template <typename R, class LD>
struct LoadRaw<R, LD, std::enable_if_t<std::is_same<typename LD::Raw, cpx<R>>::value>> {
  template <class K>
  __device__ static cpx<R> raw(const LD& ld, const K& k, unsigned sp, unsigned t) { return ld.raw(k, sp, t); }
  __device__ static cpx<R> finish(const LD& ld, const cpx<R>& v, unsigned t) { return ld.finish(v, t); }
};

template <typename R, int TM, class LD>
__device__ __forceinline__ void box_fetch(cpx<R> (&v)[TM], const LD& ld, const BoxArgs& a, unsigned b, unsigned nb,
                                          unsigned csp, unsigned ct, const int* __restrict__ tsp,
                                          const int* __restrict__ tt) {
  const BoxOrigin O = box_origin(a, box_remap(b, nb, a.xcd_group));
  const unsigned sp0 = box_sp(a, O, 0, 0, 0) + csp;
  const unsigned t0 = mul24((unsigned)O.bt, (unsigned)a.tl_mul) + ct;
  using LK = LoadKey<LD>;
  // no per-element branches: an element past the last tile re-reads the last
  // tile's (valid) element and is never stored, so all loads of the fetch
  // sit in one basic block and are issued back to back
  const int last = a.ntiles - 1;
  typename LK::Key kk[TM];
#pragma unroll
  for (int j = 0; j < TM; ++j) kk[j] = LK::key(ld, sp0 + (unsigned)tsp[j < last ? j : last]);
#pragma unroll
  for (int j = 0; j < TM; ++j) {
    const int jj = j < last ? j : last;
    v[j] = LoadRaw<R, LD>::raw(ld, kk[j], sp0 + (unsigned)tsp[jj], t0 + (unsigned)tt[jj]);
  }
}

template <typename R, class LD, class ST>
__device__ __forceinline__ void box_persistent(cpx<R>* lds, const LD& ld, const ST& st, const BoxArgs& a,
                                               const cpx<R>* __restrict__ tw, const int* __restrict__ tabs, R scale,
                                               int tid, int sign) {
  constexpr int TM = BOX_TILE_REGS<R>;
  const unsigned nb = (unsigned)a.blocks();
  unsigned c[4];
  box_tile_coord(a, (unsigned)tid, c);
  const unsigned csp = c[0] + mul24(a.Nx, c[1]) + mul24(a.Nx * a.Ny, c[2]);
  const unsigned ct = mul24(c[3], (unsigned)a.tl_str);
  const unsigned l0 = c[0] + mul24(c[1], (unsigned)a.lds_stride[1]) + mul24(c[2], (unsigned)a.lds_stride[2]) +
                      mul24(c[3], (unsigned)a.lds_stride[3]);
  const int* tsp = tabs + a.tab_off;
  const int* tt = tsp + a.ntiles;
  const int* tld = tt + a.ntiles;
  cpx<R> v[TM];
  box_twiddles_to_lds<R>(lds, a, tw, tid, BOX_THREADS);  // ordered before use by the loop's first barrier
  unsigned b = box_next_live(st, a, blockIdx.x, nb);
  if (b < nb) box_fetch<R, TM>(v, ld, a, b, nb, csp, ct, tsp, tt);
  while (b < nb) {
    {
      const BoxOrigin O = box_origin(a, box_remap(b, nb, a.xcd_group));
      const unsigned t0 = mul24((unsigned)O.bt, (unsigned)a.tl_mul) + ct;
#pragma unroll
      for (int j = 0; j < TM; ++j)
        if (j < a.ntiles) lds[l0 + (unsigned)tld[j]] = LoadRaw<R, LD>::finish(ld, v[j], t0 + (unsigned)tt[j]);
    }
    __syncthreads();
    const unsigned nxt = box_next_live(st, a, b + gridDim.x, nb);
    for (int i = 0; i < a.nsub; ++i) {
      if (i == a.pf_at && nxt < nb) box_fetch<R, TM>(v, ld, a, nxt, nb, csp, ct, tsp, tt);
      box_dispatch<R>(lds, a, box_sub_at(a, i), tw, tid, BOX_THREADS, sign);
      __syncthreads();
    }
    if (a.pf_at >= a.nsub && nxt < nb) box_fetch<R, TM>(v, ld, a, nxt, nb, csp, ct, tsp, tt);
    box_store_tiled<R>(lds, st, a, box_origin(a, box_remap(b, nb, a.xcd_group)), scale, (unsigned)tid, tabs);
    __syncthreads();  // the LDS box is rewritten by the next iteration
    b = nxt;
  }
}

template <typename R>
struct GridLoad;  // fft.h

// The spatial reversal r -> -r (each coordinate x -> (n - x) mod n) of a
// spatial index on an nx x ny x nz grid, x fastest; the multiply-high
// divisions are exact while S nx < 2^32 and ny nz ny < 2^32 (make_sprev)
struct SpRev {
  FDiv fx, fy;
  unsigned nx, ny, nz;
  __host__ __device__ __forceinline__ unsigned operator()(unsigned sp) const {
    const unsigned q = fdiv(sp, fx), x = sp - q * nx;
    const unsigned z = fdiv(q, fy), y = q - z * ny;
    return (x ? nx - x : 0u) + nx * ((y ? ny - y : 0u) + ny * (z ? nz - z : 0u));
  }
};

inline bool sprev_ok(int nx, int ny, int nz) {
  const unsigned long long S = (unsigned long long)nx * ny * nz;
  return S * (unsigned long long)nx < (1ull << 32) && (unsigned long long)ny * nz * ny < (1ull << 32);
}

inline SpRev make_sprev(int nx, int ny, int nz) {
  return SpRev{make_fdiv((unsigned)nx), make_fdiv((unsigned)ny), (unsigned)nx, (unsigned)ny, (unsigned)nz};
}

// A source of the accumulating pass: the earlier passes' output stored at p,
// read as stored, or (cr) conjugated at the reversed point (-r, -t) plus,
// at t = 0, dh[sp] (the transformed slot-0 mismatch of the two grids): the
// grid whose transform is conj X(-r, -t) + delta_t0 dh
template <typename R>
struct AccumSrcLoad {
  const cpx<R>* p;
  const cpx<R>* dh;
  SpRev rv;
  unsigned S, M;
  int cr;
  using Key = unsigned;  // the spatial index read
  __host__ __device__ Key key(unsigned sp) const { return cr ? rv(sp) : sp; }
  __host__ __device__ cpx<R> value(Key k, unsigned sp, unsigned t) const {
    const unsigned tt = (cr && t != 0) ? M - t : t;
    const cpx<R> v = p[k + mul24(S, tt)];
    if (!cr) return v;
    const cpx<R> c(v.re, -v.im);
    return t == 0 ? c + dh[sp] : c;
  }
  __host__ __device__ cpx<R> operator()(unsigned sp, unsigned t) const { return value(key(sp), sp, t); }
};

// Accumulating last pass: nsrc transforms that share this last pass (their
// earlier passes left them in src + c src_stride, c < nsrc) are summed, each
// with its own operands, into one grid.  A persistent work-group transforms
// source c = 0 .. nsrc-1 of its box in turn (the next source, or the next
// box, prefetched into registers as in box_persistent), adds each source's
// contribution to a sum kept in registers, and writes the sum once: the sum
// crosses HBM once instead of twice per source.  The accumulator functor AC
// provides prep(c, sp, t) (its operand loads for source c), contrib(c, pre,
// v) (the contribution of transformed value v), side(c, sp, t, pre, v) (its
// writes besides the sum) and finish(sp, t, sum).  MAXR caps the radices compiled in.
// srcmap (optional): source c is buffer srcmap[c] >> 1 of src, conjugated at
// (-r, -t) when srcmap[c] & 1 (rv: the spatial reversal), with the slot-0
// correction dhat + (srcmap[c] >> 1) S; else buffer c.
template <typename R, int MAXR, class AC>
__global__ void __launch_bounds__(BOX_THREADS)
box_accum_kernel(const cpx<R>* __restrict__ src, long src_stride, int nsrc, AC ac, const BoxArgs* __restrict__ ap,
                 const cpx<R>* __restrict__ tw, const int* __restrict__ tabs, R scale, int sign,
                 const int* __restrict__ srcmap, SpRev rv, const cpx<R>* __restrict__ dhat) {
  extern __shared__ unsigned char smem_raw[];
  cpx<R>* lds = reinterpret_cast<cpx<R>*>(smem_raw);
  const BoxArgs& a = *ap;
  constexpr int TM = BOX_TILE_REGS<R>;
  constexpr int U = 8;
  const int tid = threadIdx.x;
  const unsigned nb = (unsigned)a.blocks();
  unsigned c4[4];
  box_tile_coord(a, (unsigned)tid, c4);
  const unsigned csp = c4[0] + mul24(a.Nx, c4[1]) + mul24(a.Nx * a.Ny, c4[2]);
  const unsigned ct = mul24(c4[3], (unsigned)a.tl_str);
  const unsigned l0 = c4[0] + mul24(c4[1], (unsigned)a.lds_stride[1]) + mul24(c4[2], (unsigned)a.lds_stride[2]) +
                      mul24(c4[3], (unsigned)a.lds_stride[3]);
  const unsigned ls0 = (unsigned)box_outpos(a, 0, (int)c4[0]) +
                       mul24((unsigned)box_outpos(a, 1, (int)c4[1]), (unsigned)a.lds_stride[1]) +
                       mul24((unsigned)box_outpos(a, 2, (int)c4[2]), (unsigned)a.lds_stride[2]) +
                       mul24((unsigned)box_outpos(a, 3, (int)c4[3]), (unsigned)a.lds_stride[3]);
  const int* tsp = tabs + a.tab_off;
  const int* tt = tsp + a.ntiles;
  const int* tld = tt + a.ntiles;
  const int* tst = tld + a.ntiles;
  // the plan guarantees a.ntiles == TM (FFTPlan::accum_ok): every register
  // slot is a tile
  // FP32: the next source's box is prefetched into registers; FP64 (twice
  // the registers per element) loads each box straight into LDS instead
  constexpr bool PF = sizeof(R) == 4;
  constexpr int TV = PF ? TM : 1;
  cpx<R> v[TV], sum[TM];
  auto srcld = [&](int cc) {
    const int e = srcmap ? srcmap[cc] : 2 * cc;
    return AccumSrcLoad<R>{src + (long)(e >> 1) * src_stride, dhat ? dhat + (long)(e >> 1) * a.S : nullptr, rv,
                           a.S, a.M, e & 1};
  };
  box_twiddles_to_lds<R>(lds, a, tw, tid, BOX_THREADS);  // ordered before use by the loop's first barrier
  unsigned b = blockIdx.x;
  int c = 0;
  if constexpr (PF) {
    if (b < nb) box_fetch<R, TM>(v, srcld(0), a, b, nb, csp, ct, tsp, tt);
  }
  while (b < nb) {
    if constexpr (PF) {
#pragma unroll
      for (int j = 0; j < TM; ++j)
        if (j < a.ntiles) lds[l0 + (unsigned)tld[j]] = v[j];
    } else {
      const BoxOrigin O = box_origin(a, box_remap(b, nb, a.xcd_group));
      const unsigned sp0 = box_sp(a, O, 0, 0, 0) + csp;
      const unsigned t0 = mul24((unsigned)O.bt, (unsigned)a.tl_mul) + ct;
      const AccumSrcLoad<R> sc = srcld(c);
#pragma unroll
      for (int j0 = 0; j0 < TM; j0 += 8) {
        cpx<R> w[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) w[u] = sc(sp0 + (unsigned)tsp[j0 + u], t0 + (unsigned)tt[j0 + u]);
#pragma unroll
        for (int u = 0; u < 8; ++u) lds[l0 + (unsigned)tld[j0 + u]] = w[u];
      }
    }
    __syncthreads();
    const bool more = c + 1 < nsrc;
    const unsigned b2 = more ? b : b + gridDim.x;
    const int c2 = more ? c + 1 : 0;
    const AccumSrcLoad<R> ld2 = srcld(c2);
    for (int i = 0; i < a.nsub; ++i) {
      if constexpr (PF) {
        if (i == a.pf_at && b2 < nb) box_fetch<R, TM>(v, ld2, a, b2, nb, csp, ct, tsp, tt);
      }
      box_dispatch<R, MAXR>(lds, a, box_sub_at(a, i), tw, tid, BOX_THREADS, sign);
      __syncthreads();
    }
    if constexpr (PF) {
      if (a.pf_at >= a.nsub && b2 < nb) box_fetch<R, TM>(v, ld2, a, b2, nb, csp, ct, tsp, tt);
    }
    // contributions of source c
    const BoxOrigin O = box_origin(a, box_remap(b, nb, a.xcd_group));
    const unsigned sp0 = box_sp(a, O, 0, 0, 0) + csp;
    const unsigned t0 = mul24((unsigned)O.bt, (unsigned)a.ts_mul) + mul24(c4[3], (unsigned)a.ts_str);
#pragma unroll
    for (int j0 = 0; j0 < TM; j0 += U) {
      typename AC::Pre pre[U];
#pragma unroll
      for (int u = 0; u < U; ++u) pre[u] = ac.prep(c, sp0 + (unsigned)tsp[j0 + u], t0 + (unsigned)tt[j0 + u]);
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int j = j0 + u;
        const cpx<R> val = lds[ls0 + (unsigned)tst[j]] * scale;
        sum[j] = c == 0 ? ac.contrib(c, pre[u], val) : sum[j] + ac.contrib(c, pre[u], val);
        ac.side(c, sp0 + (unsigned)tsp[j], t0 + (unsigned)tt[j], pre[u], val);
      }
    }
    if (!more) {
#pragma unroll
      for (int j = 0; j < TM; ++j) ac.finish(sp0 + (unsigned)tsp[j], t0 + (unsigned)tt[j], sum[j]);
    }
    __syncthreads();  // the LDS box is rewritten by the next iteration
    b = b2;
    c = c2;
  }
}

// The pass descriptor is read from device memory (the plan's copy) rather
// than passed by value: run-time indices into a by-value kernel argument
// make the compiler copy the whole argument to scratch memory.
template <typename R, class LD, class ST>
__global__ void __launch_bounds__(BOX_THREADS)
box_fft_kernel(LD ld, ST st, const BoxArgs* __restrict__ ap, const cpx<R>* __restrict__ tw,
               const int* __restrict__ tabs, R scale, int sign) {
  extern __shared__ unsigned char smem_raw[];
  cpx<R>* lds = reinterpret_cast<cpx<R>*>(smem_raw);
  const BoxArgs& a = *ap;
  if (a.tiled) {
    box_persistent<R>(lds, ld, st, a, tw, tabs, scale, threadIdx.x, sign);
    return;
  }
  const unsigned lb = box_remap(blockIdx.x, gridDim.x, a.xcd_group);
  const BoxOrigin O = box_origin(a, lb);
  if (box_skipped(st, a, lb, O)) return;
  const int tid = threadIdx.x;
  box_twiddles_to_lds<R>(lds, a, tw, tid, BOX_THREADS);
  box_load<R>(lds, ld, a, O, tid, BOX_THREADS, tabs);
  __syncthreads();
  for (int i = 0; i < a.nsub; ++i) {
    box_dispatch<R>(lds, a, box_sub_at(a, i), tw, tid, BOX_THREADS, sign);
    __syncthreads();
  }
  box_store<R>(lds, st, a, O, scale, tw, tid, BOX_THREADS, tabs);
}

// Host-side plan of box passes (built without a device).
template <typename R>
struct BoxPlan {
  bool ok = false;
  std::vector<BoxArgs> fwd, bwd;
  std::vector<cpx<R>> tw_host[2];  // [0]: sign -1, [1]: sign +1
  std::vector<int> tab_host;       // tile tables of the tiled passes
  size_t max_lds_bytes = 0;  // largest box of this plan
  size_t lds_limit = 0;      // the device's per-work-group LDS limit the plan was built for
  // plan the transform of the axes in `mask` of an (nx,ny,nz,m) grid with
  // boxes of at most `budget` elements and `lds_bytes` of LDS
  void build(int nx, int ny, int nz, int m, int mask, long budget, size_t lds_bytes);
};

}  // namespace gmod
