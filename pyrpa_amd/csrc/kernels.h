// Grid kernels of the convolution path: irreducible -> full-grid scatter with
// the reference's symmetry unfolding, pointwise products, full -> irreducible
// gather, thresholds and reductions.
//
// k-space arrays keep the reference's Fortran layouts (column-major):
//   Gk(Nk,Nw,No,No), chi(Nk,Nw,Nchi,Nchi), hamk(No,No,Nk), eig(No,Nk),
//   uni(No,No,Nk).
// Full grids are (Nx,Ny,Nz,M=2Nw), x fastest; the spatial index of full-BZ
// point i is sp(i) = kmap(1,i) + Nx*kmap(2,i) + Nx*Ny*kmap(3,i).
#pragma once
#include "gmod_common.h"

namespace gmod {

// Device copy of the k-mesh bookkeeping (finv.f90 outputs).
struct Mesh {
  int Nx = 0, Ny = 0, Nz = 0, Nw = 0, M = 0;
  long Nkall = 0, Nk = 0;
  DBuf<int> irr;      // invk(1,i)-1
  DBuf<int> flag;     // invk(2,i)
  DBuf<int> sp;       // spatial grid index of full point i
  DBuf<int> rev;      // spatial index of -r for spatial index s
  DBuf<int> irr_sp;   // spatial index of the flag-0 full point of irreducible k (-1: none)
  DBuf<int> irr_s;    // per spatial index: irreducible k
  DBuf<int> flag_s;   // per spatial index: 0 direct / 1 reversed
  DBuf<int> kgat_s;   // per spatial index: irreducible k if this is its flag-0 point, else -1
  std::vector<int> kgat_host;  // the same on the host
  DBuf<int> kf_s;     // per spatial index: 2 irr_s + (flag_s != 0), one load for both
  DBuf<int> full_s;   // per spatial index: full-list index i of the point (data on the full zone)
  DBuf<int> nfull_s;  // per spatial index: full-list index of -k, invk(3,i)-1
  bool nfull_ok = true;  // every invk(3,i) lies in [1, Nkall]
  DBuf<double> kw;    // full-BZ multiplicity of irreducible k
  std::vector<double> kw_host;
  bool complete = true;  // every irreducible k has a direct (flag 0) full-BZ point
  long N() const { return Nkall * (long)M; }
  void build(const int64_t* kmap, const int64_t* invk, long Nkall, long Nk, int Nx, int Ny, int Nz,
             int Nw, hipStream_t s);
};

// ---- fused load / store functors for FFTPlan::exec_fused
// Address a = sp + Nkall * t on the full grid (sp: spatial index, t: slot).
struct MeshDev {
  const int *irr_s, *flag_s, *kgat_s, *rev, *kf_s;  // per spatial index
  long Nkall, Nk;
  int Nw, M;
  const int *full_s = nullptr, *nfull_s = nullptr;
  __host__ double irr_elems() const { return (double)Nk * Nw; }  // elements of one irreducible (k, w) slice
};

// Algorithmic bytes of a fused functor per pass (the roofline accounting):
// algo_bytes(N) with N = elements of the full grid.  A grid read or write
// counts N b; a read of irreducible slices counts their Nk Nw elements.
MeshDev mesh_dev(const Mesh& m);

// G_ab on the full grid (fchi.f90:563-583, fself.f90:226-236):
//   slot t <  Nw : w_t ;  k direct -> G(k)_ab, k = -k_irr -> G(k_irr)_ba
//   slot t >= Nw : -w_j with j = M-1-t ; direct -> conj G_ba, reversed -> conj G_ab
template <typename R, typename S = double>
struct GreenLoad {
  const cpx<S>* G;  // S = float: a difference formed once in k space
  int a, b, No;
  MeshDev m;
  // key / value split: the box FFT issues the index loads of all its
  // elements before their value loads (a dependent pair per element would
  // otherwise cost two memory latencies each); value() is branch-free
  using Key = int;  // packed 2 k + flag (Mesh::kf_s), decoded only in value()
  __device__ Key key(unsigned sp) const { return m.kf_s[sp]; }
  using Raw = cpx<S>;  // raw / finish as for BosonLoad
  __device__ Raw raw(Key kf, unsigned, unsigned tu) const {
    const int t = (int)tu;
    const bool lo = t < m.Nw;
    const int j = lo ? t : m.M - 1 - t;
    const bool d = ((kf & 1) == 0) == lo;
    const int p = d ? a : b, r = d ? b : a;
    const long NkNw = m.Nk * (long)m.Nw;
    return G[(kf >> 1) + m.Nk * (long)j + NkNw * (p + (long)No * r)];
  }
  __device__ cpx<R> finish(const Raw& v, unsigned tu) const { return cpx<R>((int)tu < m.Nw ? v : conj(v)); }
  __device__ cpx<R> value(Key kf, unsigned sp, unsigned t) const { return finish(raw(kf, sp, t), t); }
  __device__ cpx<R> operator()(unsigned sp, unsigned t) const { return value(key(sp), sp, t); }
  __host__ double algo_bytes(double) const { return 2.0 * m.irr_elems() * sizeof(cpx<S>); }
};

// G_ab of a spinful basis on the full grid (fchi0withsoc.f90:57-75,
// fselfwithsoc.f90:627-635): a point reached through -k reads
// s_a s_b G_{-b,-a}(k, w) at positive w and s_a s_b conj G_{-a,-b}(k, w) at
// negative w (time reversal; -a the spin-flipped orbital ai, -b = bi); a
// point of the irreducible set reads G_ab(k, w) and conj G_ba(k, w)
template <typename R>
struct SocGreenLoad {
  const zd* G;
  int a, b, ai, bi, No;
  double sg;
  MeshDev m;
  using Key = int;  // packed 2 k + flag (Mesh::kf_s)
  __device__ Key key(unsigned sp) const { return m.kf_s[sp]; }
  using Raw = zd;
  __device__ Raw raw(Key kf, unsigned, unsigned tu) const {
    const int t = (int)tu;
    const bool lo = t < m.Nw;
    const int j = lo ? t : m.M - 1 - t;
    const bool rev = (kf & 1) != 0;
    const int p = rev ? (lo ? bi : ai) : (lo ? a : b);
    const int r = rev ? (lo ? ai : bi) : (lo ? b : a);
    const long NkNw = m.Nk * (long)m.Nw;
    const zd v = G[(kf >> 1) + m.Nk * (long)j + NkNw * (p + (long)No * r)];
    return rev ? v * sg : v;
  }
  __device__ cpx<R> finish(const Raw& v, unsigned tu) const { return cpx<R>((int)tu < m.Nw ? v : conj(v)); }
  __device__ cpx<R> value(Key kf, unsigned sp, unsigned t) const { return finish(raw(kf, sp, t), t); }
  __device__ cpx<R> operator()(unsigned sp, unsigned t) const { return value(key(sp), sp, t); }
  __host__ double algo_bytes(double) const { return 2.0 * m.irr_elems() * sizeof(zd); }
};

// bosonic V_(m,l) (fself.f90:203-224, feliash.f90:543-564): slot 0 = nu_0,
// slot Nw = the nu -> infinity value `stat`, slots t and M-t = +-nu_t
template <typename R, typename S = double>
struct BosonLoad {
  const cpx<S>* V;
  long off_ml, off_lm;
  zd stat;
  MeshDev m;
  using Key = int;  // packed 2 k + flag (Mesh::kf_s), decoded only in value()
  __device__ Key key(unsigned sp) const { return m.kf_s[sp]; }
  // raw (the load alone) / finish (conjugation, the nu -> inf slot): the box
  // FFT issues the raw loads of all its elements back to back and finishes
  // them when it writes them to LDS
  using Raw = cpx<S>;
  __device__ Raw raw(Key kf, unsigned, unsigned tu) const {
    const int t = (int)tu;
    const bool hi = t > m.Nw;
    const int j = hi ? m.M - t : (t < m.Nw ? t : m.Nw - 1);  // slot Nw reads a valid element it then drops
    const long off = (((kf & 1) == 0) != hi) ? off_ml : off_lm;
    return V[(kf >> 1) + m.Nk * (long)j + off];
  }
  __device__ cpx<R> finish(const Raw& v, unsigned tu) const {
    const int t = (int)tu;
    const cpx<R> r(t > m.Nw ? conj(v) : v);
    return t == m.Nw ? cpx<R>(stat) : r;
  }
  __device__ cpx<R> value(Key kf, unsigned sp, unsigned t) const { return finish(raw(kf, sp, t), t); }
  __device__ cpx<R> operator()(unsigned sp, unsigned t) const { return value(key(sp), sp, t); }
  __host__ double algo_bytes(double) const { return 2.0 * m.irr_elems() * sizeof(cpx<S>); }  // V_ml, V_lm slices
};

// The nu_0 mismatch of a transposed pair of bosonic grids: for the stored
// transform of pair (m, l) read by pair (l, m) through the conjugate
// reversal, D(x) = B_lm(x, 0) - conj B_ml(x, 0), zero when V(k, nu_0) is
// Hermitian in the pair index (every other slot, and the nu -> infinity slot
// when the static part is symmetric, obeys B_lm(x, t) = conj B_ml(x, -t) by
// the construction of BosonLoad).  offs[2 h], offs[2 h + 1]: the V offsets of
// (m, l) and (l, m) of entry h; out[h S + sp].
template <typename R, typename Sv>
__global__ void k_slot0_diff(cpx<R>* __restrict__ out, const cpx<Sv>* __restrict__ V, const long* __restrict__ offs,
                             int nh, unsigned S, MeshDev m) {
  const long tot = (long)nh * S;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < tot; i += (long)gridDim.x * blockDim.x) {
    const int h = (int)(i / S);
    const unsigned sp = (unsigned)(i - (long)h * S);
    const long oml = offs[2 * h], olm = offs[2 * h + 1];
    const BosonLoad<R, Sv> bml{V, oml, olm, zd(0, 0), m}, blm{V, olm, oml, zd(0, 0), m};
    const int kf = m.kf_s[sp];
    out[i] = blm.value(kf, sp, 0u) - conj(bml.value(kf, sp, 0u));
  }
}

// bosonic V_(m,l) of a spinful basis (feliashwithsoc.f90:545-573): a point
// reached through -k reads s V(q)_(l',m') at positive nu and
// s conj V(q)_(m',l') at negative nu, l' / m' the spin-flipped pairs and
// s = s_l1 s_l2 s_m1 s_m2; slot Nw holds `stat`
template <typename R>
struct SocBosonLoad {
  const zd* V;
  long off_ml, off_lm, off_rp, off_rn;
  double sg;
  zd stat;
  MeshDev m;
  using Key = int;
  __device__ Key key(unsigned sp) const { return m.kf_s[sp]; }
  __device__ cpx<R> value(Key kf, unsigned, unsigned tu) const {
    const int t = (int)tu;
    if (t == m.Nw) return cpx<R>(stat);
    const bool hi = t > m.Nw;
    const int j = hi ? m.M - t : t;
    const bool rev = (kf & 1) != 0;
    const long off = rev ? (hi ? off_rn : off_rp) : (hi ? off_lm : off_ml);
    zd v = V[(kf >> 1) + m.Nk * (long)j + off];
    if (hi) v = conj(v);
    return cpx<R>(rev ? v * sg : v);
  }
  __device__ cpx<R> operator()(unsigned sp, unsigned t) const { return value(key(sp), sp, t); }
  __host__ double algo_bytes(double) const { return 2.0 * m.irr_elems() * sizeof(zd); }
};

// anomalous F_(a,b) held on the full zone (Nkall, Nw, No, No) of a spinful
// basis (feliashwithsoc.f90:575-580): slot t < Nw reads F_ab(k, w_t), slot
// M-1-j reads -F_ba(-k, w_j)
template <typename R>
struct FullAnomLoad {
  const zd* F;
  long oab, oba;
  MeshDev m;
  __device__ cpx<R> operator()(unsigned sp, unsigned tu) const {
    const int t = (int)tu;
    const long NkW = m.Nkall;
    if (t < m.Nw) return cpx<R>(F[m.full_s[sp] + NkW * (long)t + oab]);
    const int j = m.M - 1 - t;
    const zd v = F[m.nfull_s[sp] + NkW * (long)j + oba];
    return cpx<R>(zd(-v.re, -v.im));
  }
};

// scatter slots t < Nw of every spatial point to the full-zone array
// dst(Nkall, Nw) (feliashwithsoc.f90:606-612), added when acc
template <typename R>
struct FullGatherStore {
  zd* dst;
  int acc;
  MeshDev m;
  struct Pre { long o; };
  __device__ Pre prep(unsigned sp, unsigned t) const {
    return Pre{(int)t < m.Nw ? m.full_s[sp] + m.Nkall * (long)t : -1};
  }
  __device__ void commit(unsigned, unsigned, const Pre& q, cpx<R> v) const {
    if (q.o < 0) return;
    dst[q.o] = acc ? dst[q.o] + zd(v) : zd(v);
  }
  __device__ void operator()(unsigned sp, unsigned t, cpx<R> v) const { commit(sp, t, prep(sp, t), v); }
  __device__ bool skip(unsigned t) const { return (int)t >= m.Nw; }
  __device__ bool skip_box(unsigned) const { return false; }
  __host__ double read_fraction(bool t_box) const { return t_box ? 1.0 : (double)m.Nw / m.M; }
};

// anomalous F_(a,b) with parity factors (feliash.f90:566-577)
template <typename R>
struct AnomLoad {
  const zd* F;
  long oab, oba;
  double pp, sgn, dprt;
  MeshDev m;
  __device__ cpx<R> operator()(unsigned sp, unsigned tu) const {
    const int t = (int)tu;
    const long k = m.irr_s[sp];
    const int f = m.flag_s[sp];
    if (t < m.Nw) {
      zd v = F[k + m.Nk * (long)t + oab];
      return cpx<R>(f == 0 ? v : v * (pp * dprt));
    }
    const int j = m.M - 1 - t;
    if (f == 0) return cpx<R>(conj(F[k + m.Nk * (long)j + oab]) * (sgn * dprt * pp));
    return cpx<R>(F[k + m.Nk * (long)j + oba] * sgn);
  }
};

// one leg of the anomalous bubble chi^FF (ffeliash.f90:556-580): slot t < Nw
// reads F(k_irr, w_t) at orbital pair od (flag-0 point) or or_ (-k point),
// slot M-1-j reads c2 F(k_irr, w_j) at pair nd / nr; cj conjugates the leg
// (F^+14).  The four pairs are the reference's own index choices.
template <typename R>
struct FFLoad {
  const zd* F;
  long od, nd, or_, nr;  // element offsets of the (a, b) slices
  double c2;
  int cj;
  MeshDev m;
  __device__ cpx<R> operator()(unsigned sp, unsigned tu) const {
    const int t = (int)tu;
    const long k = m.irr_s[sp];
    const int f = m.flag_s[sp];
    zd v;
    if (t < m.Nw) {
      v = F[k + m.Nk * (long)t + (f == 0 ? od : or_)];
    } else {
      const int j = m.M - 1 - t;
      v = F[k + m.Nk * (long)j + (f == 0 ? nd : nr)] * c2;
    }
    return cpx<R>(cj ? conj(v) : v);
  }
};

// sum_n U_an conj(U_bn) g_n(k, +-tau_s); -k uses the conjugated coefficient
// (fchi.f90:703-729, 791-812), overflow-safe forms of the reference
template <typename R>
struct TauRefLoad {
  const double* eig;
  const zd* uni;
  int a, b, No, which;
  double mu, beta, dtau;
  MeshDev m;
  __device__ cpx<R> operator()(unsigned sp, unsigned su) const {
    const int s = (int)su;
    const long k = m.irr_s[sp];
    const int f = m.flag_s[sp];
    const double tau = (double)s * dtau;
    zd acc(0, 0);
    for (int n = 0; n < No; ++n) {
      const double xi = eig[n + (long)No * k] - mu;
      double gv;
      if (xi >= 0.0) {
        const double d = 1.0 + exp(-beta * xi);
        gv = which == 0 ? -exp(-xi * tau) / d : exp(-xi * (beta - tau)) / d;
      } else {
        const double d = 1.0 + exp(beta * xi);
        gv = which == 0 ? -exp(xi * (beta - tau)) / d : exp(xi * tau) / d;
      }
      zd c = uni[a + (long)No * (n + (long)No * k)] * conj(uni[b + (long)No * (n + (long)No * k)]);
      if (f != 0) c = conj(c);
      acc += c * gv;
    }
    return cpx<R>(acc);
  }
};

// sum of up to three products coef * A(r,t) * B(-r, -t)  (revt) or B(-r, t)
// (fchi.f90:593, 824)
constexpr int PROD_MAXT = 9;
// NT > 0: exactly NT terms, known at compile time (the loads of an element
// then carry no loop, and the box FFT issues them for several elements at
// once); NT = 0: nterm terms.
template <typename R, int NT = 0>
struct ProductLoad {
  const cpx<R>* A[PROD_MAXT];
  const cpx<R>* B[PROD_MAXT];
  R coef[PROD_MAXT];
  int revt[PROD_MAXT];
  int nterm;
  __host__ void add(const cpx<R>* a, const cpx<R>* b, R c, int rt) {
    if (nterm >= PROD_MAXT) throw Error("too many product terms");
    A[nterm] = a; B[nterm] = b; coef[nterm] = c; revt[nterm] = rt; ++nterm;
  }
  MeshDev m;
  int bmir = 0;  // 1: every B is a mirrored grid (mirror_grid, all revt = 1) read at (r, t)
  template <int N>
  __host__ ProductLoad<R, N> fixed() const {
    ProductLoad<R, N> p{};
    for (int i = 0; i < PROD_MAXT; ++i) { p.A[i] = A[i]; p.B[i] = B[i]; p.coef[i] = coef[i]; p.revt[i] = revt[i]; }
    p.nterm = nterm;
    p.m = m;
    p.bmir = bmir;
    return p;
  }
  using Key = unsigned;  // spatial index of -r (no load with mirrored B)
  __device__ Key key(unsigned sp) const { return bmir ? sp : (unsigned)m.rev[sp]; }
  __device__ cpx<R> value(Key rs, unsigned sp, unsigned t) const {
    const size_t ad = sp + (size_t)m.Nkall * t;
    const unsigned rtt = t == 0 ? 0u : (unsigned)m.M - t;
    cpx<R> v(0, 0);
#pragma unroll
    for (int i = 0; i < (NT ? NT : nterm); ++i) {
      const unsigned rt = (revt[i] && !bmir) ? rtt : t;
      v += A[i][ad] * B[i][rs + (size_t)m.Nkall * rt] * coef[i];
    }
    return v;
  }
  __device__ cpx<R> operator()(unsigned sp, unsigned t) const { return value(key(sp), sp, t); }
  __host__ double algo_bytes(double N) const { return 2.0 * nterm * N * sizeof(cpx<R>); }  // A and B per term
};

// G_ab(current) - G_ab(reference) on the full grid, formed in FP64 before
// the conversion to the grid precision (increments of the bilinear bubble)
template <typename R>
struct GreenDiffLoad {
  GreenLoad<double> cur, ref;  // on one mesh: one key serves both
  using Key = typename GreenLoad<double>::Key;
  __device__ Key key(unsigned sp) const { return cur.key(sp); }
  __device__ cpx<R> value(Key q, unsigned sp, unsigned t) const {
    return cpx<R>(cur.value(q, sp, t) - ref.value(q, sp, t));
  }
  __device__ cpx<R> operator()(unsigned sp, unsigned t) const { return value(key(sp), sp, t); }
  __host__ double algo_bytes(double N) const { return cur.algo_bytes(N) + ref.algo_bytes(N); }
};

// V(current) - V(reference): the static nu->inf slot cancels
template <typename R>
struct BosonDiffLoad {
  BosonLoad<double> cur, ref;
  using Key = typename BosonLoad<double>::Key;
  __device__ Key key(unsigned sp) const { return cur.key(sp); }
  __device__ cpx<R> value(Key q, unsigned sp, unsigned t) const {
    return cpx<R>(cur.value(q, sp, t) - ref.value(q, sp, t));
  }
  __device__ cpx<R> operator()(unsigned sp, unsigned t) const { return value(key(sp), sp, t); }
  __host__ double algo_bytes(double N) const { return cur.algo_bytes(N) + ref.algo_bytes(N); }
};

template <typename R>
struct TauRefDiffLoad {
  TauRefLoad<double> cur, ref;
  __device__ cpx<R> operator()(unsigned sp, unsigned t) const { return cpx<R>(cur(sp, t) - ref(sp, t)); }
};

// Increment of a bilinear bubble B(X) = X13(r,t) X42(-r, -t or t) between a
// reference X = Xr and X + Xd, summed over up to three terms:
//   v = sum_i coef_i [ Ad13 (Br42 + Bd42) + Ar13 Bd42 ]
// (the products of fchi.f90:593, 824 expanded to first + second order in Xd)
constexpr int INC_MAXT = 3;
template <typename R, int NT = 0>  // NT as for ProductLoad
struct IncProductLoad {
  const cpx<R>* Ar[INC_MAXT];
  const cpx<R>* Ad[INC_MAXT];
  const cpx<R>* Br[INC_MAXT];
  const cpx<R>* Bd[INC_MAXT];
  R coef[INC_MAXT];
  int revt[INC_MAXT];
  int nterm;
  MeshDev m;
  int bmir = 0;  // 1: every Br, Bd is a mirrored grid (all revt = 1) read at (r, t)
  template <int N>
  __host__ IncProductLoad<R, N> fixed() const {
    IncProductLoad<R, N> p{};
    for (int i = 0; i < INC_MAXT; ++i) {
      p.Ar[i] = Ar[i]; p.Ad[i] = Ad[i]; p.Br[i] = Br[i]; p.Bd[i] = Bd[i]; p.coef[i] = coef[i]; p.revt[i] = revt[i];
    }
    p.nterm = nterm;
    p.m = m;
    p.bmir = bmir;
    return p;
  }
  __host__ void add(const cpx<R>* ar, const cpx<R>* ad, const cpx<R>* br, const cpx<R>* bd, R c, int rt) {
    if (nterm >= INC_MAXT) throw Error("too many increment terms");
    Ar[nterm] = ar; Ad[nterm] = ad; Br[nterm] = br; Bd[nterm] = bd; coef[nterm] = c; revt[nterm] = rt; ++nterm;
  }
  using Key = unsigned;  // spatial index of -r (no load with mirrored B)
  __device__ Key key(unsigned sp) const { return bmir ? sp : (unsigned)m.rev[sp]; }
  __device__ cpx<R> operator()(unsigned sp, unsigned t) const { return value(key(sp), sp, t); }
  __device__ cpx<R> value(Key rs, unsigned sp, unsigned t) const {
    const size_t ad = sp + (size_t)m.Nkall * t;
    const unsigned rtt = t == 0 ? 0u : (unsigned)m.M - t;
    cpx<R> v(0, 0);
#pragma unroll
    for (int i = 0; i < (NT ? NT : nterm); ++i) {
      const size_t bd = rs + (size_t)m.Nkall * ((revt[i] && !bmir) ? rtt : t);
      const cpx<R> bdv = Bd[i][bd];
      v += (Ad[i][ad] * (Br[i][bd] + bdv) + Ar[i][ad] * bdv) * coef[i];
    }
    return v;
  }
  __host__ double algo_bytes(double N) const { return 4.0 * nterm * N * sizeof(cpx<R>); }  // Ar, Ad, Br, Bd per term
};

// Sigma increment per vertex pair: acc (+)= dV~ (Fr + Fd) + FVr Fd, where dV~
// is the transformed vertex increment, Fr / Fd the transformed reference /
// increment of G_ab and FVr the transformed reference vertex of the pair
// (V G - Vr Gr with V = Vr + dV, G = Gr + dG)
template <typename R>
struct SigmaIncStore {
  cpx<R>* acc;
  const cpx<R>* Fr;
  const cpx<R>* Fd;
  const cpx<R>* FVr;
  int first;
  unsigned S;
  // prep (all loads) / commit (arithmetic and stores): the box FFT issues the
  // loads of a batch of elements before any of their stores
  struct Pre { cpx<R> acc, fr, fd, fv; };
  __device__ Pre prep(unsigned sp, unsigned t) const {
    const size_t ad = sp + (size_t)S * t;
    Pre q;
    // no branch around the load: on the first pass every element reads
    // element 0 (one cached line, no HBM stream) and drops it
    const cpx<R> a0 = acc[first ? 0 : ad];
    q.acc = first ? cpx<R>(0, 0) : a0;
    q.fr = Fr[ad];
    q.fd = Fd[ad];
    q.fv = FVr[ad];
    return q;
  }
  __device__ void commit(unsigned sp, unsigned t, const Pre& q, cpx<R> v) const {
    const size_t ad = sp + (size_t)S * t;
    acc[ad] = q.acc + v * (q.fr + q.fd) + q.fv * q.fd;
  }
  __device__ void operator()(unsigned sp, unsigned t, cpx<R> v) const { commit(sp, t, prep(sp, t), v); }
  __host__ double algo_bytes(double N) const {  // acc (read unless first), Fr, Fd, FVr read; acc written
    return ((first ? 3.0 : 4.0) + 1.0) * N * sizeof(cpx<R>);
  }
  __device__ bool skip(unsigned) const { return false; }
};

// acc = (first ? 0 : acc) + v * F
template <typename R>
struct MulAccStore {
  cpx<R>* acc;
  const cpx<R>* F;
  int first;
  unsigned S;
  struct Pre { cpx<R> acc, f; };
  __device__ Pre prep(unsigned sp, unsigned t) const {
    const size_t ad = sp + (size_t)S * t;
    Pre q;
    // no branch around the load: on the first pass every element reads
    // element 0 (one cached line, no HBM stream) and drops it
    const cpx<R> a0 = acc[first ? 0 : ad];
    q.acc = first ? cpx<R>(0, 0) : a0;
    q.f = F[ad];
    return q;
  }
  __device__ void commit(unsigned sp, unsigned t, const Pre& q, cpx<R> v) const {
    acc[sp + (size_t)S * t] = q.acc + v * q.f;
  }
  __device__ void operator()(unsigned sp, unsigned t, cpx<R> v) const { commit(sp, t, prep(sp, t), v); }
  __host__ double algo_bytes(double N) const {  // acc (read unless first) and F read, acc written
    return ((first ? 1.0 : 2.0) + 1.0) * N * sizeof(cpx<R>);
  }
  __device__ bool skip(unsigned) const { return false; }
};

// Accumulator functors of FFTPlan::exec_accum_last.  The per-source operand
// pointers sit in a device table (tab), so that a run-time source index does
// not index a kernel argument.
//   SigmaIncAccum: sum_c [dV~_c (Fr_c + Fd_c) + FVr_c Fd_c]
//   (SigmaIncStore's increment summed over the sources of one output)
//   MulAccAccum:   sum_c v_c F_c  (MulAccStore's sum)
template <typename R>
struct SigmaIncAccum {
  cpx<R>* const* tab;  // [c][0..2] = Fr_c, Fd_c, FVr_c
  cpx<R>* out;
  unsigned S;
  struct Pre { cpx<R> fr, fd, fv; };
  __device__ Pre prep(int c, unsigned sp, unsigned t) const {
    const size_t ad = sp + __umul24(S, t);
    return Pre{tab[3 * c][ad], tab[3 * c + 1][ad], tab[3 * c + 2][ad]};
  }
  __device__ cpx<R> contrib(int, const Pre& q, cpx<R> v) const { return v * (q.fr + q.fd) + q.fv * q.fd; }
  __device__ void side(int, unsigned, unsigned, const Pre&, cpx<R>) const {}
  __device__ void finish(unsigned sp, unsigned t, cpx<R> sum) const { out[sp + __umul24(S, t)] = sum; }
  __host__ double src_bytes(double N) const { return 3.0 * N * sizeof(cpx<R>); }  // Fr, Fd, FVr read
  __host__ double sum_bytes(double N) const { return N * sizeof(cpx<R>); }
};
// SigmaIncAccumG: sum_c dV~_c Gn_c + Y, the Sigma increment when the
// reference-vertex term Y = sum_c FVr_c Fd_c and Gn = Fr + Fd were formed
// beforehand (sigma_ref_contract)
template <typename R>
struct SigmaIncAccumG {
  cpx<R>* const* tab;  // [c] = Gn_c
  const cpx<R>* Y;
  cpx<R>* out;
  unsigned S;
  struct Pre { cpx<R> g; };
  __device__ Pre prep(int c, unsigned sp, unsigned t) const { return Pre{tab[c][sp + __umul24(S, t)]}; }
  __device__ cpx<R> contrib(int, const Pre& q, cpx<R> v) const { return v * q.g; }
  __device__ void side(int, unsigned, unsigned, const Pre&, cpx<R>) const {}
  __device__ void finish(unsigned sp, unsigned t, cpx<R> sum) const {
    const size_t ad = sp + __umul24(S, t);
    out[ad] = sum + Y[ad];
  }
  __host__ double src_bytes(double N) const { return N * sizeof(cpx<R>); }      // Gn read
  __host__ double sum_bytes(double N) const { return 2.0 * N * sizeof(cpx<R>); }  // Y read, the sum written
};

// For every point e of the full grid and Sigma output o (nout of them):
// Y_o = sum_{ab} FVr[tab[o NAB + ab]] Fd_ab, where ab = a + No b runs over
// all No^2 orbital pairs (NAB = No^2); also Gn_ab = Fr_ab + Fd_ab.  Grids of
// N elements, the orbital-pair grids consecutive (Fr, Fd, Gn), the vertex
// grids at FVr + c N, the outputs at Y + o N.
template <typename R>
void sigma_ref_contract(cpx<R>* Y, cpx<R>* Gn, const cpx<R>* Fr, const cpx<R>* Fd, const cpx<R>* FVr,
                        const int* tab, int nout, int No, long N, hipStream_t s);

// For every point (sp, t) of the full grid (N = S M) and Sigma output o
// (nout of them): out_o = sum_ab FVr_e Fd_ab + DVt_d (Fr_ab + Fd_ab), where
// ab = a + No b runs over the No^2 orbital pairs and (e, d) = tab[2 (o No^2
// + ab)], tab[2 (o No^2 + ab) + 1] are stored-transform entries: grid e >> 1
// at (sp, t), or conjugated at (rev[sp], t) plus the correction
// dhF / dhD[(e >> 1) S + sp] (the transformed nu_0 mismatch) when e & 1.
template <typename R>
void sigma_inc_contract(cpx<R>* out, const cpx<R>* Fr, const cpx<R>* Fd, const cpx<R>* FVr, const cpx<R>* DVt,
                        const cpx<R>* dhF, const cpx<R>* dhD, const int* tab, int nout, int No, unsigned S,
                        unsigned M, const int* rev, hipStream_t s);

template <typename R>
struct MulAccAccum {
  cpx<R>* const* tab;  // [c] = F_c
  cpx<R>* out;
  unsigned S;
  struct Pre { cpx<R> f; };
  __device__ Pre prep(int c, unsigned sp, unsigned t) const { return Pre{tab[c][sp + __umul24(S, t)]}; }
  __device__ cpx<R> contrib(int, const Pre& q, cpx<R> v) const { return v * q.f; }
  __device__ void side(int, unsigned, unsigned, const Pre&, cpx<R>) const {}
  __device__ void finish(unsigned sp, unsigned t, cpx<R> sum) const { out[sp + __umul24(S, t)] = sum; }
  __host__ double src_bytes(double N) const { return N * sizeof(cpx<R>); }
  __host__ double sum_bytes(double N) const { return N * sizeof(cpx<R>); }
};

// dst(k, t) (+)= w * v + J * sawc(t) for the flag-0 point of irreducible k and
// slots t < Nw (fchi.f90:601-609, fself.f90:258-267); other slots are dropped
template <typename R>
struct GatherStore {
  zd* dst;
  double w;
  zd J;
  const zd* sawc;
  int acc;
  MeshDev m;
  const zd* base = nullptr;  // optional: dst = base + w v (FP64 reference + increment)
  // optional: the partner pair's slice, written psign conj(dst value) at the
  // same (k, w) (the reference's partner copy, fused)
  zd* pdst = nullptr;
  double psign = 1.0;
  // optional, for a last pass whose boxes span the t axis: per box (index
  // after box_remap), nonzero if it holds a kept spatial point; live_frac is
  // the fraction of boxes that do
  const unsigned char* live = nullptr;
  double live_frac = 1.0;
  struct Pre { long o; zd old, bv, js; };  // o < 0: slot or point not kept
  // branch-free: a dropped slot or point is not stored, and an operand the
  // store does not use is read at element 0 (one cached line) and dropped
  __device__ Pre prep(unsigned sp, unsigned t) const {
    Pre q;
    const long k = m.kgat_s[sp];
    const int tt = (int)t < m.Nw ? (int)t : m.Nw - 1;
    const long o = (k < 0 ? 0 : k) + m.Nk * (long)tt;
    q.o = ((int)t < m.Nw && k >= 0) ? o : -1;
    const zd old = dst[acc ? o : 0];
    const zd bv = (base ? base : dst)[base ? o : 0];
    const zd sw = (sawc ? sawc : dst)[sawc ? tt : 0];
    q.old = acc ? old : zd(0, 0);
    q.bv = base ? bv : zd(0, 0);
    q.js = sawc ? J * sw : zd(0, 0);
    return q;
  }
  __device__ void commit(unsigned, unsigned, const Pre& q, cpx<R> v) const {
    if (q.o < 0) return;
    zd x = zd(v) * w;
    if (sawc) x += q.js;
    if (base) x += q.bv;
    const zd f = acc ? q.old + x : x;
    dst[q.o] = f;
    if (pdst) pdst[q.o] = zd(psign * f.re, -psign * f.im);
  }
  __device__ void operator()(unsigned sp, unsigned t, cpx<R> v) const { commit(sp, t, prep(sp, t), v); }
  __host__ double algo_bytes(double) const {  // the kept (k, w) slice written, read too when accumulating / rebasing
    return (1.0 + (acc ? 1.0 : 0.0) + (base ? 1.0 : 0.0) + (pdst ? 1.0 : 0.0)) * m.irr_elems() * sizeof(zd);
  }
  __device__ bool skip(unsigned t) const { return (int)t >= m.Nw; }
  __device__ bool skip_box(unsigned lb) const { return live && !live[lb]; }
  // share of the pass's input the kept boxes read: boxes spanning the t
  // axis are dropped through `live`, boxes of a few slots by skip()
  __host__ double read_fraction(bool t_box) const { return t_box ? (live ? live_frac : 1.0) : (double)m.Nw / m.M; }
};

// acc (+)= A * B
template <typename R>
void mul_acc(cpx<R>* acc, const cpx<R>* A, const cpx<R>* B, bool first, long n, hipStream_t s);

// The analytic tau reference kernels of one orbital pair (a, b) written out on
// the full grid (the functions of TauRefLoad, fchi.f90:703-729, 791-812):
// ta(sp, s) = sum_n c_n g0_n(tau_s), tb(sp, s) = sum_n c_n g1_n(tau_s), with
// c_n = U_an conj(U_bn) (conjugated at -k).  diff: the differences between
// chemical potentials mu and mu2.  The exponentials advance by one factor per
// slot (No <= 8; FP64 arithmetic, stored at the grid precision).
template <typename R>
void tau_ref_grids(cpx<R>* ta, cpx<R>* tb, const double* eig, const zd* uni, int a, int b, int No, double mu,
                   double mu2, bool diff, double beta, double dtau, const MeshDev& m, hipStream_t s);

// ---- elementwise
void conj_copy(zd* dst, const zd* src, long n, hipStream_t s);
void conj_copy(zd* dst, const zd* src, long n, double sign, hipStream_t s);  // dst = sign conj(src), sign = +-1
void threshold(zd* a, long n, double eps, hipStream_t s);
// dst(r, t) = src(-r, -t) on an (Nx, Ny, Nz, M) grid (the operand a
// (-r, -tau) product reads, laid out to be read in place)
template <typename R>
void mirror_grid(cpx<R>* dst, const cpx<R>* src, int Nx, int Ny, int Nz, int M, hipStream_t s);
// dst(k, e) = src(k, w = 0, e) with threshold eps, e < nmat (the first
// frequency slice of a (Nk, Nw, nmat) array, as a (Nk, nmat) array)
void threshold_slice(zd* dst, const zd* src, long Nk, long NkNw, long nmat, double eps, hipStream_t s);
void scale(zd* a, long n, double f, hipStream_t s);
void axpy(zd* y, const zd* x, zd alpha, long n, hipStream_t s);      // y += alpha x
void copy(zd* dst, const zd* src, long n, hipStream_t s);
void lincomb(zd* y, const zd* a, double ca, const zd* b, double cb, long n, hipStream_t s);  // y = ca a + cb b
void threshold_copy(zd* dst, const zd* src, long n, double eps, hipStream_t s);              // dst = thr(src)
void diff_to_f32(cpx<float>* dst, const zd* a, const zd* b, long n, hipStream_t s);          // dst = (float)(a - b)
template <typename R>
void add_inplace(cpx<R>* y, const cpx<R>* x, long n, hipStream_t s);  // y += x

// ---- deterministic reductions (two-stage, fixed order)
// sum over k-space arrays laid out (Nk, rest): weights kw[k] (or 1)
zd dot_w(const zd* u, const zd* v, const double* kw, long Nk, long rest, hipStream_t s);  // sum kw conj(u) v
double max_rel_err(const zd* s0, const zd* sd, long n, double floor, hipStream_t s, long* where);
double max_real(const zd* a, long n, hipStream_t s);
double max_abs_diff(const zd* a, const zd* b, long n, hipStream_t s);                      // max |a - b| (b null: max |a|)
void sum_diff(zd* a, zd* b, long n, hipStream_t s);                                        // a, b <- a + b, a - b
double min_real(const zd* a, long n, hipStream_t s);

}  // namespace gmod
