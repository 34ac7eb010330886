// Convolution engine: chi0 (sharp cutoff / tail corrected), Sigma = V (*) G and
// the Eliashberg gap convolution on the full (k-mesh x tau) grid.
#pragma once
#include <array>

#include "fft.h"
#include "kernels.h"
#include "linalg.h"

namespace gmod {

struct OrbPairs {
  int No = 0, Nchi = 0;
  std::vector<std::array<int, 2>> ol;       // 0-based (o1, o2) per chi index (Fortran olist(:,1), olist(:,2))
  std::vector<std::array<int, 2>> irr;      // irreducible (l, m) pairs  (fchi.f90:470-480)
  std::vector<std::array<int, 2>> partner;  // partner[(m + Nchi*l)] = (m2, l2)  (fchi.f90:456-457)
  // spinful basis (fchi0withsoc.f90): spin-flipped orbital, spin sign per
  // orbital and the sign of each partner copy (sgnsig2; +1 without spin)
  bool soc = false;
  std::vector<int> inv;
  std::vector<double> spin;
  std::vector<double> psign;
  void from_olist(const int64_t* olist, int Nchi, int No);
  // partners chi_{1234} = s chi*_{-4-3-2-1} with spin flips (fchi0withsoc.f90:216-279)
  void from_olist_soc(const int64_t* olist, int Nchi, int No, const int64_t* invs, const int64_t* slist);
  void from_maps(const int32_t* irr_chi, const int32_t* chi_map);  // caller-provided tables
};

template <typename R>
class ConvEngine {
 public:
  ConvEngine(const Mesh& m, const OrbPairs& op);
  // Fourier transform every full-grid G_ab (a,b = 0..No-1) into Fg (or F0g)
  void transform_green(const zd* G, bool g0, hipStream_t s);
  // chi(Nk,Nw,Nchi,Nchi) <- bubble of the transformed Fg (and, if tail, minus
  // the G0 bubble plus the analytic tau reference).  Thresholds at 1e-9.
  void chi0(zd* chi, double temp, bool tail, const double* eig_d, const zd* uni_d, const double* eig_h,
            const zd* uni_h, const int64_t* invk, double mu, hipStream_t s, bool apply_threshold = true);
  // The tail-corrected chi0 of an FP32 engine: conv[G] - conv[G0] at grid
  // precision, the analytic tau reference and its sawtooth term through an
  // FP64 engine (tau_reference), combined in FP64 in (q, i nu)
  void chi0_split_tail(zd* chi, double temp, const double* eig_d, const zd* uni_d, const double* eig_h,
                       const zd* uni_h, const int64_t* invk, double mu, hipStream_t s);
  // chi(pair) += -(1/(T Nkall)) ref[G0](pair) + J(pair) sawc for every
  // irreducible pair (no partner copies), at this engine's precision
  void tau_reference(zd* chi, double temp, const double* eig_d, const zd* uni_d, double mu, const std::vector<zd>& J,
                     const zd* sawc_d, hipStream_t s);
  // sigma(Nk,Nw,No,No) <- (T/Nkall) sum_pairs V (*) G, using Fg; thresholds at 1e-9
  void sigma(zd* sig, const zd* V, const double* S_h, const double* C_h, double temp, hipStream_t s,
             bool apply_threshold = true);
  // the same with the nu -> infinity value of each pair given directly,
  // stat_h[m + Nchi*l] (-Vmat(m,l) in a spinful basis, fselfwithsoc.f90:592)
  void sigma(zd* sig, const zd* V, const double* stat_h, double temp, hipStream_t s, bool apply_threshold = true);

  // chi^FF(Nk,Nw,Nchi,Nchi) <- the anomalous bubble of F (ffeliash.f90:491-612):
  // per irreducible pair, the legs F^+14 and F23 (FFLoad) transformed, their
  // product with the (-r,-tau) reversal inverse-transformed and gathered with
  // T/Nkall, the partner slot conjugated; thresholds at 1e-9
  void chi0_ff(zd* chi, const zd* F, const double* prt, bool singlet, double temp, hipStream_t s);

  // the spatial transforms of the tau reference kernels TA, TB of pair (a, b)
  // at mu, or (diff) of their change from mu2 to mu (fchi.f90:703-812)
  void tau_grids(cpx<R>* ta, cpx<R>* tb, int a, int b, const double* eig_d, const zd* uni_d, double mu, double mu2,
                 bool diff, double temp, hipStream_t s);

  // tail constants J (per chi pair, [l + Nchi*m]) at mu and the sawtooth table (built once)
  void tail_constants(double temp, const double* eig_h, const zd* uni_h, const int64_t* invk, double mu,
                      std::vector<zd>& J, DBuf<zd>& sawc_d, hipStream_t s) const;

  // ---- increments from a fixed reference (mixed-precision FLEX): the
  // transforms of a reference iterate are held at grid precision and every
  // later iterate is reached by the transforms of its difference from it.
  // Transforms of Gref (and, with the tail, G0ref and the tau kernels at mu_ref).
  void set_reference(const zd* Gref, const zd* G0ref, bool tail, const double* eig_d, const zd* uni_d, double mu_ref,
                     double temp, hipStream_t s);
  // Transforms of the reference vertex for all Nchi^2 pairs (full grid incl.
  // the nu->inf slot), one per transposed pair in the contraction form when
  // the static slot is symmetric.
  void set_vertex_reference(const zd* Vref, const double* S_h, const double* C_h, hipStream_t s);
  // chi <- chi_base + chi0[G] - chi0[Gref] (tail: + the G0 and tau-reference
  // terms at mu vs mu_ref and dJ sawc), with the reference's partner /
  // self-partner conjugation; chi_base holds chi0[Gref].  Every slot the
  // gather keeps is written; the others keep their contents.  Leaves the
  // transformed differences in Fd (F0d, TAd, TBd).
  void chi0_increment(zd* chi, const zd* chi_base, const zd* G, const zd* Gref, const zd* G0, const zd* G0ref,
                      double temp, bool tail, const double* eig_d, const zd* uni_d, double mu, double mu_ref,
                      const std::vector<zd>& dJ, const zd* sawc_d, hipStream_t s);
  // sig <- sig_base + Sigma[V, G] - Sigma[Vref, Gref] (unthresholded); uses
  // Fr/Fd of the preceding chi0_increment and FVr.
  // dv_ready: the FP32 vertex difference V - Vref is already in dv_buffer()
  // (written by the fused vertex kernel); else it is formed here from V and Vref
  void sigma_increment(zd* sig, const zd* sig_base, const zd* V, const zd* Vref, double temp, hipStream_t s,
                       bool dv_ready = false);
  cpx<float>* dv_buffer(long n) {
    if (dvbuf.n != (size_t)n) dvbuf.alloc(n);
    return dvbuf.p;
  }
  // Eliashberg: cache FFT[V_Delta] of the qualifying pairs (feliash.f90:540)
  void prepare_gap(const zd* Vd, const double* S_h, const double* C_h, bool singlet, hipStream_t s);
  // newdelta (un-weighted, Hermitised) <- sum V_Delta (*) F  (feliash.f90:486-635)
  void gap(zd* nd, const zd* F, const double* prt, bool singlet, hipStream_t s);
  // Eliashberg in a spinful basis (feliashwithsoc.f90:520-612): the pairs
  // mkdelta_soc convolves for gap_sym and FFT[V_Delta] of each (cached when
  // memory allows); gap_soc: nd(Nkall,Nw,No,No) <- the sum over them of
  // V (*) F, F on the full zone, before the spin symmetrisation
  void prepare_gap_soc(const zd* Vd, const double* Vmat_h, const int64_t* invschi, long gap_sym, hipStream_t s);
  void gap_soc(zd* nd, const zd* F, hipStream_t s);

  const Mesh& mesh;
  const OrbPairs& op;
  long N;
  const FFTPlan<R>* p4;
  const FFTPlan<R>* p3;
  DBuf<cpx<R>> Fg, F0g, TA, TB, work, tmp, acc, Vft, Ff;
  DBuf<cpx<R>> Fr, Fd, F0r, F0d, TAr, TBr, TAd, TBd, FVr;  // increment mode (r: reference, d: difference)
  DBuf<cpx<R>> Mg;  // mirrored Fg (G(-r,-t)): the second factor of the bubble, read in place
  DBuf<cpx<R>> Lff;  // the two legs of one anomalous-bubble pair
  // Sigma through the accumulating last pass: the sources of one output
  // after their first passes, and the device table of their operand grids
  DBuf<cpx<R>> srcs;
  DBuf<cpx<R>*> accum_tab;
  // the grouped Sigma increment: Y of every output, then Gn = Fr + Fd of every
  // orbital pair (the contraction form: the outputs themselves); ytab_d: per
  // output and pair the reference-vertex grid
  DBuf<cpx<R>> ysum;
  DBuf<int> ytab_d;
  std::vector<cpx<R>*> accum_tab_h;
  DBuf<cpx<float>> dvbuf;                                  // vertex difference, k space
  DBuf<cpx<float>> dgbuf;                                  // G - Gref (then G0 - G0ref), k space
  // boxes of p4's forward last pass (it spans the t axis) that hold a kept
  // spatial point: a spatial-first gather skips the others
  DBuf<unsigned char> tlive;
  double tlive_frac = 1.0;
  bool tlive_built = false;
  // the gather of a spatial-first inverse transform
  GatherStore<R> gather_sf(zd* dst, double w, zd J, const zd* sawc, int acc, hipStream_t s);
  std::vector<std::array<int, 2>> gap_pairs;
  std::vector<SocBosonLoad<R>> soc_vload;  // per gap pair of a spinful basis
  bool vft_cached = false;
  const zd* Vd_dev = nullptr;
  std::vector<double> S_cache, C_cache;
  cpx<R>* grid(DBuf<cpx<R>>& b, int a, int c) { return b.p + (long)(a + op.No * c) * N; }
  // The Sigma outputs in the order their accumulating passes run: an output
  // (m1, l1) and its transpose (l1, m1) back to back, sharing the head passes
  // (the spatial transforms) of their vertex sources when `share`: the
  // bosonic grid of pair (l, m) is conj B_ml(x, -t) in every slot but slot
  // 0, so its head passes are conj X_ml(-r, -t) plus, at t = 0, the spatial
  // transform of the slot-0 mismatch (slot0_corr), read through the
  // accumulating pass's source map.
  struct SigmaPlan {
    std::vector<int> outs;                               // m1 + No l1
    std::vector<std::vector<std::array<int, 3>>> src;    // per output: (m, l, source-map entry)
    std::vector<std::vector<std::array<int, 3>>> heads;  // per output, run before it: (m, l, slot)
    int nmax = 0;                                        // head slots
  };
  SigmaPlan plan_sigma(bool share) const;
  DBuf<int> srcmap_d;
  // The contraction form of the Sigma increment: the full transforms of the
  // vertex differences (DVt) and of the reference vertex (FVr) are stored
  // once per pair of transposed vertex pairs (the transform of pair (l, m)
  // is conj X_ml(-r, t) plus the transformed slot-0 mismatch, the same for
  // every t, after all passes), and one streaming kernel forms
  // every output out_o = sum_ab FVr_o,ab Fd_ab + DVt_o,ab (Fr_ab + Fd_ab).
  // grouped_table: per output o and orbital pair ab the vertex pair
  // m + Nchi l that joins them (false unless each (o, ab) has exactly one);
  // share_heads: a slot per stored transform, entry[m + Nchi l] = 2 slot
  // (+1: read conjugated at -r); contract_form: whether this form applies.
  bool grouped_table(std::vector<int>& ytab) const;
  std::vector<std::array<int, 2>> share_heads(std::vector<int>& entry) const;
  bool contract_form() const;
  // The conjugate reversal is exact except in slot 0, where it needs
  // V(k, nu_0) Hermitian in the pair index: slot0_corr writes, for nh
  // stored transforms of pairs (m, l) read by (l, m), the spatial transform
  // of D = B_lm(x, 0) - conj B_ml(x, 0) to dh[h S + sp] (offs_d: the V
  // offsets of (m, l) and (l, m) per h).  head_offsets appends those offsets.
  template <typename Sv>
  void slot0_corr(cpx<R>* dh, const cpx<Sv>* V, const long* offs_d, int nh, hipStream_t s);
  void head_offsets(std::vector<long>& offs, int m, int l) const;
  DBuf<cpx<R>> dhF, dhD, dh_tmp;
  DBuf<long> dhF_offs, dh_offs;
  std::vector<int> fv_entry;  // FVr entry per vertex pair (set_vertex_reference)
  bool fv_shared = false;     // FVr holds one transform per transposed pair
  DBuf<cpx<R>> DVt;
  DBuf<int> ctab_d;
  // an inverse transform whose first pass forms a product of 1-3 terms (PL:
  // ProductLoad or IncProductLoad), the term count fixed at compile time so
  // the operand pointers stay in scalar registers (a run-time term index puts
  // the functor in scratch memory)
  template <class PL, class ST>
  void product_pass(const PL& pl, const ST& st, hipStream_t s);
  void fft(cpx<R>* g, int sign, bool spatial_only, hipStream_t s);
};

}  // namespace gmod
