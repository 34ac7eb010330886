// Convolution engine (see engine.h).
//
// Restructure relative to the reference (same mathematics):
//  * The reference transforms two full grids per orbital pair and per call
//    (fchi.f90:584-585, fself.f90:239-240).  Here every full-grid G_ab is
//    transformed once (No^2 FFTs) and reused by all chi pairs and by Sigma.
//  * Sigma accumulates V*G products per output (m1,l1) in (r,tau) space
//    before one inverse FFT, instead of one inverse FFT per vertex pair
//    (fself.f90:257): No^2 instead of Nchi^2 inverse transforms.
//  * The tail correction conv[G] - conv[G0] + ref[G0] (fchi.f90:852-892) is
//    combined in (r,tau) space before a single inverse FFT per pair.
#include <cmath>
#include <cstdlib>

#include <array>
#include <map>

#include "engine.h"

namespace gmod {

constexpr double PI_E = 3.141592653589793238462643383279;

static bool sigma_unfused() {
  const char* e = getenv("GMOD_SIGMA_UNFUSED");
  return e && *e && *e != '0';
}
// GMOD_SIGMA_UNGROUPED=1: the Sigma increment's accumulating pass reads the
// three operands of every source instead of the grouped form
static bool sigma_ungrouped() {
  const char* e = getenv("GMOD_SIGMA_UNGROUPED");
  return e && *e && *e != '0';
}
// GMOD_SIGMA_ACCUM=1: the Sigma increment in the grouped accumulating-pass
// form instead of the contraction form
static bool sigma_accum_form() {
  const char* e = getenv("GMOD_SIGMA_ACCUM");
  return e && *e && *e != '0';
}

void OrbPairs::from_olist(const int64_t* olist, int Nchi_, int No_) {
  Nchi = Nchi_;
  No = No_;
  ol.resize(Nchi);
  for (int l = 0; l < Nchi; ++l) {
    ol[l] = {(int)olist[l] - 1, (int)olist[l + Nchi] - 1};
    if (ol[l][0] < 0 || ol[l][0] >= No || ol[l][1] < 0 || ol[l][1] >= No)
      throw Error("olist entry outside [1,Norb]");
  }
  // exchange-symmetry partners chi_{1234} = chi*_{4321} (fchi.f90:421-481)
  partner.assign(Nchi * Nchi, {-1, -1});
  psign.assign(Nchi * Nchi, 1.0);
  std::vector<char> cirr(Nchi * Nchi, 0);
  for (int l1 = 0; l1 < Nchi; ++l1)
    for (int m1 = 0; m1 < Nchi; ++m1) {
      bool found = false;
      for (int l2 = 0; l2 < Nchi && !found; ++l2)
        for (int m2 = 0; m2 < Nchi; ++m2) {
          if (ol[l1][0] == ol[m2][1] && ol[l1][1] == ol[m2][0] && ol[m1][0] == ol[l2][1] &&
              ol[m1][1] == ol[l2][0]) {
            partner[m1 + Nchi * l1] = {m2, l2};
            if (!cirr[m2 + Nchi * l2]) cirr[m1 + Nchi * l1] = 1;
            found = true;
            break;
          }
        }
    }
  irr.clear();
  for (int l1 = 0; l1 < Nchi; ++l1)
    for (int m1 = 0; m1 < Nchi; ++m1)
      if (cirr[m1 + Nchi * l1]) irr.push_back({l1, m1});
  const size_t P = (size_t)Nchi * (Nchi + 1) / 2;
  if (irr.size() > P) irr.resize(P);
}

void OrbPairs::from_olist_soc(const int64_t* olist, int Nchi_, int No_, const int64_t* invs, const int64_t* slist) {
  Nchi = Nchi_;
  No = No_;
  soc = true;
  ol.resize(Nchi);
  inv.resize(No);
  spin.resize(No);
  for (int a = 0; a < No; ++a) {
    inv[a] = (int)invs[a] - 1;
    spin[a] = (double)slist[a];
    if (inv[a] < 0 || inv[a] >= No) throw Error("invs entry outside [1,Norb]");
    if (slist[a] != 1 && slist[a] != -1) throw Error("slist entries must be +1 or -1");
  }
  for (int l = 0; l < Nchi; ++l) {
    ol[l] = {(int)olist[l] - 1, (int)olist[l + Nchi] - 1};
    if (ol[l][0] < 0 || ol[l][0] >= No || ol[l][1] < 0 || ol[l][1] >= No)
      throw Error("olist entry outside [1,Norb]");
  }
  // the pair (l, m) and its spin-flipped exchange partner (m2, l2):
  // (l1, l2, m1, m2) = (-m2_2, -m2_1, -l2_2, -l2_1); the first of the two
  // in (l, m) order is the one computed
  partner.assign(Nchi * Nchi, {-1, -1});
  psign.assign(Nchi * Nchi, 1.0);
  std::vector<char> cirr(Nchi * Nchi, 0);
  for (int l1 = 0; l1 < Nchi; ++l1)
    for (int m1 = 0; m1 < Nchi; ++m1) {
      bool found = false;
      for (int l2 = 0; l2 < Nchi && !found; ++l2)
        for (int m2 = 0; m2 < Nchi; ++m2) {
          if (ol[l1][0] == inv[ol[m2][1]] && ol[l1][1] == inv[ol[m2][0]] && ol[m1][0] == inv[ol[l2][1]] &&
              ol[m1][1] == inv[ol[l2][0]]) {
            partner[m1 + Nchi * l1] = {m2, l2};
            if (!cirr[m2 + Nchi * l2]) cirr[m1 + Nchi * l1] = 1;
            found = true;
            break;
          }
        }
      psign[m1 + Nchi * l1] = spin[ol[l1][0]] * spin[ol[l1][1]] * spin[ol[m1][0]] * spin[ol[m1][1]];
    }
  irr.clear();
  for (int l1 = 0; l1 < Nchi; ++l1)
    for (int m1 = 0; m1 < Nchi; ++m1)
      if (cirr[m1 + Nchi * l1]) irr.push_back({l1, m1});
  const size_t P = (size_t)Nchi * (Nchi + 1) / 2;
  if (irr.size() > P) irr.resize(P);
}

void OrbPairs::from_maps(const int32_t* irr_chi, const int32_t* chi_map) {
  const int P = Nchi * (Nchi + 1) / 2;
  irr.resize(P);
  for (int i = 0; i < P; ++i) irr[i] = {irr_chi[i] - 1, irr_chi[i + P] - 1};
  partner.assign(Nchi * Nchi, {-1, -1});
  if (psign.size() != (size_t)Nchi * Nchi) psign.assign(Nchi * Nchi, 1.0);
  for (int l = 0; l < Nchi; ++l)
    for (int m = 0; m < Nchi; ++m)
      partner[m + Nchi * l] = {chi_map[m + Nchi * l] - 1, chi_map[m + Nchi * l + Nchi * Nchi] - 1};
}

template <typename R>
ConvEngine<R>::ConvEngine(const Mesh& m, const OrbPairs& p) : mesh(m), op(p) {
  N = m.N();
  p4 = &fft_plan<R>(m.Nx, m.Ny, m.Nz, m.M, 15);
  p3 = &fft_plan<R>(m.Nx, m.Ny, m.Nz, m.M, 7);
  work.alloc(N);
  tmp.alloc(N);
}

template <typename R>
void ConvEngine<R>::fft(cpx<R>* g, int sign, bool spatial_only, hipStream_t s) {
  const FFTPlan<R>* p = spatial_only ? p3 : p4;
  R sc = sign > 0 ? (R)(1.0 / (double)p->transformed_points()) : (R)1;
  p->exec(g, tmp.p, sign, sc, s);
}

template <typename R>
void ConvEngine<R>::transform_green(const zd* G, bool g0, hipStream_t s) {
  DBuf<cpx<R>>& dst = g0 ? F0g : Fg;
  const int No = op.No;
  if (dst.n != (size_t)(No * No) * N) dst.alloc((size_t)No * No * N);
  const MeshDev md = mesh_dev(mesh);
  for (int b = 0; b < No; ++b)
    for (int a = 0; a < No; ++a) {
      cpx<R>* g = grid(dst, a, b);
      if (op.soc) {
        const SocGreenLoad<R> ld{G, a, b, op.inv[a], op.inv[b], No, op.spin[a] * op.spin[b], md};
        p4->exec_fused(ld, GridStore<R>{g, (unsigned)mesh.Nkall}, g, tmp.p, -1, (R)1, s);
      } else {
        p4->exec_fused(GreenLoad<R>{G, a, b, No, md}, GridStore<R>{g, (unsigned)mesh.Nkall}, g, tmp.p, -1, (R)1, s);
      }
    }
}

template <typename R>
template <class PL, class ST>
void ConvEngine<R>::product_pass(const PL& pl, const ST& st, hipStream_t s) {
  const R invN = (R)(1.0 / (double)p4->transformed_points());
  switch (pl.nterm) {
    case 1: p4->exec_fused(pl.template fixed<1>(), st, work.p, tmp.p, +1, invN, s, true); break;
    case 2: p4->exec_fused(pl.template fixed<2>(), st, work.p, tmp.p, +1, invN, s, true); break;
    case 3: p4->exec_fused(pl.template fixed<3>(), st, work.p, tmp.p, +1, invN, s, true); break;
    default: p4->exec_fused(pl, st, work.p, tmp.p, +1, invN, s, true); break;
  }
}

// TA, TB written out by an elementwise kernel (the exponentials advance by a
// factor per slot, at full occupancy) and then transformed; above 8 orbitals
// the functions are evaluated inside the transform's first pass instead
template <typename R>
void ConvEngine<R>::tau_grids(cpx<R>* ta, cpx<R>* tb, int a, int b, const double* eig_d, const zd* uni_d, double mu,
                              double mu2, bool diff, double temp, hipStream_t s) {
  const int No = op.No;
  const double beta = 1.0 / temp, dtau = beta / (double)mesh.M;
  const MeshDev md = mesh_dev(mesh);
  const unsigned S = (unsigned)mesh.Nkall;
  if (No <= 8) {
    tau_ref_grids<R>(ta, tb, eig_d, uni_d, a, b, No, mu, mu2, diff, beta, dtau, md, s);
    p3->exec(ta, tmp.p, -1, (R)1, s);
    p3->exec(tb, tmp.p, -1, (R)1, s);
    return;
  }
  for (int which = 0; which < 2; ++which) {
    cpx<R>* t = which == 0 ? ta : tb;
    if (diff) {
      const TauRefLoad<double> cur{eig_d, uni_d, a, b, No, which, mu, beta, dtau, md};
      const TauRefLoad<double> prev{eig_d, uni_d, a, b, No, which, mu2, beta, dtau, md};
      p3->exec_fused(TauRefDiffLoad<R>{cur, prev}, GridStore<R>{t, S}, t, tmp.p, -1, (R)1, s);
    } else {
      p3->exec_fused(TauRefLoad<R>{eig_d, uni_d, a, b, No, which, mu, beta, dtau, md}, GridStore<R>{t, S}, t, tmp.p, -1,
                     (R)1, s);
    }
  }
}

// Tail constants of chi0 at chemical potential mu (fchi.f90:731-772):
// J[l + Nchi*m] = delta_{a13 b13} rhobar_{a42 b42} - delta_{a42 b42} rhobar_{a13 b13}
// with rhobar the BZ-averaged free density matrix, and the sawtooth correction
// sawc(nu_j) of the discrete tau reference.
template <typename R>
void ConvEngine<R>::tail_constants(double temp, const double* eig_h, const zd* uni_h, const int64_t* invk, double mu,
                                   std::vector<zd>& J, DBuf<zd>& sawc_d, hipStream_t s) const {
  const int No = op.No, Nchi = op.Nchi;
  const long Nk = mesh.Nk, Nw = mesh.Nw, M = mesh.M, Nkall = mesh.Nkall;
  const double beta = 1.0 / temp;
  std::vector<zd> rho((size_t)No * No * Nk);
  for (long k = 0; k < Nk; ++k) {
    std::vector<double> f(No);
    for (int n = 0; n < No; ++n) {
      double xi = eig_h[n + No * k] - mu;
      f[n] = xi >= 0 ? std::exp(-beta * xi) / (1.0 + std::exp(-beta * xi)) : 1.0 / (1.0 + std::exp(beta * xi));
    }
    for (int l = 0; l < No; ++l)
      for (int m = 0; m < No; ++m) {
        zd acc(0, 0);
        for (int n = 0; n < No; ++n) acc += uni_h[m + No * (n + No * k)] * conj(uni_h[l + No * (n + No * k)]) * f[n];
        rho[m + No * (l + No * k)] = acc;
      }
  }
  std::vector<zd> rb(No * No, zd(0, 0));
  for (long i = 0; i < Nkall; ++i) {
    long kk = invk[3 * i] - 1;
    if (invk[3 * i + 1] == 0) {
      for (int l = 0; l < No; ++l)
        for (int m = 0; m < No; ++m) rb[m + No * l] += rho[m + No * (l + No * kk)];
    } else if (invk[3 * i + 1] == 1) {
      for (int l = 0; l < No; ++l)
        for (int m = 0; m < No; ++m) rb[m + No * l] += rho[l + No * (m + No * kk)];
    }
  }
  for (auto& v : rb) v = v * (1.0 / (double)Nkall);
  J.assign((size_t)Nchi * Nchi, zd(0, 0));
  for (int l = 0; l < Nchi; ++l)
    for (int m = 0; m < Nchi; ++m) {
      const int a13 = op.ol[l][0], b13 = op.ol[m][0], a42 = op.ol[m][1], b42 = op.ol[l][1];
      zd v(0, 0);
      if (a13 == b13) v += rb[a42 + No * b42];
      if (a42 == b42) v -= rb[a13 + No * b13];
      J[l + Nchi * m] = v;
    }
  if (sawc_d.n == (size_t)Nw) return;
  std::vector<zd> sawc(Nw);
  const double dtau = beta / (double)M;
  for (long j = 1; j <= Nw; ++j) {
    zd v(0, 0);
    if (j > 1) v = zd(0, 1.0 / (2.0 * PI_E * (double)(j - 1) * temp));
    for (long sidx = 0; sidx < M; ++sidx) {
      double ang = 2.0 * PI_E * (double)(j - 1) * (double)sidx / (double)M;
      zd ph(std::cos(ang), std::sin(ang));
      v -= ph * (dtau * (0.5 - (double)sidx / (double)M));
    }
    sawc[j - 1] = v;
  }
  sawc_d.alloc(Nw);
  sawc_d.upload(sawc.data(), Nw, s);
}

template <typename R>
GatherStore<R> ConvEngine<R>::gather_sf(zd* dst, double w, zd J, const zd* sawc, int acc, hipStream_t s) {
  if (!tlive_built) {
    const std::vector<unsigned char> h = p4->fwd_last_live(mesh.kgat_host, &tlive_frac);
    if (!h.empty()) {
      tlive.alloc(h.size());
      tlive.upload(h.data(), h.size(), s);
      GMOD_HIP(hipStreamSynchronize(s));  // h is released at return
    }
    tlive_built = true;
  }
  GatherStore<R> g{dst, w, J, sawc, acc, mesh_dev(mesh)};
  if (tlive.n) {
    g.live = tlive.p;
    g.live_frac = tlive_frac;
  }
  return g;
}

template <typename R>
void ConvEngine<R>::chi0(zd* chi, double temp, bool tail, const double* eig_d, const zd* uni_d,
                         const double* eig_h, const zd* uni_h, const int64_t* invk, double mu, hipStream_t s,
                         bool apply_threshold) {
  const int No = op.No, Nchi = op.Nchi;
  const long Nk = mesh.Nk, Nw = mesh.Nw, M = mesh.M, Nkall = mesh.Nkall;
  const long NkNw = Nk * Nw;
  const double w = temp / (double)Nkall;
  if (tail && sizeof(R) < sizeof(double)) {
    // the tail subtraction stays FP64 when the transforms are FP32
    chi0_split_tail(chi, temp, eig_d, uni_d, eig_h, uni_h, invk, mu, s);
    if (apply_threshold) threshold(chi, NkNw * Nchi * Nchi, 1e-9, s);
    return;
  }
  std::vector<zd> rb;
  DBuf<zd> sawc_d;
  if (tail) {
    DBuf<zd> G0((size_t)NkNw * No * No);
    green0(G0.p, eig_d, uni_d, mu, temp, Nk, (int)Nw, No, s);
    transform_green(G0.p, true, s);
    if (TA.n != (size_t)(No * No) * N) { TA.alloc((size_t)No * No * N); TB.alloc((size_t)No * No * N); }
    for (int b = 0; b < No; ++b)
      for (int a = 0; a < No; ++a) tau_grids(grid(TA, a, b), grid(TB, a, b), a, b, eig_d, uni_d, mu, mu, false, temp, s);
    std::vector<zd> Jall;
    tail_constants(temp, eig_h, uni_h, invk, mu, Jall, sawc_d, s);
    rb = Jall;
  }
  const double wref = 1.0 / (temp * (double)Nkall);
  const MeshDev md = mesh_dev(mesh);
  const R invN = (R)(1.0 / (double)p4->transformed_points());
  if (!mesh.complete) GMOD_HIP(hipMemsetAsync(chi, 0, NkNw * Nchi * Nchi * sizeof(zd), s));
  // without the tail every product is G13(r,t) G42(-r,-t): the mirrored G
  // grids are formed once (coalesced) and then read in place.  The inverse
  // transforms run spatial-first, so a (-r,-t) operand that is not mirrored
  // (the tail terms) is read as whole reversed planes.
  if (!tail) {
    if (Mg.n != Fg.n) Mg.alloc(Fg.n);
    for (int b = 0; b < No; ++b)
      for (int a = 0; a < No; ++a) mirror_grid(grid(Mg, a, b), grid(Fg, a, b), mesh.Nx, mesh.Ny, mesh.Nz, mesh.M, s);
  }
  for (auto& lm : op.irr) {
    const int l = lm[0], m = lm[1];
    const int a13 = op.ol[l][0], b13 = op.ol[m][0], a42 = op.ol[m][1], b42 = op.ol[l][1];
    auto pr = op.partner[m + Nchi * l];
    // A self-partner slice (partner == itself) is conjugated in place after each
    // of the reference's three accumulation passes (fchi.f90:610, 846); the net
    // result is conj(conv[G] + ref) - conv[G0].  It equals conj of the plain sum
    // whenever TRS holds; keep the reference's order so results match it on any
    // input.
    const bool self_tail = tail && pr[0] == m && pr[1] == l;
    ProductLoad<R> pl{};
    pl.m = md;
    if (!tail) {
      pl.nterm = 1;
      pl.A[0] = grid(Fg, a13, b13); pl.B[0] = grid(Mg, a42, b42); pl.coef[0] = (R)-1; pl.revt[0] = 1;
      pl.bmir = 1;
    } else {
      int t = 0;
      pl.A[t] = grid(Fg, a13, b13); pl.B[t] = grid(Fg, a42, b42); pl.coef[t] = (R)(-w); pl.revt[t++] = 1;
      if (!self_tail) {
        pl.A[t] = grid(F0g, a13, b13); pl.B[t] = grid(F0g, a42, b42); pl.coef[t] = (R)w; pl.revt[t++] = 1;
      }
      pl.A[t] = grid(TA, a13, b13); pl.B[t] = grid(TB, a42, b42); pl.coef[t] = (R)(-wref); pl.revt[t++] = 0;
      pl.nterm = t;
    }
    const zd J = tail ? rb[l + Nchi * m] : zd(0, 0);
    zd* dst = chi + NkNw * (m + (long)Nchi * l);
    GatherStore<R> gst = gather_sf(dst, tail ? 1.0 : w, J, tail ? sawc_d.p : nullptr, 0, s);
    // the partner copy written by the gather (a complete mesh: every slot of
    // the slice is written)
    const bool fuse_pr = !self_tail && pr[0] >= 0 && mesh.complete;
    if (fuse_pr) {
      gst.pdst = chi + NkNw * (pr[0] + (long)Nchi * pr[1]);
      gst.psign = op.psign[m + Nchi * l];
    }
    product_pass(pl, gst, s);
    if (self_tail) {
      conj_copy(dst, dst, NkNw, s);
      ProductLoad<R> p0{};
      p0.m = md;
      p0.nterm = 1;
      p0.A[0] = grid(F0g, a13, b13); p0.B[0] = grid(F0g, a42, b42); p0.coef[0] = (R)w; p0.revt[0] = 1;
      p4->exec_fused(p0.template fixed<1>(), gather_sf(dst, 1.0, zd(0, 0), nullptr, 1, s), work.p, tmp.p, +1, invN, s,
                     true);
      continue;
    }
    // the spinful partner carries the sign s_l1 s_l2 s_m1 s_m2 (fchi0withsoc.f90:107)
    if (pr[0] >= 0 && !fuse_pr)
      conj_copy(chi + NkNw * (pr[0] + (long)Nchi * pr[1]), dst, NkNw, op.psign[m + Nchi * l], s);
  }
  if (apply_threshold) threshold(chi, NkNw * Nchi * Nchi, 1e-9, s);
}

template <typename R>
void ConvEngine<R>::chi0_ff(zd* chi, const zd* F, const double* prt, bool singlet, double temp, hipStream_t s) {
  const int No = op.No, Nchi = op.Nchi;
  const long NkNw = mesh.Nk * (long)mesh.Nw;
  const double w = temp / (double)mesh.Nkall;
  // F(-k) = sgn F^T(k) and the gap's frequency parity dprt: both +1 for a
  // singlet, -1 for a triplet (ffeliash.f90:541-547)
  const double sgn = singlet ? 1.0 : -1.0, dprt = sgn;
  if (Lff.n != (size_t)2 * N) Lff.alloc((size_t)2 * N);
  cpx<R>* A = Lff.p;
  cpx<R>* B = Lff.p + N;
  const MeshDev md = mesh_dev(mesh);
  const unsigned S = (unsigned)mesh.Nkall;
  auto off = [&](int a, int b) { return NkNw * (long)(a + No * b); };
  if (!mesh.complete) GMOD_HIP(hipMemsetAsync(chi, 0, NkNw * Nchi * Nchi * sizeof(zd), s));
  for (auto& lm : op.irr) {
    const int l = lm[0], m = lm[1];
    const int l1 = op.ol[l][0], l2 = op.ol[l][1], m1 = op.ol[m][0], m2 = op.ol[m][1];
    const FFLoad<R> f14{F, off(m2, l1), off(l1, m2), off(m1, l1), off(l1, m2), prt[l1] * prt[m2] * sgn * dprt, 1, md};
    const FFLoad<R> f23{F, off(l2, m1), off(m1, l2), off(l2, m2), off(l2, m1), prt[m1] * prt[l2] * sgn * dprt, 0, md};
    p4->exec_fused(f14, GridStore<R>{A, S}, A, tmp.p, -1, (R)1, s);
    p4->exec_fused(f23, GridStore<R>{B, S}, B, tmp.p, -1, (R)1, s);
    ProductLoad<R> pl{};
    pl.m = md;
    pl.add(A, B, (R)-1, 1);
    zd* dst = chi + NkNw * (m + (long)Nchi * l);
    product_pass(pl, gather_sf(dst, w, zd(0, 0), nullptr, 0, s), s);
    auto pr = op.partner[m + Nchi * l];
    if (pr[0] >= 0) conj_copy(chi + NkNw * (pr[0] + (long)Nchi * pr[1]), dst, NkNw, s);
  }
  threshold(chi, NkNw * Nchi * Nchi, 1e-9, s);
}

template <typename R>
void ConvEngine<R>::tau_reference(zd* chi, double temp, const double* eig_d, const zd* uni_d, double mu,
                                  const std::vector<zd>& J, const zd* sawc_d, hipStream_t s) {
  const int No = op.No, Nchi = op.Nchi;
  const long NkNw = mesh.Nk * (long)mesh.Nw;
  if (TA.n != (size_t)(No * No) * N) { TA.alloc((size_t)No * No * N); TB.alloc((size_t)No * No * N); }
  const MeshDev md = mesh_dev(mesh);
  for (int b = 0; b < No; ++b)
    for (int a = 0; a < No; ++a) tau_grids(grid(TA, a, b), grid(TB, a, b), a, b, eig_d, uni_d, mu, mu, false, temp, s);
  const double wref = 1.0 / (temp * (double)mesh.Nkall);
  const R invN = (R)(1.0 / (double)p4->transformed_points());
  for (auto& lm : op.irr) {
    const int l = lm[0], m = lm[1];
    const int a13 = op.ol[l][0], b13 = op.ol[m][0], a42 = op.ol[m][1], b42 = op.ol[l][1];
    ProductLoad<R> pl{};
    pl.m = md;
    pl.add(grid(TA, a13, b13), grid(TB, a42, b42), (R)(-wref), 0);
    zd* dst = chi + NkNw * (m + (long)Nchi * l);
    p4->exec_fused(pl.template fixed<1>(), gather_sf(dst, 1.0, J[l + Nchi * m], sawc_d, 1, s), work.p, tmp.p, +1,
                   invN, s, true);
  }
}

template <typename R>
void ConvEngine<R>::chi0_split_tail(zd* chi, double temp, const double* eig_d, const zd* uni_d, const double* eig_h,
                                    const zd* uni_h, const int64_t* invk, double mu, hipStream_t s) {
  const int No = op.No, Nchi = op.Nchi;
  const long Nk = mesh.Nk, Nw = mesh.Nw, NkNw = Nk * Nw;
  const double w = temp / (double)mesh.Nkall;
  {
    DBuf<zd> G0((size_t)NkNw * No * No);
    green0(G0.p, eig_d, uni_d, mu, temp, Nk, (int)Nw, No, s);
    transform_green(G0.p, true, s);
    GMOD_HIP(hipStreamSynchronize(s));  // G0 is released at the end of the scope
  }
  std::vector<zd> J;
  DBuf<zd> sawc_d;
  tail_constants(temp, eig_h, uni_h, invk, mu, J, sawc_d, s);
  const MeshDev md = mesh_dev(mesh);
  const R invN = (R)(1.0 / (double)p4->transformed_points());
  if (!mesh.complete) GMOD_HIP(hipMemsetAsync(chi, 0, NkNw * Nchi * Nchi * sizeof(zd), s));
  auto is_self = [&](int l, int m) {
    auto pr = op.partner[m + Nchi * l];
    return pr[0] == m && pr[1] == l;
  };
  // 1) -w conv[G] + w conv[G0] at grid precision (a self-partner pair: -w conv[G] only)
  for (auto& lm : op.irr) {
    const int l = lm[0], m = lm[1];
    const int a13 = op.ol[l][0], b13 = op.ol[m][0], a42 = op.ol[m][1], b42 = op.ol[l][1];
    ProductLoad<R> pl{};
    pl.m = md;
    pl.add(grid(Fg, a13, b13), grid(Fg, a42, b42), (R)(-w), 1);
    if (!is_self(l, m)) pl.add(grid(F0g, a13, b13), grid(F0g, a42, b42), (R)w, 1);
    zd* dst = chi + NkNw * (m + (long)Nchi * l);
    const GatherStore<R> gst = gather_sf(dst, 1.0, zd(0, 0), nullptr, 0, s);
    product_pass(pl, gst, s);
  }
  // 2) + the tau reference and the sawtooth term, FP64
  {
    ConvEngine<double> e64(mesh, op);
    e64.tau_reference(chi, temp, eig_d, uni_d, mu, J, sawc_d.p, s);
    GMOD_HIP(hipStreamSynchronize(s));  // e64's grids are released at the end of the scope
  }
  // 3) the reference's order for a self-partner pair, conj(conv[G] + ref) -
  //    conv[G0] (fchi.f90:610, 846, 886); partner copies of the others
  for (auto& lm : op.irr) {
    const int l = lm[0], m = lm[1];
    const int a13 = op.ol[l][0], b13 = op.ol[m][0];
    const int a42 = op.ol[m][1], b42 = op.ol[l][1];
    zd* dst = chi + NkNw * (m + (long)Nchi * l);
    if (is_self(l, m)) {
      conj_copy(dst, dst, NkNw, s);
      ProductLoad<R> p0{};
      p0.m = md;
      p0.add(grid(F0g, a13, b13), grid(F0g, a42, b42), (R)w, 1);
      p4->exec_fused(p0.template fixed<1>(), gather_sf(dst, 1.0, zd(0, 0), nullptr, 1, s), work.p, tmp.p, +1, invN,
                     s, true);
      continue;
    }
    auto pr = op.partner[m + Nchi * l];
    if (pr[0] >= 0) conj_copy(chi + NkNw * (pr[0] + (long)Nchi * pr[1]), dst, NkNw, s);
  }
}

template <typename R>
bool ConvEngine<R>::grouped_table(std::vector<int>& ytab) const {
  const int No = op.No, Nchi = op.Nchi, NAB = No * No;
  ytab.assign((size_t)NAB * NAB, -1);
  for (int l = 0; l < Nchi; ++l)
    for (int m = 0; m < Nchi; ++m) {
      const int o = op.ol[m][0] + No * op.ol[l][0], ab = op.ol[m][1] + No * op.ol[l][1];
      int& slot = ytab[(size_t)o * NAB + ab];
      if (slot >= 0) return false;
      slot = m + Nchi * l;
    }
  for (int v : ytab)
    if (v < 0) return false;
  return true;
}

template <typename R>
std::vector<std::array<int, 2>> ConvEngine<R>::share_heads(std::vector<int>& entry) const {
  const int Nchi = op.Nchi;
  std::vector<std::array<int, 2>> heads;
  entry.assign((size_t)Nchi * Nchi, -1);
  for (int l = 0; l < Nchi; ++l)
    for (int m = 0; m < Nchi; ++m) {
      const int t = l + Nchi * m;  // the transposed pair
      if (entry[t] >= 0) {
        entry[m + Nchi * l] = (entry[t] & ~1) + 1;
      } else {
        entry[m + Nchi * l] = 2 * (int)heads.size();
        heads.push_back({m, l});
      }
    }
  return heads;
}

template <typename R>
void ConvEngine<R>::head_offsets(std::vector<long>& offs, int m, int l) const {
  const long NC = mesh.Nk * (long)mesh.Nw;
  offs.push_back(NC * (m + (long)op.Nchi * l));
  offs.push_back(NC * (l + (long)op.Nchi * m));
}

template <typename R>
template <typename Sv>
void ConvEngine<R>::slot0_corr(cpx<R>* dh, const cpx<Sv>* V, const long* offs_d, int nh, hipStream_t s) {
  if (nh <= 0) return;
  const unsigned S = (unsigned)mesh.Nkall;
  const size_t n = (size_t)nh * S;
  if (dh_tmp.n < n) dh_tmp.alloc(n);
  const MeshDev md = mesh_dev(mesh);
  const unsigned grid = (unsigned)std::min<long>(((long)n + 255) / 256, 256L * 64);
  hipLaunchKernelGGL((k_slot0_diff<R, Sv>), dim3(grid), dim3(256), 0, s, dh, V, offs_d, nh, S, md);
  GMOD_CHECK_LAUNCH();
  fft_plan<R>(mesh.Nx, mesh.Ny, mesh.Nz, nh, 7).exec(dh, dh_tmp.p, -1, (R)1, s);
}

template <typename R>
bool ConvEngine<R>::contract_form() const {
  if (sigma_unfused() || sigma_ungrouped() || sigma_accum_form()) return false;
  if (op.No > 4 || !p4->reversal_ok()) return false;
  std::vector<int> ytab;
  return grouped_table(ytab);
}

template <typename R>
typename ConvEngine<R>::SigmaPlan ConvEngine<R>::plan_sigma(bool share) const {
  const int No = op.No, Nchi = op.Nchi;
  SigmaPlan P;
  std::vector<char> done((size_t)No * No, 0);
  for (int l1 = 0; l1 < No; ++l1)
    for (int m1 = 0; m1 < No; ++m1) {
      const int o = m1 + No * l1;
      if (done[o]) continue;
      std::vector<int> batch{o};
      if (share && m1 != l1) batch.push_back(l1 + No * m1);
      std::map<std::pair<int, int>, int> slot;  // (m, l) -> head slot of this batch
      for (int ob : batch) {
        const int bm1 = ob % No, bl1 = ob / No;
        std::vector<std::array<int, 3>> srcl, heads;
        for (int l = 0; l < Nchi; ++l)
          for (int m = 0; m < Nchi; ++m) {
            if (op.ol[m][0] != bm1 || op.ol[l][0] != bl1) continue;
            int e;
            auto it = slot.find({m, l});
            if (it != slot.end()) {
              e = 2 * it->second;
            } else if (share && (it = slot.find({l, m})) != slot.end()) {
              e = 2 * it->second + 1;
            } else {
              const int sl = (int)slot.size();
              slot[{m, l}] = sl;
              heads.push_back({m, l, sl});
              e = 2 * sl;
            }
            srcl.push_back({m, l, e});
          }
        P.outs.push_back(ob);
        P.src.push_back(srcl);
        P.heads.push_back(heads);
        done[ob] = 1;
      }
      P.nmax = std::max(P.nmax, (int)slot.size());
    }
  return P;
}

template <typename R>
void ConvEngine<R>::sigma(zd* sig, const zd* V, const double* S_h, const double* C_h, double temp, hipStream_t s,
                          bool apply_threshold) {
  // the nu -> infinity slot of V_sigma is 1.5 S - 0.5 C (fself.f90:208)
  const int Nchi = op.Nchi;
  std::vector<double> stat((size_t)Nchi * Nchi);
  for (size_t e = 0; e < stat.size(); ++e) stat[e] = 1.5 * S_h[e] - 0.5 * C_h[e];
  sigma(sig, V, stat.data(), temp, s, apply_threshold);
}

template <typename R>
void ConvEngine<R>::sigma(zd* sig, const zd* V, const double* stat_h, double temp, hipStream_t s,
                          bool apply_threshold) {
  const int No = op.No, Nchi = op.Nchi;
  const long NkNw = mesh.Nk * (long)mesh.Nw;
  if (acc.n != (size_t)N) acc.alloc(N);
  const MeshDev md = mesh_dev(mesh);
  const R invN = (R)(1.0 / (double)p4->transformed_points());
  const long NC = NkNw;
  // the accumulating last pass sums the pairs of one output in registers
  // (GMOD_SIGMA_UNFUSED=1: one pass per pair, accumulating through memory)
  const bool fused = p4->accum_ok(-1) && !sigma_unfused();
  auto gather_out = [&](int m1, int l1, bool any) {
    zd* dst = sig + NkNw * (m1 + (long)No * l1);
    if (!any) {
      GMOD_HIP(hipMemsetAsync(dst, 0, NkNw * sizeof(zd), s));
      return;
    }
    if (!mesh.complete) GMOD_HIP(hipMemsetAsync(dst, 0, NkNw * sizeof(zd), s));
    p4->exec_fused(GridLoad<R>{acc.p, (unsigned)mesh.Nkall},
                   GatherStore<R>{dst, temp / (double)mesh.Nkall, zd(0, 0), nullptr, 0, md}, acc.p, tmp.p, +1, invN, s);
  };
  if (fused) {
    // transposed outputs share their sources' head passes when the static
    // slot is symmetric (the conjugate-reversal relation of the sources, with
    // the slot-0 correction)
    bool share = p4->accum_last_is_t(-1) && p4->reversal_ok();
    for (int l = 0; l < Nchi && share; ++l)
      for (int m = 0; m < Nchi; ++m)
        if (stat_h[m + Nchi * l] != stat_h[l + Nchi * m]) { share = false; break; }
    const SigmaPlan P = plan_sigma(share);
    accum_tab_h.clear();
    std::vector<int> smap;
    for (size_t i = 0; i < P.outs.size(); ++i)
      for (auto& e : P.src[i]) {
        accum_tab_h.push_back(grid(Fg, op.ol[e[0]][1], op.ol[e[1]][1]));
        smap.push_back(e[2]);
      }
    if (srcs.n != (size_t)P.nmax * N) srcs.alloc((size_t)P.nmax * N);
    accum_tab.alloc(accum_tab_h.size());
    accum_tab.upload(accum_tab_h.data(), accum_tab_h.size(), s);
    srcmap_d.alloc(smap.size());
    srcmap_d.upload(smap.data(), smap.size(), s);
    std::vector<long> hoffs;
    for (auto& hs : P.heads)
      for (auto& h : hs) head_offsets(hoffs, h[0], h[1]);
    if (dh_offs.n < hoffs.size()) dh_offs.alloc(hoffs.size());
    if (!hoffs.empty()) dh_offs.upload(hoffs.data(), hoffs.size(), s);
    if (dhD.n < (size_t)P.nmax * mesh.Nkall) dhD.alloc((size_t)P.nmax * mesh.Nkall);
    GMOD_HIP(hipStreamSynchronize(s));
    long off = 0, hoff = 0;
    for (size_t i = 0; i < P.outs.size(); ++i) {
      for (auto& h : P.heads[i]) {
        const int m = h[0], l = h[1];
        const zd stat(stat_h[m + Nchi * l], 0);
        p4->exec_head(BosonLoad<R>{V, NC * (m + (long)Nchi * l), NC * (l + (long)Nchi * m), stat, md},
                      srcs.p + (long)h[2] * N, tmp.p, -1, s);
      }
      // the heads of one output take consecutive slots
      if (!P.heads[i].empty())
        slot0_corr(dhD.p + (size_t)P.heads[i][0][2] * mesh.Nkall, V, dh_offs.p + 2 * hoff, (int)P.heads[i].size(),
                   s);
      hoff += (long)P.heads[i].size();
      const int nsrc = (int)P.src[i].size();
      if (nsrc > 0)
        p4->exec_accum_last(srcs.p, N, nsrc, MulAccAccum<R>{accum_tab.p + off, acc.p, (unsigned)mesh.Nkall}, -1,
                            (R)1, s, srcmap_d.p + off, dhD.p);
      off += nsrc;
      gather_out(P.outs[i] % No, P.outs[i] / No, nsrc > 0);
    }
  } else {
    for (int l1 = 0; l1 < No; ++l1)
      for (int m1 = 0; m1 < No; ++m1) {
        bool first = true;
        for (int l = 0; l < Nchi; ++l)
          for (int m = 0; m < Nchi; ++m) {
            if (op.ol[m][0] != m1 || op.ol[l][0] != l1) continue;
            const zd stat(stat_h[m + Nchi * l], 0);
            BosonLoad<R> bl{V, NC * (m + (long)Nchi * l), NC * (l + (long)Nchi * m), stat, md};
            p4->exec_fused(bl, MulAccStore<R>{acc.p, grid(Fg, op.ol[m][1], op.ol[l][1]), first ? 1 : 0,
                                              (unsigned)mesh.Nkall},
                           work.p, tmp.p, -1, (R)1, s);
            first = false;
          }
        gather_out(m1, l1, !first);
      }
  }
  if (apply_threshold) threshold(sig, NkNw * No * No, 1e-9, s);
}

template <typename R>
void ConvEngine<R>::prepare_gap(const zd* Vd, const double* S_h, const double* C_h, bool singlet, hipStream_t s) {
  const int Nchi = op.Nchi;
  const double sgn = singlet ? 1.0 : -1.0;
  gap_pairs.clear();
  for (int l = 0; l < Nchi; ++l)
    for (int m = 0; m < Nchi; ++m)
      if (op.ol[l][0] >= op.ol[m][1]) gap_pairs.push_back({l, m});
  S_cache.assign(S_h, S_h + Nchi * Nchi);
  C_cache.assign(C_h, C_h + Nchi * Nchi);
  Vd_dev = Vd;
  size_t free_b = 0, total_b = 0;
  GMOD_HIP(hipMemGetInfo(&free_b, &total_b));
  const double need = (double)gap_pairs.size() * N * sizeof(cpx<R>);
  vft_cached = need < 0.5 * (double)free_b;
  if (!vft_cached) return;
  Vft.alloc(gap_pairs.size() * (size_t)N);
  const MeshDev md = mesh_dev(mesh);
  const long NC = mesh.Nk * (long)mesh.Nw;
  for (size_t q = 0; q < gap_pairs.size(); ++q) {
    int l = gap_pairs[q][0], m = gap_pairs[q][1];
    zd stat(0.5 * (S_h[m + Nchi * l] + sgn * C_h[m + Nchi * l]), 0);
    cpx<R>* g = Vft.p + q * N;
    p4->exec_fused(BosonLoad<R>{Vd, NC * (m + (long)Nchi * l), NC * (l + (long)Nchi * m), stat, md}, GridStore<R>{g, (unsigned)mesh.Nkall},
                   g, tmp.p, -1, (R)1, s);
  }
}

template <typename R>
void ConvEngine<R>::gap(zd* nd, const zd* F, const double* prt, bool singlet, hipStream_t s) {
  const int No = op.No, Nchi = op.Nchi;
  const long NkNw = mesh.Nk * (long)mesh.Nw;
  const double sgn = singlet ? 1.0 : -1.0, dprt = sgn;
  if (Ff.n != (size_t)(No * No) * N) Ff.alloc((size_t)No * No * N);
  if (acc.n != (size_t)N) acc.alloc(N);
  const MeshDev md = mesh_dev(mesh);
  const R invN = (R)(1.0 / (double)p4->transformed_points());
  for (int b = 0; b < No; ++b)
    for (int a = 0; a < No; ++a) {
      cpx<R>* g = grid(Ff, a, b);
      AnomLoad<R> al{F, NkNw * (a + (long)No * b), NkNw * (b + (long)No * a), prt[a] * prt[b], sgn, dprt, md};
      p4->exec_fused(al, GridStore<R>{g, (unsigned)mesh.Nkall}, g, tmp.p, -1, (R)1, s);
    }
  GMOD_HIP(hipMemsetAsync(nd, 0, NkNw * No * No * sizeof(zd), s));
  for (int m2 = 0; m2 < No; ++m2)
    for (int l1 = 0; l1 < No; ++l1) {
      bool first = true;
      for (size_t q = 0; q < gap_pairs.size(); ++q) {
        int l = gap_pairs[q][0], m = gap_pairs[q][1];
        if (op.ol[l][0] != l1 || op.ol[m][1] != m2) continue;
        const cpx<R>* fg = grid(Ff, op.ol[l][1], op.ol[m][0]);
        if (vft_cached) {
          mul_acc<R>(acc.p, Vft.p + q * N, fg, first, N, s);
        } else {
          zd stat(0.5 * (S_cache[m + Nchi * l] + sgn * C_cache[m + Nchi * l]), 0);
          BosonLoad<R> bl{Vd_dev, NkNw * (m + (long)Nchi * l), NkNw * (l + (long)Nchi * m), stat, md};
          p4->exec_fused(bl, MulAccStore<R>{acc.p, fg, first ? 1 : 0, (unsigned)mesh.Nkall}, work.p, tmp.p, -1, (R)1, s);
        }
        first = false;
      }
      if (first) continue;
      p4->exec_fused(GridLoad<R>{acc.p, (unsigned)mesh.Nkall}, GatherStore<R>{nd + NkNw * (l1 + (long)No * m2), 1.0, zd(0, 0), nullptr, 0, md},
                     acc.p, tmp.p, +1, invN, s);
    }
  hermitize_lower(nd, mesh.Nk, mesh.Nw, No, s);
}

template <typename R>
void ConvEngine<R>::prepare_gap_soc(const zd* Vd, const double* Vmat_h, const int64_t* invschi, long gap_sym,
                                    hipStream_t s) {
  if (!op.soc) throw Error("prepare_gap_soc: the orbital pairs carry no spin");
  if (!mesh.nfull_ok) throw Error("lin_eliash_soc: invk(3,:) must give the -k point of every k");
  const int Nchi = op.Nchi;
  const long NC = mesh.Nk * (long)mesh.Nw;
  const MeshDev md = mesh_dev(mesh);
  // the pairs of feliashwithsoc.f90:538-540: every spin for a triplet gap,
  // opposite spins of (l1, m2) for a singlet; of those the up-l1 ones (or,
  // for the parity-mixed gap_sym = -10, every opposite-spin one), and of the
  // opposite-spin ones the upper triangle l1 >= -m2
  gap_pairs.clear();
  soc_vload.clear();
  for (int l = 0; l < Nchi; ++l)
    for (int m = 0; m < Nchi; ++m) {
      const int l1 = op.ol[l][0], m2 = op.ol[m][1];
      const double ss = op.spin[l1] * op.spin[m2];
      if (!(gap_sym < 0 || (gap_sym > 0 && ss < 0))) continue;
      if (!((gap_sym == -10 && ss < 0) || op.spin[l1] > 0)) continue;
      if (!(op.spin[m2] > 0 || l1 >= op.inv[m2])) continue;
      const long li = invschi[l] - 1, mi = invschi[m] - 1;
      if (li < 0 || li >= Nchi || mi < 0 || mi >= Nchi) throw Error("lin_eliash_soc: invschi entry outside [1,Nchi]");
      gap_pairs.push_back({l, m});
      soc_vload.push_back(SocBosonLoad<R>{Vd, NC * (m + (long)Nchi * l), NC * (l + (long)Nchi * m),
                                          NC * (li + (long)Nchi * mi), NC * (mi + (long)Nchi * li),
                                          op.psign[m + Nchi * l], zd(-Vmat_h[m + Nchi * l], 0), md});
    }
  size_t free_b = 0, total_b = 0;
  GMOD_HIP(hipMemGetInfo(&free_b, &total_b));
  const double need = (double)gap_pairs.size() * N * sizeof(cpx<R>);
  vft_cached = need < 0.5 * (double)free_b;
  if (!vft_cached) return;
  Vft.alloc(gap_pairs.size() * (size_t)N);
  for (size_t q = 0; q < gap_pairs.size(); ++q) {
    cpx<R>* g = Vft.p + q * N;
    p4->exec_fused(soc_vload[q], GridStore<R>{g, (unsigned)mesh.Nkall}, g, tmp.p, -1, (R)1, s);
  }
}

template <typename R>
void ConvEngine<R>::gap_soc(zd* nd, const zd* F, hipStream_t s) {
  const int No = op.No;
  const long NW = mesh.Nkall * (long)mesh.Nw;
  if (Ff.n != (size_t)(No * No) * N) Ff.alloc((size_t)No * No * N);
  if (acc.n != (size_t)N) acc.alloc(N);
  const MeshDev md = mesh_dev(mesh);
  const R invN = (R)(1.0 / (double)p4->transformed_points());
  std::vector<char> need((size_t)No * No, 0);
  for (auto& lm : gap_pairs) need[op.ol[lm[0]][1] + No * op.ol[lm[1]][0]] = 1;
  for (int b = 0; b < No; ++b)
    for (int a = 0; a < No; ++a) {
      if (!need[a + No * b]) continue;
      cpx<R>* g = grid(Ff, a, b);
      p4->exec_fused(FullAnomLoad<R>{F, NW * (a + (long)No * b), NW * (b + (long)No * a), md},
                     GridStore<R>{g, (unsigned)mesh.Nkall}, g, tmp.p, -1, (R)1, s);
    }
  GMOD_HIP(hipMemsetAsync(nd, 0, NW * No * No * sizeof(zd), s));
  for (int m2 = 0; m2 < No; ++m2)
    for (int l1 = 0; l1 < No; ++l1) {
      bool first = true;
      for (size_t q = 0; q < gap_pairs.size(); ++q) {
        const int l = gap_pairs[q][0], m = gap_pairs[q][1];
        if (op.ol[l][0] != l1 || op.ol[m][1] != m2) continue;
        const cpx<R>* fg = grid(Ff, op.ol[l][1], op.ol[m][0]);
        if (vft_cached)
          mul_acc<R>(acc.p, Vft.p + q * N, fg, first, N, s);
        else
          p4->exec_fused(soc_vload[q], MulAccStore<R>{acc.p, fg, first ? 1 : 0, (unsigned)mesh.Nkall}, work.p,
                         tmp.p, -1, (R)1, s);
        first = false;
      }
      if (first) continue;
      p4->exec_fused(GridLoad<R>{acc.p, (unsigned)mesh.Nkall}, FullGatherStore<R>{nd + NW * (l1 + (long)No * m2), 0, md},
                     acc.p, tmp.p, +1, invN, s);
    }
}

// ====================================================================== increments
// Rolling increments for the mixed-precision FLEX loop.  Between iterations
// n-1 and n the bubble and the self-energy change by
//   dB = B(Gr + dG) - B(Gr)  = dG13 (Gr42 + dG42) + Gr13 dG42      (per product)
//   dS = S(Vr + dV, Gr + dG) - S(Vr, Gr) = dV (Gr + dG) + Vr dG
// with Gr, Vr the previous iterate.  dG and dV are formed in FP64 in k-space
// and only the increments are transformed at the grid precision, so the
// rounding error of an update is proportional to the update itself and
// vanishes as the loop converges; the transforms of the previous iterate
// (Fr, F0r, TAr, TBr, FVr) advance by the transformed increments.
template <typename R>
void ConvEngine<R>::set_reference(const zd* Gref, const zd* G0ref, bool tail, const double* eig_d, const zd* uni_d,
                                  double mu_ref, double temp, hipStream_t s) {
  const int No = op.No;
  const size_t nb = (size_t)No * No * N;
  const MeshDev md = mesh_dev(mesh);
  const unsigned S = (unsigned)mesh.Nkall;
  auto ensure = [&](DBuf<cpx<R>>& b) { if (b.n != nb) b.alloc(nb); };
  ensure(Fr); ensure(Fd);
  for (int b = 0; b < No; ++b)
    for (int a = 0; a < No; ++a) {
      cpx<R>* g = grid(Fr, a, b);
      p4->exec_fused(GreenLoad<R>{Gref, a, b, No, md}, GridStore<R>{g, S}, g, tmp.p, -1, (R)1, s);
    }
  if (!tail) return;
  ensure(F0r); ensure(F0d); ensure(TAr); ensure(TBr); ensure(TAd); ensure(TBd);
  for (int b = 0; b < No; ++b)
    for (int a = 0; a < No; ++a) {
      cpx<R>* g = grid(F0r, a, b);
      p4->exec_fused(GreenLoad<R>{G0ref, a, b, No, md}, GridStore<R>{g, S}, g, tmp.p, -1, (R)1, s);
      tau_grids(grid(TAr, a, b), grid(TBr, a, b), a, b, eig_d, uni_d, mu_ref, mu_ref, false, temp, s);
    }
}

template <typename R>
void ConvEngine<R>::set_vertex_reference(const zd* Vref, const double* S_h, const double* C_h, hipStream_t s) {
  const int Nchi = op.Nchi;
  const long NC = mesh.Nk * (long)mesh.Nw;
  const MeshDev md = mesh_dev(mesh);
  const unsigned S = (unsigned)mesh.Nkall;
  // the static slot 3/2 S - 1/2 C must be symmetric for the transposed pair
  // to be the conjugate reversal
  bool share = contract_form();
  for (int l = 0; l < Nchi && share; ++l)
    for (int m = 0; m < Nchi; ++m)
      if (S_h[m + Nchi * l] != S_h[l + Nchi * m] || C_h[m + Nchi * l] != C_h[l + Nchi * m]) {
        share = false;
        break;
      }
  std::vector<std::array<int, 2>> heads;
  if (share) {
    heads = share_heads(fv_entry);
  } else {
    fv_entry.assign((size_t)Nchi * Nchi, 0);
    for (int l = 0; l < Nchi; ++l)
      for (int m = 0; m < Nchi; ++m) {
        fv_entry[m + Nchi * l] = 2 * (m + Nchi * l);
        heads.push_back({m, l});
      }
  }
  fv_shared = share;
  const size_t nb = heads.size() * (size_t)N;
  if (FVr.n != nb) FVr.alloc(nb);
  if (share) {
    std::vector<long> hoffs;
    for (auto& h : heads) head_offsets(hoffs, h[0], h[1]);
    if (dhF_offs.n != hoffs.size()) dhF_offs.alloc(hoffs.size());
    dhF_offs.upload(hoffs.data(), hoffs.size(), s);
    if (dhF.n != heads.size() * (size_t)mesh.Nkall) dhF.alloc(heads.size() * (size_t)mesh.Nkall);
    slot0_corr(dhF.p, Vref, dhF_offs.p, (int)heads.size(), s);
    GMOD_HIP(hipStreamSynchronize(s));  // hoffs is released here
  }
  for (size_t h = 0; h < heads.size(); ++h) {
    const int m = heads[h][0], l = heads[h][1];
    zd stat(1.5 * S_h[m + Nchi * l] - 0.5 * C_h[m + Nchi * l], 0);
    cpx<R>* g = FVr.p + h * N;
    p4->exec_fused(BosonLoad<R>{Vref, NC * (m + (long)Nchi * l), NC * (l + (long)Nchi * m), stat, md},
                   GridStore<R>{g, S}, g, tmp.p, -1, (R)1, s);
  }
}

template <typename R>
void ConvEngine<R>::chi0_increment(zd* chi, const zd* chi_base, const zd* G, const zd* Gref, const zd* G0,
                                   const zd* G0ref, double temp, bool tail, const double* eig_d, const zd* uni_d,
                                   double mu, double mu_ref, const std::vector<zd>& dJ, const zd* sawc_d,
                                   hipStream_t s) {
  const int No = op.No, Nchi = op.Nchi;
  const long NkNw = mesh.Nk * (long)mesh.Nw;
  const double w = temp / (double)mesh.Nkall;
  const double wref = 1.0 / (temp * (double)mesh.Nkall);
  const MeshDev md = mesh_dev(mesh);
  const unsigned S = (unsigned)mesh.Nkall;
  const R invN = (R)(1.0 / (double)p4->transformed_points());
  // the Green-function differences in FP64 and rounded to the grid's FP32
  // once in k space (the values the difference load forms per grid point),
  // then transformed from a quarter of the bytes
  const long nG = NkNw * No * No;
  auto diff_pass = [&](const zd* X, const zd* Xref, DBuf<cpx<R>>& F) {
    if constexpr (sizeof(R) == 4) {
      if (dgbuf.n != (size_t)nG) dgbuf.alloc(nG);
      diff_to_f32(dgbuf.p, X, Xref, nG, s);
      for (int b = 0; b < No; ++b)
        for (int a = 0; a < No; ++a) {
          cpx<R>* g = grid(F, a, b);
          p4->exec_fused(GreenLoad<R, float>{dgbuf.p, a, b, No, md}, GridStore<R>{g, S}, g, tmp.p, -1, (R)1, s);
        }
    } else {
      for (int b = 0; b < No; ++b)
        for (int a = 0; a < No; ++a) {
          cpx<R>* g = grid(F, a, b);
          GreenDiffLoad<R> dl{GreenLoad<double>{X, a, b, No, md}, GreenLoad<double>{Xref, a, b, No, md}};
          p4->exec_fused(dl, GridStore<R>{g, S}, g, tmp.p, -1, (R)1, s);
        }
    }
  };
  diff_pass(G, Gref, Fd);
  DBuf<zd> slice;
  if (tail) {
    diff_pass(G0, G0ref, F0d);
    for (int b = 0; b < No; ++b)
      for (int a = 0; a < No; ++a)
        tau_grids(grid(TAd, a, b), grid(TBd, a, b), a, b, eig_d, uni_d, mu, mu_ref, true, temp, s);
  }
  for (auto& lm : op.irr) {
    auto pr = op.partner[lm[1] + Nchi * lm[0]];
    if (pr[0] == lm[1] && pr[1] == lm[0] && slice.n != (size_t)NkNw) slice.alloc(NkNw);
  }

  for (auto& lm : op.irr) {
    const int l = lm[0], m = lm[1];
    const int a13 = op.ol[l][0], b13 = op.ol[m][0], a42 = op.ol[m][1], b42 = op.ol[l][1];
    auto pr = op.partner[m + Nchi * l];
    const bool self = pr[0] == m && pr[1] == l;
    const bool self_tail = tail && self;
    IncProductLoad<R> pl{};
    pl.m = md;
    // (mirrored copies as in chi0 measured no faster here: the 2 No^2 copies
    // per iteration cost what the in-place reads save, and the spatial-first
    // inverse reads the (-r,-t) operands as whole reversed planes)
    pl.add(grid(Fr, a13, b13), grid(Fd, a13, b13), grid(Fr, a42, b42), grid(Fd, a42, b42), (R)(tail ? -w : -1.0), 1);
    if (tail) {
      if (!self_tail)
        pl.add(grid(F0r, a13, b13), grid(F0d, a13, b13), grid(F0r, a42, b42), grid(F0d, a42, b42), (R)w, 1);
      pl.add(grid(TAr, a13, b13), grid(TAd, a13, b13), grid(TBr, a42, b42), grid(TBd, a42, b42), (R)(-wref), 0);
    }
    const zd J = tail ? dJ[l + Nchi * m] : zd(0, 0);
    zd* dst = chi + NkNw * (m + (long)Nchi * l);
    const zd* base = chi_base + NkNw * (m + (long)Nchi * l);
    if (self_tail) {
      // increment of conj(conv[G] + ref) - conv[G0], formed apart and then added
      if (!mesh.complete) GMOD_HIP(hipMemsetAsync(slice.p, 0, NkNw * sizeof(zd), s));
      product_pass(pl, gather_sf(slice.p, 1.0, J, sawc_d, 0, s), s);
      conj_copy(slice.p, slice.p, NkNw, s);
      IncProductLoad<R> p0{};
      p0.m = md;
      p0.add(grid(F0r, a13, b13), grid(F0d, a13, b13), grid(F0r, a42, b42), grid(F0d, a42, b42), (R)w, 1);
      p4->exec_fused(p0.template fixed<1>(), gather_sf(slice.p, 1.0, zd(0, 0), nullptr, 1, s), work.p, tmp.p, +1, invN,
                     s, true);
      lincomb(dst, base, 1.0, slice.p, 1.0, NkNw, s);
      continue;
    }
    if (self) {
      // the reference conjugates a self-partner pair in place after the
      // accumulation (fchi.f90:610): the stored value is conj(conv[G]), so the
      // increment is conjugated before it is added
      if (!mesh.complete) GMOD_HIP(hipMemsetAsync(slice.p, 0, NkNw * sizeof(zd), s));
      product_pass(pl, gather_sf(slice.p, w, zd(0, 0), nullptr, 0, s), s);
      conj_copy(slice.p, slice.p, NkNw, s);
      lincomb(dst, base, 1.0, slice.p, 1.0, NkNw, s);
      continue;
    }
    GatherStore<R> gst = gather_sf(dst, tail ? 1.0 : w, J, tail ? sawc_d : nullptr, 0, s);
    gst.base = base;
    // the partner copy written by the gather (a complete mesh: every slot of
    // the slice is written)
    const bool fuse_pr = pr[0] >= 0 && mesh.complete;
    if (fuse_pr) gst.pdst = chi + NkNw * (pr[0] + (long)Nchi * pr[1]);
    product_pass(pl, gst, s);
    if (pr[0] >= 0 && !fuse_pr) conj_copy(chi + NkNw * (pr[0] + (long)Nchi * pr[1]), dst, NkNw, s);
  }
}

template <typename R>
void ConvEngine<R>::sigma_increment(zd* sig, const zd* sig_base, const zd* V, const zd* Vref, double temp,
                                    hipStream_t s, bool dv_ready) {
  const int No = op.No, Nchi = op.Nchi;
  const long NkNw = mesh.Nk * (long)mesh.Nw;
  if (acc.n != (size_t)N) acc.alloc(N);
  const MeshDev md = mesh_dev(mesh);
  const unsigned S = (unsigned)mesh.Nkall;
  const R invN = (R)(1.0 / (double)p4->transformed_points());
  const long NC = NkNw;
  // the vertex difference once in k space at grid precision: each element is
  // then read (from a quarter of the bytes) by the four grid points that
  // unfold it, instead of two FP64 reads per grid point
  const long nC = NkNw * Nchi * Nchi;
  if (!dv_ready || dvbuf.n != (size_t)nC) {
    if (dvbuf.n != (size_t)nC) dvbuf.alloc(nC);
    diff_to_f32(dvbuf.p, V, Vref, nC, s);
  }
  // the inverse transform of output (m1, l1) from grid g, gathered onto sig_base + sig
  auto gather_grid = [&](cpx<R>* g, int m1, int l1) {
    const long o = NkNw * (m1 + (long)No * l1);
    GatherStore<R> gst{sig + o, temp / (double)mesh.Nkall, zd(0, 0), nullptr, 0, md};
    gst.base = sig_base + o;
    p4->exec_fused(GridLoad<R>{g, S}, gst, g, tmp.p, +1, invN, s);
  };
  auto gather_out = [&](int m1, int l1) { gather_grid(acc.p, m1, l1); };
  std::vector<int> ytab;
  if (contract_form()) {
    // contraction form: the full transform of every vertex difference once
    // per transposed pair, then all outputs in one streaming kernel
    grouped_table(ytab);
    std::vector<int> dent;
    const std::vector<std::array<int, 2>> heads = share_heads(dent);
    const size_t nb = heads.size() * (size_t)N;
    if (DVt.n != nb) DVt.alloc(nb);
    for (size_t h = 0; h < heads.size(); ++h) {
      const int m = heads[h][0], l = heads[h][1];
      const long oml = NC * (m + (long)Nchi * l), olm = NC * (l + (long)Nchi * m);
      cpx<R>* g = DVt.p + h * N;
      p4->exec_fused(BosonLoad<R, float>{dvbuf.p, oml, olm, zd(0, 0), md}, GridStore<R>{g, S}, g, tmp.p, -1, (R)1,
                     s);
    }
    {
      std::vector<long> hoffs;
      for (auto& h : heads) head_offsets(hoffs, h[0], h[1]);
      if (dh_offs.n < hoffs.size()) dh_offs.alloc(hoffs.size());
      dh_offs.upload(hoffs.data(), hoffs.size(), s);
      if (dhD.n < heads.size() * (size_t)mesh.Nkall) dhD.alloc(heads.size() * (size_t)mesh.Nkall);
      slot0_corr(dhD.p, (const cpx<float>*)dvbuf.p, dh_offs.p, (int)heads.size(), s);
      GMOD_HIP(hipStreamSynchronize(s));  // hoffs is released here
    }
    const int NAB = No * No;
    if (fv_entry.size() != (size_t)Nchi * Nchi) throw Error("sigma_increment: no reference vertex transforms");
    std::vector<int> ct(2 * ytab.size());
    for (size_t i = 0; i < ytab.size(); ++i) {
      ct[2 * i] = fv_entry[ytab[i]];
      ct[2 * i + 1] = dent[ytab[i]];
    }
    if (ctab_d.n != ct.size()) ctab_d.alloc(ct.size());
    ctab_d.upload(ct.data(), ct.size(), s);
    if (ysum.n < (size_t)NAB * N) ysum.alloc((size_t)NAB * N);
    {
      // per point: two vertex grids and Fd per (output, pair), Fr, one write per output
      const double per = 2.0 * NAB * NAB + 2.0 * NAB + NAB;
      ProfScope ps(sizeof(R) == 4 ? PROF_FFT32 : PROF_FFT, per * (double)N * sizeof(cpx<R>), s);
      sigma_inc_contract<R>(ysum.p, Fr.p, Fd.p, FVr.p, DVt.p, dhF.p, dhD.p, ctab_d.p, NAB, No, S, (unsigned)(N / S),
                            mesh.rev.p, s);
    }
    GMOD_HIP(hipStreamSynchronize(s));  // ct is released at return
    for (int o = 0; o < NAB; ++o) gather_grid(ysum.p + (size_t)o * N, o % No, o / No);
    return;
  }
  if (fv_shared) throw Error("sigma_increment: the reference vertex transforms are stored in the contraction form");
  const bool fused = p4->accum_ok(-1) && !sigma_unfused();
  // grouped form (every output has one source per orbital pair (a, b), No <= 4):
  // the reference-vertex term Y_o = sum FVr Fd and Gn = Fr + Fd in one
  // streaming kernel, then the accumulating pass with one operand per source
  bool grouped = fused && No <= 4 && !sigma_ungrouped() && grouped_table(ytab);
  // ysum: Y of the No^2 outputs, then Gn of the No^2 orbital pairs
  if (grouped && ysum.n != (size_t)2 * No * No * N) ysum.alloc((size_t)2 * No * No * N);
  if (!fused) {
    for (int l1 = 0; l1 < No; ++l1)
      for (int m1 = 0; m1 < No; ++m1) {
        bool first = true;
        for (int l = 0; l < Nchi; ++l)
          for (int m = 0; m < Nchi; ++m) {
            if (op.ol[m][0] != m1 || op.ol[l][0] != l1) continue;
            const long oml = NC * (m + (long)Nchi * l), olm = NC * (l + (long)Nchi * m);
            const int a = op.ol[m][1], b = op.ol[l][1];
            SigmaIncStore<R> st{acc.p, grid(Fr, a, b), grid(Fd, a, b), FVr.p + (size_t)(m + Nchi * l) * N,
                                first ? 1 : 0, S};
            p4->exec_fused(BosonLoad<R, float>{dvbuf.p, oml, olm, zd(0, 0), md}, st, work.p, tmp.p, -1, (R)1, s);
            first = false;
          }
        if (!first) gather_out(m1, l1);
      }
    return;
  }
  // transposed outputs share the head passes of their sources (the vertex
  // differences' static slot is zero, so the conjugate-reversal relation
  // holds, with the slot-0 correction)
  const SigmaPlan P = plan_sigma(p4->accum_last_is_t(-1) && p4->reversal_ok());
  accum_tab_h.clear();
  std::vector<int> smap;
  for (size_t i = 0; i < P.outs.size(); ++i)
    for (auto& e : P.src[i]) {
      const int m = e[0], l = e[1], a = op.ol[m][1], b = op.ol[l][1];
      if (grouped) {
        accum_tab_h.push_back(ysum.p + (size_t)No * No * N + (size_t)(a + No * b) * N);  // Gn_ab
      } else {
        accum_tab_h.push_back(grid(Fr, a, b));
        accum_tab_h.push_back(grid(Fd, a, b));
        accum_tab_h.push_back(FVr.p + (size_t)(m + Nchi * l) * N);
      }
      smap.push_back(e[2]);
    }
  if (srcs.n != (size_t)P.nmax * N) srcs.alloc((size_t)P.nmax * N);
  accum_tab.alloc(accum_tab_h.size());
  accum_tab.upload(accum_tab_h.data(), accum_tab_h.size(), s);
  srcmap_d.alloc(smap.size());
  srcmap_d.upload(smap.data(), smap.size(), s);
  std::vector<long> hoffs;
  for (auto& hs : P.heads)
    for (auto& h : hs) head_offsets(hoffs, h[0], h[1]);
  if (dh_offs.n < hoffs.size()) dh_offs.alloc(hoffs.size());
  if (!hoffs.empty()) dh_offs.upload(hoffs.data(), hoffs.size(), s);
  if (dhD.n < (size_t)P.nmax * mesh.Nkall) dhD.alloc((size_t)P.nmax * mesh.Nkall);
  if (grouped) {
    if (ytab_d.n != ytab.size()) ytab_d.alloc(ytab.size());
    ytab_d.upload(ytab.data(), ytab.size(), s);
  }
  GMOD_HIP(hipStreamSynchronize(s));
  if (grouped) {
    const double NAB = (double)No * No;
    ProfScope ps(sizeof(R) == 4 ? PROF_FFT32 : PROF_FFT, (NAB * NAB + 4.0 * NAB) * (double)N * sizeof(cpx<R>), s);
    sigma_ref_contract<R>(ysum.p, ysum.p + (size_t)No * No * N, Fr.p, Fd.p, FVr.p, ytab_d.p, No * No, No, N, s);
  }
  long off = 0, hoff = 0;
  for (size_t i = 0; i < P.outs.size(); ++i) {
    for (auto& h : P.heads[i]) {
      const int m = h[0], l = h[1];
      const long oml = NC * (m + (long)Nchi * l), olm = NC * (l + (long)Nchi * m);
      p4->exec_head(BosonLoad<R, float>{dvbuf.p, oml, olm, zd(0, 0), md}, srcs.p + (long)h[2] * N, tmp.p, -1, s);
    }
    // the heads of one output take consecutive slots
    if (!P.heads[i].empty())
      slot0_corr(dhD.p + (size_t)P.heads[i][0][2] * mesh.Nkall, (const cpx<float>*)dvbuf.p, dh_offs.p + 2 * hoff,
                 (int)P.heads[i].size(), s);
    hoff += (long)P.heads[i].size();
    const int nsrc = (int)P.src[i].size();
    const int m1 = P.outs[i] % No, l1 = P.outs[i] / No;
    if (nsrc == 0) continue;
    if (grouped) {
      const cpx<R>* Y = ysum.p + (size_t)P.outs[i] * N;
      p4->exec_accum_last(srcs.p, N, nsrc, SigmaIncAccumG<R>{accum_tab.p + off, Y, acc.p, S}, -1, (R)1, s,
                          srcmap_d.p + off, dhD.p);
    } else {
      p4->exec_accum_last(srcs.p, N, nsrc, SigmaIncAccum<R>{accum_tab.p + 3 * off, acc.p, S}, -1, (R)1, s,
                          srcmap_d.p + off, dhD.p);
    }
    off += nsrc;
    gather_out(m1, l1);
  }
}

template class ConvEngine<double>;
template class ConvEngine<float>;

}  // namespace gmod
