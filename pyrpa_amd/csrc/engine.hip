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

#include "engine.h"

namespace gmod {

constexpr double PI_E = 3.141592653589793238462643383279;

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

void OrbPairs::from_maps(const int32_t* irr_chi, const int32_t* chi_map) {
  const int P = Nchi * (Nchi + 1) / 2;
  irr.resize(P);
  for (int i = 0; i < P; ++i) irr[i] = {irr_chi[i] - 1, irr_chi[i + P] - 1};
  partner.assign(Nchi * Nchi, {-1, -1});
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
  for (int b = 0; b < No; ++b)
    for (int a = 0; a < No; ++a) {
      {
        ProfScope ps(PROF_SCATTER, (double)N * sizeof(cpx<R>), s);
        scatter_green<R>(grid(dst, a, b), G, a, b, No, mesh, s);
      }
      fft(grid(dst, a, b), -1, false, s);
    }
}

template <typename R>
void ConvEngine<R>::chi0(zd* chi, double temp, bool tail, const double* eig_d, const zd* uni_d,
                         const double* eig_h, const zd* uni_h, const int64_t* invk, double mu, hipStream_t s) {
  const int No = op.No, Nchi = op.Nchi;
  const long Nk = mesh.Nk, Nw = mesh.Nw, M = mesh.M, Nkall = mesh.Nkall;
  const long NkNw = Nk * Nw;
  const double w = temp / (double)Nkall;
  std::vector<zd> rb(No * No, zd(0, 0));
  DBuf<zd> sawc_d;
  if (tail) {
    DBuf<zd> G0((size_t)NkNw * No * No);
    green0(G0.p, eig_d, uni_d, mu, temp, Nk, (int)Nw, No, s);
    transform_green(G0.p, true, s);
    if (TA.n != (size_t)(No * No) * N) { TA.alloc((size_t)No * No * N); TB.alloc((size_t)No * No * N); }
    for (int b = 0; b < No; ++b)
      for (int a = 0; a < No; ++a) {
        scatter_tau_ref<R>(grid(TA, a, b), eig_d, uni_d, a, b, No, mu, temp, 0, mesh, s);
        fft(grid(TA, a, b), -1, true, s);
        scatter_tau_ref<R>(grid(TB, a, b), eig_d, uni_d, a, b, No, mu, temp, 1, mesh, s);
        fft(grid(TB, a, b), -1, true, s);
      }
    // BZ-averaged free density matrix (fchi.f90:731-755)
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
          for (int n = 0; n < No; ++n)
            acc += uni_h[m + No * (n + No * k)] * conj(uni_h[l + No * (n + No * k)]) * f[n];
          rho[m + No * (l + No * k)] = acc;
        }
    }
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
    // sawtooth correction (fchi.f90:757-772)
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
  const double wref = 1.0 / (temp * (double)Nkall);
  for (auto& lm : op.irr) {
    const int l = lm[0], m = lm[1];
    const int a13 = op.ol[l][0], b13 = op.ol[m][0], a42 = op.ol[m][1], b42 = op.ol[l][1];
    auto pr = op.partner[m + Nchi * l];
    // A self-partner slice (partner == itself) is conjugated in place after each
    // of the reference's three accumulation passes (fchi.f90:610, 846); the net
    // result is conj(conv[G] + ref) - conv[G0].  It equals conj of the plain sum
    // whenever TRS holds; keep the reference's order so results match it
    // on any input.
    const bool self_tail = tail && pr[0] == m && pr[1] == l;
    {
      ProfScope ps(PROF_PRODUCT, (double)N * sizeof(cpx<R>) * (tail ? 7 : 3), s);
      if (!tail) {
        product_rev<R>(work.p, grid(Fg, a13, b13), grid(Fg, a42, b42), (R)-1, true, false, mesh, s);
      } else {
        product_rev<R>(work.p, grid(Fg, a13, b13), grid(Fg, a42, b42), (R)(-w), true, false, mesh, s);
        if (!self_tail) product_rev<R>(work.p, grid(F0g, a13, b13), grid(F0g, a42, b42), (R)(w), true, true, mesh, s);
        product_rev<R>(work.p, grid(TA, a13, b13), grid(TB, a42, b42), (R)(-wref), false, true, mesh, s);
      }
    }
    fft(work.p, +1, false, s);
    zd J(0, 0);
    if (tail) {
      if (a13 == b13) J += rb[a42 + No * b42];
      if (a42 == b42) J -= rb[a13 + No * b13];
    }
    zd* dst = chi + NkNw * (m + (long)Nchi * l);
    {
      ProfScope ps(PROF_GATHER, (double)NkNw * (sizeof(cpx<R>) + sizeof(zd)), s);
      gather<R>(dst, work.p, tail ? 1.0 : w, false, J, tail ? sawc_d.p : nullptr, mesh, s);
    }
    if (self_tail) {
      conj_copy(dst, dst, NkNw, s);
      product_rev<R>(work.p, grid(F0g, a13, b13), grid(F0g, a42, b42), (R)(w), true, false, mesh, s);
      fft(work.p, +1, false, s);
      gather<R>(dst, work.p, 1.0, true, zd(0, 0), nullptr, mesh, s);
      continue;
    }
    if (pr[0] >= 0) conj_copy(chi + NkNw * (pr[0] + (long)Nchi * pr[1]), dst, NkNw, s);
  }
  threshold(chi, NkNw * Nchi * Nchi, 1e-9, s);
}

template <typename R>
void ConvEngine<R>::sigma(zd* sig, const zd* V, const double* S_h, const double* C_h, double temp, hipStream_t s) {
  const int No = op.No, Nchi = op.Nchi;
  const long NkNw = mesh.Nk * (long)mesh.Nw;
  if (acc.n != (size_t)N) acc.alloc(N);
  for (int l1 = 0; l1 < No; ++l1)
    for (int m1 = 0; m1 < No; ++m1) {
      bool first = true;
      for (int l = 0; l < Nchi; ++l)
        for (int m = 0; m < Nchi; ++m) {
          if (op.ol[m][0] != m1 || op.ol[l][0] != l1) continue;
          zd stat(1.5 * S_h[m + Nchi * l] - 0.5 * C_h[m + Nchi * l], 0);
          {
            ProfScope ps(PROF_SCATTER, (double)N * sizeof(cpx<R>), s);
            scatter_boson<R>(work.p, V, m, l, Nchi, stat, mesh, s);
          }
          fft(work.p, -1, false, s);
          {
            ProfScope ps(PROF_PRODUCT, (double)N * sizeof(cpx<R>) * (first ? 3 : 4), s);
            mul_acc<R>(acc.p, work.p, grid(Fg, op.ol[m][1], op.ol[l][1]), first, N, s);
          }
          first = false;
        }
      zd* dst = sig + NkNw * (m1 + (long)No * l1);
      if (first) {
        GMOD_HIP(hipMemsetAsync(dst, 0, NkNw * sizeof(zd), s));
        continue;
      }
      fft(acc.p, +1, false, s);
      ProfScope ps(PROF_GATHER, (double)NkNw * (sizeof(cpx<R>) + sizeof(zd)), s);
      gather<R>(dst, acc.p, temp / (double)mesh.Nkall, false, zd(0, 0), nullptr, mesh, s);
    }
  threshold(sig, NkNw * No * No, 1e-9, s);
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
  for (size_t q = 0; q < gap_pairs.size(); ++q) {
    int l = gap_pairs[q][0], m = gap_pairs[q][1];
    zd stat(0.5 * (S_h[m + Nchi * l] + sgn * C_h[m + Nchi * l]), 0);
    scatter_boson<R>(Vft.p + q * N, Vd, m, l, Nchi, stat, mesh, s);
    fft(Vft.p + q * N, -1, false, s);
  }
}

template <typename R>
void ConvEngine<R>::gap(zd* nd, const zd* F, const double* prt, bool singlet, hipStream_t s) {
  const int No = op.No, Nchi = op.Nchi;
  const long NkNw = mesh.Nk * (long)mesh.Nw;
  const double sgn = singlet ? 1.0 : -1.0, dprt = sgn;
  if (Ff.n != (size_t)(No * No) * N) Ff.alloc((size_t)No * No * N);
  if (acc.n != (size_t)N) acc.alloc(N);
  for (int b = 0; b < No; ++b)
    for (int a = 0; a < No; ++a) {
      scatter_anom<R>(grid(Ff, a, b), F, a, b, No, prt[a] * prt[b], sgn, dprt, mesh, s);
      fft(grid(Ff, a, b), -1, false, s);
    }
  GMOD_HIP(hipMemsetAsync(nd, 0, NkNw * No * No * sizeof(zd), s));
  for (int m2 = 0; m2 < No; ++m2)
    for (int l1 = 0; l1 < No; ++l1) {
      bool first = true;
      for (size_t q = 0; q < gap_pairs.size(); ++q) {
        int l = gap_pairs[q][0], m = gap_pairs[q][1];
        if (op.ol[l][0] != l1 || op.ol[m][1] != m2) continue;
        const cpx<R>* v;
        if (vft_cached) {
          v = Vft.p + q * N;
        } else {
          zd stat(0.5 * (S_cache[m + Nchi * l] + sgn * C_cache[m + Nchi * l]), 0);
          scatter_boson<R>(work.p, Vd_dev, m, l, Nchi, stat, mesh, s);
          fft(work.p, -1, false, s);
          v = work.p;
        }
        mul_acc<R>(acc.p, v, grid(Ff, op.ol[l][1], op.ol[m][0]), first, N, s);
        first = false;
      }
      if (first) continue;
      fft(acc.p, +1, false, s);
      gather<R>(nd + NkNw * (l1 + (long)No * m2), acc.p, 1.0, false, zd(0, 0), nullptr, mesh, s);
    }
  hermitize_lower(nd, mesh.Nk, mesh.Nw, No, s);
}

template class ConvEngine<double>;
template class ConvEngine<float>;

}  // namespace gmod
