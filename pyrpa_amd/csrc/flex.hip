// C-ABI entry points of the hot path and the device-resident drivers:
// chi0 (sharp / tail), RPA vertices and susceptibilities, Green functions,
// the FLEX self-energy loop (mkself) and the linearized Eliashberg solver.
#include <chrono>
#include <cmath>
#include <fstream>
#include <memory>

#include "../../include/gmod.h"
#include "engine.h"

namespace gmod {

using clk = std::chrono::steady_clock;
static double secs(clk::time_point a) { return std::chrono::duration<double>(clk::now() - a).count(); }

// ---------------------------------------------------------------- shared state
struct Seed {
  bool valid = false;
  std::vector<zd> sigma;
  double mu = 0, mu_old = 0;
};
static Seed g_seed;
static double g_stats[16];

static void check_err(DBuf<int>& err, hipStream_t s, const char* where) {
  int h = 0;
  err.download(&h, 1, s);
  GMOD_HIP(hipStreamSynchronize(s));
  if (h) {
    std::string w = where;
    if (h & 1) throw Error(w + ": zgesv failed (singular RPA denominator)");
    if (h & 2) throw Error(w + ": eigenvalue iteration did not converge");
    if (h & 4) throw Error(w + ": singular matrix in the Dyson inverse (zgetrf)");
    if (h & 8) throw Error(w + ": singular matrix in n(mu)");
    throw Error(w + ": device error");
  }
}

struct Upload {
  hipStream_t s;
  template <typename T>
  void operator()(DBuf<T>& d, const T* h, size_t n) {
    d.alloc(n);
    d.upload(h, n, s);
  }
};

// Stoner factor: max_q max Re eig(chi0(q,0) S) (fchi.f90:352-399); thr > 0:
// of the thresholded chi0 (the first frequency slice is copied with it)
// (spinful basis: max Re eig(-chi0(q,0) V), fchi0withsoc.f90:156-198, alpha = -1)
static double stoner_factor(const zd* chi, long Nk, long Nw, int Nchi, const double* S_d, hipStream_t s,
                            double thr = 0.0, double alpha = 1.0) {
  DBuf<double> out(Nk);
  DBuf<int> err(1);
  err.zero(s);
  DBuf<zd> slice;
  if (thr > 0.0) {
    slice.alloc((size_t)Nk * Nchi * Nchi);
    threshold_slice(slice.p, chi, Nk, Nk * Nw, (long)Nchi * Nchi, thr, s);
    max_re_eig(out.p, slice.p, BatchLayout{1, Nk}, Nk, Nchi, S_d, alpha, err.p, s);
  } else {
    max_re_eig(out.p, chi, BatchLayout{1, Nk * Nw}, Nk, Nchi, S_d, alpha, err.p, s);
  }
  check_err(err, s, "ckchi");
  std::vector<double> h(Nk);
  out.download(h.data(), Nk, s);
  GMOD_HIP(hipStreamSynchronize(s));
  double best = -1.0e5;
  for (long i = 0; i < Nk; ++i)
    if (h[i] > best) best = h[i];
  return best;
}

template <typename R>
static void chi0_driver(zd* chi_h, const zd* Gk, const int64_t* kmap, const int64_t* invk, OrbPairs& op,
                        double temp, long Nx, long Ny, long Nz, long Nw, long Nk, long Nkall, int No, bool tail,
                        const double* eig, const zd* uni, double mu, const double* S, double* stoner,
                        double alpha = 1.0) {
  hipStream_t s = ctx().stream;
  Mesh mesh;
  mesh.build(kmap, invk, Nkall, Nk, (int)Nx, (int)Ny, (int)Nz, (int)Nw, s);
  const long NkNw = Nk * Nw;
  const int Nchi = op.Nchi;
  Upload up{s};
  DBuf<zd> G, chi(NkNw * Nchi * Nchi), uni_d;
  DBuf<double> eig_d, S_d;
  up(G, Gk, NkNw * No * No);
  if (tail) { up(eig_d, eig, (size_t)No * Nk); up(uni_d, uni, (size_t)No * No * Nk); }
  ConvEngine<R> eng(mesh, op);
  eng.transform_green(G.p, false, s);
  eng.chi0(chi.p, temp, tail, eig_d.p, uni_d.p, eig, uni, invk, mu, s);
  if (stoner) {
    up(S_d, S, (size_t)Nchi * Nchi);
    *stoner = stoner_factor(chi.p, Nk, Nw, Nchi, S_d.p, s, 0.0, alpha);
  }
  chi.download(chi_h, NkNw * Nchi * Nchi, s);
  GMOD_HIP(hipStreamSynchronize(s));
}

// ---------------------------------------------------------------- sigma.bin (fself.f90:742-780)
static void write_record(std::ofstream& f, const void* data, size_t bytes) {
  // gfortran-style records; > 2^31-1 bytes split into sub-records whose
  // length marker is negative except for the last one
  const char* p = static_cast<const char*>(data);
  const size_t maxrec = 2147483639;
  do {
    size_t chunk = std::min(bytes, maxrec);
    bool more = bytes > chunk;
    int32_t mk = more ? -(int32_t)chunk : (int32_t)chunk;
    f.write(reinterpret_cast<const char*>(&mk), 4);
    f.write(p, chunk);
    f.write(reinterpret_cast<const char*>(&mk), 4);
    p += chunk;
    bytes -= chunk;
  } while (bytes > 0);
}

static bool read_record(std::ifstream& f, std::vector<char>& out) {
  out.clear();
  while (true) {
    int32_t mk = 0, mk2 = 0;
    if (!f.read(reinterpret_cast<char*>(&mk), 4)) return !out.empty();
    size_t n = (size_t)(mk < 0 ? -(long)mk : mk);
    size_t old = out.size();
    out.resize(old + n);
    if (!f.read(out.data() + old, n)) return false;
    f.read(reinterpret_cast<char*>(&mk2), 4);
    if (mk >= 0) return true;
  }
}

// ---------------------------------------------------------------- FLEX loop
struct FlexIn {
  const double *Smat, *Cmat, *eig;
  const int64_t *kmap, *invk, *olist;
  const zd *hamk, *uni;
  double mu_init, rfill, temp, pp, eps, sigma_scale;
  long scf_loop, Nkall, Nk, Nw, Nx, Ny, Nz, sub_sigma, m_diis;
  int No, Nchi;
  bool sw_out, sw_in, sw_rescale, sw_tail;
  // spinful basis (fselfwithsoc.f90): the vertex Vmat replaces S and C
  bool soc = false;
  const double* Vmat = nullptr;
  const int64_t *invs = nullptr, *slist = nullptr;
};

// Precision modes: R = double, mixed = false: FP64 throughout (the reference's
// arithmetic).  R = float, mixed = false: FP32 transforms of the full fields.
// R = double, mixed = true: the first `head` iterations in FP64, then chi0 and
// Sigma as FP64 values at a reference iterate plus FP32 transforms of the
// differences from it (ConvEngine::chi0_increment / sigma_increment); vertex,
// Dyson, mu search, mixing and the convergence test stay FP64.  The FP32
// rounding of a difference scales with its size, so the reference moves to
// the current iterate once the convergence measure comes within
// GMOD_REBASE (default 1e-5) times the distance of Sigma from the reference.
template <typename R>
static void mkself_driver(const FlexIn& A, zd* sig_out, double* mu_out, bool mixed) {
  auto t_start = clk::now();
  hipStream_t s = ctx().stream;
  const int verbose = ctx().verbose;
  const int No = A.No, Nchi = A.Nchi;
  const long Nk = A.Nk, Nw = A.Nw, NkNw = Nk * Nw;
  const size_t nS = (size_t)NkNw * No * No, nC = (size_t)NkNw * Nchi * Nchi;
  const double T = A.temp;
  if (A.m_diis < 1) throw Error("m_diis must be >= 1");
  if (A.soc && (mixed || A.sw_tail)) throw Error("mkself_soc: the spinful loop runs without the tail and increments");
  Mesh mesh;
  mesh.build(A.kmap, A.invk, A.Nkall, Nk, (int)A.Nx, (int)A.Ny, (int)A.Nz, (int)Nw, s);
  OrbPairs op;
  if (A.soc) op.from_olist_soc(A.olist, Nchi, No, A.invs, A.slist);
  else op.from_olist(A.olist, Nchi, No);
  Upload up{s};
  DBuf<double> S_d, C_d, eig_d;
  DBuf<zd> ham_d, uni_d;
  // spinful: S_d holds Vmat, and the nu -> infinity value of pair (m, l) is
  // -Vmat(m, l) (fselfwithsoc.f90:592)
  std::vector<double> soc_stat;
  if (A.soc) {
    up(S_d, A.Vmat, (size_t)Nchi * Nchi);
    soc_stat.resize((size_t)Nchi * Nchi);
    for (size_t e = 0; e < soc_stat.size(); ++e) soc_stat[e] = -A.Vmat[e];
  } else {
    up(S_d, A.Smat, (size_t)Nchi * Nchi);
    up(C_d, A.Cmat, (size_t)Nchi * Nchi);
  }
  up(eig_d, A.eig, (size_t)No * Nk);
  up(ham_d, A.hamk, (size_t)No * No * Nk);
  up(uni_d, A.uni, (size_t)No * No * Nk);
  DBuf<zd> G(nS), sig(nS), sig0(nS), sd(nS), chi(nC);
  DBuf<int> err(1);
  err.zero(s);
  std::vector<DBuf<zd>> xout(A.m_diis), res(A.m_diis);  // DIIS history (unused for m_diis = 1)
  if (A.m_diis > 1)
    for (long i = 0; i < A.m_diis; ++i) { xout[i].alloc(nS); res[i].alloc(nS); }
  std::unique_ptr<ConvEngine<R>> eng(new ConvEngine<R>(mesh, op));
  std::unique_ptr<ConvEngine<float>> e32;
  // mixed-precision state: chi0 and Sigma of the current iterate (chi_acc,
  // sig_acc) and the reference iterate (chi_ref, sig_ref, Vref, Gref, G0ref,
  // mu_ref, Jref).  In the increment iterations the vertex overwrites
  // chi_acc in place, and chi_ref takes over the buffer `chi` held.
  DBuf<zd> chi_acc, chi_ref, Vref, sig_acc, sig_ref, Gref, G0, G0ref, sawc_d;
  std::vector<zd> Jref, Jcur, dJ;
  double mu_ref = 0.0;
  long head = 1;
  if (const char* e = getenv("GMOD_MIXED_HEAD")) head = std::max(1L, atol(e));
  double rebase_c = 1.0e-5;
  if (const char* e = getenv("GMOD_REBASE")) rebase_c = atof(e);
  double dist_ref = 0.0;  // max relative distance of Sigma from sig_ref
  if (mixed) {
    chi_acc.alloc(nC);
    Vref.alloc(nC);
    sig_acc.alloc(nS);
    sig_ref.alloc(nS);
    Gref.alloc(nS);
    if (A.sw_tail) { G0.alloc(nS); G0ref.alloc(nS); }
  }
  int n_inc = 0, n_rebase = 0;

  double mu = A.mu_init, mu_old = 0.0;
  // device-resident solve time: from the inputs resident in HBM to the
  // converged Sigma and mu (excludes the host<->device copies of the ABI)
  GMOD_HIP(hipStreamSynchronize(s));
  auto t_dev = clk::now();
  if (A.sw_in) {
    std::vector<zd> h;
    if (g_seed.valid) {
      if (g_seed.sigma.size() != nS) throw Error("in-memory sigma seed has the wrong size");
      h = g_seed.sigma;
      mu = g_seed.mu;
      mu_old = g_seed.mu_old;
      g_seed.valid = false;
    } else {
      std::ifstream f("sigma.bin", std::ios::binary);
      if (!f) throw Error("sw_in: cannot open sigma.bin");
      std::vector<char> r;
      if (!read_record(f, r) || r.size() != 8) throw Error("sigma.bin: bad mu record");
      memcpy(&mu, r.data(), 8);
      if (!read_record(f, r) || r.size() != 8) throw Error("sigma.bin: bad mu_OLD record");
      memcpy(&mu_old, r.data(), 8);
      if (!read_record(f, r) || r.size() < nS * sizeof(zd)) throw Error("sigma.bin: sigma record too short");
      h.resize(nS);
      memcpy(h.data(), r.data(), nS * sizeof(zd));
      if (read_record(f, r) && r.size() == 32) {  // metadata footer (fself.f90:761-780)
        double meta[4];
        memcpy(meta, r.data(), 32);
        if (std::lround(meta[1]) != Nw || std::lround(meta[2]) != No || std::lround(meta[3]) != Nk)
          throw Error("io_sigma: sigma.bin was written on a different mesh");
      }
    }
    sig.upload(h.data(), nS, s);
    GMOD_HIP(hipStreamSynchronize(s));
    t_dev = clk::now();
    if (A.sigma_scale != 1.0) scale(sig.p, nS, A.sigma_scale, s);
    copy(sig0.p, sig.p, nS, s);
    dyson(G.p, sig.p, ham_d.p, mu, T, Nk, (int)Nw, No, err.p, s);
  } else {
    mu_old = mu * 1.2;
    GMOD_HIP(hipMemsetAsync(sig0.p, 0, nS * sizeof(zd), s));
    green0(G.p, eig_d.p, uni_d.p, mu, T, Nk, (int)Nw, No, s);
  }

  // ---- mixing state (compare_sigma, fself.f90:449-542)
  int n_hist = 0, i_hist = 0;
  const int m = (int)A.m_diis;
  double esterr = 0.0;
  // ---- chemical potential state (renew_mu, fself.f90:544-685)
  bool bracket_valid = false;
  double muS_cache = 0.0, muL_cache = 0.0;
  // n(mu) (fself.f90:687-714) through the characteristic polynomials of
  // H + Sigma, formed once per renew_mu (Sigma is fixed while mu moves)
  DBuf<zd> cpoly;
  auto get_rn = [&](double rmu) {
    double parts[2];
    filling_parts_cp(parts, cpoly.p, eig_d.p, rmu, T, Nk, (int)Nw, No, s);
    return (parts[0] + 2.0 * T * parts[1]) / (double)Nk;
  };
  const double rfill = A.rfill;
  auto renew_mu = [&]() {
    if (cpoly.n != (size_t)(NkNw * No)) cpoly.alloc((size_t)NkNw * No);
    filling_charpoly(cpoly.p, sig.p, ham_d.p, Nk, (int)Nw, No, s);
    double eps = esterr > 1.0e-2 ? 1.0e-8 : esterr * 0.1;
    double dmu = std::fabs(mu - mu_old) * 2.0;
    if (dmu < eps * 4.0) dmu = eps * 4.0;
    mu_old = mu;
    double rnS = 0, rnL = 0, muS = 0, muL = 0;
    bool found = false;
    if (bracket_valid) {
      rnS = get_rn(muS_cache);
      rnL = get_rn(muL_cache);
      if (rnS < rfill && rnL > rfill) { muS = muS_cache; muL = muL_cache; found = true; }
    }
    if (!found) {
      muL = mu + dmu;
      muS = mu - dmu;
      for (int it = 1; it <= 100; ++it) {
        mu = muL;
        rnL = get_rn(mu);
        if (rnL < rfill) {
          if (std::fabs(rfill - rnL) > 0.5) muL += 1.0;
          else if (std::fabs(rfill - rnL) > dmu) muL += 0.5;
          else muL += dmu;
        } else break;
        if (it == 100) throw Error("renew_mu: Too many (upper bracket)");
      }
      for (int it = 1; it <= 100; ++it) {
        mu = muS;
        rnS = get_rn(mu);
        if (rnS > rfill) {
          if (std::fabs(rnS - rfill) > 0.5) muS -= 1.0;
          else if (std::fabs(rnS - rfill) > dmu) muS -= 0.5;
          else muS -= dmu;
        } else break;
        if (it == 100) throw Error("renew_mu: Too many (lower bracket)");
      }
    }
    muS_cache = muS;
    muL_cache = muL;
    bracket_valid = true;
    rnL -= rfill;
    rnS -= rfill;
    double rnc = rnS, muc = muS, mud = 0.0, muM = 0.0, rnM = 0.0;
    bool flag = false;
    for (int it = 1; it <= 100; ++it) {
      if (rnc != rnS && rnc != rnL)
        muM = (muL * rnS * rnc * (rnS - rnc) + muS * rnL * rnc * (rnc - rnL) + muc * rnL * rnS * (rnL - rnS)) /
              ((rnL - rnS) * (rnc - rnL) * (rnc - rnS));
      else
        muM = muL - rnL * (muL - muS) / (rnL - rnS);
      if ((0.25 * (3.0 * muS + muL) > muM || muM > muL) ||
          (flag && std::fabs(muM - muL) >= std::fabs(muL - muc) * 0.5) ||
          (!flag && std::fabs(muM - muL) >= std::fabs(muc - mud) * 0.5) ||
          (flag && std::fabs(muL - muc) < 1.0e-8) || (!flag && std::fabs(muc - mud) < 1.0e-8)) {
        muM = (muL + muS) * 0.5;
        flag = true;
      } else {
        flag = false;
      }
      if (std::fabs(muL - muM) < eps) break;
      mu = muM;
      rnM = get_rn(muM);
      mud = muc;
      muc = muL;
      rnc = rnL;
      if (rnS * (rnM - rfill) < 0) { muL = muM; rnL = rnM - rfill; }
      else { muS = muM; rnS = rnM - rfill; }
      if (std::fabs(rnS) < std::fabs(rnL)) {
        std::swap(muL, muS);
        std::swap(rnL, rnS);
      }
      if (it == 100) throw Error("renew_mu: Too many loop!");
    }
    if (rnL == rnS) mu = (muS + muL) * 0.5;
    else mu = (muS * rnL - muL * rnS) / (rnL - rnS);
    double fill = get_rn(mu);
    mu_old = mu;
    if (verbose) printf("mu  =%8.4f filling =%8.4f\n", mu, fill);
  };

  // the reference of the increment iterations: the current iterate, whose
  // chi0 (chi0_n), vertex (V_n) and Sigma (sig_acc) are FP64 values
  auto take_reference = [&](const zd* chi0_n, const zd* V_n) {
    copy(Gref.p, G.p, nS, s);
    copy(Vref.p, V_n, nC, s);
    copy(sig_ref.p, sig_acc.p, nS, s);
    if (chi0_n != chi_ref.p) copy(chi_ref.p, chi0_n, nC, s);
    mu_ref = mu;
    if (A.sw_tail) {
      green0(G0ref.p, eig_d.p, uni_d.p, mu, T, Nk, (int)Nw, No, s);
      Jref = Jcur;
    }
    e32->set_reference(Gref.p, G0ref.p, A.sw_tail, eig_d.p, uni_d.p, mu_ref, T, s);
    e32->set_vertex_reference(Vref.p, A.Smat, A.Cmat, s);
    dist_ref = 0.0;
  };

  double stoner = 0.0;
  double t_chi = 0, t_vert = 0, t_sig = 0, t_mu = 0;
  int iters = 0;
  for (long it = 1; it <= A.scf_loop; ++it) {
    iters = (int)it;
    if (verbose) printf("iter=%5ld\n", it);
    auto t0 = clk::now();
    const bool inc = mixed && e32 != nullptr;
    // chi0 leaves the accumulation untresholded (chi_acc in the mixed loop,
    // chi in place in the FP64 loop); the reference's 1e-9 threshold
    // (fself.f90:375) is applied where chi0 is read: by the Stoner slice and
    // by the vertex kernel
    const zd* chi_src = mixed ? chi_acc.p : chi.p;
    // the increment iterations move the reference to this iterate when the
    // last convergence measure came near the rounding of the differences
    const bool rebase = inc && rebase_c > 0.0 && it > 1 && esterr < rebase_c * dist_ref;
    if (!inc) {
      eng->transform_green(G.p, false, s);
      eng->chi0(mixed ? chi_acc.p : chi.p, T, A.sw_tail, eig_d.p, uni_d.p, A.eig, A.uni, A.invk, mu, s, false);
    } else {
      if (A.sw_tail) {
        green0(G0.p, eig_d.p, uni_d.p, mu, T, Nk, (int)Nw, No, s);
        e32->tail_constants(T, A.eig, A.uni, A.invk, mu, Jcur, sawc_d, s);
        dJ.resize(Jcur.size());
        for (size_t i = 0; i < Jcur.size(); ++i) dJ[i] = Jcur[i] - Jref[i];
      }
      // slots the gather does not keep hold the last vertex: restore them
      if (!mesh.complete) copy(chi_acc.p, chi_ref.p, nC, s);
      e32->chi0_increment(chi_acc.p, chi_ref.p, G.p, Gref.p, G0.p, G0ref.p, T, A.sw_tail, eig_d.p, uni_d.p, mu,
                          mu_ref, dJ, sawc_d.p, s);
      if (rebase) copy(chi_ref.p, chi_acc.p, nC, s);  // chi0 of the new reference, before the vertex overwrites it
    }
    stoner = stoner_factor(chi_src, Nk, Nw, Nchi, S_d.p, s, 1e-9, A.soc ? -1.0 : 1.0);
    t_chi += secs(t0);
    if (verbose) printf("SDW max eig(chi0 S) = %12.8f\n", stoner);
    const bool rescale = A.sw_rescale && stoner >= 1.0;
    // the vertex: into `chi`, or in place over chi_acc in the increment
    // iterations (chi0 is recomputed from chi_ref in the next one)
    zd* Vout = inc ? chi_acc.p : chi.p;
    if (rescale) {
      const double rescale_t = 1.0 - 1.0e-4;  // fself.f90:371-380
      if (verbose) printf("[FLEX] Stoner factor=%10.6f >= 1: rescaling chi0 to%7.4f\n", stoner, rescale_t);
      threshold_copy(Vout, chi_src, nC, 1e-9, s);
      scale(Vout, nC, rescale_t / stoner, s);
    }
    t0 = clk::now();
    // the FP32 vertex difference V - Vref comes out of the vertex kernel in
    // the increment iterations (unless chi0 was rescaled)
    const bool dv_fused = inc && !rescale;
    {
      ProfScope ps(PROF_VERTEX, (double)nC * 2 * sizeof(zd), s);
      if (A.soc)  // chi already thresholded (and rescaled) when rescale
        soc_vertex(rescale ? chi.p : chi_src, chi.p, NkNw, BatchLayout{1, NkNw}, Nchi, S_d.p, rescale ? 0.0 : 1e-9,
                   err.p, s);
      else if (rescale)
        rpa_vertex(Vout, NkNw, BatchLayout{1, NkNw}, Nchi, S_d.p, C_d.p, 0, err.p, s);
      else
        // an increment iteration that keeps its reference reads only V - Vref
        rpa_vertex_fused(chi_src, Vout, NkNw, BatchLayout{1, NkNw}, Nchi, S_d.p, C_d.p, 1e-9,
                         dv_fused ? Vref.p : nullptr, dv_fused ? e32->dv_buffer(nC) : nullptr, err.p, s,
                         !dv_fused || rebase || verbose);
    }
    check_err(err, s, A.soc ? "get_V_soc_flex" : "get_vsigma_flex_nosoc");
    t_vert += secs(t0);
    if (verbose)
      printf("Re V_sigma: max:%12.4E min:%12.4E\n", max_real(Vout, nC, s), min_real(Vout, nC, s));
    t0 = clk::now();
    if (!inc) {
      if (A.soc) eng->sigma(sig.p, chi.p, soc_stat.data(), T, s, true);
      else eng->sigma(mixed ? sig_acc.p : sig.p, chi.p, A.Smat, A.Cmat, T, s, !mixed);
      if (mixed) {
        threshold_copy(sig.p, sig_acc.p, nS, 1e-9, s);
        if (it >= head) {
          if (verbose) printf("[FLEX] switching to FP32 increments\n");
          if (A.sw_tail) eng->tail_constants(T, A.eig, A.uni, A.invk, mu, Jcur, sawc_d, s);
          eng.reset();  // FP64 grids are no longer needed
          e32.reset(new ConvEngine<float>(mesh, op));
          // chi_ref takes the buffer of the vertex, which is copied to Vref first
          std::swap(chi_ref, chi);
          take_reference(chi_acc.p, chi_ref.p);
        }
      }
    } else {
      e32->sigma_increment(sig_acc.p, sig_ref.p, Vout, Vref.p, T, s, dv_fused);
      threshold_copy(sig.p, sig_acc.p, nS, 1e-9, s);
      if (rebase) {
        if (verbose) printf("[FLEX] reference moved to this iterate (esterr %9.2E, distance %9.2E)\n", esterr, dist_ref);
        take_reference(chi_ref.p, Vout);
        ++n_rebase;
      } else {
        dist_ref = max_rel_err(sig_ref.p, sig_acc.p, nS, 1.0e-10, s, nullptr);
      }
      ++n_inc;
    }
    if (A.sub_sigma > 0) subtract_static(sig.p, Nk, (int)Nw, No, (int)A.sub_sigma, s);
    // ---- compare_sigma
    if (m == 1) {
      // one history entry: the DIIS system [[r.r + 1e-12, 1], [1, 0]] c =
      // [0, 1] gives c0 = 1 exactly (either pivot order divides a number by
      // itself or by 1), so the extrapolated Sigma is Sigma itself and the
      // step is the linear mixing; the same values without the history
      // copy, the residual, its norm and the scaling
      esterr = max_rel_err(sig0.p, sig.p, nS, 1.0e-10, s, nullptr);
      lincomb(sig.p, sig.p, A.pp, sig0.p, 1.0 - A.pp, nS, s);
    } else {
      i_hist = i_hist % m + 1;
      n_hist = std::min(n_hist + 1, m);
      copy(xout[i_hist - 1].p, sig.p, nS, s);
      lincomb(res[i_hist - 1].p, sig.p, 1.0, sig0.p, -1.0, nS, s);
      const int n = n_hist;
      std::vector<int> idx(n);
      for (int ih = 1; ih <= n; ++ih) idx[ih - 1] = ((i_hist - n + ih - 1) % m + m) % m;
      std::vector<double> B((n + 1) * (n + 1), 0.0), c(n + 1, 0.0);
      for (int a = 0; a < n; ++a) {
        for (int b = a; b < n; ++b) {
          double v = dot_w(res[idx[a]].p, res[idx[b]].p, nullptr, (long)nS, 1, s).re;
          B[a + (n + 1) * b] = B[b + (n + 1) * a] = v;
        }
        B[a + (n + 1) * n] = B[n + (n + 1) * a] = 1.0;
      }
      for (int a = 0; a < n; ++a) B[a + (n + 1) * a] += 1.0e-12;
      c[n] = 1.0;
      if (!host_dgesv(B, n + 1, c)) {
        if (verbose) printf("DIIS: dgesv failed, fallback to most-recent entry\n");
        std::fill(c.begin(), c.end(), 0.0);
        c[n - 1] = 1.0;
      }
      copy(sd.p, xout[idx[0]].p, nS, s);
      scale(sd.p, nS, c[0], s);
      for (int a = 1; a < n; ++a) axpy(sd.p, xout[idx[a]].p, zd(c[a], 0), nS, s);
      esterr = max_rel_err(sig0.p, sd.p, nS, 1.0e-10, s, nullptr);
      if (n >= 2) copy(sig.p, sd.p, nS, s);
      else lincomb(sig.p, sd.p, A.pp, sig0.p, 1.0 - A.pp, nS, s);
    }
    t_sig += secs(t0);
    if (verbose) printf("esterr=%9.2E\n", esterr);
    if (esterr < A.eps) break;
    t0 = clk::now();
    renew_mu();
    {
      ProfScope ps(PROF_DYSON, (double)nS * 2 * sizeof(zd), s);
      dyson(G.p, sig.p, ham_d.p, mu, T, Nk, (int)Nw, No, err.p, s);
    }
    check_err(err, s, "getinv");
    copy(sig0.p, sig.p, nS, s);
    t_mu += secs(t0);
  }
  // sigma.bin records the mu that built the last G and its mu_OLD: the
  // reference writes the file before its final renew_mu (fself.f90:443-446)
  const double mu_file = mu, mu_old_file = mu_old;
  {
    auto t0 = clk::now();
    renew_mu();
    t_mu += secs(t0);
  }
  check_err(err, s, "mkself");
  const double t_solve = secs(t_dev);
  if (stoner >= 1.0 && verbose) printf("[FLEX] WARNING: Stoner factor >= 1 at the final iteration\n");
  std::vector<zd> h(nS);
  sig.download(h.data(), nS, s);
  GMOD_HIP(hipStreamSynchronize(s));
  memcpy(sig_out, h.data(), nS * sizeof(zd));
  *mu_out = mu;
  if (A.sw_out) {
    std::ofstream f("sigma.bin", std::ios::binary);
    write_record(f, &mu_file, 8);
    write_record(f, &mu_old_file, 8);
    write_record(f, h.data(), nS * sizeof(zd));
    double meta[4] = {T, (double)Nw, (double)No, (double)Nk};
    write_record(f, meta, 32);
    if (!f) throw Error("io_sigma: cannot write sigma.bin");
  }
  double st[12] = {(double)iters, esterr, stoner, mu,    secs(t_start), t_chi, t_vert, t_sig, t_mu, (double)n_inc,
                   t_solve,       (double)n_rebase};
  std::copy(st, st + 12, g_stats);
}

// ---------------------------------------------------------------- Eliashberg
__global__ void k_broadcast_w(zd* D, const zd* D0, long Nk, long Nw, int No) {
  const long NkNw = Nk * Nw;
  const long tot = NkNw * No * No;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
    long k = e % Nk;
    long pq = e / NkNw;
    D[e] = D0[k + Nk * pq];
  }
}

// orbital-basis initial gap (feliash.f90:637-731) into device D (Nk,Nw,No,No)
static void initial_delta(zd* D, const double* init_delta, const zd* uni, const Mesh& mesh, int No,
                          long gap_sym, hipStream_t s) {
  const long Nk = mesh.Nk, Nw = mesh.Nw;
  std::vector<zd> D0((size_t)Nk * No * No, zd(0, 0));
  if (gap_sym == 0) {
    for (long k = 0; k < Nk; ++k)
      for (int l = 0; l < No; ++l) D0[k + Nk * (l + No * l)] = zd(1.0, 0);
  } else {
    for (long k = 0; k < Nk; ++k) {
      for (int l = 0; l < No; ++l)
        for (int mm = 0; mm < No; ++mm) {
          zd acc(0, 0);
          for (int n = 0; n < No; ++n)
            acc += uni[l + No * (n + No * k)] * init_delta[k + Nk * n] * conj(uni[mm + No * (n + No * k)]);
          D0[k + Nk * (l + No * mm)] = acc;
        }
      for (int l = 0; l < No; ++l) {
        D0[k + Nk * (l + No * l)].im = 0.0;
        for (int mm = l + 1; mm < No; ++mm) D0[k + Nk * (l + No * mm)] = conj(D0[k + Nk * (mm + No * l)]);
      }
    }
  }
  DBuf<zd> d0(D0.size());
  d0.upload(D0.data(), D0.size(), s);
  hipLaunchKernelGGL(k_broadcast_w, dim3(4096), dim3(256), 0, s, D, d0.p, Nk, Nw, No);
  GMOD_CHECK_LAUNCH();
  const long rest = Nw * No * No;
  double nrm = dot_w(D, D, mesh.kw.p, Nk, rest, s).re;
  if (nrm <= 0.0) throw Error("get_initial_delta: norm <= 0");
  scale(D, Nk * rest, 1.0 / std::sqrt(2.0 * nrm), s);
}

struct EliashIn {
  zd* chi;
  const zd *Gk, *uni;
  const double *init_delta, *Smat, *Cmat, *prt;
  const int64_t *olist, *kmap, *invk;
  double temp, eps;
  long Nkall, Nk, Nw, Nx, Ny, Nz, itemax, gap_sym, arnoldi_m;
  int No, Nchi;
};

// Two-pass Arnoldi of the linearized gap equation (feliash.f90 and
// feliashwithsoc.f90:190-387, solve_arnoldi): pass 1 builds an md-dimensional
// Krylov basis from the initial gap (modified Gram-Schmidt) and takes the
// most negative Ritz value lambda_-; unless a Ritz value >= 0.1 already
// appeared, pass 2 runs on (A + shift) with shift = -lambda_-, projected off
// the pass-1 Ritz vector, and takes its largest Ritz value.  apply(w, v):
// w = A v; norm, inner: the gap's norm and inner product; init(v): the unit
// initial gap.  D receives the unit eigenvector; returns lambda.
template <class Apply, class Norm, class Inner, class Init>
static double arnoldi_two_pass(int md, size_t n, zd* D, Apply apply, Norm norm, Inner inner, Init init,
                               hipStream_t s, int verbose) {
  std::vector<DBuf<zd>> V(md);
  for (auto& v : V) v.alloc(n);
  DBuf<zd> d1(n), nd(n);
  auto arnoldi = [&](double shift, bool defl, std::vector<zd>& Hm, int& m_act) {
    Hm.assign((size_t)(md + 1) * md, zd(0, 0));
    m_act = md;
    for (int j = 0; j < md; ++j) {
      apply(nd.p, V[j].p);
      if (defl) {
        axpy(nd.p, V[j].p, zd(shift, 0), n, s);
        zd h = inner(d1.p, nd.p);
        axpy(nd.p, d1.p, zd(0, 0) - h, n, s);
      }
      for (int ii = 0; ii <= j; ++ii) {
        zd h = inner(V[ii].p, nd.p);
        Hm[ii + (md + 1) * j] = h;
        axpy(nd.p, V[ii].p, zd(0, 0) - h, n, s);
      }
      double beta = norm(nd.p);
      Hm[j + 1 + (md + 1) * j] = zd(beta, 0);
      if (beta < 1.0e-14) { m_act = j + 1; break; }
      if (j < md - 1) {
        copy(V[j + 1].p, nd.p, n, s);
        scale(V[j + 1].p, n, 1.0 / beta, s);
      }
    }
  };
  auto ritz = [&](const std::vector<zd>& Hm, int m_act, std::vector<zd>& ev) {
    std::vector<zd> Am((size_t)m_act * m_act);
    for (int q = 0; q < m_act; ++q)
      for (int p = 0; p < m_act; ++p) Am[p + m_act * q] = Hm[p + (md + 1) * q];
    if (host_eigvals(Am, m_act, ev)) throw Error("ZGEEV (Hessenberg QR) failed in Arnoldi");
    return Am;
  };
  auto build = [&](zd* out, const std::vector<zd>& Am, int m_act, zd lam) {
    std::vector<zd> y = host_eigvec(Am, m_act, lam);
    GMOD_HIP(hipMemsetAsync(out, 0, n * sizeof(zd), s));
    for (int ii = 0; ii < m_act; ++ii) axpy(out, V[ii].p, y[ii], n, s);
    scale(out, n, 1.0 / norm(out), s);
  };
  init(V[0].p);
  std::vector<zd> Hm, ev;
  int m_act;
  arnoldi(0.0, false, Hm, m_act);
  auto Am = ritz(Hm, m_act, ev);
  int idx = 0;
  for (int i = 1; i < m_act; ++i)
    if (ev[i].re < ev[idx].re) idx = i;
  const double lam1 = ev[idx].re;
  double maxre = ev[0].re;
  int imax = 0;
  for (int i = 1; i < m_act; ++i)
    if (ev[i].re > maxre) { maxre = ev[i].re; imax = i; }
  if (verbose) {
    printf("Pass 1 Ritz values (real):\n");
    for (int i = 0; i < m_act; ++i) printf("%4d%14.6f\n", i + 1, ev[i].re);
  }
  if (maxre >= 0.1) {
    build(D, Am, m_act, ev[imax]);
    return ev[imax].re;
  }
  build(d1.p, Am, m_act, ev[idx]);
  const double shift = -lam1;
  init(V[0].p);
  zd h = inner(d1.p, V[0].p);
  axpy(V[0].p, d1.p, zd(0, 0) - h, n, s);
  scale(V[0].p, n, 1.0 / norm(V[0].p), s);
  arnoldi(shift, true, Hm, m_act);
  Am = ritz(Hm, m_act, ev);
  imax = 0;
  for (int i = 1; i < m_act; ++i)
    if (ev[i].re > ev[imax].re) imax = i;
  build(D, Am, m_act, ev[imax]);
  return ev[imax].re - shift;
}

template <typename R>
static double eliash_driver(const EliashIn& A, zd* delta_out) {
  hipStream_t s = ctx().stream;
  const int verbose = ctx().verbose;
  const int No = A.No, Nchi = A.Nchi;
  const long Nk = A.Nk, Nw = A.Nw, NkNw = Nk * Nw;
  const size_t nS = (size_t)NkNw * No * No, nC = (size_t)NkNw * Nchi * Nchi;
  const bool singlet = A.gap_sym >= 0;
  const double weight = A.temp / (double)A.Nkall;
  Mesh mesh;
  mesh.build(A.kmap, A.invk, A.Nkall, Nk, (int)A.Nx, (int)A.Ny, (int)A.Nz, (int)Nw, s);
  OrbPairs op;
  op.from_olist(A.olist, Nchi, No);
  Upload up{s};
  DBuf<zd> chi, G, Fk(nS), nd(nS);
  DBuf<double> S_d, C_d;
  DBuf<int> err(1);
  err.zero(s);
  up(chi, A.chi, nC);
  up(G, A.Gk, nS);
  up(S_d, A.Smat, (size_t)Nchi * Nchi);
  up(C_d, A.Cmat, (size_t)Nchi * Nchi);
  rpa_vertex(chi.p, NkNw, BatchLayout{1, NkNw}, Nchi, S_d.p, C_d.p, singlet ? 1 : 2, err.p, s);
  check_err(err, s, "get_V_delta_nsoc_flex");
  chi.download(A.chi, nC, s);  // chi is inout: it leaves holding V_Delta (feliash.f90:42,61)
  ConvEngine<R> eng(mesh, op);
  eng.prepare_gap(chi.p, A.Smat, A.Cmat, singlet, s);
  const long rest = Nw * No * No;
  auto apply = [&](zd* w_out, const zd* v_in) {
    anomalous_f(Fk.p, G.p, v_in, Nk, (int)Nw, No, s);
    eng.gap(w_out, Fk.p, A.prt, singlet, s);
    scale(w_out, nS, weight, s);
  };
  auto norm = [&](const zd* f) { return std::sqrt(2.0 * dot_w(f, f, mesh.kw.p, Nk, rest, s).re); };
  auto inner = [&](const zd* u, const zd* v) { return dot_w(u, v, mesh.kw.p, Nk, rest, s) * 2.0; };
  DBuf<zd> D(nS);
  double lam_out = 0.0;
  if (A.arnoldi_m > 0) {
    auto init = [&](zd* v) { initial_delta(v, A.init_delta, A.uni, mesh, No, A.gap_sym, s); };
    lam_out = arnoldi_two_pass((int)A.arnoldi_m, nS, D.p, apply, norm, inner, init, s, verbose);
  } else {
    // power method with shift + deflation (feliash.f90:67-140)
    double lam1 = 0.0, lam_phys = 0.0, lam_rq = 0.0, inorm = 1.0;
    DBuf<zd> d1(nS);
    for (int i_eig = 1; i_eig <= 2; ++i_eig) {
      initial_delta(D.p, A.init_delta, A.uni, mesh, No, A.gap_sym, s);
      if (i_eig == 2) {
        double proj = inner(d1.p, D.p).re;
        axpy(D.p, d1.p, zd(-proj, 0), nS, s);
        scale(D.p, nS, 1.0 / norm(D.p), s);
      }
      double lam_prev = 0.0;
      for (long it = 1; it <= A.itemax; ++it) {
        apply(nd.p, D.p);
        if (i_eig == 2) {
          axpy(nd.p, D.p, zd(-lam1, 0), nS, s);
          double proj = inner(d1.p, nd.p).re;
          axpy(nd.p, d1.p, zd(-proj, 0), nS, s);
        }
        double nrm = norm(nd.p);
        if (nrm <= 0.0) throw Error("Error: norm <= 0 in iter_loop");
        inorm = 1.0 / nrm;
        lam_rq = inner(D.p, nd.p).re;
        lam_phys = lam_rq + lam1;
        // vec_err = sqrt(2 min(|nd/|nd| - D|^2, |nd/|nd| + D|^2))
        double nn = 2.0 * dot_w(nd.p, nd.p, mesh.kw.p, Nk, rest, s).re;
        double dd = 2.0 * dot_w(D.p, D.p, mesh.kw.p, Nk, rest, s).re;
        double xr = 2.0 * dot_w(nd.p, D.p, mesh.kw.p, Nk, rest, s).re;
        double tpos = 0.5 * (nn * inorm * inorm + dd - 2.0 * inorm * xr);
        double tneg = 0.5 * (nn * inorm * inorm + dd + 2.0 * inorm * xr);
        double verr = std::sqrt(2.0 * std::max(0.0, std::min(tpos, tneg)));
        if (verbose) printf("%4ld lambda_elsh=%16.8E%16.8E  vec_err=%12.4E\n", it, lam_phys, std::fabs(lam_phys - lam_prev), verr);
        if (lam_phys < 0.0 && verr < A.eps) break;
        if (verr < A.eps && std::fabs(lam_phys - lam_prev) < A.eps * (std::fabs(lam_phys) + 1.0e-10)) break;
        lam_prev = lam_phys;
        copy(D.p, nd.p, nS, s);
        scale(D.p, nS, inorm, s);
      }
      if (i_eig == 1) {
        if (lam_phys > 0.0) { lam_out = lam_phys; break; }
        lam1 = lam_rq;
        copy(d1.p, nd.p, nS, s);
        scale(d1.p, nS, inorm, s);
      } else {
        lam_out = lam_phys;
      }
    }
    double nrm = norm(nd.p);
    if (nrm <= 0.0) throw Error("Error: final norm <= 0");
    copy(D.p, nd.p, nS, s);
    scale(D.p, nS, 1.0 / nrm, s);
  }
  D.download(delta_out, nS, s);
  GMOD_HIP(hipStreamSynchronize(s));
  return lam_out;
}

// ---------------------------------------------------------------- spinful Eliashberg
struct EliashSocIn {
  zd* chi;
  const zd *Gk, *uni;
  const double *init_delta, *Vmat, *sgnsig;
  const int64_t *olist, *slist, *kmap, *invk, *invs, *invschi;
  double temp, eps;
  long Nkall, Nk, Nw, Nx, Ny, Nz, itemax, gap_sym, arnoldi_m;
  int No, Nchi;
};
static double g_lambda_soc = 0.0;

// lin_eliash_soc (feliashwithsoc.f90:1-424): the gap lives on the full zone,
// norms and inner products are 2 x the plain sums over it
template <typename R>
static double eliash_soc_driver(const EliashSocIn& A, zd* delta_out) {
  hipStream_t s = ctx().stream;
  const int verbose = ctx().verbose;
  const int No = A.No, Nchi = A.Nchi;
  const long Nk = A.Nk, Nw = A.Nw, NkNw = Nk * Nw, Nkall = A.Nkall, NW = Nkall * Nw;
  const size_t nF = (size_t)NW * No * No, nC = (size_t)NkNw * Nchi * Nchi, nG = (size_t)NkNw * No * No;
  const double weight = A.temp / (double)Nkall;
  Mesh mesh;
  mesh.build(A.kmap, A.invk, Nkall, Nk, (int)A.Nx, (int)A.Ny, (int)A.Nz, (int)Nw, s);
  OrbPairs op;
  op.from_olist_soc(A.olist, Nchi, No, A.invs, A.slist);
  Upload up{s};
  DBuf<zd> chi, G, uni_d, Fk(nF), nd(nF), D(nF);
  DBuf<double> V_d, sg_d, init_d;
  DBuf<int64_t> invk_d;
  DBuf<int> sl_d, inv_d, err(1);
  err.zero(s);
  std::vector<int> slh(No), invh(No);
  for (int a = 0; a < No; ++a) { slh[a] = (int)A.slist[a]; invh[a] = op.inv[a]; }
  up(chi, A.chi, nC);
  up(G, A.Gk, nG);
  up(uni_d, A.uni, (size_t)No * No * Nk);
  up(V_d, A.Vmat, (size_t)Nchi * Nchi);
  up(sg_d, A.sgnsig, (size_t)No * No);
  up(init_d, A.init_delta, (size_t)Nk * No);
  up(invk_d, A.invk, (size_t)3 * Nkall);
  up(sl_d, slh.data(), (size_t)No);
  up(inv_d, invh.data(), (size_t)No);
  soc_vertex(chi.p, chi.p, NkNw, BatchLayout{1, NkNw}, Nchi, V_d.p, 0.0, err.p, s);
  check_err(err, s, "get_V_soc_flex");
  chi.download(A.chi, nC, s);  // chi is inout: it leaves holding V_Delta (feliashwithsoc.f90:55)
  if (verbose)
    printf("V_delta max is %16.8E\nV_delta min is %16.8E\n", max_real(chi.p, nC, s), min_real(chi.p, nC, s));
  ConvEngine<R> eng(mesh, op);
  eng.prepare_gap_soc(chi.p, A.Vmat, A.invschi, A.gap_sym, s);
  auto apply = [&](zd* w_out, const zd* v_in) {
    anomalous_f_soc(Fk.p, G.p, v_in, invk_d.p, sg_d.p, sl_d.p, inv_d.p, Nkall, Nk, (int)Nw, No, A.gap_sym, s);
    eng.gap_soc(w_out, Fk.p, s);
    delta_sym_soc(w_out, sl_d.p, inv_d.p, NW, No, A.gap_sym, s);
    scale(w_out, nF, -weight, s);
  };
  const long rest = No * No;
  auto dot = [&](const zd* u, const zd* v) { return dot_w(u, v, nullptr, NW, rest, s); };
  auto norm = [&](const zd* f) { return std::sqrt(2.0 * dot(f, f).re); };
  auto inner = [&](const zd* u, const zd* v) { return dot(u, v) * 2.0; };
  auto init = [&](zd* v) {
    initial_delta_soc(v, init_d.p, uni_d.p, invk_d.p, sl_d.p, inv_d.p, Nkall, Nk, (int)Nw, No, A.gap_sym, s);
    const double n2 = dot(v, v).re;
    if (n2 <= 0.0) throw Error("get_initial_delta_soc: norm <= 0");
    scale(v, nF, 1.0 / std::sqrt(2.0 * n2), s);
  };
  double lam = 0.0;
  if (A.arnoldi_m > 0) {
    lam = arnoldi_two_pass((int)A.arnoldi_m, nF, D.p, apply, norm, inner, init, s, verbose);
  } else {
    // shifted power method (feliashwithsoc.f90:59-110): newdelta = A delta +
    // norm2 delta, norm2 = 0 for the first eigenvalue, then the norm of the
    // first run's last iterate when that eigenvalue is not positive
    double norm2 = 0.0, lam_rq = 0.0, nrm = 0.0;
    for (int i_eig = 1; i_eig <= 2; ++i_eig) {
      init(D.p);
      int count = 0;
      for (long it = 1; it <= A.itemax; ++it) {
        apply(nd.p, D.p);
        if (norm2 != 0.0) axpy(nd.p, D.p, zd(norm2, 0), nF, s);
        nrm = norm(nd.p);
        if (nrm <= 0.0) throw Error("lin_eliash_soc: norm <= 0 in iter_loop");
        const double inorm = 1.0 / nrm;
        lam_rq = inner(D.p, nd.p).re;
        // vec_err = sqrt(2 min(sum|nd/|nd| - D|^2, sum|nd/|nd| + D|^2))
        const double nn = dot(nd.p, nd.p).re, dd = dot(D.p, D.p).re, xr = dot(nd.p, D.p).re;
        const double tpos = nn * inorm * inorm + dd - 2.0 * inorm * xr;
        const double tneg = nn * inorm * inorm + dd + 2.0 * inorm * xr;
        const double verr = std::sqrt(2.0 * std::max(0.0, std::min(tpos, tneg)));
        if (verbose) printf("%3ld lambda_elsh=%16.8E  vec_err=%12.4E\n", it, lam_rq - norm2, verr);
        if (verr < A.eps) {
          if (lam_rq - norm2 > 1.0e-1) break;
          if (lam_rq - norm2 < 0.0) break;
          if (verr < A.eps * 1.0e-2) ++count;
          if (count > 30) break;
        }
        copy(D.p, nd.p, nF, s);
        scale(D.p, nF, inorm, s);
      }
      lam = lam_rq - norm2;
      if (verbose) printf("eliash=%16.8E\n", lam);
      if (i_eig == 1 && lam > 0.0) break;
      norm2 = nrm;
    }
    nrm = norm(nd.p);
    if (nrm <= 0.0) throw Error("lin_eliash_soc: final norm <= 0");
    copy(D.p, nd.p, nF, s);
    scale(D.p, nF, 1.0 / nrm, s);
  }
  D.download(delta_out, nF, s);
  GMOD_HIP(hipStreamSynchronize(s));
  return lam;
}

// ---------------------------------------------------------------- nonlinear Eliashberg
struct NonlinIn {
  const zd *Gk, *hamk;
  const double *Smat, *Cmat, *prt;
  const int64_t *olist, *kmap, *invk;
  double mu, temp, eps, gap_min;
  long Nkall, Nk, Nw, Nx, Ny, Nz, itemax, gap_sym, m_diis;
  bool sw_sigma, sw_Vconst, sw_amp_newton;
  int No, Nchi;
};
static double g_nonlin_stats[6];  // iterations, status (0 itemax, 1 converged, 2 collapsed), anchors, last rel. change

// Non-linear FLEX-Eliashberg loop (ffeliash.f90:2-313).  Per iteration: the
// gap equation (the Eliashberg gap convolution), the convergence and
// gap-collapse tests, the amplitude phase (Delta = a v0, a from a probe or the
// pitchfork fit g(a) = lambda0 - c a^2 of the v0-projected gain) or a
// Pulay-DIIS / linear-mixing (pp = 0.3) step, the FLEX Sigma and the dressed
// G0, one step of the superconducting Dyson equations for G and F, then (unless
// the vertex is held constant) chi0_S = chi^GG + chi^FF, chi0_C = chi^GG -
// chi^FF and the vertices V_sigma, V_Delta.  Delta (in/out), Sigma and the
// final normal G are returned; every field stays on the device.
template <typename R>
static void nonlin_eliash_driver(const NonlinIn& A, zd* delta_io, zd* sigma_out, zd* Gk_io) {
  hipStream_t s = ctx().stream;
  const int verbose = ctx().verbose;
  const int No = A.No, Nchi = A.Nchi;
  const long Nk = A.Nk, Nw = A.Nw, NkNw = Nk * Nw;
  const size_t nS = (size_t)NkNw * No * No, nC = (size_t)NkNw * Nchi * Nchi;
  const bool singlet = A.gap_sym >= 0;
  const double T = A.temp, weight = T / (double)A.Nkall;
  const double pp = 0.3, eps_reg = 1.0e-12, amp_tol = 2.0e-2;  // ffeliash.f90:76-77
  const int n_polish_max = 10;
  Mesh mesh;
  mesh.build(A.kmap, A.invk, A.Nkall, Nk, (int)A.Nx, (int)A.Ny, (int)A.Nz, (int)Nw, s);
  OrbPairs op;
  op.from_olist(A.olist, Nchi, No);
  Upload up{s};
  DBuf<zd> D, G, ham, G0d(nS), F(nS), nd(nS), sig(nS), v0(nS), VS(nC), VD(nC);
  DBuf<double> S_d, C_d;
  DBuf<int> err(1);
  err.zero(s);
  up(D, delta_io, nS);
  up(G, Gk_io, nS);
  up(ham, A.hamk, (size_t)No * No * Nk);
  up(S_d, A.Smat, (size_t)Nchi * Nchi);
  up(C_d, A.Cmat, (size_t)Nchi * Nchi);
  ConvEngine<R> eng(mesh, op);
  const int m_eff = (int)std::max<long>(A.m_diis, 1);
  std::vector<DBuf<zd>> xout(m_eff), res(m_eff);
  for (int i = 0; i < m_eff; ++i) { xout[i].alloc(nS); res[i].alloc(nS); }
  auto sumsq = [&](const zd* a) { return dot_w(a, a, nullptr, (long)nS, 1, s).re; };
  // VS <- chi0_S, VD <- chi0_C from G and F (ffeliash.f90:459-489)
  auto chi0sc = [&]() {
    eng.transform_green(G.p, false, s);
    eng.chi0(VS.p, T, false, nullptr, nullptr, nullptr, nullptr, nullptr, 0.0, s, true);
    eng.chi0_ff(VD.p, F.p, A.prt, singlet, T, s);
    sum_diff(VS.p, VD.p, (long)nC, s);
  };
  auto vertices = [&]() {
    if (verbose) printf("SDW max eig(chi0_S S) = %12.8f\n", stoner_factor(VS.p, Nk, Nw, Nchi, S_d.p, s, 0.0, 1.0));
    sc_vertex(VS.p, VD.p, NkNw, BatchLayout{1, NkNw}, Nchi, S_d.p, C_d.p, singlet, err.p, s);
    check_err(err, s, "mkV_flex_nosoc");
    eng.prepare_gap(VD.p, A.Smat, A.Cmat, singlet, s);
  };
  // amplitude state (ffeliash.f90:88-101)
  int amp_phase = A.sw_amp_newton ? 1 : 2, npts = 0, n_polish = 0, n_anchor = 0;
  double a1 = 0, a2 = 0, g1 = 0, g2 = 0;
  double a_cur = std::sqrt(sumsq(D.p)) + 1.0e-300;
  copy(v0.p, D.p, nS, s);
  scale(v0.p, nS, 1.0 / a_cur, s);
  int n_hist = 0, i_hist = 0;
  if (verbose) printf("%s\n", singlet ? "singlet" : "triplet");
  copy(G0d.p, G.p, nS, s);
  const double delta0_max = max_abs_diff(D.p, nullptr, (long)nS, s) + 1.0e-30;
  if (A.sw_Vconst) {
    GMOD_HIP(hipMemsetAsync(F.p, 0, nS * sizeof(zd), s));
    chi0sc();
    anomalous_f(F.p, G.p, D.p, Nk, (int)Nw, No, s);
  } else {
    anomalous_f(F.p, G.p, D.p, Nk, (int)Nw, No, s);
    chi0sc();
  }
  vertices();
  GMOD_HIP(hipMemsetAsync(sig.p, 0, nS * sizeof(zd), s));
  long it_done = 0;
  int status = 0;
  double rel = 0.0;
  for (long it = 1; it <= A.itemax; ++it) {
    it_done = it;
    if (verbose) printf("iter=%5ld\n", it);
    eng.gap(nd.p, F.p, A.prt, singlet, s);
    scale(nd.p, nS, weight, s);
    const double diff = max_abs_diff(nd.p, D.p, (long)nS, s);
    const double dnorm = max_abs_diff(D.p, nullptr, (long)nS, s) + 1.0e-30;
    const double new_max = max_abs_diff(nd.p, nullptr, (long)nS, s);
    rel = diff / dnorm;
    if (verbose) printf("  |D_new-D|max=%12.4E  rel=%12.4E  |D_new|max=%12.4E\n", diff, rel, new_max);
    if (rel < A.eps) {
      copy(D.p, nd.p, nS, s);
      status = 1;
      if (verbose) printf("converged at iter=%5ld\n", it);
      break;
    }
    if (A.gap_min > 0.0 && new_max < A.gap_min * delta0_max) {
      copy(D.p, nd.p, nS, s);
      status = 2;
      if (verbose) printf("gap collapsed: |D_new|max=%12.4E < %12.4E, stopping\n", new_max, A.gap_min * delta0_max);
      break;
    }
    if (A.sw_amp_newton && amp_phase == 2 && n_polish >= n_polish_max) {
      a_cur = std::sqrt(sumsq(D.p)) + 1.0e-300;
      copy(v0.p, D.p, nS, s);
      scale(v0.p, nS, 1.0 / a_cur, s);
      npts = 0; a1 = a2 = g1 = g2 = 0.0;
      amp_phase = 1;
      ++n_anchor;
      if (verbose) printf("  re-anchor #%d\n", n_anchor);
    }
    if (amp_phase == 1) {
      const double proj = dot_w(v0.p, nd.p, nullptr, (long)nS, 1, s).re;
      const double g_cur = proj / a_cur;
      a2 = a1; g2 = g1; a1 = a_cur; g1 = g_cur;
      npts = std::min(npts + 1, 2);
      double a_nxt;
      if (npts >= 2 && std::fabs(a1 * a1 - a2 * a2) > 1.0e-30) {
        const double cfit = (g1 - g2) / (a2 * a2 - a1 * a1);
        const double lam0 = g1 + cfit * a1 * a1;
        if (cfit > 1.0e-12 && lam0 > 1.0) {
          a_nxt = std::sqrt((lam0 - 1.0) / cfit);
          a_nxt = std::min(std::max(a_nxt, 0.3 * a_cur), 3.0 * a_cur);
        } else {
          a_nxt = 0.5 * a_cur;
        }
      } else {
        a_nxt = std::fabs(g_cur - 1.0) < 0.2 ? 0.9 * a_cur : 0.6 * a_cur;
      }
      if (verbose) printf("  amp: g=%9.4f  a=%11.4E -> %11.4E\n", g_cur, a_cur, a_nxt);
      if (std::fabs(g_cur - 1.0) < amp_tol) {
        amp_phase = 2;
        a_nxt = a_cur;
        n_polish = 0;
        n_hist = i_hist = 0;
      }
      copy(D.p, v0.p, nS, s);
      scale(D.p, nS, a_nxt, s);
      a_cur = a_nxt;
    } else {
      i_hist = i_hist % m_eff + 1;
      n_hist = std::min(n_hist + 1, m_eff);
      const int n = n_hist;
      copy(xout[i_hist - 1].p, nd.p, nS, s);
      lincomb(res[i_hist - 1].p, nd.p, 1.0, D.p, -1.0, nS, s);
      bool pulay = n >= 2 && m_eff >= 2;
      std::vector<double> c(n + 1, 0.0);
      std::vector<int> idx(n);
      for (int ih = 1; ih <= n; ++ih) idx[ih - 1] = ((i_hist - n + ih - 1) % m_eff + m_eff) % m_eff;
      if (pulay) {
        std::vector<double> B((n + 1) * (n + 1), 0.0);
        for (int a = 0; a < n; ++a) {
          for (int b = a; b < n; ++b) {
            const double v = dot_w(res[idx[a]].p, res[idx[b]].p, nullptr, (long)nS, 1, s).re;
            B[a + (n + 1) * b] = B[b + (n + 1) * a] = v;
          }
          B[a + (n + 1) * a] += eps_reg;
          B[a + (n + 1) * n] = B[n + (n + 1) * a] = 1.0;
        }
        c[n] = 1.0;
        if (!host_dgesv(B, n + 1, c)) {
          if (verbose) printf("DIIS: dgesv failed, falling back to linear mixing pp=%4.2f.\n", pp);
          pulay = false;
        }
      }
      if (pulay) {
        GMOD_HIP(hipMemsetAsync(D.p, 0, nS * sizeof(zd), s));
        for (int a = 0; a < n; ++a) axpy(D.p, xout[idx[a]].p, zd(c[a], 0), nS, s);
      } else {
        lincomb(D.p, nd.p, pp, D.p, 1.0 - pp, nS, s);
      }
      ++n_polish;
    }
    if (A.sw_sigma) {
      eng.transform_green(G.p, false, s);
      eng.sigma(sig.p, VS.p, A.Smat, A.Cmat, T, s, true);
      dyson(G0d.p, sig.p, ham.p, A.mu, T, Nk, (int)Nw, No, err.p, s);
      check_err(err, s, "getinv");
    }
    sc_green_update(G.p, F.p, G0d.p, D.p, Nk, (int)Nw, No, s);
    if (!A.sw_Vconst) {
      chi0sc();
      vertices();
    }
  }
  D.download(delta_io, nS, s);
  sig.download(sigma_out, nS, s);
  G.download(Gk_io, nS, s);
  GMOD_HIP(hipStreamSynchronize(s));
  g_nonlin_stats[0] = (double)it_done;
  g_nonlin_stats[1] = (double)status;
  g_nonlin_stats[2] = (double)n_anchor;
  g_nonlin_stats[3] = rel;
}

}  // namespace gmod

using namespace gmod;

#define GMOD_TRY try {
#define GMOD_CATCH                                       \
  }                                                      \
  catch (std::exception & e) { set_error(e.what()); }    \
  catch (...) { set_error("unknown error"); }

static std::lock_guard<std::mutex> api_lock() { return std::lock_guard<std::mutex>(ctx().lock); }

// ====================================================================== chi0
extern "C" void get_chi0(gmod_z* chi, const double* Smat, const double* Cmat, const gmod_z* Gk, const int64_t* kmap,
                         const int64_t* invk, const int64_t* olist, const double* temp, const int64_t* Nx,
                         const int64_t* Ny, const int64_t* Nz, const int64_t* Nw, const int64_t* Nk,
                         const int64_t* Nkall, const int64_t* Norb, const int64_t* Nchi, double* maxchi0s_out) {
  GMOD_TRY
  auto g = api_lock();
  (void)Cmat;
  OrbPairs op;
  op.from_olist(olist, (int)*Nchi, (int)*Norb);
  if (ctx().precision >= 1)
    chi0_driver<float>((zd*)chi, (const zd*)Gk, kmap, invk, op, *temp, *Nx, *Ny, *Nz, *Nw, *Nk, *Nkall, (int)*Norb,
                       false, nullptr, nullptr, 0.0, Smat, maxchi0s_out);
  else
    chi0_driver<double>((zd*)chi, (const zd*)Gk, kmap, invk, op, *temp, *Nx, *Ny, *Nz, *Nw, *Nk, *Nkall, (int)*Norb,
                        false, nullptr, nullptr, 0.0, Smat, maxchi0s_out);
  GMOD_CATCH
}

extern "C" void get_chi0_tail(gmod_z* chi, const double* Smat, const double* Cmat, const gmod_z* Gk,
                              const double* eig, const gmod_z* uni, const int64_t* kmap, const int64_t* invk,
                              const int64_t* olist, const double* mu, const double* temp, const int64_t* Nx,
                              const int64_t* Ny, const int64_t* Nz, const int64_t* Nw, const int64_t* Nk,
                              const int64_t* Nkall, const int64_t* Norb, const int64_t* Nchi, double* maxchi0s_out) {
  GMOD_TRY
  auto g = api_lock();
  (void)Cmat;
  OrbPairs op;
  op.from_olist(olist, (int)*Nchi, (int)*Norb);
  if (ctx().precision >= 1)
    chi0_driver<float>((zd*)chi, (const zd*)Gk, kmap, invk, op, *temp, *Nx, *Ny, *Nz, *Nw, *Nk, *Nkall, (int)*Norb,
                       true, eig, (const zd*)uni, *mu, Smat, maxchi0s_out);
  else
    chi0_driver<double>((zd*)chi, (const zd*)Gk, kmap, invk, op, *temp, *Nx, *Ny, *Nz, *Nw, *Nk, *Nkall, (int)*Norb,
                        true, eig, (const zd*)uni, *mu, Smat, maxchi0s_out);
  GMOD_CATCH
}

extern "C" void get_chi0_conv_(gmod_z* chi, const gmod_z* Gk, const int64_t* kmap, const int64_t* invk,
                               const int32_t* irr_chi, const int32_t* chi_map, const int64_t* olist,
                               const double* temp, const int64_t* Nx, const int64_t* Ny, const int64_t* Nz,
                               const int64_t* Nw, const int64_t* Nk, const int64_t* Nkall, const int64_t* Norb,
                               const int64_t* Nchi) {
  GMOD_TRY
  auto g = api_lock();
  OrbPairs op;
  op.from_olist(olist, (int)*Nchi, (int)*Norb);
  op.from_maps(irr_chi, chi_map);
  if (ctx().precision >= 1)
    chi0_driver<float>((zd*)chi, (const zd*)Gk, kmap, invk, op, *temp, *Nx, *Ny, *Nz, *Nw, *Nk, *Nkall, (int)*Norb,
                       false, nullptr, nullptr, 0.0, nullptr, nullptr);
  else
    chi0_driver<double>((zd*)chi, (const zd*)Gk, kmap, invk, op, *temp, *Nx, *Ny, *Nz, *Nw, *Nk, *Nkall, (int)*Norb,
                        false, nullptr, nullptr, 0.0, nullptr, nullptr);
  GMOD_CATCH
}

// ====================================================================== vertices / RPA
extern "C" void get_vsigma_flex_nosoc_(gmod_z* chi, const double* Smat, const double* Cmat, const int64_t* Nk,
                                       const int64_t* Nw, const int64_t* Nchi) {
  GMOD_TRY
  auto g = api_lock();
  hipStream_t s = ctx().stream;
  const long NkNw = *Nk * *Nw;
  const int n = (int)*Nchi;
  Upload up{s};
  DBuf<zd> c;
  DBuf<double> S, C;
  DBuf<int> err(1);
  err.zero(s);
  up(c, (const zd*)chi, (size_t)NkNw * n * n);
  up(S, Smat, (size_t)n * n);
  up(C, Cmat, (size_t)n * n);
  rpa_vertex(c.p, NkNw, BatchLayout{1, NkNw}, n, S.p, C.p, 0, err.p, s);
  check_err(err, s, "get_vsigma_flex_nosoc");
  c.download((zd*)chi, (size_t)NkNw * n * n, s);
  GMOD_HIP(hipStreamSynchronize(s));
  GMOD_CATCH
}

// the spinful vertex of fUJ.f90:367-399 in place on chi(Nk,Nw,Nchi,Nchi)
// (internal to mkself_soc in the reference, exported here for parity tests)
extern "C" void gmod_v_soc_flex(gmod_z* chi, const double* Vmat, const int64_t* Nk, const int64_t* Nw,
                                const int64_t* Nchi) {
  GMOD_TRY
  auto g = api_lock();
  hipStream_t s = ctx().stream;
  const long NkNw = *Nk * *Nw;
  const int n = (int)*Nchi;
  Upload up{s};
  DBuf<zd> c;
  DBuf<double> V;
  DBuf<int> err(1);
  err.zero(s);
  up(c, (const zd*)chi, (size_t)NkNw * n * n);
  up(V, Vmat, (size_t)n * n);
  soc_vertex(c.p, c.p, NkNw, BatchLayout{1, NkNw}, n, V.p, 0.0, err.p, s);
  check_err(err, s, "get_V_soc_flex");
  c.download((zd*)chi, (size_t)NkNw * n * n, s);
  GMOD_HIP(hipStreamSynchronize(s));
  GMOD_CATCH
}

extern "C" void get_chis(gmod_z* chis, const gmod_z* chi0, const double* Smat, const int64_t* Nchi,
                         const int64_t* Nw) {
  GMOD_TRY
  auto g = api_lock();
  hipStream_t s = ctx().stream;
  const int n = (int)*Nchi;
  const long nw = *Nw;
  Upload up{s};
  DBuf<zd> a, o((size_t)n * n * nw);
  DBuf<double> S;
  DBuf<int> err(1);
  err.zero(s);
  up(a, (const zd*)chi0, (size_t)n * n * nw);
  up(S, Smat, (size_t)n * n);
  rpa_solve(o.p, BatchLayout{(long)n * n, 1}, a.p, BatchLayout{(long)n * n, 1}, nw, n, S.p, -1.0, err.p, s);
  check_err(err, s, "get_chis");
  o.download((zd*)chis, (size_t)n * n * nw, s);
  GMOD_HIP(hipStreamSynchronize(s));
  GMOD_CATCH
}

extern "C" void get_chis_chic(gmod_z* chis, gmod_z* chic, const gmod_z* chi, const double* Smat, const double* Cmat,
                              const int64_t* Nk, const int64_t* Nw, const int64_t* Nchi) {
  GMOD_TRY
  auto g = api_lock();
  hipStream_t s = ctx().stream;
  const int n = (int)*Nchi;
  const long nk = *Nk, NkNw = *Nk * *Nw;
  Upload up{s};
  DBuf<zd> a, os((size_t)nk * n * n), oc((size_t)nk * n * n);
  DBuf<double> S, C;
  DBuf<int> err(1);
  err.zero(s);
  // only the static slice chi(:,1,:,:) is read: copy it with the full layout
  up(a, (const zd*)chi, (size_t)NkNw * n * n);
  up(S, Smat, (size_t)n * n);
  up(C, Cmat, (size_t)n * n);
  rpa_solve(os.p, BatchLayout{1, nk}, a.p, BatchLayout{1, NkNw}, nk, n, S.p, -1.0, err.p, s);
  rpa_solve(oc.p, BatchLayout{1, nk}, a.p, BatchLayout{1, NkNw}, nk, n, C.p, 1.0, err.p, s);
  check_err(err, s, "get_chis_chic");
  os.download((zd*)chis, (size_t)nk * n * n, s);
  oc.download((zd*)chic, (size_t)nk * n * n, s);
  GMOD_HIP(hipStreamSynchronize(s));
  GMOD_CATCH
}

// chi_s and chi_c at every (q, i nu): the algebra of get_chis (fchi.f90:319-350)
// and get_chis_chic (fself.f90:783-816) over the whole (Nk, Nw) batch
extern "C" void gmod_get_chis_chic_all(gmod_z* chis, gmod_z* chic, const gmod_z* chi, const double* Smat,
                                       const double* Cmat, const int64_t* Nk, const int64_t* Nw,
                                       const int64_t* Nchi) {
  GMOD_TRY
  auto g = api_lock();
  hipStream_t s = ctx().stream;
  const int n = (int)*Nchi;
  const long NkNw = *Nk * *Nw;
  const size_t tot = (size_t)NkNw * n * n;
  if (!chis && !chic) return;
  Upload up{s};
  DBuf<zd> a, o(tot);
  DBuf<double> S, C;
  DBuf<int> err(1);
  err.zero(s);
  up(a, (const zd*)chi, tot);
  if (chis) {
    up(S, Smat, (size_t)n * n);
    rpa_solve(o.p, BatchLayout{1, NkNw}, a.p, BatchLayout{1, NkNw}, NkNw, n, S.p, -1.0, err.p, s);
    check_err(err, s, "gmod_get_chis_chic_all (chi_s)");
    o.download((zd*)chis, tot, s);
    GMOD_HIP(hipStreamSynchronize(s));
  }
  if (chic) {
    up(C, Cmat, (size_t)n * n);
    rpa_solve(o.p, BatchLayout{1, NkNw}, a.p, BatchLayout{1, NkNw}, NkNw, n, C.p, 1.0, err.p, s);
    check_err(err, s, "gmod_get_chis_chic_all (chi_c)");
    o.download((zd*)chic, tot, s);
    GMOD_HIP(hipStreamSynchronize(s));
  }
  GMOD_CATCH
}

__global__ void k_trace_map(zd* out, const zd* chi, const int64_t* invk, long Nkall, long Nk, int n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < Nkall; i += (long)gridDim.x * blockDim.x) {
    long k = invk[3 * i] - 1;
    zd t(0, 0);
    for (int l = 0; l < n; ++l) t += chi[k + Nk * (l + (long)n * l)];
    out[i] = t;
  }
}
__global__ void k_eig_map(zd* out, const double* v, const int64_t* invk, long Nkall) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < Nkall; i += (long)gridDim.x * blockDim.x)
    out[i] = zd(v[invk[3 * i] - 1], 0.0);
}

extern "C" void get_eig_or_tr_chi(gmod_z* chiqout, const gmod_z* chi, const int64_t* invk, const int64_t* Nkall,
                                  const int64_t* Nk, const int64_t* Nchi, const bool* sw_eig) {
  GMOD_TRY
  auto g = api_lock();
  hipStream_t s = ctx().stream;
  const int n = (int)*Nchi;
  const long nk = *Nk, nka = *Nkall;
  Upload up{s};
  DBuf<zd> a, o(nka);
  DBuf<int64_t> iv;
  DBuf<int> err(1);
  err.zero(s);
  up(a, (const zd*)chi, (size_t)nk * n * n);
  up(iv, invk, (size_t)3 * nka);
  if (*sw_eig) {
    DBuf<double> v(nk);
    max_re_eig(v.p, a.p, BatchLayout{1, nk}, nk, n, nullptr, 1.0, err.p, s);
    check_err(err, s, "get_eig_or_tr_chi");
    hipLaunchKernelGGL(k_eig_map, dim3(256), dim3(256), 0, s, o.p, v.p, iv.p, nka);
  } else {
    hipLaunchKernelGGL(k_trace_map, dim3(256), dim3(256), 0, s, o.p, a.p, iv.p, nka, nk, n);
  }
  GMOD_CHECK_LAUNCH();
  o.download((zd*)chiqout, nka, s);
  GMOD_HIP(hipStreamSynchronize(s));
  GMOD_CATCH
}

// ---------------------------------------------------------------------- allocation-free variant
// A file-scope device array holds the workspace of gmod_eig_or_tr_chi_noalloc
// (its error flag, the per-k maxima and the interleaved Hessenberg matrices of
// k_max_re_eig).  In host code `scratch` names a host-side stand-in,
// so the kernels receive the device address that hipGetSymbolAddress returns,
// looked up once per device and kept.  Every call on a device shares that
// device's array: calls that overlap on one device race.
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

static unsigned char* noalloc_scratch_ptr() {
  static std::mutex mu;
  static std::vector<unsigned char*> per_dev;
  int dev = 0;
  GMOD_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> g(mu);
  if ((int)per_dev.size() <= dev) per_dev.resize(dev + 1, nullptr);
  if (!per_dev[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw Error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    per_dev[dev] = static_cast<unsigned char*>(p);
  }
  return per_dev[dev];
}

__global__ void k_zero_int(int* p) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0;
}

extern "C" int gmod_eig_or_tr_chi_noalloc(gmod_z* chiqout, const gmod_z* chi, const int64_t* invk, int64_t Nkall,
                                          int64_t Nk, int64_t Nchi, int sw_eig) {
  GMOD_TRY
  auto g = api_lock();
  hipStream_t s = ctx().stream;
  const int n = (int)Nchi;
  if (Nk < 1 || Nkall < 1 || n < 1) throw Error("gmod_eig_or_tr_chi_noalloc: empty batch");
  if (!chiqout || !chi || !invk) throw Error("gmod_eig_or_tr_chi_noalloc: null device pointer");
  // workspace: error flag | per-k maxima (Nk doubles) | Nk (n^2 + n) complex
  const size_t off_v = 256, off_h = off_v + (((size_t)Nk * sizeof(double) + 255) / 256) * 256;
  const size_t need = sw_eig ? off_h + (size_t)Nk * (size_t)(n * n + n) * sizeof(zd) : off_v;
  if (need > NOALLOC_SCRATCH_BYTES)
    throw Error("gmod_eig_or_tr_chi_noalloc: Nk = " + std::to_string(Nk) + ", Nchi = " + std::to_string(n) +
                " needs " + std::to_string(need) + " bytes of workspace, more than its " +
                std::to_string(NOALLOC_SCRATCH_BYTES) + "-byte maximum");
  unsigned char* ws = noalloc_scratch_ptr();
  int* err = reinterpret_cast<int*>(ws);
  if (sw_eig) {
    double* v = reinterpret_cast<double*>(ws + off_v);
    zd* h = reinterpret_cast<zd*>(ws + off_h);
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, s, err);
    GMOD_CHECK_LAUNCH();
    max_re_eig_ws(v, (const zd*)chi, BatchLayout{1, Nk}, Nk, n, nullptr, 1.0, h, err, s);
    hipLaunchKernelGGL(k_eig_map, dim3(256), dim3(256), 0, s, (zd*)chiqout, v, invk, (long)Nkall);
    GMOD_CHECK_LAUNCH();
    int herr = 0;
    GMOD_HIP(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, s));
    GMOD_HIP(hipStreamSynchronize(s));
    if (herr) throw Error("gmod_eig_or_tr_chi_noalloc: eigenvalue iteration did not converge");
  } else {
    hipLaunchKernelGGL(k_trace_map, dim3(256), dim3(256), 0, s, (zd*)chiqout, (const zd*)chi, invk, (long)Nkall,
                       (long)Nk, n);
    GMOD_CHECK_LAUNCH();
    GMOD_HIP(hipStreamSynchronize(s));
  }
  return 0;
  }
  catch (std::exception& e) { set_error(e.what()); return 1; }
  catch (...) { set_error("unknown error"); return 1; }
}

// ====================================================================== Green functions
extern "C" void gen_green0_(gmod_z* Gk, const double* eig, const gmod_z* uni, const double* mu, const double* temp,
                            const int64_t* Nk, const int64_t* Nw, const int64_t* Norb) {
  GMOD_TRY
  auto g = api_lock();
  hipStream_t s = ctx().stream;
  const int No = (int)*Norb;
  const size_t n = (size_t)*Nk * *Nw * No * No;
  Upload up{s};
  DBuf<double> e;
  DBuf<zd> u, G(n);
  up(e, eig, (size_t)No * *Nk);
  up(u, (const zd*)uni, (size_t)No * No * *Nk);
  green0(G.p, e.p, u.p, *mu, *temp, *Nk, (int)*Nw, No, s);
  G.download((zd*)Gk, n, s);
  GMOD_HIP(hipStreamSynchronize(s));
  GMOD_CATCH
}

extern "C" void gen_green_inv_(gmod_z* Gk, const gmod_z* self, const gmod_z* hamk, const double* mu,
                               const double* temp, const int64_t* Nk, const int64_t* Nw, const int64_t* Norb) {
  GMOD_TRY
  auto g = api_lock();
  hipStream_t s = ctx().stream;
  const int No = (int)*Norb;
  const size_t n = (size_t)*Nk * *Nw * No * No;
  Upload up{s};
  DBuf<zd> sg, h, G(n);
  up(sg, (const zd*)self, n);
  up(h, (const zd*)hamk, (size_t)No * No * *Nk);
  green_inv(G.p, sg.p, h.p, *mu, *temp, *Nk, (int)*Nw, No, s);
  G.download((zd*)Gk, n, s);
  GMOD_HIP(hipStreamSynchronize(s));
  GMOD_CATCH
}

extern "C" void getinv_(gmod_z* Gk, const int64_t* Nk, const int64_t* Nw, const int64_t* Norb) {
  GMOD_TRY
  auto g = api_lock();
  hipStream_t s = ctx().stream;
  const int No = (int)*Norb;
  const long tot = *Nk * *Nw;
  const size_t n = (size_t)tot * No * No;
  Upload up{s};
  DBuf<zd> G;
  DBuf<int> err(1);
  err.zero(s);
  up(G, (const zd*)Gk, n);
  batched_inv(G.p, tot, No, err.p, s);
  check_err(err, s, "getinv");
  G.download((zd*)Gk, n, s);
  GMOD_HIP(hipStreamSynchronize(s));
  GMOD_CATCH
}

// ====================================================================== FLEX
extern "C" void mkself(gmod_z* sigmak, double* mu, const double* Smat, const double* Cmat, const int64_t* kmap,
                       const int64_t* invk, const int64_t* olist, const gmod_z* hamk, const double* eig,
                       const gmod_z* uni, const double* mu_init, const double* rfill, const double* temp,
                       const int64_t* scf_loop, const double* pp, const double* eps, const int64_t* Nkall,
                       const int64_t* Nk, const int64_t* Nw, const int64_t* Norb, const int64_t* Nchi,
                       const int64_t* Nx, const int64_t* Ny, const int64_t* Nz, const int64_t* sub_sigma,
                       const bool* sw_out, const bool* sw_in, const int64_t* m_diis, const bool* sw_rescale,
                       const bool* sw_tail, const double* sigma_scale) {
  GMOD_TRY
  auto g = api_lock();
  FlexIn A;
  A.Smat = Smat; A.Cmat = Cmat; A.eig = eig;
  A.kmap = kmap; A.invk = invk; A.olist = olist;
  A.hamk = (const zd*)hamk; A.uni = (const zd*)uni;
  A.mu_init = *mu_init; A.rfill = *rfill; A.temp = *temp; A.pp = *pp; A.eps = *eps; A.sigma_scale = *sigma_scale;
  A.scf_loop = *scf_loop; A.Nkall = *Nkall; A.Nk = *Nk; A.Nw = *Nw; A.Nx = *Nx; A.Ny = *Ny; A.Nz = *Nz;
  A.sub_sigma = *sub_sigma; A.m_diis = *m_diis;
  A.No = (int)*Norb; A.Nchi = (int)*Nchi;
  A.sw_out = *sw_out; A.sw_in = *sw_in; A.sw_rescale = *sw_rescale; A.sw_tail = *sw_tail;
  if (ctx().precision == 1) mkself_driver<double>(A, (zd*)sigmak, mu, true);
  else if (ctx().precision == 2) mkself_driver<float>(A, (zd*)sigmak, mu, false);
  else mkself_driver<double>(A, (zd*)sigmak, mu, false);
  GMOD_CATCH
}

// ====================================================================== spinful basis (SOC)
// fchi0withsoc.f90:111-214: chi0 of a spinful basis with the sign tables
// sgnsig(Norb,Norb), sgnsig2(Nchi,Nchi) and invschi(Nchi) it returns
extern "C" void get_chi0_soc(gmod_z* chi, double* sgnsig, double* sgnsig2, int64_t* invschi, const double* Vmat,
                             const gmod_z* Gk, const int64_t* kmap, const int64_t* invk, const int64_t* invs,
                             const int64_t* olist, const int64_t* slist, const double* temp, const int64_t* Nx,
                             const int64_t* Ny, const int64_t* Nz, const int64_t* Nw, const int64_t* Nk,
                             const int64_t* Nkall, const int64_t* Nchi, const int64_t* Norb) {
  GMOD_TRY
  auto g = api_lock();
  const int No = (int)*Norb, n = (int)*Nchi;
  OrbPairs op;
  op.from_olist_soc(olist, n, No, invs, slist);
  for (int j = 0; j < No; ++j)
    for (int i = 0; i < No; ++i) sgnsig[i + No * j] = op.spin[i] * op.spin[j];
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) sgnsig2[i + n * j] = op.psign[i + n * j];
  for (int l = 0; l < n; ++l) {
    invschi[l] = 0;
    for (int m = 0; m < n; ++m)
      if (op.ol[m][0] == op.inv[op.ol[l][0]] && op.ol[m][1] == op.inv[op.ol[l][1]]) {
        invschi[l] = m + 1;
        break;
      }
  }
  double stoner = 0.0;
  double* st = ctx().verbose ? &stoner : nullptr;
  if (ctx().precision >= 1)
    chi0_driver<float>((zd*)chi, (const zd*)Gk, kmap, invk, op, *temp, *Nx, *Ny, *Nz, *Nw, *Nk, *Nkall, No, false,
                       nullptr, nullptr, 0.0, Vmat, st, -1.0);
  else
    chi0_driver<double>((zd*)chi, (const zd*)Gk, kmap, invk, op, *temp, *Nx, *Ny, *Nz, *Nw, *Nk, *Nkall, No, false,
                        nullptr, nullptr, 0.0, Vmat, st, -1.0);
  if (st) printf("SDW/CDW max eig(-chi0 V) = %12.8f\n", stoner);
  GMOD_CATCH
}

// fchi0withsoc.f90:281-326: charge, longitudinal and transverse spin
// susceptibilities (Nk, Nchi/4, Nchi/4) at the first boson frequency from
// X = (I + chi0 V)^-1 chi0, summed over the pairs orb_list maps together.
// Like the reference, the sums are added to the arrays passed in, which are
// then scaled (chic 1/2, chiszz 1/4).
extern "C" void get_chis_chic_soc(gmod_z* chic, gmod_z* chiszz, gmod_z* chispm, const gmod_z* chi,
                                  const double* Vmat, const int64_t* orb_list, const int64_t* olist,
                                  const int64_t* slist, const int64_t* invs, const int64_t* Nk, const int64_t* Nw,
                                  const int64_t* Nchi, const int64_t* Norb) {
  GMOD_TRY
  auto g = api_lock();
  (void)invs;
  (void)Norb;
  hipStream_t s = ctx().stream;
  const int n = (int)*Nchi, n4 = n / 4;
  const long nk = *Nk, NkNw = nk * *Nw;
  const size_t tot = (size_t)nk * n * n;
  for (int l = 0; l < n; ++l)
    if (orb_list[l] < 1 || orb_list[l] > n4) throw Error("get_chis_chic_soc: orb_list entry outside [1, Nchi/4]");
  Upload up{s};
  DBuf<zd> a, o(tot);
  DBuf<double> V;
  DBuf<int> err(1);
  err.zero(s);
  up(a, (const zd*)chi, (size_t)NkNw * n * n);
  up(V, Vmat, (size_t)n * n);
  rpa_solve(o.p, BatchLayout{1, nk}, a.p, BatchLayout{1, NkNw}, nk, n, V.p, 1.0, err.p, s);
  check_err(err, s, "get_chis_chic_soc");
  std::vector<zd> X(tot);
  o.download(X.data(), tot, s);
  GMOD_HIP(hipStreamSynchronize(s));
  auto sp = [&](int orb) { return slist[orb - 1]; };
  zd* cc = (zd*)chic;
  zd* zz = (zd*)chiszz;
  zd* pm = (zd*)chispm;
  for (long q = 0; q < nk; ++q)
    for (int l = 0; l < n; ++l) {
      const int64_t l1 = olist[l], l2 = olist[l + n];
      for (int m = 0; m < n; ++m) {
        const int64_t m1 = olist[m], m2 = olist[m + n];
        const zd x = X[q + nk * (m + (long)n * l)];
        const long o = q + nk * ((orb_list[l] - 1) + (long)n4 * (orb_list[m] - 1));
        if (sp(l1) == sp(m1) && sp(l2) == sp(m2) && sp(l1) == -sp(l2) && sp(l1) == 1) pm[o] += x;
        if (sp(l1) == sp(l2) && sp(m1) == sp(m2)) {
          cc[o] += x;
          if (sp(l1) == sp(m1)) zz[o] += x;
          else zz[o] -= x;
        }
      }
    }
  for (long e = 0; e < nk * n4 * n4; ++e) {
    cc[e] = cc[e] * 0.5;
    zz[e] = zz[e] * 0.25;
  }
  GMOD_CATCH
}

// fselfwithsoc.f90:1-539: the FLEX loop in a spinful basis.  Precision 0 runs
// FP64 throughout; 1 and 2 run the transforms in FP32 (full fields: the
// increment loop around a reference iterate covers the spinless basis only).
extern "C" void mkself_soc(gmod_z* sigmak, double* mu, const double* Vmat, const int64_t* kmap, const int64_t* invk,
                           const int64_t* invs, const int64_t* olist, const int64_t* slist, const gmod_z* hamk,
                           const double* eig, const gmod_z* uni, const double* mu_init, const double* rfill,
                           const double* temp, const int64_t* scf_loop, const double* pp, const double* eps,
                           const int64_t* Nkall, const int64_t* Nk, const int64_t* Nw, const int64_t* Norb,
                           const int64_t* Nchi, const int64_t* Nx, const int64_t* Ny, const int64_t* Nz,
                           const int64_t* sub_sigma, const bool* sw_out, const bool* sw_in, const int64_t* m_diis,
                           const bool* sw_rescale, const double* sigma_scale) {
  GMOD_TRY
  auto g = api_lock();
  FlexIn A;
  A.Smat = nullptr; A.Cmat = nullptr; A.eig = eig;
  A.kmap = kmap; A.invk = invk; A.olist = olist;
  A.hamk = (const zd*)hamk; A.uni = (const zd*)uni;
  A.mu_init = *mu_init; A.rfill = *rfill; A.temp = *temp; A.pp = *pp; A.eps = *eps; A.sigma_scale = *sigma_scale;
  A.scf_loop = *scf_loop; A.Nkall = *Nkall; A.Nk = *Nk; A.Nw = *Nw; A.Nx = *Nx; A.Ny = *Ny; A.Nz = *Nz;
  A.sub_sigma = *sub_sigma; A.m_diis = *m_diis;
  A.No = (int)*Norb; A.Nchi = (int)*Nchi;
  A.sw_out = *sw_out; A.sw_in = *sw_in; A.sw_rescale = *sw_rescale; A.sw_tail = false;
  A.soc = true; A.Vmat = Vmat; A.invs = invs; A.slist = slist;
  if (ctx().precision >= 1) mkself_driver<float>(A, (zd*)sigmak, mu, false);
  else mkself_driver<double>(A, (zd*)sigmak, mu, false);
  GMOD_CATCH
}

// feliashwithsoc.f90:1-424: the linearized gap equation in a spinful basis.
// delta(Nkall,Nw,Norb,Norb) on the full zone; chi leaves holding V_Delta.
// The eigenvalue (printed by the reference) is returned by gmod_last_lambda.
extern "C" void lin_eliash_soc(gmod_z* delta, gmod_z* chi, const gmod_z* Gk, const gmod_z* uni,
                               const double* init_delta, const double* Vmat, const double* sgnsig,
                               const double* sgnsig2, const double* prt, const int64_t* olist, const int64_t* slist,
                               const int64_t* kmap, const int64_t* invk, const int64_t* invs, const int64_t* invschi,
                               const double* temp, const double* eps, const int64_t* Nkall, const int64_t* Nk,
                               const int64_t* Nw, const int64_t* Nchi, const int64_t* Norb, const int64_t* Nx,
                               const int64_t* Ny, const int64_t* Nz, const int64_t* itemax, const int64_t* gap_sym,
                               const int64_t* arnoldi_m) {
  GMOD_TRY
  auto g = api_lock();
  (void)sgnsig2;  // the partner signs follow from slist
  (void)prt;
  EliashSocIn A;
  A.chi = (zd*)chi; A.Gk = (const zd*)Gk; A.uni = (const zd*)uni;
  A.init_delta = init_delta; A.Vmat = Vmat; A.sgnsig = sgnsig;
  A.olist = olist; A.slist = slist; A.kmap = kmap; A.invk = invk; A.invs = invs; A.invschi = invschi;
  A.temp = *temp; A.eps = *eps;
  A.Nkall = *Nkall; A.Nk = *Nk; A.Nw = *Nw; A.Nx = *Nx; A.Ny = *Ny; A.Nz = *Nz;
  A.itemax = *itemax; A.gap_sym = *gap_sym; A.arnoldi_m = *arnoldi_m;
  A.No = (int)*Norb; A.Nchi = (int)*Nchi;
  g_lambda_soc = ctx().precision >= 1 ? eliash_soc_driver<float>(A, (zd*)delta)
                                      : eliash_soc_driver<double>(A, (zd*)delta);
  GMOD_CATCH
}

extern "C" double gmod_last_lambda(void) { return g_lambda_soc; }

// feliashwithsoc.f90:426-518 (bind name mkfk_trs_soc_): F(Nkall,Nw,Norb,Norb)
extern "C" void mkfk_trs_soc_(gmod_z* fk, const gmod_z* Gk, const gmod_z* delta, const double* sgnsig,
                              const int64_t* slist, const int64_t* invk, const int64_t* invs, const int64_t* Nkall,
                              const int64_t* Nk, const int64_t* Nw, const int64_t* Norb, const int64_t* gap_sym) {
  GMOD_TRY
  auto g = api_lock();
  hipStream_t s = ctx().stream;
  const int No = (int)*Norb;
  const long nkall = *Nkall, nk = *Nk, nw = *Nw;
  const size_t nF = (size_t)nkall * nw * No * No;
  std::vector<int> slh(No), invh(No);
  for (int a = 0; a < No; ++a) {
    slh[a] = (int)slist[a];
    invh[a] = (int)invs[a] - 1;
    if (invh[a] < 0 || invh[a] >= No) throw Error("mkfk_trs_soc: invs entry outside [1,Norb]");
  }
  for (long i = 0; i < nkall; ++i)
    if (invk[3 * i] < 1 || invk[3 * i] > nk) throw Error("mkfk_trs_soc: invk(1,:) outside [1,Nk]");
  Upload up{s};
  DBuf<zd> G, D, F(nF);
  DBuf<double> sg;
  DBuf<int64_t> iv;
  DBuf<int> sl, inv;
  up(G, (const zd*)Gk, (size_t)nk * nw * No * No);
  up(D, (const zd*)delta, nF);
  up(sg, sgnsig, (size_t)No * No);
  up(iv, invk, (size_t)3 * nkall);
  up(sl, slh.data(), (size_t)No);
  up(inv, invh.data(), (size_t)No);
  anomalous_f_soc(F.p, G.p, D.p, iv.p, sg.p, sl.p, inv.p, nkall, nk, (int)nw, No, *gap_sym, s);
  F.download((zd*)fk, nF, s);
  GMOD_HIP(hipStreamSynchronize(s));
  GMOD_CATCH
}

// feliashwithsoc.f90:750-794: band-basis gap at the first frequency,
// deltab(Nkall,Norb,Norb) (host; output formatting)
extern "C" void conv_delta_orb_to_band_soc(gmod_z* deltab, const gmod_z* delta, const gmod_z* uni,
                                           const int64_t* invk, const int64_t* invs, const int64_t* slist,
                                           const int64_t* Norb, const int64_t* Nkall, const int64_t* Nk,
                                           const int64_t* Nw) {
  GMOD_TRY
  const int No = (int)*Norb;
  const long nkall = *Nkall;
  (void)Nw;
  const zd* D = (const zd*)delta;
  const zd* U = (const zd*)uni;
  zd* out = (zd*)deltab;
  std::vector<zd> L((size_t)No * No), Rm((size_t)No * No), T((size_t)No * No);
  for (long i = 0; i < nkall; ++i) {
    const long k = invk[3 * i] - 1, f = invk[3 * i + 1];
    if (f != 0 && f != 1) {
      for (int q = 0; q < No; ++q)
        for (int p = 0; p < No; ++p) out[i + nkall * (p + (long)No * q)] = zd(0, 0);
      continue;
    }
    if (k < 0 || k >= *Nk) throw Error("conv_delta_orb_to_band_soc: invk(1,:) outside [1,Nk]");
    auto u = [&](long a, long n) { return U[a + No * (n + (long)No * k)]; };
    for (int r = 0; r < No; ++r)
      for (int c = 0; c < No; ++c) {
        const long inv_c = invs[c] - 1, inv_r = invs[r] - 1;
        if (f == 0) {
          L[r + No * c] = conj(u(c, r));                  // U^dagger
          Rm[r + No * c] = u(inv_r, c) * (double)slist[r];  // s_n U(-n, :)
        } else {
          L[r + No * c] = u(inv_c, r) * (-(double)slist[c]);  // column n: -s_n U(-n, :)
          Rm[r + No * c] = conj(u(r, c));
        }
      }
    // deltab = L delta(i, w_1) Rm; delta(i, 1, a, q) at i + Nkall Nw (a + No q)
    for (int q = 0; q < No; ++q)
      for (int p = 0; p < No; ++p) {
        zd t(0, 0);
        for (int a = 0; a < No; ++a) t += L[p + No * a] * D[i + nkall * *Nw * (a + (long)No * q)];
        T[p + No * q] = t;
      }
    for (int q = 0; q < No; ++q)
      for (int p = 0; p < No; ++p) {
        zd t(0, 0);
        for (int b = 0; b < No; ++b) t += T[p + No * b] * Rm[b + No * q];
        out[i + nkall * (p + (long)No * q)] = t;
      }
  }
  GMOD_CATCH
}

extern "C" void gmod_set_sigma_seed(const gmod_z* sigma, double mu, double mu_old, int64_t count) {
  GMOD_TRY
  g_seed.sigma.assign((const zd*)sigma, (const zd*)sigma + count);
  g_seed.mu = mu;
  g_seed.mu_old = mu_old;
  g_seed.valid = true;
  GMOD_CATCH
}

extern "C" int gmod_flex_stats(double* out, int n) {
  int m = std::min(n, 12);
  std::copy(g_stats, g_stats + m, out);
  return m;
}

// ====================================================================== Eliashberg
extern "C" void lin_eliash(gmod_z* delta, gmod_z* chi, const gmod_z* Gk, const gmod_z* uni, const double* init_delta,
                           const double* Smat, const double* Cmat, const int64_t* olist, const double* prt,
                           const int64_t* kmap, const int64_t* invk, const double* temp, const double* eps,
                           const int64_t* Nkall, const int64_t* Nk, const int64_t* Nw, const int64_t* Nchi,
                           const int64_t* Norb, const int64_t* Nx, const int64_t* Ny, const int64_t* Nz,
                           const int64_t* itemax, const int64_t* gap_sym, const int64_t* arnoldi_m,
                           double* lambda_out) {
  GMOD_TRY
  auto g = api_lock();
  EliashIn A;
  A.chi = (zd*)chi; A.Gk = (const zd*)Gk; A.uni = (const zd*)uni;
  A.init_delta = init_delta; A.Smat = Smat; A.Cmat = Cmat; A.prt = prt;
  A.olist = olist; A.kmap = kmap; A.invk = invk;
  A.temp = *temp; A.eps = *eps;
  A.Nkall = *Nkall; A.Nk = *Nk; A.Nw = *Nw; A.Nx = *Nx; A.Ny = *Ny; A.Nz = *Nz;
  A.itemax = *itemax; A.gap_sym = *gap_sym; A.arnoldi_m = *arnoldi_m;
  A.No = (int)*Norb; A.Nchi = (int)*Nchi;
  *lambda_out = 0.0;
  if (ctx().precision >= 1) *lambda_out = eliash_driver<float>(A, (zd*)delta);
  else *lambda_out = eliash_driver<double>(A, (zd*)delta);
  GMOD_CATCH
}

extern "C" void eliashberg(gmod_z* delta, gmod_z* sigmak, gmod_z* Gk, const gmod_z* hamk, const double* Smat,
                           const double* Cmat, const int64_t* olist, const double* prt, const int64_t* kmap,
                           const int64_t* invk, const double* mu, const double* temp, const double* eps,
                           const int64_t* Nkall, const int64_t* Nk, const int64_t* Nw, const int64_t* Nchi,
                           const int64_t* Norb, const int64_t* Nx, const int64_t* Ny, const int64_t* Nz,
                           const int64_t* itemax, const int64_t* gap_sym, const bool* sw_sigma,
                           const bool* sw_Vconst, const int64_t* m_diis, const double* gap_min,
                           const bool* sw_amp_newton) {
  GMOD_TRY
  auto g = api_lock();
  NonlinIn A;
  A.Gk = (const zd*)Gk; A.hamk = (const zd*)hamk;
  A.Smat = Smat; A.Cmat = Cmat; A.prt = prt; A.olist = olist; A.kmap = kmap; A.invk = invk;
  A.mu = *mu; A.temp = *temp; A.eps = *eps; A.gap_min = *gap_min;
  A.Nkall = *Nkall; A.Nk = *Nk; A.Nw = *Nw; A.Nx = *Nx; A.Ny = *Ny; A.Nz = *Nz;
  A.itemax = *itemax; A.gap_sym = *gap_sym; A.m_diis = *m_diis;
  A.sw_sigma = *sw_sigma; A.sw_Vconst = *sw_Vconst; A.sw_amp_newton = *sw_amp_newton;
  A.No = (int)*Norb; A.Nchi = (int)*Nchi;
  if (ctx().precision >= 1) nonlin_eliash_driver<float>(A, (zd*)delta, (zd*)sigmak, (zd*)Gk);
  else nonlin_eliash_driver<double>(A, (zd*)delta, (zd*)sigmak, (zd*)Gk);
  GMOD_CATCH
}

extern "C" int gmod_eliashberg_stats(double* out, int n) {
  const int m = std::min(n, 4);
  std::copy(g_nonlin_stats, g_nonlin_stats + m, out);
  return m;
}

// the superconducting bubbles chi0_S = chi^GG + chi^FF, chi0_C = chi^GG - chi^FF
// (ffeliash.f90:459-489; its irr_chi / chi_map tables as get_chi0_conv_ takes them)
extern "C" void get_chi0sc_(gmod_z* Vsigma, gmod_z* Vdelta, const gmod_z* Gk, const gmod_z* fk, const int64_t* kmap,
                            const int64_t* invk, const int32_t* irr_chi, const int32_t* chi_map,
                            const int64_t* olist, const double* prt, const double* temp, const int64_t* Nx,
                            const int64_t* Ny, const int64_t* Nz, const int64_t* Nw, const int64_t* Nk,
                            const int64_t* Nkall, const int64_t* Norb, const int64_t* Nchi, const bool* sw_pair) {
  GMOD_TRY
  auto g = api_lock();
  hipStream_t s = ctx().stream;
  const int No = (int)*Norb, Nc = (int)*Nchi;
  const size_t nS = (size_t)*Nk * *Nw * No * No, nC = (size_t)*Nk * *Nw * Nc * Nc;
  OrbPairs op;
  op.from_olist(olist, Nc, No);
  op.from_maps(irr_chi, chi_map);
  Mesh mesh;
  mesh.build(kmap, invk, *Nkall, *Nk, (int)*Nx, (int)*Ny, (int)*Nz, (int)*Nw, s);
  Upload up{s};
  DBuf<zd> G, F, cs(nC), cc(nC);
  up(G, (const zd*)Gk, nS);
  up(F, (const zd*)fk, nS);
  auto run = [&](auto eng) {
    eng->transform_green(G.p, false, s);
    eng->chi0(cs.p, *temp, false, nullptr, nullptr, nullptr, nullptr, nullptr, 0.0, s, true);
    eng->chi0_ff(cc.p, F.p, prt, *sw_pair, *temp, s);
  };
  if (ctx().precision >= 1) run(std::make_unique<ConvEngine<float>>(mesh, op));
  else run(std::make_unique<ConvEngine<double>>(mesh, op));
  sum_diff(cs.p, cc.p, (long)nC, s);
  cs.download((zd*)Vsigma, nC, s);
  cc.download((zd*)Vdelta, nC, s);
  GMOD_HIP(hipStreamSynchronize(s));
  GMOD_CATCH
}

// the vertices of the superconducting loop in place (ffeliash.f90:386-457):
// Vsigma chi0_S -> V_sigma, Vdelta chi0_C -> V_Delta
extern "C" void mkv_flex_nosoc_(gmod_z* Vdelta, gmod_z* Vsigma, const double* Smat, const double* Cmat,
                                const int64_t* Nk, const int64_t* Nw, const int64_t* Nchi, const bool* sw_pair) {
  GMOD_TRY
  auto g = api_lock();
  hipStream_t s = ctx().stream;
  const long NkNw = *Nk * *Nw;
  const int n = (int)*Nchi;
  const size_t nC = (size_t)NkNw * n * n;
  Upload up{s};
  DBuf<zd> vs, vd;
  DBuf<double> S, C;
  DBuf<int> err(1);
  err.zero(s);
  up(vs, (const zd*)Vsigma, nC);
  up(vd, (const zd*)Vdelta, nC);
  up(S, Smat, (size_t)n * n);
  up(C, Cmat, (size_t)n * n);
  sc_vertex(vs.p, vd.p, NkNw, BatchLayout{1, NkNw}, n, S.p, C.p, *sw_pair, err.p, s);
  check_err(err, s, "mkV_flex_nosoc");
  vs.download((zd*)Vsigma, nC, s);
  vd.download((zd*)Vdelta, nC, s);
  GMOD_HIP(hipStreamSynchronize(s));
  GMOD_CATCH
}

extern "C" void mkfk_trs_nsoc_(gmod_z* fk, const gmod_z* Gk, const gmod_z* delta, const int64_t* Nk,
                               const int64_t* Nw, const int64_t* Norb) {
  GMOD_TRY
  auto g = api_lock();
  hipStream_t s = ctx().stream;
  const int No = (int)*Norb;
  const size_t n = (size_t)*Nk * *Nw * No * No;
  Upload up{s};
  DBuf<zd> G, D, F(n);
  up(G, (const zd*)Gk, n);
  up(D, (const zd*)delta, n);
  anomalous_f(F.p, G.p, D.p, *Nk, (int)*Nw, No, s);
  F.download((zd*)fk, n, s);
  GMOD_HIP(hipStreamSynchronize(s));
  GMOD_CATCH
}

extern "C" void get_initial_delta_(gmod_z* delta, const double* init_delta, const gmod_z* uni, const int64_t* kmap,
                                   const int64_t* invk, const int64_t* Nkall, const int64_t* Nk, const int64_t* Nw,
                                   const int64_t* Norb, const int64_t* gap_sym) {
  GMOD_TRY
  auto g = api_lock();
  hipStream_t s = ctx().stream;
  const int No = (int)*Norb;
  // only the multiplicities are needed from the mesh
  Mesh mesh;
  long Nkall_ = *Nkall;
  long nx = Nkall_, ny = 1, nz = 1;
  std::vector<int64_t> km(3 * Nkall_);
  for (long i = 0; i < Nkall_; ++i) { km[3 * i] = i; km[3 * i + 1] = 0; km[3 * i + 2] = 0; }
  (void)kmap;
  mesh.build(km.data(), invk, Nkall_, *Nk, (int)nx, (int)ny, (int)nz, (int)*Nw, s);
  const size_t n = (size_t)*Nk * *Nw * No * No;
  DBuf<zd> D(n);
  initial_delta(D.p, init_delta, (const zd*)uni, mesh, No, *gap_sym, s);
  D.download((zd*)delta, n, s);
  GMOD_HIP(hipStreamSynchronize(s));
  GMOD_CATCH
}

// band-basis gap at w_0 on the full BZ (feliash.f90:733-785): output formatting, host
extern "C" void conv_delta_orb_to_band(gmod_z* deltab, const gmod_z* delta, const gmod_z* uni, const double* prt,
                                       const int64_t* invk, const int64_t* Norb, const int64_t* Nkall,
                                       const int64_t* Nk, const int64_t* Nw, const int64_t* gap_sym) {
  GMOD_TRY
  const int No = (int)*Norb;
  const long nka = *Nkall, nk = *Nk, NkNw = *Nk * *Nw;
  const double dprt = *gap_sym < 0 ? -1.0 : 1.0;
  const zd* D = (const zd*)delta;
  const zd* U = (const zd*)uni;
  zd* out = (zd*)deltab;
  std::vector<zd> t(No * No);
  for (long i = 0; i < nka; ++i) {
    long k = invk[3 * i] - 1;
    auto Uk = [&](int a, int b) { return U[a + No * (b + (long)No * k)]; };
    auto Dk = [&](int a, int b) { return D[k + NkNw * (a + (long)No * b)]; };
    for (int p = 0; p < No; ++p)
      for (int q = 0; q < No; ++q) out[i + nka * (p + (long)No * q)] = zd(0, 0);
    if (invk[3 * i + 1] == 0) {
      for (int p = 0; p < No; ++p)
        for (int q = 0; q < No; ++q) {
          zd acc(0, 0);
          for (int r = 0; r < No; ++r) acc += conj(Uk(r, p)) * Dk(r, q);
          t[p + No * q] = acc;
        }
      for (int p = 0; p < No; ++p)
        for (int q = 0; q < No; ++q) {
          zd acc(0, 0);
          for (int r = 0; r < No; ++r) acc += t[p + No * r] * Uk(r, q);
          out[i + nka * (p + (long)No * q)] = acc;
        }
    } else if (invk[3 * i + 1] == 1) {
      for (int p = 0; p < No; ++p)
        for (int q = 0; q < No; ++q) {
          zd acc(0, 0);
          for (int r = 0; r < No; ++r) acc += Uk(r, p) * prt[r] * Dk(r, q);
          t[p + No * q] = acc * (dprt * prt[q]);
        }
      for (int p = 0; p < No; ++p)
        for (int q = 0; q < No; ++q) {
          zd acc(0, 0);
          for (int r = 0; r < No; ++r) acc += t[p + No * r] * conj(Uk(r, q));
          out[i + nka * (p + (long)No * q)] = acc;
        }
    }
  }
  (void)nk;
  GMOD_CATCH
}

extern "C" void remap_delta_(gmod_z* delta, const gmod_z* delta0, const double* prt, const int64_t* invk,
                             const int64_t* Nkall, const int64_t* Nk, const int64_t* Nw, const int64_t* Norb,
                             const int64_t* gap_sym) {
  GMOD_TRY
  const int No = (int)*Norb;
  const long nka = *Nkall, nk = *Nk, nw = *Nw;
  const double dprt = *gap_sym < 0 ? -1.0 : 1.0;
  const zd* D0 = (const zd*)delta0;
  zd* D = (zd*)delta;
  for (int l = 0; l < No; ++l)
    for (int m = 0; m < No; ++m)
      for (long j = 0; j < nw; ++j)
        for (long i = 0; i < nka; ++i) {
          long k = invk[3 * i] - 1;
          zd v = D0[k + nk * (j + nw * (m + (long)No * l))];
          if (invk[3 * i + 1] == 1) v = v * (dprt * prt[m] * prt[l]);
          if (invk[3 * i + 1] == 0 || invk[3 * i + 1] == 1) D[i + nka * (j + nw * (m + (long)No * l))] = v;
        }
  GMOD_CATCH
}
