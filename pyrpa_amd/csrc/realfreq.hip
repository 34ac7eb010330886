// Real-frequency susceptibility maps and Pade analytic continuation on the
// device: the band-sum chi0(q, w + i delta) of libs/src/fchi.f90:94-317
// (calc_chi, get_chi_irr, chiq_map, get_tr_chi), its k + q table
// (get_qshift, fchi.f90:1-92), and the continued-fraction Pade of
// libs/src/fpade.f90 (get_a, get_QP, pade_analytic_continuation_arrays,
// pade_with_trace).
//
// chi0 at a point (q, w) is a sum over k and the band pairs (l at k+q, m at
// k) of rank-one terms weight * B (x) A with B_i = conj(U_{o_i1,l}(k+q))
// U_{o_i2,m}(k), A_j = U_{o_j1,l}(k+q) conj(U_{o_j2,m}(k)).  One work-group
// owns one point: per k it stages the eigenvectors of k and k+q in LDS,
// forms weight * B and A for every band pair once, and each thread adds the
// products into the chi_{ij} elements it owns, in registers.  Points are
// (one q, many w) for get_chi_irr and (many q, one w) for chiq_map.
#include <cmath>
#include <unordered_map>

#include "../../include/gmod.h"
#include "gmod_common.h"
#include "linalg.h"

namespace gmod {

constexpr int RF_THREADS = 256;
constexpr int RF_MAXE = 4;      // chi elements per thread: Nchi^2 <= 1024
constexpr int RF_MAXNO = 8;     // orbitals
constexpr long RF_QSCALE = 1048576;  // k coordinates quantised to 2^20 steps (fchi.f90:22-23)

// ------------------------------------------------------------------ k + q table
// Key of a fractional k point: each coordinate int(x 2^20 + 0.5) mod 2^20.
static inline long long rf_key(const double x[3]) {
  long long k = 0, mul = 1;
  for (int d = 0; d < 3; ++d) {
    long long i = (long long)(x[d] * (double)RF_QSCALE + 0.5);
    i %= RF_QSCALE;
    if (i < 0) i += RF_QSCALE;
    k += i * mul;
    mul *= RF_QSCALE;
  }
  return k;
}

// 0-based index of k + q (periodic in [0,1)) for every k of klist (3, Nk);
// returns the number of shifted points not on the mesh (index 0 then, the
// reference's fallback)
static long qshift_host(const double* klist, const double* q, long Nk, int64_t* out0) {
  std::unordered_map<long long, long> at;
  at.reserve((size_t)Nk * 2);
  for (long j = 0; j < Nk; ++j) at.emplace(rf_key(klist + 3 * j), j);
  long miss = 0;
  for (long i = 0; i < Nk; ++i) {
    double t[3];
    for (int d = 0; d < 3; ++d) {
      t[d] = klist[3 * i + d] + q[d];
      if (t[d] >= 1.0) t[d] -= 1.0;
      else if (t[d] < 0.0) t[d] += 1.0;
    }
    auto it = at.find(rf_key(t));
    if (it == at.end()) { ++miss; out0[i] = 0; }
    else out0[i] = it->second;
  }
  return miss;
}

// ------------------------------------------------------------------ band sum
// out[pt][i + Nchi j] = chi_{ij}(q_pt, w_pt) / Nk.  qs[qi[pt] Nk + k]: 0-based
// index of k + q.  uni (No, No, Nk) = uni(orb, band, k); eig, ff (No, Nk);
// ol (Nchi, 2) 0-based.
__global__ void __launch_bounds__(RF_THREADS)
k_chi_band(zd* out, const zd* __restrict__ uni, const double* __restrict__ eig, const double* __restrict__ ff,
           const int* __restrict__ qs, const int* __restrict__ qi, const double* __restrict__ wpt,
           const int* __restrict__ ol, int Nchi, int No, long Nk, double idelta, double eps, double temp) {
  extern __shared__ __align__(16) unsigned char rf_smem[];
  const int NN = No * No, nc2 = Nchi * Nchi;
  zd* Uq = reinterpret_cast<zd*>(rf_smem);         // U(k+q): NN
  zd* Uk = Uq + NN;                                 // U(k):   NN
  zd* WB = Uk + NN;                                 // weight B_i per band pair: NN Nchi
  zd* AV = WB + NN * Nchi;                          // A_j per band pair:        NN Nchi
  double* Eq = reinterpret_cast<double*>(AV + NN * Nchi);
  double* Ek = Eq + No;
  double* Fq = Ek + No;
  double* Fk = Fq + No;
  int* ols = reinterpret_cast<int*>(Fk + No);       // 2 Nchi
  const int pt = blockIdx.x, tid = threadIdx.x;
  const double w = wpt[pt];
  const int* qrow = qs + (long)qi[pt] * Nk;
  const double temp_safe = fmax(temp, 1.0e-12), w_eps = 1.0e-12;
  for (int e = tid; e < 2 * Nchi; e += RF_THREADS) ols[e] = ol[e];
  zd acc[RF_MAXE];
#pragma unroll
  for (int r = 0; r < RF_MAXE; ++r) acc[r] = zd(0, 0);
  for (long k = 0; k < Nk; ++k) {
    const long kq = qrow[k];
    __syncthreads();  // the previous k's terms are consumed
    for (int e = tid; e < NN; e += RF_THREADS) {
      Uq[e] = uni[e + (long)NN * kq];
      Uk[e] = uni[e + (long)NN * k];
    }
    for (int e = tid; e < No; e += RF_THREADS) {
      Eq[e] = eig[e + (long)No * kq];
      Ek[e] = eig[e + (long)No * k];
      Fq[e] = ff[e + (long)No * kq];
      Fk[e] = ff[e + (long)No * k];
    }
    __syncthreads();
    // weight B and A of every band pair (l at k+q, m at k), fchi.f90:131-155
    for (int e = tid; e < NN * Nchi; e += RF_THREADS) {
      const int c = e % Nchi, lm = e / Nchi;
      const int l = lm % No, m = lm / No;
      zd wt(0, 0);
      if (fabs(w) < w_eps && fabs(Ek[m] - Eq[l]) < 1.0e-9) {
        wt = zd(Fk[m] * (1.0 - Fk[m]) / temp_safe, 0.0);
      } else if (fabs(Fq[l] - Fk[m]) > eps) {
        wt = zd(Fq[l] - Fk[m], 0.0) / zd(w + Ek[m] - Eq[l], idelta);
      }
      const int o1 = ols[c], o2 = ols[c + Nchi];
      const zd uq = Uq[o1 + No * l], uk = Uk[o2 + No * m];
      WB[lm * Nchi + c] = wt * (conj(uq) * uk);
      AV[lm * Nchi + c] = uq * conj(uk);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RF_MAXE; ++r) {
      const int e = tid + r * RF_THREADS;
      if (e < nc2) {
        const int i = e % Nchi, j = e / Nchi;
        zd s = acc[r];
        for (int lm = 0; lm < NN; ++lm) s += WB[lm * Nchi + i] * AV[lm * Nchi + j];
        acc[r] = s;
      }
    }
  }
  const double inv = 1.0 / (double)Nk;
#pragma unroll
  for (int r = 0; r < RF_MAXE; ++r) {
    const int e = tid + r * RF_THREADS;
    if (e < nc2) out[(long)pt * nc2 + e] = acc[r] * inv;
  }
}

static size_t chi_band_lds(int Nchi, int No) {
  const int NN = No * No;
  return (size_t)(2 * NN + 2 * NN * Nchi) * sizeof(zd) + 4 * No * sizeof(double) + 2 * Nchi * sizeof(int);
}

struct BandInputs {
  DBuf<zd> uni;
  DBuf<double> eig, ff;
  DBuf<int> ol;
};

static void check_band_shapes(long Nk, int No, int Nchi) {
  if (Nk < 1 || No < 1 || Nchi < 1) throw Error("band sum: empty input");
  if (No > RF_MAXNO) throw Error("band sum: more than " + std::to_string(RF_MAXNO) + " orbitals");
  if ((long)Nchi * Nchi > (long)RF_THREADS * RF_MAXE) throw Error("band sum: Nchi too large");
  if (chi_band_lds(Nchi, No) > 64 * 1024) throw Error("band sum: LDS staging too large");
}

static void upload_band(BandInputs& b, const zd* uni, const double* eig, const double* ff, const int64_t* ol,
                        long Nk, int No, int Nchi, hipStream_t s) {
  b.uni.alloc((size_t)No * No * Nk);
  b.uni.upload(uni, (size_t)No * No * Nk, s);
  b.eig.alloc((size_t)No * Nk);
  b.eig.upload(eig, (size_t)No * Nk, s);
  b.ff.alloc((size_t)No * Nk);
  b.ff.upload(ff, (size_t)No * Nk, s);
  std::vector<int> o(2 * Nchi);
  for (int c = 0; c < 2 * Nchi; ++c) {
    o[c] = (int)ol[c] - 1;
    if (o[c] < 0 || o[c] >= No) throw Error("band sum: olist entry outside [1, Norb]");
  }
  b.ol.alloc(2 * Nchi);
  b.ol.upload(o.data(), 2 * Nchi, s);
  GMOD_HIP(hipStreamSynchronize(s));  // o is released at return
}

static void chi_band(zd* out_d, const BandInputs& b, const DBuf<int>& qs, const std::vector<int>& qi,
                     const std::vector<double>& w, int Nchi, int No, long Nk, double idelta, double eps, double temp,
                     hipStream_t s) {
  const long npt = (long)w.size();
  DBuf<int> qi_d(npt);
  DBuf<double> w_d(npt);
  qi_d.upload(qi.data(), npt, s);
  w_d.upload(w.data(), npt, s);
  hipLaunchKernelGGL(k_chi_band, dim3((unsigned)npt), dim3(RF_THREADS), chi_band_lds(Nchi, No), s, out_d, b.uni.p,
                     b.eig.p, b.ff.p, qs.p, qi_d.p, w_d.p, b.ol.p, Nchi, No, Nk, idelta, eps, temp);
  GMOD_CHECK_LAUNCH();
  GMOD_HIP(hipStreamSynchronize(s));  // qi, w are released at return
}

// sum over the orbital-diagonal pairs (o1 == o2) of both indices of a batch of
// chi matrices (fchi.f90:296-310): tr[pt] = sum_{l,m diagonal} chi(m, l)
__global__ void k_diag_trace(zd* tr, const zd* chi, long npt, int Nchi, const int* __restrict__ ol) {
  for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < npt; p += (long)gridDim.x * blockDim.x) {
    zd t(0, 0);
    for (int l = 0; l < Nchi; ++l) {
      if (ol[l] != ol[l + Nchi]) continue;
      for (int m = 0; m < Nchi; ++m)
        if (ol[m] == ol[m + Nchi]) t += chi[p * Nchi * Nchi + m + (long)Nchi * l];
    }
    tr[p] = t;
  }
}

// ------------------------------------------------------------------ Pade
// One wavefront per array: the continued-fraction coefficients a_i by the
// recursion of fpade.f90:1-42 (all data points in parallel per step, in
// LDS), then P_N / Q_N at every target frequency by the three-term
// recursion of fpade.f90:44-85, one frequency per lane.  in(Nk, Np) array i
// at in[i + Nk n]; out[w + ow i] (+)= P/Q (acc: add).
constexpr int PADE_MAXP = 1024;

__global__ void __launch_bounds__(64)
k_pade(zd* out, long ow, long oi, int acc, const zd* __restrict__ in, long in_stride_n, long in_off,
       const zd* __restrict__ xn, const zd* __restrict__ wl, long Nk, int Np, long Nw, zd* coef_out) {
  __shared__ double2 g_[PADE_MAXP], a_[PADE_MAXP], x_[PADE_MAXP];
  zd* g = reinterpret_cast<zd*>(g_);
  zd* a = reinterpret_cast<zd*>(a_);
  zd* x = reinterpret_cast<zd*>(x_);
  const long i = blockIdx.x;
  if (i >= Nk) return;
  const int lane = threadIdx.x;
  for (int j = lane; j < Np; j += 64) {
    g[j] = in[in_off + i + in_stride_n * j];
    x[j] = xn[j];
  }
  __syncthreads();
  if (lane == 0) a[0] = g[0];
  for (int s = 1; s < Np; ++s) {
    const zd gp = g[s - 1];
    __syncthreads();  // every lane has read g[s-1] before any lane rewrites g
    for (int j = lane; j < Np; j += 64) {
      const zd gj = g[j];
      const bool zero = gj.re == 0.0 && gj.im == 0.0;
      g[j] = zero ? zd(0, 0) : (gp - gj) / ((x[j] - x[s - 1]) * gj);
    }
    __syncthreads();
    if (lane == 0) a[s] = g[s];
  }
  __syncthreads();
  if (coef_out)
    for (int j = lane; j < Np; j += 64) coef_out[j] = a[j];
  for (long w = lane; w < Nw; w += 64) {
    const zd z = wl[w];
    zd P0(0, 0), P1 = a[0], Q0(1, 0), Q1(1, 0), P = P1, Q = Q1;
    for (int s = 1; s < Np; ++s) {
      const zd f = a[s] * (z - x[s - 1]);
      P = P1 + f * P0;
      Q = Q1 + f * Q0;
      P0 = P1; P1 = P; Q0 = Q1; Q1 = Q;
    }
    zd* o = out + w * ow + i * oi;
    const zd v = P / Q;
    *o = acc ? *o + v : v;
  }
}

}  // namespace gmod

using namespace gmod;

#define RF_TRY try { std::lock_guard<std::mutex> g_(ctx().lock);
#define RF_CATCH } catch (std::exception& e) { set_error(e.what()); } catch (...) { set_error("unknown error"); }

// ====================================================================== k + q
extern "C" void get_qshift_(const double* qpoint, const double* klist, int64_t* qshift, const int64_t* Nk) {
  try {
    const long nk = *Nk;
    std::vector<int64_t> q0(nk);
    const long miss = qshift_host(klist, qpoint, nk, q0.data());
    for (long i = 0; i < nk; ++i) qshift[i] = q0[i] + 1;
    if (miss > 0)
      fprintf(stderr, "WARNING(get_qshift): %ld of %ld k+q points not found on the k-mesh; they fall back to "
                      "index 1\n", miss, nk);
  } catch (std::exception& e) { set_error(e.what()); }
}

// ====================================================================== chi0(q, w)
extern "C" void get_chi_irr(gmod_z* chi, const gmod_z* uni, const double* eig, const double* ffermi,
                            const int64_t* qshift, const int64_t* ol, const double* wl, const int64_t* Nchi,
                            const int64_t* Norb, const int64_t* Nk, const int64_t* Nw, const double* idelta,
                            const double* eps, const double* temp) {
  RF_TRY
  hipStream_t s = ctx().stream;
  const int nchi = (int)*Nchi, no = (int)*Norb;
  const long nk = *Nk, nw = *Nw;
  check_band_shapes(nk, no, nchi);
  if (nw < 1) return;
  BandInputs b;
  upload_band(b, (const zd*)uni, eig, ffermi, ol, nk, no, nchi, s);
  std::vector<int> q0(nk);
  for (long k = 0; k < nk; ++k) {
    q0[k] = (int)(qshift[k] - 1);
    if (q0[k] < 0 || q0[k] >= nk) throw Error("get_chi_irr: qshift entry outside [1, Nk]");
  }
  DBuf<int> qs(nk);
  qs.upload(q0.data(), nk, s);
  DBuf<zd> out((size_t)nw * nchi * nchi);
  chi_band(out.p, b, qs, std::vector<int>(nw, 0), std::vector<double>(wl, wl + nw), nchi, no, nk, *idelta, *eps,
           *temp, s);
  out.download((zd*)chi, (size_t)nw * nchi * nchi, s);
  GMOD_HIP(hipStreamSynchronize(s));
  RF_CATCH
}

extern "C" void chiq_map(gmod_z* trchis, gmod_z* trchi, const gmod_z* uni, const double* eig, const double* ffermi,
                         const double* klist, const double* Smat, const int64_t* ol, const double* temp,
                         const double* ecut, const double* idelta, const double* eps, const int64_t* Nx,
                         const int64_t* Ny, const int64_t* Nk, const int64_t* Norb, const int64_t* Nchi) {
  RF_TRY
  hipStream_t s = ctx().stream;
  const int nchi = (int)*Nchi, no = (int)*Norb;
  const long nk = *Nk, nx = *Nx, ny = *Ny, npt = nx * ny;
  check_band_shapes(nk, no, nchi);
  if (npt < 1) return;
  BandInputs b;
  upload_band(b, (const zd*)uni, eig, ffermi, ol, nk, no, nchi, s);
  // k + q of every q of the (qx, qy) plane, q = ((i-1)/Nx, (j-1)/Ny, 0);
  // point p = (i-1) Ny + (j-1), the memory order of the outputs (Ny, Nx)
  std::vector<int> qsh((size_t)npt * nk);
  std::vector<int64_t> row(nk);
  long miss = 0;
  for (long i = 0; i < nx; ++i)
    for (long j = 0; j < ny; ++j) {
      const double q[3] = {(double)i / (double)nx, (double)j / (double)ny, 0.0};
      miss += qshift_host(klist, q, nk, row.data());
      for (long k = 0; k < nk; ++k) qsh[(size_t)(i * ny + j) * nk + k] = (int)row[k];
    }
  if (miss > 0)
    fprintf(stderr, "WARNING(get_qshift): %ld k+q points of the q map not found on the k-mesh\n", miss);
  DBuf<int> qs(qsh.size());
  qs.upload(qsh.data(), qsh.size(), s);
  std::vector<int> qi(npt);
  for (long p = 0; p < npt; ++p) qi[p] = (int)p;
  DBuf<zd> c0((size_t)npt * nchi * nchi), cs((size_t)npt * nchi * nchi);
  chi_band(c0.p, b, qs, qi, std::vector<double>(npt, *ecut), nchi, no, nk, *idelta, *eps, *temp, s);
  // RPA: chi_s = (I - chi0 S)^-1 chi0 per q (fchi.f90:286-294)
  DBuf<double> S_d((size_t)nchi * nchi);
  S_d.upload(Smat, (size_t)nchi * nchi, s);
  DBuf<int> err(1);
  err.zero(s);
  const long nn = (long)nchi * nchi;
  rpa_solve(cs.p, BatchLayout{nn, 1}, c0.p, BatchLayout{nn, 1}, npt, nchi, S_d.p, -1.0, err.p, s);
  int herr = 0;
  err.download(&herr, 1, s);
  GMOD_HIP(hipStreamSynchronize(s));
  if (herr) throw Error("chiq_map: zgesv failed (singular RPA denominator)");
  DBuf<zd> t0(npt), ts(npt);
  hipLaunchKernelGGL(k_diag_trace, dim3(256), dim3(256), 0, s, t0.p, c0.p, npt, nchi, b.ol.p);
  hipLaunchKernelGGL(k_diag_trace, dim3(256), dim3(256), 0, s, ts.p, cs.p, npt, nchi, b.ol.p);
  GMOD_CHECK_LAUNCH();
  t0.download((zd*)trchi, npt, s);
  ts.download((zd*)trchis, npt, s);
  GMOD_HIP(hipStreamSynchronize(s));
  RF_CATCH
}

// output formatting of get_chi_irr results (fchi.f90:160-204), host
extern "C" void get_tr_chi(gmod_z* trchis, gmod_z* trchi0, gmod_z* chis_orb, const gmod_z* chis, const gmod_z* chi0,
                           const int64_t* olist, const int64_t* Nw, const int64_t* Nchi, const int64_t* Norb) {
  try {
    const long nw = *Nw, n = *Nchi, no = *Norb;
    const zd* S = (const zd*)chis;
    const zd* C = (const zd*)chi0;
    zd* ts = (zd*)trchis;
    zd* t0 = (zd*)trchi0;
    zd* so = (zd*)chis_orb;
    for (long i = 0; i < nw; ++i) { ts[i] = zd(0, 0); t0[i] = zd(0, 0); }
    for (long e = 0; e < (no + 2) * nw; ++e) so[e] = zd(0, 0);
    for (long i = 0; i < nw; ++i)
      for (long j = 0; j < n; ++j) {
        if (olist[j] != olist[j + n]) continue;
        for (long k = 0; k < n; ++k) {
          if (olist[k] != olist[k + n]) continue;
          const long e = k + n * (j + n * i);
          ts[i] += S[e];
          t0[i] += C[e];
          if (olist[j] == olist[k]) so[(olist[j] - 1) + (no + 2) * i] = S[e];
          if (olist[k] == 2 && olist[j] == 3) so[no + (no + 2) * i] = S[e];
          if (olist[k] == 2 && olist[j] == 4) so[no + 1 + (no + 2) * i] = S[e];
        }
      }
  } catch (std::exception& e) { set_error(e.what()); }
}

// ====================================================================== Pade
static void pade_run(zd* out_h, long ow, long oi, bool acc_orbs, const zd* in_h, long Nk, long Np, long Nw,
                     long norb, const zd* xn_h, const zd* wl_h, zd* coef_h) {
  hipStream_t s = ctx().stream;
  if (Np < 1 || Np > PADE_MAXP) throw Error("pade: the number of data points must lie in [1, 1024]");
  if (Nk < 1 || Nw < 1) return;
  const size_t nin = (size_t)Nk * Np * norb * norb;
  DBuf<zd> in(nin), xn(Np), wl(Nw), out((size_t)Nk * Nw), coef(coef_h ? Np : 0);
  in.upload(in_h, nin, s);
  xn.upload(xn_h, Np, s);
  wl.upload(wl_h, Nw, s);
  for (long o = 0; o < norb; ++o) {
    // arrayin(Nk, Np, o, o): element (i, n) at i + Nk n + Nk Np (o + norb o)
    const long off = (long)Nk * Np * (o + norb * o);
    hipLaunchKernelGGL(k_pade, dim3((unsigned)Nk), dim3(64), 0, s, out.p, ow, oi, (acc_orbs && o > 0) ? 1 : 0, in.p,
                       (long)Nk, off, xn.p, wl.p, Nk, (int)Np, Nw, coef_h ? coef.p : nullptr);
    GMOD_CHECK_LAUNCH();
  }
  out.download(out_h, (size_t)Nk * Nw, s);
  if (coef_h) coef.download(coef_h, Np, s);
  GMOD_HIP(hipStreamSynchronize(s));
}

extern "C" void get_a_(gmod_z* a, const gmod_z* xn, const gmod_z* inp_data, const int64_t* Np) {
  RF_TRY
  // one array; the coefficients are the output (the evaluation at xn[0] is discarded)
  zd dummy;
  pade_run(&dummy, 1, 1, false, (const zd*)inp_data, 1, *Np, 1, 1, (const zd*)xn, (const zd*)xn, (zd*)a);
  RF_CATCH
}

extern "C" void get_qp_(gmod_z* P, gmod_z* Q, const gmod_z* a, const gmod_z* xn, const gmod_z* wlist,
                        const int64_t* Nw, const int64_t* Np) {
  try {
    // the three-term recursion of fpade.f90:44-85 for given coefficients (host)
    const long nw = *Nw, np = *Np;
    const zd* A = (const zd*)a;
    const zd* X = (const zd*)xn;
    const zd* W = (const zd*)wlist;
    for (long j = 0; j < nw; ++j) {
      zd P0(0, 0), P1 = np > 0 ? A[0] : zd(0, 0), Q0(1, 0), Q1(1, 0), Pv = P1, Qv = Q1;
      for (long i = 1; i < np; ++i) {
        const zd f = A[i] * (W[j] - X[i - 1]);
        Pv = P1 + f * P0;
        Qv = Q1 + f * Q0;
        P0 = P1; P1 = Pv; Q0 = Q1; Q1 = Qv;
      }
      ((zd*)P)[j] = Pv;
      ((zd*)Q)[j] = Qv;
    }
  } catch (std::exception& e) { set_error(e.what()); }
}

extern "C" void pade_analytic_continuation_arrays_(const gmod_z* arrayin, gmod_z* arrayout, const gmod_z* iwlist,
                                                   const gmod_z* wlist, const int64_t* Nk, const int64_t* Niw,
                                                   const int64_t* Nw) {
  RF_TRY
  // arrayout(Nk, Nw): element (i, w) at i + Nk w
  pade_run((zd*)arrayout, *Nk, 1, false, (const zd*)arrayin, *Nk, *Niw, *Nw, 1, (const zd*)iwlist,
           (const zd*)wlist, nullptr);
  RF_CATCH
}

extern "C" void pade_with_trace(const gmod_z* arrayin, gmod_z* arrayout, const gmod_z* iwlist, const gmod_z* wlist,
                                const int64_t* Nk, const int64_t* Niw, const int64_t* Nw, const int64_t* Norb) {
  RF_TRY
  // arrayout(Nw, Nk): element (w, i) at w + Nw i; the orbital-diagonal
  // continuations summed
  pade_run((zd*)arrayout, 1, *Nw, true, (const zd*)arrayin, *Nk, *Niw, *Nw, *Norb, (const zd*)iwlist,
           (const zd*)wlist, nullptr);
  RF_CATCH
}
