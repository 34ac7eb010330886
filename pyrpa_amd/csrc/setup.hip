// Setup on the device: H(k) from the hopping integrals (fmod.f90:162-207) and
// the band eigenpairs of every H(k) (fmod.f90:209-248, zheev('V','U') per k),
// so that the bands a solve starts from are formed where the solve runs.
//
// The eigen-solver is the cyclic complex Jacobi method, one thread per k
// point with its matrix in a global workspace (any Norb): the same
// eigenvalues as zheev and eigenvectors equal up to a phase per column;
// every consumer on the path uses phase-free combinations U_an conj(U_bn).
#include <cmath>

#include "../../include/gmod.h"
#include "gmod_common.h"

namespace gmod {

constexpr double PI_SETUP = 3.141592653589793238462643383279;  // fmod.f90:4

static inline unsigned setup_blocks(long n, int tpb) {
  long b = (n + tpb - 1) / tpb;
  if (b > (1L << 20)) b = 1L << 20;
  return (unsigned)(b < 1 ? 1 : b);
}

// one thread per (k, m >= l): H(k)_ml = sum_R H(R)_ml exp(-2 pi i k.R) in the
// reference's order over R, then the diagonal made real and H_lm = conj H_ml
__global__ void k_gen_ham(zd* hk, const double* __restrict__ kl, const zd* __restrict__ hr,
                          const double* __restrict__ rv, long Nk, int Nr, int No) {
  const int np = No * (No + 1) / 2;
  const long tot = Nk * np;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
    const long i = e / np;
    int p = (int)(e - i * np);
    int l = 0;
    while (p >= No - l) {
      p -= No - l;
      ++l;
    }
    const int m = l + p;
    zd acc(0, 0);
    for (int j = 0; j < Nr; ++j) {
      const double ph = 2 * PI_SETUP * (kl[3 * i] * rv[3 * j] + kl[3 * i + 1] * rv[3 * j + 1] +
                                        kl[3 * i + 2] * rv[3 * j + 2]);
      const double c = cos(ph), s = sin(ph);
      const zd h = hr[m + No * (l + (long)No * j)];
      acc += zd(h.re * c + h.im * s, h.im * c - h.re * s);  // H(R) (cos - i sin)
    }
    zd* H = hk + (long)No * No * i;
    if (m == l) {
      H[m + No * l] = zd(acc.re, 0.0);
    } else {
      H[m + No * l] = acc;
      H[l + No * m] = conj(acc);
    }
  }
}

// Cyclic Jacobi for the Hermitian matrix A (n x n, column-major, the upper
// triangle is read as zheev('U') reads it), eigenvectors accumulated in V;
// then eigenvalues ascending (stable) into w and the matching columns of V
// into U.  A and V are this thread's workspace.
__device__ void herm_jacobi_dev(zd* A, zd* V, int n, double* w, zd* U) {
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) V[i + n * j] = zd(i == j ? 1.0 : 0.0, 0.0);
  for (int j = 0; j < n; ++j)
    for (int i = j + 1; i < n; ++i) A[i + n * j] = conj(A[j + n * i]);
  for (int i = 0; i < n; ++i) A[i + n * i].im = 0.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0, tot = 0;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        const zd a = A[i + n * j];
        const double a2 = a.re * a.re + a.im * a.im;
        tot += a2;
        if (i != j) off += a2;
      }
    if (off <= 1e-32 * tot || off == 0.0) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        const zd apq = A[p + n * q];
        const double mag = hypot(apq.re, apq.im);
        if (mag < 1e-300) continue;
        const double app = A[p + n * p].re, aqq = A[q + n * q].re;
        const zd ph(apq.re / mag, apq.im / mag);  // A_pq = |A_pq| e^{i phi}
        const double theta = 0.5 * atan2(2 * mag, aqq - app);
        const double c = cos(theta), s = sin(theta);
        const zd sph(s * ph.re, s * ph.im), sphc(s * ph.re, -s * ph.im);
        // rotation J = [[c, s ph], [-s conj(ph), c]] on (p, q)
        for (int k = 0; k < n; ++k) {  // A <- A J
          const zd akp = A[k + n * p], akq = A[k + n * q];
          A[k + n * p] = akp * c - sphc * akq;
          A[k + n * q] = sph * akp + akq * c;
        }
        for (int k = 0; k < n; ++k) {  // A <- J^H A
          const zd apk = A[p + n * k], aqk = A[q + n * k];
          A[p + n * k] = apk * c - sph * aqk;
          A[q + n * k] = sphc * apk + aqk * c;
        }
        for (int k = 0; k < n; ++k) {  // V <- V J
          const zd vkp = V[k + n * p], vkq = V[k + n * q];
          V[k + n * p] = vkp * c - sphc * vkq;
          V[k + n * q] = sph * vkp + vkq * c;
        }
        A[p + n * q] = A[q + n * p] = zd(0, 0);
      }
  }
  // eigenvalues ascending, ties in index order (a stable selection)
  unsigned long long taken = 0;  // n <= 64
  for (int c = 0; c < n; ++c) {
    int best = -1;
    for (int i = 0; i < n; ++i) {
      if (taken >> i & 1ull) continue;
      if (best < 0 || A[i + n * i].re < A[best + n * best].re) best = i;
    }
    taken |= 1ull << best;
    w[c] = A[best + n * best].re;
    for (int r = 0; r < n; ++r) U[r + n * c] = V[r + n * best];
  }
}

__global__ void k_herm_eig(double* eig, zd* uni, zd* A, zd* V, long Nk, int n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < Nk; i += (long)gridDim.x * blockDim.x) {
    const long o = (long)n * n * i;
    herm_jacobi_dev(A + o, V + o, n, eig + (long)n * i, uni + o);
  }
}

}  // namespace gmod

using namespace gmod;

#define GMOD_TRY try {
#define GMOD_CATCH                                       \
  }                                                      \
  catch (std::exception & e) { set_error(e.what()); }    \
  catch (...) { set_error("unknown error"); }

// fmod.f90:162-207: ham_k(Norb,Norb,Nk) from ham_r(Norb,Norb,Nr) at klist(3,Nk), rvec(3,Nr)
extern "C" void gen_ham(gmod_z* ham_k, const double* klist, const gmod_z* ham_r, const double* rvec,
                        const int64_t* Nk_, const int64_t* Nr_, const int64_t* Norb_) {
  GMOD_TRY
  std::lock_guard<std::mutex> g(ctx().lock);
  hipStream_t s = ctx().stream;
  const long Nk = *Nk_, Nr = *Nr_;
  const int No = (int)*Norb_;
  if (Nk <= 0 || No <= 0) return;
  DBuf<double> kl(3 * Nk), rv(3 * (size_t)std::max<long>(Nr, 1));
  DBuf<zd> hr((size_t)No * No * std::max<long>(Nr, 1)), hk((size_t)No * No * Nk);
  kl.upload(klist, 3 * Nk, s);
  if (Nr > 0) {
    rv.upload(rvec, 3 * Nr, s);
    hr.upload((const zd*)ham_r, (size_t)No * No * Nr, s);
  }
  const long tot = Nk * (long)(No * (No + 1) / 2);
  hipLaunchKernelGGL(k_gen_ham, dim3(setup_blocks(tot, 256)), dim3(256), 0, s, hk.p, kl.p, hr.p, rv.p, Nk, (int)Nr,
                     No);
  GMOD_CHECK_LAUNCH();
  hk.download((zd*)ham_k, (size_t)No * No * Nk, s);
  GMOD_HIP(hipStreamSynchronize(s));
  GMOD_CATCH
}

// fmod.f90:209-248: eig(Norb,Nk) ascending and uni(Norb,Norb,Nk) (eigenvectors in columns)
extern "C" void get_eig(double* eig, gmod_z* uni, const gmod_z* ham_k, const int64_t* Nk_, const int64_t* Norb_) {
  GMOD_TRY
  std::lock_guard<std::mutex> g(ctx().lock);
  hipStream_t s = ctx().stream;
  const long Nk = *Nk_;
  const int n = (int)*Norb_;
  if (Nk <= 0 || n <= 0) return;
  if (n > 64) throw Error("get_eig: Norb > 64");
  const size_t nm = (size_t)n * n * Nk;
  DBuf<zd> A(nm), V(nm), U(nm);
  DBuf<double> w((size_t)n * Nk);
  A.upload((const zd*)ham_k, nm, s);
  hipLaunchKernelGGL(k_herm_eig, dim3(setup_blocks(Nk, 64)), dim3(64), 0, s, w.p, U.p, A.p, V.p, Nk, n);
  GMOD_CHECK_LAUNCH();
  w.download(eig, (size_t)n * Nk, s);
  U.download((zd*)uni, nm, s);
  GMOD_HIP(hipStreamSynchronize(s));
  GMOD_CATCH
}
