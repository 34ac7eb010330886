// Grid kernels: scatter / product / gather / elementwise / reductions.
#include <algorithm>
#include <cmath>

#include "kernels.h"

namespace gmod {

constexpr int TPB = 256;
static inline unsigned nblocks(long n, int tpb = TPB, long cap = 1L << 20) {
  long b = (n + tpb - 1) / tpb;
  if (b > cap) b = cap;
  if (b < 1) b = 1;
  return (unsigned)b;
}

// ------------------------------------------------------------------ mesh
void Mesh::build(const int64_t* kmap, const int64_t* invk, long Nkall_, long Nk_, int Nx_, int Ny_,
                 int Nz_, int Nw_, hipStream_t s) {
  Nx = Nx_; Ny = Ny_; Nz = Nz_; Nw = Nw_; M = 2 * Nw_;
  Nkall = Nkall_; Nk = Nk_;
  if ((long)Nx * Ny * Nz != Nkall) throw Error("Nx*Ny*Nz != Nkall");
  std::vector<int> hirr(Nkall), hflag(Nkall), hsp(Nkall), hrev(Nkall), hirrsp(Nk, -1);
  kw_host.assign(Nk, 0.0);
  std::vector<char> seen(Nkall, 0);
  for (long i = 0; i < Nkall; ++i) {
    long x = kmap[3 * i], y = kmap[3 * i + 1], z = kmap[3 * i + 2];
    if (x < 0 || x >= Nx || y < 0 || y >= Ny || z < 0 || z >= Nz) throw Error("kmap entry outside the mesh");
    long sp_ = x + (long)Nx * (y + (long)Ny * z);
    if (seen[sp_]) throw Error("kmap is not a permutation of the mesh");
    seen[sp_] = 1;
    long k = invk[3 * i] - 1;
    if (k < 0 || k >= Nk) throw Error("invk(1,:) outside [1,Nk]");
    int f = (int)invk[3 * i + 1];
    if (f != 0 && f != 1) throw Error("invk(2,:) must be 0 or 1");
    hirr[i] = (int)k;
    hflag[i] = f;
    hsp[i] = (int)sp_;
    kw_host[k] += 1.0;
    if (f == 0 && hirrsp[k] < 0) hirrsp[k] = (int)sp_;
  }
  for (long z = 0; z < Nz; ++z)
    for (long y = 0; y < Ny; ++y)
      for (long x = 0; x < Nx; ++x) {
        long sp_ = x + Nx * (y + Ny * z);
        long rx = (Nx - x) % Nx, ry = (Ny - y) % Ny, rz = (Nz - z) % Nz;
        hrev[sp_] = (int)(rx + Nx * (ry + Ny * rz));
      }
  std::vector<int> hirr_s(Nkall), hflag_s(Nkall), hkgat_s(Nkall, -1);
  for (long i = 0; i < Nkall; ++i) {
    hirr_s[hsp[i]] = hirr[i];
    hflag_s[hsp[i]] = hflag[i];
  }
  complete = true;
  for (long k = 0; k < Nk; ++k) {
    if (hirrsp[k] >= 0) hkgat_s[hirrsp[k]] = (int)k;
    else complete = false;
  }
  irr.alloc(Nkall); flag.alloc(Nkall); sp.alloc(Nkall); rev.alloc(Nkall); irr_sp.alloc(Nk); kw.alloc(Nk);
  irr_s.alloc(Nkall); flag_s.alloc(Nkall); kgat_s.alloc(Nkall); kf_s.alloc(Nkall);
  irr_s.upload(hirr_s.data(), Nkall, s);
  flag_s.upload(hflag_s.data(), Nkall, s);
  std::vector<int> hkf_s(Nkall);
  for (long i = 0; i < Nkall; ++i) hkf_s[i] = (hirr_s[i] << 1) | (hflag_s[i] != 0 ? 1 : 0);
  kf_s.upload(hkf_s.data(), Nkall, s);
  kgat_s.upload(hkgat_s.data(), Nkall, s);
  kgat_host = hkgat_s;
  irr.upload(hirr.data(), Nkall, s);
  flag.upload(hflag.data(), Nkall, s);
  sp.upload(hsp.data(), Nkall, s);
  rev.upload(hrev.data(), Nkall, s);
  irr_sp.upload(hirrsp.data(), Nk, s);
  kw.upload(kw_host.data(), Nk, s);
  std::vector<int> hfull(Nkall), hnfull(Nkall);
  nfull_ok = true;
  for (long i = 0; i < Nkall; ++i) {
    hfull[hsp[i]] = (int)i;
    const long nf = invk[3 * i + 2] - 1;
    if (nf < 0 || nf >= Nkall) nfull_ok = false;
    hnfull[hsp[i]] = (nf < 0 || nf >= Nkall) ? 0 : (int)nf;
  }
  full_s.alloc(Nkall);
  nfull_s.alloc(Nkall);
  full_s.upload(hfull.data(), Nkall, s);
  nfull_s.upload(hnfull.data(), Nkall, s);
  GMOD_HIP(hipStreamSynchronize(s));
}

MeshDev mesh_dev(const Mesh& m) {
  MeshDev d{m.irr_s.p, m.flag_s.p, m.kgat_s.p, m.rev.p, m.kf_s.p, m.Nkall, m.Nk, m.Nw, m.M};
  d.full_s = m.full_s.p;
  d.nfull_s = m.nfull_s.p;
  return d;
}

template <typename R>
__global__ void k_mul_acc(cpx<R>* __restrict__ acc, const cpx<R>* __restrict__ A, const cpx<R>* __restrict__ B,
                          int first, long n) {
  for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < n; g += (long)gridDim.x * blockDim.x) {
    cpx<R> v = A[g] * B[g];
    acc[g] = first ? v : acc[g] + v;
  }
}

template <typename R>
void mul_acc(cpx<R>* acc, const cpx<R>* A, const cpx<R>* B, bool first, long n, hipStream_t s) {
  hipLaunchKernelGGL((k_mul_acc<R>), dim3(nblocks(n)), dim3(TPB), 0, s, acc, A, B, (int)first, n);
  GMOD_CHECK_LAUNCH();
}

// One thread per spatial point and chunk of L slots.  For band n with
// a = |e_n - mu|: E1(s) = exp(-a tau_s) falls by r = exp(-a dtau) per slot,
// E2(s) = exp(-a (beta - tau_s)) rises by 1/r; g0 = -(xi >= 0 ? E1 : E2) / d,
// g1 = (xi >= 0 ? E2 : E1) / d, d = 1 + exp(-beta a).  E2 is evaluated
// directly where its chunk start would underflow (deep bands, low T).
template <typename R, int NO>
__global__ void __launch_bounds__(256)
k_tau_ref(cpx<R>* ta, cpx<R>* tb, const double* __restrict__ eig, const zd* __restrict__ U, int a, int b, double mu,
          double mu2, int nv, double beta, double dtau, MeshDev m, int L) {
  const long nch = (m.M + L - 1) / L;
  const long tot = m.Nkall * nch;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
    const long sp = e % m.Nkall;
    const int s0 = (int)(e / m.Nkall) * L;
    const int s1 = s0 + L < m.M ? s0 + L : m.M;
    const long k = m.irr_s[sp];
    const bool rev = m.flag_s[sp] != 0;
    zd c[NO];
    double E1[NO][2], E2[NO][2], r[NO][2], ir[NO][2], id[NO][2], ad[NO][2];
    bool pos[NO][2], direct[NO][2];
#pragma unroll
    for (int n = 0; n < NO; ++n) {
      zd cn = U[a + NO * (n + (long)NO * k)] * conj(U[b + NO * (n + (long)NO * k)]);
      c[n] = rev ? conj(cn) : cn;
#pragma unroll
      for (int v = 0; v < 2; ++v) {
        const double xi = eig[n + (long)NO * k] - (v == 0 ? mu : mu2);
        const double aa = fabs(xi);
        pos[n][v] = xi >= 0.0;
        id[n][v] = 1.0 / (1.0 + exp(-beta * aa));
        ad[n][v] = aa * dtau;
        direct[n][v] = ad[n][v] * (double)(m.M - s0) > 600.0 || ad[n][v] > 300.0;
        E1[n][v] = exp(-ad[n][v] * (double)s0);
        E2[n][v] = direct[n][v] ? 0.0 : exp(-ad[n][v] * (double)(m.M - s0));
        r[n][v] = exp(-ad[n][v]);
        ir[n][v] = direct[n][v] ? 0.0 : exp(ad[n][v]);
      }
    }
    for (int s = s0; s < s1; ++s) {
      zd A(0, 0), B(0, 0);
#pragma unroll
      for (int n = 0; n < NO; ++n) {
#pragma unroll
        for (int v = 0; v < 2; ++v) {
          if (v >= nv) continue;
          const double e2 = direct[n][v] ? exp(-ad[n][v] * (double)(m.M - s)) : E2[n][v];
          const double g0 = -(pos[n][v] ? E1[n][v] : e2) * id[n][v];
          const double g1 = (pos[n][v] ? e2 : E1[n][v]) * id[n][v];
          const double sg = v == 0 ? 1.0 : -1.0;
          A += c[n] * (sg * g0);
          B += c[n] * (sg * g1);
          E1[n][v] *= r[n][v];
          E2[n][v] *= ir[n][v];
        }
      }
      ta[sp + m.Nkall * (long)s] = cpx<R>(A);
      tb[sp + m.Nkall * (long)s] = cpx<R>(B);
    }
  }
}

template <typename R>
void tau_ref_grids(cpx<R>* ta, cpx<R>* tb, const double* eig, const zd* uni, int a, int b, int No, double mu,
                   double mu2, bool diff, double beta, double dtau, const MeshDev& m, hipStream_t s) {
  const int L = 64;
  const long tot = m.Nkall * (long)((m.M + L - 1) / L);
  const int nv = diff ? 2 : 1;
  const dim3 g((unsigned)std::min<long>((tot + 255) / 256, 1L << 20)), blk(256);
  switch (No) {
#define GMOD_TAU_CASE(NOV)                                                                                           \
  case NOV:                                                                                                          \
    hipLaunchKernelGGL((k_tau_ref<R, NOV>), g, blk, 0, s, ta, tb, eig, uni, a, b, mu, mu2, nv, beta, dtau, m, L); \
    break;
    GMOD_TAU_CASE(1) GMOD_TAU_CASE(2) GMOD_TAU_CASE(3) GMOD_TAU_CASE(4)
    GMOD_TAU_CASE(5) GMOD_TAU_CASE(6) GMOD_TAU_CASE(7) GMOD_TAU_CASE(8)
#undef GMOD_TAU_CASE
    default: throw Error("tau reference: Norb > 8 is not supported");
  }
  GMOD_CHECK_LAUNCH();
}

template void tau_ref_grids<float>(cpx<float>*, cpx<float>*, const double*, const zd*, int, int, int, double, double,
                                   bool, double, double, const MeshDev&, hipStream_t);
template void tau_ref_grids<double>(cpx<double>*, cpx<double>*, const double*, const zd*, int, int, int, double,
                                    double, bool, double, double, const MeshDev&, hipStream_t);

// ------------------------------------------------------------------ elementwise
__global__ void k_conj_copy(zd* dst, const zd* src, long n) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
    dst[e] = conj(src[e]);
}
void conj_copy(zd* dst, const zd* src, long n, hipStream_t s) {
  hipLaunchKernelGGL(k_conj_copy, dim3(nblocks(n)), dim3(TPB), 0, s, dst, src, n);
  GMOD_CHECK_LAUNCH();
}
__global__ void k_conj_copy_neg(zd* dst, const zd* src, long n) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const zd v = src[e];
    dst[e] = zd(-v.re, v.im);
  }
}
void conj_copy(zd* dst, const zd* src, long n, double sign, hipStream_t s) {
  if (sign > 0) {
    conj_copy(dst, src, n, s);
    return;
  }
  hipLaunchKernelGGL(k_conj_copy_neg, dim3(nblocks(n)), dim3(TPB), 0, s, dst, src, n);
  GMOD_CHECK_LAUNCH();
}

// |Re| < eps -> Re = 0, then |Im| < eps -> Im = 0 (fchi.f90:499-500)
__global__ void k_threshold(zd* a, long n, double eps) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    zd v = a[e];
    if (fabs(v.re) < eps) v.re = 0.0;
    if (fabs(v.im) < eps) v.im = 0.0;
    a[e] = v;
  }
}
void threshold(zd* a, long n, double eps, hipStream_t s) {
  hipLaunchKernelGGL(k_threshold, dim3(nblocks(n)), dim3(TPB), 0, s, a, n, eps);
  GMOD_CHECK_LAUNCH();
}

__global__ void k_scale(zd* a, long n, double f) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
    a[e] = a[e] * f;
}
void scale(zd* a, long n, double f, hipStream_t s) {
  hipLaunchKernelGGL(k_scale, dim3(nblocks(n)), dim3(TPB), 0, s, a, n, f);
  GMOD_CHECK_LAUNCH();
}

__global__ void k_axpy(zd* y, const zd* x, zd al, long n) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
    y[e] = y[e] + al * x[e];
}
void axpy(zd* y, const zd* x, zd alpha, long n, hipStream_t s) {
  hipLaunchKernelGGL(k_axpy, dim3(nblocks(n)), dim3(TPB), 0, s, y, x, alpha, n);
  GMOD_CHECK_LAUNCH();
}

void copy(zd* dst, const zd* src, long n, hipStream_t s) {
  GMOD_HIP(hipMemcpyAsync(dst, src, n * sizeof(zd), hipMemcpyDeviceToDevice, s));
}

__global__ void k_threshold_copy(zd* dst, const zd* src, long n, double eps) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    zd v = src[e];
    if (fabs(v.re) < eps) v.re = 0.0;
    if (fabs(v.im) < eps) v.im = 0.0;
    dst[e] = v;
  }
}
void threshold_copy(zd* dst, const zd* src, long n, double eps, hipStream_t s) {
  hipLaunchKernelGGL(k_threshold_copy, dim3(nblocks(n)), dim3(TPB), 0, s, dst, src, n, eps);
  GMOD_CHECK_LAUNCH();
}

// dst(r, t) = src(-r, -t) on the full grid (t -> (M - t) mod M): an output
// row of x reads one input row in reverse order, so both sides coalesce
template <typename R>
__global__ void k_mirror_grid(cpx<R>* __restrict__ dst, const cpx<R>* __restrict__ src, int Nx, int Ny, int Nz,
                              int M) {
  const long S = (long)Nx * Ny * Nz, n = S * M;
  for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < n; g += (long)gridDim.x * blockDim.x) {
    const long sp = g % S, t = g / S;
    const int x = (int)(sp % Nx), y = (int)((sp / Nx) % Ny), z = (int)(sp / ((long)Nx * Ny));
    const long rs = (x ? Nx - x : 0) + (long)Nx * ((y ? Ny - y : 0) + (long)Ny * (z ? Nz - z : 0));
    dst[g] = src[rs + S * (t ? M - t : 0)];
  }
}
template <typename R>
void mirror_grid(cpx<R>* dst, const cpx<R>* src, int Nx, int Ny, int Nz, int M, hipStream_t s) {
  const long n = (long)Nx * Ny * Nz * M;
  hipLaunchKernelGGL((k_mirror_grid<R>), dim3(nblocks(n)), dim3(TPB), 0, s, dst, src, Nx, Ny, Nz, M);
  GMOD_CHECK_LAUNCH();
}
template void mirror_grid<float>(cpx<float>*, const cpx<float>*, int, int, int, int, hipStream_t);
template void mirror_grid<double>(cpx<double>*, const cpx<double>*, int, int, int, int, hipStream_t);

__global__ void k_threshold_slice(zd* dst, const zd* src, long Nk, long NkNw, long nmat, double eps) {
  const long n = Nk * nmat;
  for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < n; g += (long)gridDim.x * blockDim.x) {
    const long k = g % Nk, e = g / Nk;
    zd v = src[k + NkNw * e];
    if (fabs(v.re) < eps) v.re = 0.0;
    if (fabs(v.im) < eps) v.im = 0.0;
    dst[g] = v;
  }
}
void threshold_slice(zd* dst, const zd* src, long Nk, long NkNw, long nmat, double eps, hipStream_t s) {
  const long n = Nk * nmat;
  hipLaunchKernelGGL(k_threshold_slice, dim3(nblocks(n)), dim3(TPB), 0, s, dst, src, Nk, NkNw, nmat, eps);
  GMOD_CHECK_LAUNCH();
}

__global__ void k_diff_to_f32(cpx<float>* dst, const zd* a, const zd* b, long n) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
    dst[e] = cpx<float>(a[e] - b[e]);
}
void diff_to_f32(cpx<float>* dst, const zd* a, const zd* b, long n, hipStream_t s) {
  hipLaunchKernelGGL(k_diff_to_f32, dim3(nblocks(n)), dim3(TPB), 0, s, dst, a, b, n);
  GMOD_CHECK_LAUNCH();
}

template <typename R>
__global__ void k_add_inplace(cpx<R>* y, const cpx<R>* x, long n) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
    y[e] = y[e] + x[e];
}
template <typename R>
void add_inplace(cpx<R>* y, const cpx<R>* x, long n, hipStream_t s) {
  hipLaunchKernelGGL(k_add_inplace<R>, dim3(nblocks(n)), dim3(TPB), 0, s, y, x, n);
  GMOD_CHECK_LAUNCH();
}

__global__ void k_lincomb(zd* y, const zd* a, double ca, const zd* b, double cb, long n) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
    y[e] = a[e] * ca + b[e] * cb;
}
void lincomb(zd* y, const zd* a, double ca, const zd* b, double cb, long n, hipStream_t s) {
  hipLaunchKernelGGL(k_lincomb, dim3(nblocks(n)), dim3(TPB), 0, s, y, a, ca, b, cb, n);
  GMOD_CHECK_LAUNCH();
}

// a <- a + b, b <- a - b (with the old a)
__global__ void k_sum_diff(zd* a, zd* b, long n) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const zd x = a[e], y = b[e];
    a[e] = x + y;
    b[e] = x - y;
  }
}
void sum_diff(zd* a, zd* b, long n, hipStream_t s) {
  hipLaunchKernelGGL(k_sum_diff, dim3(nblocks(n)), dim3(TPB), 0, s, a, b, n);
  GMOD_CHECK_LAUNCH();
}

// ------------------------------------------------------------------ reductions
constexpr int RED_BLOCKS = 1024;

__device__ inline double block_sum(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}

__global__ void k_dot_w(const zd* u, const zd* v, const double* kw, long Nk, long n, double* part) {
  __shared__ double sh[TPB];
  double re = 0, im = 0;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    zd p = conj(u[e]) * v[e];
    double w = kw ? kw[e % Nk] : 1.0;
    re += w * p.re;
    im += w * p.im;
  }
  re = block_sum(re, sh);
  im = block_sum(im, sh);
  if (threadIdx.x == 0) { part[2 * blockIdx.x] = re; part[2 * blockIdx.x + 1] = im; }
}

__global__ void k_finish_sum(const double* part, int nb, int width, double* out) {
  // one thread per component, fixed order
  int c = threadIdx.x;
  if (c < width) {
    double s = 0;
    for (int b = 0; b < nb; ++b) s += part[width * b + c];
    out[c] = s;
  }
}

zd dot_w(const zd* u, const zd* v, const double* kw, long Nk, long rest, hipStream_t s) {
  const long n = Nk * rest;
  DBuf<double> part(2 * RED_BLOCKS + 2);
  hipLaunchKernelGGL(k_dot_w, dim3(RED_BLOCKS), dim3(TPB), 0, s, u, v, kw, Nk, n, part.p);
  GMOD_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_finish_sum, dim3(1), dim3(64), 0, s, part.p, RED_BLOCKS, 2, part.p + 2 * RED_BLOCKS);
  GMOD_CHECK_LAUNCH();
  double h[2];
  GMOD_HIP(hipMemcpyAsync(h, part.p + 2 * RED_BLOCKS, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  GMOD_HIP(hipStreamSynchronize(s));
  return zd(h[0], h[1]);
}

__global__ void k_max_rel(const zd* s0, const zd* sd, long n, double floor, double* part) {
  __shared__ double sh[TPB];
  double best = 0.0;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    zd d = sd[e];
    double ad = hypot(d.re, d.im);
    if (ad > floor) {
      zd df = s0[e] - d;
      double r = hypot(df.re, df.im) / ad;
      if (r > best) best = r;
    }
  }
  sh[threadIdx.x] = best;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void k_minmax_re(const zd* a, long n, int want_max, double* part) {
  __shared__ double sh[TPB];
  double best = want_max ? -HUGE_VAL : HUGE_VAL;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
    best = want_max ? fmax(best, a[e].re) : fmin(best, a[e].re);
  sh[threadIdx.x] = best;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o)
      sh[threadIdx.x] = want_max ? fmax(sh[threadIdx.x], sh[threadIdx.x + o]) : fmin(sh[threadIdx.x], sh[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

static double host_reduce(const double* dpart, int nb, hipStream_t s, int mode) {
  std::vector<double> h(nb);
  GMOD_HIP(hipMemcpyAsync(h.data(), dpart, nb * sizeof(double), hipMemcpyDeviceToHost, s));
  GMOD_HIP(hipStreamSynchronize(s));
  double r = h[0];
  for (int i = 1; i < nb; ++i) r = mode == 0 ? std::max(r, h[i]) : std::min(r, h[i]);
  return r;
}

double max_rel_err(const zd* s0, const zd* sd, long n, double floor, hipStream_t s, long* where) {
  DBuf<double> part(RED_BLOCKS);
  hipLaunchKernelGGL(k_max_rel, dim3(RED_BLOCKS), dim3(TPB), 0, s, s0, sd, n, floor, part.p);
  GMOD_CHECK_LAUNCH();
  if (where) *where = -1;
  return host_reduce(part.p, RED_BLOCKS, s, 0);
}

// max |a - b| (b null: max |a|)
__global__ void k_max_absdiff(const zd* a, const zd* b, long n, double* part) {
  __shared__ double sh[TPB];
  double best = 0.0;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const zd d = b ? a[e] - b[e] : a[e];
    const double r = hypot(d.re, d.im);
    if (r > best) best = r;
  }
  sh[threadIdx.x] = best;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

double max_abs_diff(const zd* a, const zd* b, long n, hipStream_t s) {
  DBuf<double> part(RED_BLOCKS);
  hipLaunchKernelGGL(k_max_absdiff, dim3(RED_BLOCKS), dim3(TPB), 0, s, a, b, n, part.p);
  GMOD_CHECK_LAUNCH();
  return host_reduce(part.p, RED_BLOCKS, s, 0);
}

double max_real(const zd* a, long n, hipStream_t s) {
  DBuf<double> part(RED_BLOCKS);
  hipLaunchKernelGGL(k_minmax_re, dim3(RED_BLOCKS), dim3(TPB), 0, s, a, n, 1, part.p);
  GMOD_CHECK_LAUNCH();
  return host_reduce(part.p, RED_BLOCKS, s, 0);
}

double min_real(const zd* a, long n, hipStream_t s) {
  DBuf<double> part(RED_BLOCKS);
  hipLaunchKernelGGL(k_minmax_re, dim3(RED_BLOCKS), dim3(TPB), 0, s, a, n, 0, part.p);
  GMOD_CHECK_LAUNCH();
  return host_reduce(part.p, RED_BLOCKS, s, 1);
}

// ------------------------------------------------------------------ Sigma reference-vertex term
// One thread per grid point: the No^2 increments Fd_ab in registers (read
// once), then every output's sum over its No^2 reference-vertex grids; the
// vertex grids are read once each, coalesced along the point index.
template <typename R, int NAB>
__global__ void __launch_bounds__(256)
k_sigma_ref_contract(cpx<R>* __restrict__ Y, cpx<R>* __restrict__ Gn, const cpx<R>* __restrict__ Fr,
                     const cpx<R>* __restrict__ Fd, const cpx<R>* __restrict__ FVr, const int* __restrict__ tab,
                     int nout, long N) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < N; e += (long)gridDim.x * blockDim.x) {
    cpx<R> fd[NAB];
#pragma unroll
    for (int ab = 0; ab < NAB; ++ab) {
      fd[ab] = Fd[ab * N + e];
      Gn[ab * N + e] = Fr[ab * N + e] + fd[ab];
    }
    for (int o = 0; o < nout; ++o) {
      const int* t = tab + o * NAB;
      cpx<R> acc(0, 0);
#pragma unroll
      for (int ab = 0; ab < NAB; ++ab) acc += FVr[(long)t[ab] * N + e] * fd[ab];
      Y[(long)o * N + e] = acc;
    }
  }
}

template <typename R>
void sigma_ref_contract(cpx<R>* Y, cpx<R>* Gn, const cpx<R>* Fr, const cpx<R>* Fd, const cpx<R>* FVr,
                        const int* tab, int nout, int No, long N, hipStream_t s) {
  const unsigned grid = (unsigned)std::min<long>((N + 255) / 256, 256L * 32);
  switch (No) {
    case 1: hipLaunchKernelGGL((k_sigma_ref_contract<R, 1>), dim3(grid), dim3(256), 0, s, Y, Gn, Fr, Fd, FVr, tab, nout, N); break;
    case 2: hipLaunchKernelGGL((k_sigma_ref_contract<R, 4>), dim3(grid), dim3(256), 0, s, Y, Gn, Fr, Fd, FVr, tab, nout, N); break;
    case 3: hipLaunchKernelGGL((k_sigma_ref_contract<R, 9>), dim3(grid), dim3(256), 0, s, Y, Gn, Fr, Fd, FVr, tab, nout, N); break;
    case 4: hipLaunchKernelGGL((k_sigma_ref_contract<R, 16>), dim3(grid), dim3(256), 0, s, Y, Gn, Fr, Fd, FVr, tab, nout, N); break;
    default: throw Error("sigma_ref_contract: Norb > 4");
  }
  GMOD_CHECK_LAUNCH();
}

// One thread per spatial point and a stride of slots: the reversed point of
// a thread is fixed, so its loads run along the grid (descending) as the
// direct ones do.  The orbital-pair grids stay in registers across outputs.
template <typename R, int NAB>
__global__ void __launch_bounds__(256)
k_sigma_inc_contract(cpx<R>* __restrict__ out, const cpx<R>* __restrict__ Fr, const cpx<R>* __restrict__ Fd,
                     const cpx<R>* __restrict__ FVr, const cpx<R>* __restrict__ DVt,
                     const cpx<R>* __restrict__ dhF, const cpx<R>* __restrict__ dhD, const int* __restrict__ tab,
                     int nout, unsigned S, unsigned M, const int* __restrict__ rev) {
  const unsigned sp = blockIdx.x * blockDim.x + threadIdx.x;
  if (sp >= S) return;
  const long N = (long)S * M;
  const unsigned rs = (unsigned)rev[sp];
  for (unsigned t = blockIdx.y; t < M; t += gridDim.y) {
    const long e = sp + (long)S * t, er = rs + (long)S * t;
    cpx<R> fd[NAB], gn[NAB];
#pragma unroll
    for (int ab = 0; ab < NAB; ++ab) {
      fd[ab] = Fd[ab * N + e];
      gn[ab] = Fr[ab * N + e] + fd[ab];
    }
    for (int o = 0; o < nout; ++o) {
      const int* tb = tab + 2 * o * NAB;
      cpx<R> acc(0, 0);
#pragma unroll
      for (int ab = 0; ab < NAB; ++ab) {
        const int ef = tb[2 * ab], ed = tb[2 * ab + 1];
        cpx<R> fv = FVr[(long)(ef >> 1) * N + ((ef & 1) ? er : e)];
        cpx<R> dv = DVt[(long)(ed >> 1) * N + ((ed & 1) ? er : e)];
        if (ef & 1) fv = conj(fv) + dhF[(long)(ef >> 1) * S + sp];
        if (ed & 1) dv = conj(dv) + dhD[(long)(ed >> 1) * S + sp];
        acc += fv * fd[ab] + dv * gn[ab];
      }
      out[(long)o * N + e] = acc;
    }
  }
}

template <typename R>
void sigma_inc_contract(cpx<R>* out, const cpx<R>* Fr, const cpx<R>* Fd, const cpx<R>* FVr, const cpx<R>* DVt,
                        const cpx<R>* dhF, const cpx<R>* dhD, const int* tab, int nout, int No, unsigned S,
                        unsigned M, const int* rev, hipStream_t s) {
  const unsigned gx = (S + 255) / 256;
  const unsigned gy = (unsigned)std::max<long>(1, std::min<long>(M, (256L * 32 + gx - 1) / gx));
  const dim3 g(gx, gy);
  switch (No) {
    case 1: hipLaunchKernelGGL((k_sigma_inc_contract<R, 1>), g, dim3(256), 0, s, out, Fr, Fd, FVr, DVt, dhF, dhD, tab, nout, S, M, rev); break;
    case 2: hipLaunchKernelGGL((k_sigma_inc_contract<R, 4>), g, dim3(256), 0, s, out, Fr, Fd, FVr, DVt, dhF, dhD, tab, nout, S, M, rev); break;
    case 3: hipLaunchKernelGGL((k_sigma_inc_contract<R, 9>), g, dim3(256), 0, s, out, Fr, Fd, FVr, DVt, dhF, dhD, tab, nout, S, M, rev); break;
    case 4: hipLaunchKernelGGL((k_sigma_inc_contract<R, 16>), g, dim3(256), 0, s, out, Fr, Fd, FVr, DVt, dhF, dhD, tab, nout, S, M, rev); break;
    default: throw Error("sigma_inc_contract: Norb > 4");
  }
  GMOD_CHECK_LAUNCH();
}

// ------------------------------------------------------------------ instantiations
#define INST(R) \
  template void mul_acc<R>(cpx<R>*, const cpx<R>*, const cpx<R>*, bool, long, hipStream_t); \
  template void add_inplace<R>(cpx<R>*, const cpx<R>*, long, hipStream_t); \
  template void sigma_ref_contract<R>(cpx<R>*, cpx<R>*, const cpx<R>*, const cpx<R>*, const cpx<R>*, const int*, int, \
                                      int, long, hipStream_t); \
  template void sigma_inc_contract<R>(cpx<R>*, const cpx<R>*, const cpx<R>*, const cpx<R>*, const cpx<R>*, \
                                      const cpx<R>*, const cpx<R>*, const int*, int, int, unsigned, unsigned, const int*, \
                                      hipStream_t);
INST(double)
INST(float)

}  // namespace gmod
