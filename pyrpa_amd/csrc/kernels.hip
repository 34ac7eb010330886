// Grid kernels: scatter / product / gather / elementwise / reductions.
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
  irr.alloc(Nkall); flag.alloc(Nkall); sp.alloc(Nkall); rev.alloc(Nkall); irr_sp.alloc(Nk); kw.alloc(Nk);
  irr.upload(hirr.data(), Nkall, s);
  flag.upload(hflag.data(), Nkall, s);
  sp.upload(hsp.data(), Nkall, s);
  rev.upload(hrev.data(), Nkall, s);
  irr_sp.upload(hirrsp.data(), Nk, s);
  kw.upload(kw_host.data(), Nk, s);
  GMOD_HIP(hipStreamSynchronize(s));
}

struct MeshView {
  const int *irr, *flag, *sp, *rev, *irr_sp;
  long Nkall, Nk;
  int Nw, M;
};
static MeshView view(const Mesh& m) {
  return {m.irr.p, m.flag.p, m.sp.p, m.rev.p, m.irr_sp.p, m.Nkall, m.Nk, m.Nw, m.M};
}

// ------------------------------------------------------------------ scatters
// G_ab on the full grid (fchi.f90:563-583, fself.f90:226-236):
//   slot t <  Nw : w_t ;  k direct -> G(k)_ab, k = -k_irr -> G(k_irr)_ba
//   slot t >= Nw : -w_j with j = M-1-t ; direct -> conj G_ba, reversed -> conj G_ab
template <typename R>
__global__ void k_scatter_green(cpx<R>* __restrict__ grid, const zd* __restrict__ G, int a, int b, int No,
                                MeshView m) {
  const long N = m.Nkall * m.M;
  const long NkNw = m.Nk * (long)m.Nw;
  for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < N; g += (long)gridDim.x * blockDim.x) {
    long i = g % m.Nkall;
    int t = (int)(g / m.Nkall);
    int k = m.irr[i];
    int f = m.flag[i];
    zd v;
    if (t < m.Nw) {
      int p = f == 0 ? a : b, q = f == 0 ? b : a;
      v = G[k + m.Nk * (long)t + NkNw * (p + (long)No * q)];
    } else {
      int j = m.M - 1 - t;
      int p = f == 0 ? b : a, q = f == 0 ? a : b;
      v = conj(G[k + m.Nk * (long)j + NkNw * (p + (long)No * q)]);
    }
    grid[m.sp[i] + m.Nkall * (long)t] = cpx<R>(v);
  }
}

template <typename R>
void scatter_green(cpx<R>* grid, const zd* G, int a, int b, int No, const Mesh& m, hipStream_t s) {
  hipLaunchKernelGGL((k_scatter_green<R>), dim3(nblocks(m.N())), dim3(TPB), 0, s, grid, G, a, b, No, view(m));
  GMOD_CHECK_LAUNCH();
}

// bosonic V_(m,l) (fself.f90:203-224, feliash.f90:543-564): slot 0 = nu_0,
// slot Nw = nu->inf value `stat`, slots t and M-t = +-nu_t
template <typename R>
__global__ void k_scatter_boson(cpx<R>* __restrict__ grid, const zd* __restrict__ V, int mi, int li, int Nchi,
                                zd stat, MeshView m) {
  const long N = m.Nkall * m.M;
  const long NkNw = m.Nk * (long)m.Nw;
  const long off_ml = NkNw * (mi + (long)Nchi * li);
  const long off_lm = NkNw * (li + (long)Nchi * mi);
  for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < N; g += (long)gridDim.x * blockDim.x) {
    long i = g % m.Nkall;
    int t = (int)(g / m.Nkall);
    int k = m.irr[i];
    int f = m.flag[i];
    zd v;
    if (t == m.Nw) {
      v = stat;
    } else if (t < m.Nw) {
      v = V[k + m.Nk * (long)t + (f == 0 ? off_ml : off_lm)];
    } else {
      int j = m.M - t;
      v = conj(V[k + m.Nk * (long)j + (f == 0 ? off_lm : off_ml)]);
    }
    grid[m.sp[i] + m.Nkall * (long)t] = cpx<R>(v);
  }
}

template <typename R>
void scatter_boson(cpx<R>* grid, const zd* V, int mi, int li, int Nchi, zd stat, const Mesh& m, hipStream_t s) {
  hipLaunchKernelGGL((k_scatter_boson<R>), dim3(nblocks(m.N())), dim3(TPB), 0, s, grid, V, mi, li, Nchi, stat,
                     view(m));
  GMOD_CHECK_LAUNCH();
}

// anomalous F_(a,b) with parity (feliash.f90:566-577)
template <typename R>
__global__ void k_scatter_anom(cpx<R>* __restrict__ grid, const zd* __restrict__ F, int a, int b, int No,
                               double pp, double sgn, double dprt, MeshView m) {
  const long N = m.Nkall * m.M;
  const long NkNw = m.Nk * (long)m.Nw;
  const long oab = NkNw * (a + (long)No * b), oba = NkNw * (b + (long)No * a);
  for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < N; g += (long)gridDim.x * blockDim.x) {
    long i = g % m.Nkall;
    int t = (int)(g / m.Nkall);
    int k = m.irr[i];
    int f = m.flag[i];
    zd v;
    if (t < m.Nw) {
      v = F[k + m.Nk * (long)t + oab];
      if (f != 0) v = v * (pp * dprt);
    } else {
      int j = m.M - 1 - t;
      if (f == 0) v = conj(F[k + m.Nk * (long)j + oab]) * (sgn * dprt * pp);
      else v = F[k + m.Nk * (long)j + oba] * sgn;
    }
    grid[m.sp[i] + m.Nkall * (long)t] = cpx<R>(v);
  }
}

template <typename R>
void scatter_anom(cpx<R>* grid, const zd* F, int a, int b, int No, double pp, double sgn, double dprt,
                  const Mesh& m, hipStream_t s) {
  hipLaunchKernelGGL((k_scatter_anom<R>), dim3(nblocks(m.N())), dim3(TPB), 0, s, grid, F, a, b, No, pp, sgn, dprt,
                     view(m));
  GMOD_CHECK_LAUNCH();
}

// sum_n U_an conj(U_bn) g_n(k, +-tau_s); -k uses the conjugated coefficient
// (fchi.f90:703-729, 791-812).  g(+tau) = -e^{-xi tau}(1-f), g(-tau) = e^{xi tau} f,
// both written in the overflow-safe form of the reference.
template <typename R>
__global__ void k_scatter_tau_ref(cpx<R>* __restrict__ grid, const double* __restrict__ eig,
                                  const zd* __restrict__ uni, int a, int b, int No, double mu, double temp,
                                  int which, MeshView m) {
  const long N = m.Nkall * m.M;
  const double beta = 1.0 / temp;
  const double dtau = beta / (double)m.M;
  for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < N; g += (long)gridDim.x * blockDim.x) {
    long i = g % m.Nkall;
    int s = (int)(g / m.Nkall);
    int k = m.irr[i];
    int f = m.flag[i];
    double tau = (double)s * dtau;
    zd acc(0, 0);
    for (int n = 0; n < No; ++n) {
      double xi = eig[n + (long)No * k] - mu;
      double gv;
      if (xi >= 0.0) {
        double d = 1.0 + exp(-beta * xi);
        gv = which == 0 ? -exp(-xi * tau) / d : exp(-xi * (beta - tau)) / d;
      } else {
        double d = 1.0 + exp(beta * xi);
        gv = which == 0 ? -exp(xi * (beta - tau)) / d : exp(xi * tau) / d;
      }
      zd ua = uni[a + (long)No * (n + (long)No * k)];
      zd ub = uni[b + (long)No * (n + (long)No * k)];
      zd c = ua * conj(ub);
      if (f != 0) c = conj(c);
      acc += c * gv;
    }
    grid[m.sp[i] + m.Nkall * (long)s] = cpx<R>(acc);
  }
}

template <typename R>
void scatter_tau_ref(cpx<R>* grid, const double* eig, const zd* uni, int a, int b, int No, double mu,
                     double temp, int which, const Mesh& m, hipStream_t s) {
  hipLaunchKernelGGL((k_scatter_tau_ref<R>), dim3(nblocks(m.N())), dim3(TPB), 0, s, grid, eig, uni, a, b, No, mu,
                     temp, which, view(m));
  GMOD_CHECK_LAUNCH();
}

// ------------------------------------------------------------------ products
template <typename R>
__global__ void k_product_rev(cpx<R>* __restrict__ out, const cpx<R>* __restrict__ A, const cpx<R>* __restrict__ B,
                              R coef, int rev_t, int acc, MeshView m) {
  const long N = m.Nkall * m.M;
  for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < N; g += (long)gridDim.x * blockDim.x) {
    long sp_ = g % m.Nkall;
    long t = g / m.Nkall;
    long rt = rev_t ? (t == 0 ? 0 : m.M - t) : t;
    cpx<R> v = A[g] * B[m.rev[sp_] + m.Nkall * rt] * coef;
    out[g] = acc ? out[g] + v : v;
  }
}

template <typename R>
void product_rev(cpx<R>* out, const cpx<R>* A, const cpx<R>* B, R coef, bool rev_t, bool accumulate,
                 const Mesh& m, hipStream_t s) {
  hipLaunchKernelGGL((k_product_rev<R>), dim3(nblocks(m.N())), dim3(TPB), 0, s, out, A, B, coef, (int)rev_t,
                     (int)accumulate, view(m));
  GMOD_CHECK_LAUNCH();
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

// ------------------------------------------------------------------ gather
template <typename R>
__global__ void k_gather(zd* __restrict__ dst, const cpx<R>* __restrict__ grid, double w, int acc, zd J,
                         const zd* __restrict__ sawc, MeshView m) {
  const long n = m.Nk * (long)m.Nw;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    long k = e % m.Nk;
    long j = e / m.Nk;
    int sp_ = m.irr_sp[k];
    zd v = sp_ >= 0 ? zd(grid[sp_ + m.Nkall * j]) * w : zd(0, 0);
    if (sawc) v += J * sawc[j];
    dst[e] = acc ? dst[e] + v : v;
  }
}

template <typename R>
void gather(zd* dst, const cpx<R>* grid, double weight, bool accumulate, zd J, const zd* sawc, const Mesh& m,
            hipStream_t s) {
  hipLaunchKernelGGL((k_gather<R>), dim3(nblocks(m.Nk * (long)m.Nw)), dim3(TPB), 0, s, dst, grid, weight,
                     (int)accumulate, J, sawc, view(m));
  GMOD_CHECK_LAUNCH();
}

// ------------------------------------------------------------------ elementwise
__global__ void k_conj_copy(zd* dst, const zd* src, long n) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
    dst[e] = conj(src[e]);
}
void conj_copy(zd* dst, const zd* src, long n, hipStream_t s) {
  hipLaunchKernelGGL(k_conj_copy, dim3(nblocks(n)), dim3(TPB), 0, s, dst, src, n);
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

__global__ void k_lincomb(zd* y, const zd* a, double ca, const zd* b, double cb, long n) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
    y[e] = a[e] * ca + b[e] * cb;
}
void lincomb(zd* y, const zd* a, double ca, const zd* b, double cb, long n, hipStream_t s) {
  hipLaunchKernelGGL(k_lincomb, dim3(nblocks(n)), dim3(TPB), 0, s, y, a, ca, b, cb, n);
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

// ------------------------------------------------------------------ instantiations
#define INST(R)                                                                                          \
  template void scatter_green<R>(cpx<R>*, const zd*, int, int, int, const Mesh&, hipStream_t);           \
  template void scatter_boson<R>(cpx<R>*, const zd*, int, int, int, zd, const Mesh&, hipStream_t);       \
  template void scatter_anom<R>(cpx<R>*, const zd*, int, int, int, double, double, double, const Mesh&,  \
                                hipStream_t);                                                            \
  template void scatter_tau_ref<R>(cpx<R>*, const double*, const zd*, int, int, int, double, double, int, \
                                   const Mesh&, hipStream_t);                                            \
  template void product_rev<R>(cpx<R>*, const cpx<R>*, const cpx<R>*, R, bool, bool, const Mesh&,        \
                               hipStream_t);                                                             \
  template void mul_acc<R>(cpx<R>*, const cpx<R>*, const cpx<R>*, bool, long, hipStream_t);              \
  template void gather<R>(zd*, const cpx<R>*, double, bool, zd, const zd*, const Mesh&, hipStream_t);
INST(double)
INST(float)

}  // namespace gmod
