// Surrogate network of the FLEX warm start: a fully connected network
// (SiLU hidden layers, linear output) trained and evaluated by this library's
// own kernels.  GEMMs run on the matrix cores with BF16 operands and FP32
// accumulation (v_mfma_f32_16x16x16_bf16); the master weights, biases, Adam
// moments, pre-activations and the loss stay FP32.  No reference counterpart.
#include <cmath>
#include <memory>
#include <random>

#include "../../include/gmod.h"
#include "gmod_common.h"

namespace gmod {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__host__ __device__ inline uint16_t f2bf(float x) {  // round to nearest even
  uint32_t u;
  memcpy(&u, &x, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t h) {
  uint32_t u = ((uint32_t)h) << 16;
  float x;
  memcpy(&x, &u, 4);
  return x;
}

constexpr int GT = 64;  // output tile (4 waves x 32 x 32)
constexpr int GK = 16;  // k step = one MFMA
constexpr int GP = GK + 4;

// C[m,n] = alpha * sum_k A(m,k) B(k,n) + beta C[m,n] + bias[n]
// A(m,k) = A[m sam + k sak], B(k,n) = B[k sbk + n sbn]  (BF16), C FP32 row-major
__global__ void __launch_bounds__(256)
k_gemm_bf16(const uint16_t* __restrict__ A, long sam, long sak, const uint16_t* __restrict__ B, long sbk, long sbn,
            float* __restrict__ C, long ldc, int M, int N, int K, float alpha, float beta,
            const float* __restrict__ bias) {
  __shared__ uint16_t As[GT][GP];   // [m][k]
  __shared__ uint16_t Bs[GT][GP];   // [n][k]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f32x4 acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  const bool a_k_fast = sak == 1, b_k_fast = sbk == 1;
  for (int k0 = 0; k0 < K; k0 += GK) {
    for (int i = tid; i < GT * GK; i += 256) {
      int m, k;
      if (a_k_fast) { m = i / GK; k = i % GK; } else { k = i / GT; m = i % GT; }
      const int gm = m0 + m, gk = k0 + k;
      As[m][k] = (gm < M && gk < K) ? A[gm * sam + gk * sak] : (uint16_t)0;
      int n, kb;
      if (b_k_fast) { n = i / GK; kb = i % GK; } else { kb = i / GT; n = i % GT; }
      const int gn = n0 + n, gkb = k0 + kb;
      Bs[n][kb] = (gn < N && gkb < K) ? B[gkb * sbk + gn * sbn] : (uint16_t)0;
    }
    __syncthreads();
    const int r = lane & 15, kq = (lane >> 4) * 4;
#pragma unroll
    for (int ti = 0; ti < 2; ++ti) {
      bf16x4 a;
#pragma unroll
      for (int e = 0; e < 4; ++e) a[e] = (short)As[wm + ti * 16 + r][kq + e];
#pragma unroll
      for (int tj = 0; tj < 2; ++tj) {
        bf16x4 b;
#pragma unroll
        for (int e = 0; e < 4; ++e) b[e] = (short)Bs[wn + tj * 16 + r][kq + e];
        acc[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[ti][tj], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  const int col_l = lane & 15, row_q = (lane >> 4) * 4;
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int row = m0 + wm + ti * 16 + row_q + e, col = n0 + wn + tj * 16 + col_l;
        if (row < M && col < N) {
          float v = alpha * acc[ti][tj][e];
          if (beta != 0.f) v += beta * C[row * ldc + col];
          if (bias) v += bias[col];
          C[row * ldc + col] = v;
        }
      }
}

static void gemm(const uint16_t* A, long sam, long sak, const uint16_t* B, long sbk, long sbn, float* C, long ldc,
                 int M, int N, int K, float alpha, float beta, const float* bias, hipStream_t s) {
  dim3 grid((N + GT - 1) / GT, (M + GT - 1) / GT);
  hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(256), 0, s, A, sam, sak, B, sbk, sbn, C, ldc, M, N, K, alpha, beta, bias);
  GMOD_CHECK_LAUNCH();
}

static unsigned ew_blocks(long n) { return (unsigned)std::max<long>(1, std::min<long>((n + 255) / 256, 1L << 16)); }

__global__ void k_to_bf16(uint16_t* out, const float* in, long n) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
    out[e] = f2bf(in[e]);
}

__device__ inline float sigm(float z) { return 1.f / (1.f + __expf(-z)); }

// H = bf16(silu(Z))
__global__ void k_silu_fwd(uint16_t* H, const float* Z, long n) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const float z = Z[e];
    H[e] = f2bf(z * sigm(z));
  }
}
// dZ = dH * silu'(Z) (dH FP32 in place), and its BF16 copy
__global__ void k_silu_bwd(float* dZ, uint16_t* dZ16, const float* Z, long n) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const float z = Z[e], s = sigm(z);
    const float g = dZ[e] * (s * (1.f + z * (1.f - s)));
    dZ[e] = g;
    dZ16[e] = f2bf(g);
  }
}
// column sums of a (rows x cols) row-major matrix (deterministic: one thread per column)
__global__ void k_colsum(float* out, const float* a, int rows, int cols) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < cols; c += gridDim.x * blockDim.x) {
    float s = 0.f;
    for (int r = 0; r < rows; ++r) s += a[(long)r * cols + c];
    out[c] = s;
  }
}
// dY = 2 (Y - T) / n; per-block partial sums of (Y - T)^2
__global__ void k_mse(float* dY, const float* Y, const float* T, long n, double* part) {
  __shared__ double red[256];
  double acc = 0.0;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const float d = Y[e] - T[e];
    dY[e] = 2.f * d / (float)n;
    acc += (double)d * (double)d;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}
__global__ void k_sum_parts(double* out, const double* part, int np, double scale) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < np; ++i) s += part[i];
    *out = s * scale;
  }
}
// Adam (FP32 master weights) and the BF16 copy of the weights
__global__ void k_adam(float* w, uint16_t* w16, const float* g, float* m, float* v, long n, float lr, float b1,
                       float b2, float eps, float c1, float c2) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const float ge = g[e];
    const float me = b1 * m[e] + (1.f - b1) * ge;
    const float ve = b2 * v[e] + (1.f - b2) * ge * ge;
    m[e] = me;
    v[e] = ve;
    const float wn = w[e] - lr * (me / c1) / (sqrtf(ve / c2) + eps);
    w[e] = wn;
    if (w16) w16[e] = f2bf(wn);
  }
}
// rows idx[0..B) of X (n x din) and T (n x dout) -> H0 (BF16) and Tb
__global__ void k_gather_rows(uint16_t* H0, float* Tb, const float* X, const float* T, const int* idx, int B, int din,
                              int dout) {
  const long nx = (long)B * din, nt = (long)B * dout;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < nx + nt; e += (long)gridDim.x * blockDim.x) {
    if (e < nx) {
      const long r = e / din, c = e % din;
      H0[e] = f2bf(X[(long)idx[r] * din + c]);
    } else {
      const long f = e - nx;
      const long r = f / dout, c = f % dout;
      Tb[f] = T[(long)idx[r] * dout + c];
    }
  }
}

struct MLP {
  int L = 0;
  std::vector<int> dims;
  std::vector<DBuf<float>> W, b, gW, gb, mW, vW, mb, vb, Z;
  std::vector<DBuf<uint16_t>> W16, H, dZ16;
  DBuf<float> dZ, Tb;
  DBuf<double> part, lossv;
  DBuf<float> Xall, Tall;
  DBuf<int> idx;
  long ndata = 0, step = 0;
  int cap = 0;  // batch capacity of the activation buffers
  long nparams() const {
    long n = 0;
    for (int i = 0; i < L; ++i) n += (long)dims[i] * dims[i + 1] + dims[i + 1];
    return n;
  }
  void reserve(int B) {
    if (B <= cap) return;
    cap = B;
    Z.clear(); H.clear(); dZ16.clear();
    Z.resize(L); H.resize(L + 1); dZ16.resize(L);
    int widest = 0;
    for (int i = 0; i <= L; ++i) {
      H[i].alloc((size_t)B * dims[i]);
      widest = std::max(widest, dims[i]);
    }
    for (int i = 0; i < L; ++i) {
      Z[i].alloc((size_t)B * dims[i + 1]);
      dZ16[i].alloc((size_t)B * dims[i + 1]);
    }
    dZ.alloc((size_t)B * widest);
    Tb.alloc((size_t)B * dims[L]);
    idx.alloc(B);
  }
  void forward(int B, hipStream_t s) {
    for (int i = 0; i < L; ++i) {
      const int din = dims[i], dout = dims[i + 1];
      gemm(H[i].p, din, 1, W16[i].p, dout, 1, Z[i].p, dout, B, dout, din, 1.f, 0.f, b[i].p, s);
      if (i + 1 < L) {
        hipLaunchKernelGGL(k_silu_fwd, dim3(ew_blocks((long)B * dout)), dim3(256), 0, s, H[i + 1].p, Z[i].p,
                           (long)B * dout);
        GMOD_CHECK_LAUNCH();
      }
    }
  }
  // MSE against Tb; gradients into gW, gb; loss into lossv[slot]
  void backward(int B, int slot, hipStream_t s) {
    const long nout = (long)B * dims[L];
    const int np = (int)ew_blocks(nout);
    if (part.n < (size_t)np) part.alloc(np);
    hipLaunchKernelGGL(k_mse, dim3(np), dim3(256), 0, s, dZ.p, Z[L - 1].p, Tb.p, nout, part.p);
    GMOD_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(1), 0, s, lossv.p + slot, part.p, np, 1.0 / (double)nout);
    GMOD_CHECK_LAUNCH();
    for (int i = L - 1; i >= 0; --i) {
      const int din = dims[i], dout = dims[i + 1];
      const long nz = (long)B * dout;
      if (i == L - 1) {
        hipLaunchKernelGGL(k_to_bf16, dim3(ew_blocks(nz)), dim3(256), 0, s, dZ16[i].p, dZ.p, nz);
        GMOD_CHECK_LAUNCH();
      }
      // gb = colsum(dZ), gW = H^T dZ
      hipLaunchKernelGGL(k_colsum, dim3((dout + 255) / 256), dim3(256), 0, s, gb[i].p, dZ.p, B, dout);
      GMOD_CHECK_LAUNCH();
      gemm(H[i].p, 1, din, dZ16[i].p, dout, 1, gW[i].p, dout, din, dout, B, 1.f, 0.f, nullptr, s);
      if (i > 0) {
        // dH = dZ W^T, then dZ_{i-1} = dH silu'(Z_{i-1})
        gemm(dZ16[i].p, dout, 1, W16[i].p, 1, dout, dZ.p, din, B, din, dout, 1.f, 0.f, nullptr, s);
        hipLaunchKernelGGL(k_silu_bwd, dim3(ew_blocks((long)B * din)), dim3(256), 0, s, dZ.p, dZ16[i - 1].p,
                           Z[i - 1].p, (long)B * din);
        GMOD_CHECK_LAUNCH();
      }
    }
  }
  void adam(double lr, hipStream_t s) {
    ++step;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.f - std::pow(b1, (float)step), c2 = 1.f - std::pow(b2, (float)step);
    for (int i = 0; i < L; ++i) {
      const long nw = (long)dims[i] * dims[i + 1], nb = dims[i + 1];
      hipLaunchKernelGGL(k_adam, dim3(ew_blocks(nw)), dim3(256), 0, s, W[i].p, W16[i].p, gW[i].p, mW[i].p, vW[i].p,
                         nw, (float)lr, b1, b2, eps, c1, c2);
      GMOD_CHECK_LAUNCH();
      hipLaunchKernelGGL(k_adam, dim3(ew_blocks(nb)), dim3(256), 0, s, b[i].p, (uint16_t*)nullptr, gb[i].p, mb[i].p,
                         vb[i].p, nb, (float)lr, b1, b2, eps, c1, c2);
      GMOD_CHECK_LAUNCH();
    }
  }
  void refresh_bf16(hipStream_t s) {
    for (int i = 0; i < L; ++i) {
      const long nw = (long)dims[i] * dims[i + 1];
      hipLaunchKernelGGL(k_to_bf16, dim3(ew_blocks(nw)), dim3(256), 0, s, W16[i].p, W[i].p, nw);
      GMOD_CHECK_LAUNCH();
    }
  }
  void upload_input(const float* X, int B, hipStream_t s) {
    DBuf<float> xf((size_t)B * dims[0]);
    xf.upload(X, (size_t)B * dims[0], s);
    hipLaunchKernelGGL(k_to_bf16, dim3(ew_blocks((long)B * dims[0])), dim3(256), 0, s, H[0].p, xf.p,
                       (long)B * dims[0]);
    GMOD_CHECK_LAUNCH();
    GMOD_HIP(hipStreamSynchronize(s));  // xf is released at return
  }
};

static std::vector<std::unique_ptr<MLP>>& nets() {
  static std::vector<std::unique_ptr<MLP>> v;
  return v;
}
static MLP& net(int64_t h) {
  auto& v = nets();
  if (h < 0 || h >= (int64_t)v.size() || !v[h]) throw Error("invalid network handle");
  return *v[h];
}

}  // namespace gmod

using namespace gmod;

#define NN_TRY try { std::lock_guard<std::mutex> g_(ctx().lock);
#define NN_CATCH(ret) } catch (std::exception & e) { set_error(e.what()); return ret; }

extern "C" int64_t gmod_nn_create(int64_t nlayers, const int64_t* dims, uint64_t seed) {
  NN_TRY
  if (nlayers < 1) throw Error("network needs at least one layer");
  auto p = std::make_unique<MLP>();
  MLP& m = *p;
  m.L = (int)nlayers;
  for (int i = 0; i <= m.L; ++i) {
    if (dims[i] < 1) throw Error("layer width must be positive");
    m.dims.push_back((int)dims[i]);
  }
  hipStream_t s = ctx().stream;
  std::mt19937_64 rng(seed);
  m.W.resize(m.L); m.b.resize(m.L); m.gW.resize(m.L); m.gb.resize(m.L);
  m.mW.resize(m.L); m.vW.resize(m.L); m.mb.resize(m.L); m.vb.resize(m.L); m.W16.resize(m.L);
  for (int i = 0; i < m.L; ++i) {
    const int din = m.dims[i], dout = m.dims[i + 1];
    const long nw = (long)din * dout;
    // uniform in +-sqrt(6 / (din + dout)) (Glorot), zero biases
    const double a = std::sqrt(6.0 / (double)(din + dout));
    std::uniform_real_distribution<double> U(-a, a);
    std::vector<float> h(nw);
    for (auto& x : h) x = (float)U(rng);
    m.W[i].alloc(nw); m.W[i].upload(h.data(), nw, s);
    std::vector<float> z(std::max<long>(nw, dout), 0.f);
    m.b[i].alloc(dout); m.b[i].upload(z.data(), dout, s);
    m.gW[i].alloc(nw); m.gb[i].alloc(dout);
    m.mW[i].alloc(nw); m.mW[i].upload(z.data(), nw, s);
    m.vW[i].alloc(nw); m.vW[i].upload(z.data(), nw, s);
    m.mb[i].alloc(dout); m.mb[i].upload(z.data(), dout, s);
    m.vb[i].alloc(dout); m.vb[i].upload(z.data(), dout, s);
    m.W16[i].alloc(nw);
    GMOD_HIP(hipStreamSynchronize(s));
  }
  m.refresh_bf16(s);
  m.lossv.alloc(1);
  GMOD_HIP(hipStreamSynchronize(s));
  nets().push_back(std::move(p));
  return (int64_t)nets().size() - 1;
  NN_CATCH(-1)
}

extern "C" void gmod_nn_destroy(int64_t h) {
  NN_TRY
  net(h);
  nets()[h].reset();
  NN_CATCH()
}

extern "C" int64_t gmod_nn_num_params(int64_t h) {
  NN_TRY
  return net(h).nparams();
  NN_CATCH(-1)
}

// parameters / gradients as one flat FP32 vector: W0 (din x dout, row-major), b0, W1, b1, ...
static void flat_io(MLP& m, float* host, bool get, bool grads) {
  hipStream_t s = ctx().stream;
  long off = 0;
  for (int i = 0; i < m.L; ++i) {
    const long nw = (long)m.dims[i] * m.dims[i + 1], nb = m.dims[i + 1];
    DBuf<float>& w = grads ? m.gW[i] : m.W[i];
    DBuf<float>& b = grads ? m.gb[i] : m.b[i];
    if (get) { w.download(host + off, nw, s); b.download(host + off + nw, nb, s); }
    else { w.upload(host + off, nw, s); b.upload(host + off + nw, nb, s); }
    off += nw + nb;
  }
  if (!get) m.refresh_bf16(s);
  GMOD_HIP(hipStreamSynchronize(s));
}

extern "C" void gmod_nn_get_params(int64_t h, float* out) {
  NN_TRY
  flat_io(net(h), out, true, false);
  NN_CATCH()
}
extern "C" void gmod_nn_set_params(int64_t h, const float* in) {
  NN_TRY
  flat_io(net(h), const_cast<float*>(in), false, false);
  NN_CATCH()
}

extern "C" void gmod_nn_forward(int64_t h, const float* X, int64_t B, float* Y) {
  NN_TRY
  MLP& m = net(h);
  hipStream_t s = ctx().stream;
  m.reserve((int)B);
  m.upload_input(X, (int)B, s);
  m.forward((int)B, s);
  m.Z[m.L - 1].download(Y, (size_t)B * m.dims[m.L], s);
  GMOD_HIP(hipStreamSynchronize(s));
  NN_CATCH()
}

extern "C" double gmod_nn_grad(int64_t h, const float* X, const float* T, int64_t B, float* grad) {
  NN_TRY
  MLP& m = net(h);
  hipStream_t s = ctx().stream;
  m.reserve((int)B);
  m.upload_input(X, (int)B, s);
  m.Tb.upload(T, (size_t)B * m.dims[m.L], s);
  m.forward((int)B, s);
  m.backward((int)B, 0, s);
  double loss = 0.0;
  m.lossv.download(&loss, 1, s);
  GMOD_HIP(hipStreamSynchronize(s));
  if (grad) flat_io(m, grad, true, true);
  return loss;
  NN_CATCH(-1.0)
}

extern "C" void gmod_nn_set_data(int64_t h, const float* X, const float* T, int64_t n) {
  NN_TRY
  MLP& m = net(h);
  hipStream_t s = ctx().stream;
  m.Xall.alloc((size_t)n * m.dims[0]);
  m.Tall.alloc((size_t)n * m.dims[m.L]);
  m.Xall.upload(X, (size_t)n * m.dims[0], s);
  m.Tall.upload(T, (size_t)n * m.dims[m.L], s);
  m.ndata = n;
  GMOD_HIP(hipStreamSynchronize(s));
  NN_CATCH()
}

// one pass over order[0..n_used) in mini-batches (the last, partial batch
// included); Adam step per batch; returns the mean of the batch losses
extern "C" double gmod_nn_train_epoch(int64_t h, const int64_t* order, int64_t n_used, int64_t batch, double lr) {
  NN_TRY
  MLP& m = net(h);
  hipStream_t s = ctx().stream;
  if (batch < 1 || n_used < 1) throw Error("empty training epoch");
  m.reserve((int)batch);
  const long nb = (n_used + batch - 1) / batch;
  DBuf<double> losses(nb);
  std::swap(m.lossv, losses);
  std::vector<int> hidx(n_used);
  for (long i = 0; i < n_used; ++i) {
    if (order[i] < 0 || order[i] >= m.ndata) throw Error("training index out of range");
    hidx[i] = (int)order[i];
  }
  DBuf<int> didx(n_used);
  didx.upload(hidx.data(), n_used, s);
  for (long bi = 0; bi < nb; ++bi) {
    const int B = (int)std::min<long>(batch, n_used - bi * batch);
    const long tot = (long)B * (m.dims[0] + m.dims[m.L]);
    hipLaunchKernelGGL(k_gather_rows, dim3(ew_blocks(tot)), dim3(256), 0, s, m.H[0].p, m.Tb.p, m.Xall.p, m.Tall.p,
                       didx.p + bi * batch, B, m.dims[0], m.dims[m.L]);
    GMOD_CHECK_LAUNCH();
    m.forward(B, s);
    m.backward(B, (int)bi, s);
    m.adam(lr, s);
  }
  std::vector<double> hl(nb);
  m.lossv.download(hl.data(), nb, s);
  GMOD_HIP(hipStreamSynchronize(s));
  std::swap(m.lossv, losses);
  double acc = 0.0;
  for (double v : hl) acc += v;
  return acc / (double)nb;
  NN_CATCH(-1.0)
}
