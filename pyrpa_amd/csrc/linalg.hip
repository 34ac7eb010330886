// Batched small dense complex linear algebra (FP64) for gfx950.
//
// * rpa_vertex / rpa_solve: one 64-lane wavefront per (q, i nu) matrix.  The
//   matrix and both right-hand sides live in LDS; the LU with partial pivoting
//   (pivot = max |Re|+|Im|, as LAPACK's izamax) updates the trailing block with
//   one lane per element.
// * dyson / filling_parts / green0 / anomalous_f: one lane per (k, i w) with
//   the No x No matrix held in registers (No <= 8).
// * max_re_eig: one lane per matrix; Hessenberg reduction + shifted complex QR.
#include <cfloat>
#include <cmath>

#include "linalg.h"
#include "kernels.h"

namespace gmod {

constexpr double PI_D = 3.141592653589793238462643383279;  // fmod.f90:4

static inline unsigned nblk(long n, int tpb, long cap = 1L << 20) {
  long b = (n + tpb - 1) / tpb;
  if (b > cap) b = cap;
  return (unsigned)(b < 1 ? 1 : b);
}

// ====================================================================== LDS LU
__device__ bool lds_lu_solve(zd* Mx, zd* X, int n, int nrhs, int* shi) {
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int k = 0; k < n; ++k) {
    if (tid == 0) {
      int p = k;
      double best = abs1(Mx[k + n * k]);
      for (int r = k + 1; r < n; ++r) {
        double v = abs1(Mx[r + n * k]);
        if (v > best) { best = v; p = r; }
      }
      shi[0] = p;
      shi[1] = best == 0.0 ? 1 : 0;
    }
    __syncthreads();
    const int p = shi[0];
    if (shi[1]) return false;
    if (p != k) {
      for (int c = tid; c < n; c += nt) { zd t = Mx[k + n * c]; Mx[k + n * c] = Mx[p + n * c]; Mx[p + n * c] = t; }
      for (int c = tid; c < nrhs; c += nt) { zd t = X[k + n * c]; X[k + n * c] = X[p + n * c]; X[p + n * c] = t; }
    }
    __syncthreads();
    const zd d = Mx[k + n * k];
    for (int r = k + 1 + tid; r < n; r += nt) Mx[r + n * k] = Mx[r + n * k] / d;
    __syncthreads();
    const int rows = n - k - 1;
    const int cols = rows + nrhs;
    for (int e = tid; e < rows * cols; e += nt) {
      int r = k + 1 + e % rows;
      int cc = e / rows;
      zd l = Mx[r + n * k];
      if (cc < rows) {
        int c = k + 1 + cc;
        Mx[r + n * c] -= l * Mx[k + n * c];
      } else {
        int c = cc - rows;
        X[r + n * c] -= l * X[k + n * c];
      }
    }
    __syncthreads();
  }
  for (int c = tid; c < nrhs; c += nt) {
    for (int k = n - 1; k >= 0; --k) {
      zd x = X[k + n * c];
      for (int j = k + 1; j < n; ++j) x -= Mx[k + n * j] * X[j + n * c];
      X[k + n * c] = x / Mx[k + n * k];
    }
  }
  __syncthreads();
  return true;
}

__global__ void __launch_bounds__(64)
k_rpa_vertex(zd* chi, BatchLayout L, long batch, int n, const double* S, const double* C, int mode, int* err) {
  extern __shared__ unsigned char sm[];
  const int nn = n * n;
  double* Ss = reinterpret_cast<double*>(sm);
  double* Cs = Ss + nn;
  zd* A = reinterpret_cast<zd*>(Cs + nn);
  zd* M1 = A + nn;
  zd* X1 = M1 + nn;
  zd* M2 = X1 + nn;
  zd* X2 = M2 + nn;
  __shared__ int shi[2];
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int e = tid; e < nn; e += nt) { Ss[e] = S[e]; Cs[e] = C[e]; }
  for (long b = blockIdx.x; b < batch; b += gridDim.x) {
    zd* base = chi + b * L.bs;
    for (int e = tid; e < nn; e += nt) A[e] = base[e * L.es];
    __syncthreads();
    for (int e = tid; e < nn; e += nt) {
      int p = e % n, q = e / n;
      zd as(0, 0), ac(0, 0);
      for (int r = 0; r < n; ++r) {
        zd a = A[p + n * r];
        as += a * Ss[r + n * q];
        ac += a * Cs[r + n * q];
      }
      double dl = (p == q) ? 1.0 : 0.0;
      M1[e] = zd(dl, 0) - as;
      X1[e] = as;
      M2[e] = zd(dl, 0) + ac;
      X2[e] = ac;
    }
    __syncthreads();
    if (mode == 0) {
      for (int e = tid; e < nn; e += nt) A[e] = X1[e] + X2[e];  // chi0 S + chi0 C
      __syncthreads();
    }
    bool ok = lds_lu_solve(M1, X1, n, n, shi);
    ok = ok && lds_lu_solve(M2, X2, n, n, shi);
    if (!ok) {
      if (tid == 0) atomicOr(err, 1);
      __syncthreads();
      continue;
    }
    for (int e = tid; e < nn; e += nt) {
      int p = e % n, q = e / n;
      zd v;
      if (mode == 0) {
        v = zd(1.5 * Ss[e] - 0.5 * Cs[e], 0);
        for (int r = 0; r < n; ++r) {
          double s = Ss[p + n * r], c = Cs[p + n * r];
          v += X1[r + n * q] * (1.5 * s) + X2[r + n * q] * (0.5 * c) - A[r + n * q] * (0.25 * (s + c));
        }
      } else if (mode == 1) {
        v = zd(0.5 * (Ss[e] + Cs[e]), 0);
        for (int r = 0; r < n; ++r)
          v += X1[r + n * q] * (1.5 * Ss[p + n * r]) - X2[r + n * q] * (0.5 * Cs[p + n * r]);
      } else {
        v = zd(0.5 * (Ss[e] - Cs[e]), 0);
        for (int r = 0; r < n; ++r)
          v -= X1[r + n * q] * (0.5 * Ss[p + n * r]) + X2[r + n * q] * (0.5 * Cs[p + n * r]);
      }
      base[e * L.es] = v;
    }
    __syncthreads();
  }
}

// Vertices of the superconducting loop (ffeliash.f90:386-457) per (q, i nu):
// on entry vs holds chi0_S and vd chi0_C; on exit vs holds V_sigma and vd
// V_Delta (singlet or triplet).  chi_s, chi_c by pivoted LUs in LDS as in
// k_rpa_vertex, the S system on chi0_S and the C system on chi0_C.
__global__ void __launch_bounds__(64)
k_sc_vertex(zd* vs, zd* vd, BatchLayout L, long batch, int n, const double* S, const double* C, int singlet,
            int* err) {
  extern __shared__ unsigned char sm[];
  const int nn = n * n;
  double* Ss = reinterpret_cast<double*>(sm);
  double* Cs = Ss + nn;
  zd* A = reinterpret_cast<zd*>(Cs + nn);
  zd* M1 = A + nn;
  zd* X1 = M1 + nn;
  zd* M2 = X1 + nn;
  zd* X2 = M2 + nn;
  __shared__ int shi[2];
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int e = tid; e < nn; e += nt) { Ss[e] = S[e]; Cs[e] = C[e]; }
  for (long b = blockIdx.x; b < batch; b += gridDim.x) {
    zd* bs = vs + b * L.bs;
    zd* bd = vd + b * L.bs;
    for (int e = tid; e < nn; e += nt) A[e] = bs[e * L.es];  // chi0_S
    __syncthreads();
    for (int e = tid; e < nn; e += nt) {
      int p = e % n, q = e / n;
      zd as(0, 0);
      for (int r = 0; r < n; ++r) as += A[p + n * r] * Ss[r + n * q];
      M1[e] = zd(p == q ? 1.0 : 0.0, 0) - as;
      X1[e] = as;
    }
    __syncthreads();
    for (int e = tid; e < nn; e += nt) A[e] = bd[e * L.es];  // chi0_C
    __syncthreads();
    for (int e = tid; e < nn; e += nt) {
      int p = e % n, q = e / n;
      zd ac(0, 0);
      for (int r = 0; r < n; ++r) ac += A[p + n * r] * Cs[r + n * q];
      M2[e] = zd(p == q ? 1.0 : 0.0, 0) + ac;
      X2[e] = ac;
    }
    __syncthreads();
    for (int e = tid; e < nn; e += nt) A[e] = X1[e] + X2[e];  // chi0_S S + chi0_C C
    __syncthreads();
    bool ok = lds_lu_solve(M1, X1, n, n, shi);
    ok = ok && lds_lu_solve(M2, X2, n, n, shi);
    if (!ok) {
      if (tid == 0) atomicOr(err, 1);
      __syncthreads();
      continue;
    }
    for (int e = tid; e < nn; e += nt) {
      int p = e % n, q = e / n;
      zd v = zd(1.5 * Ss[e] - 0.5 * Cs[e], 0);
      zd d = singlet ? zd(0.5 * (Ss[e] + Cs[e]), 0) : zd(0.5 * (Ss[e] - Cs[e]), 0);
      for (int r = 0; r < n; ++r) {
        const double sv = Ss[p + n * r], cv = Cs[p + n * r];
        const zd xs = X1[r + n * q], xc = X2[r + n * q];
        v += xs * (1.5 * sv) + xc * (0.5 * cv) - A[r + n * q] * (0.25 * (sv + cv));
        d += singlet ? xs * (1.5 * sv) - xc * (0.5 * cv) : zd(0, 0) - xs * (0.5 * sv) - xc * (0.5 * cv);
      }
      bs[e * L.es] = v;
      bd[e * L.es] = d;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------- register LU
// Column-per-lane variant for Nchi = N in {1,4,9,16,25}.  Each solve uses 2N
// lanes: lane c < N holds column c of M = I -+ chi0 K (K = S or C), lane
// N + c holds column c of the right-hand side chi0 K.  The LU with partial
// pivoting runs on the M lanes (multipliers broadcast through LDS); L, U and
// the pivots are then published in LDS and every RHS lane solves its column
// locally.  When 4N <= 64 both solves of one (q, i nu) run side by side in a
// wavefront (several matrices per wavefront for small N); otherwise the two
// solves run one after the other.  Matrices of consecutive (q, i nu) are
// staged through LDS so the global loads and stores are coalesced.
template <int N>
struct VCfg {
  static constexpr bool PAR = 4 * N <= 64;
  static constexpr int LPM = PAR ? 4 * N : 2 * N;     // lanes per matrix
  static constexpr int MPW = PAR ? 64 / (4 * N) : 1;  // matrices per wavefront
  static constexpr int WAVES = PAR ? 4 : 2;
  static constexpr int MPB = MPW * WAVES;             // matrices per block
  static constexpr int THREADS = 64 * WAVES;
};

template <int N>
__global__ void __launch_bounds__(VCfg<N>::THREADS, 2)
k_vertex_reg(zd* __restrict__ chi, long batch, long es, const double* __restrict__ S, const double* __restrict__ C,
             int mode, int* err) {
  using Cf = VCfg<N>;
  constexpr int NN = N * N;
  constexpr int MPB = Cf::MPB;
  __shared__ double2 Abuf_[MPB * NN];
  __shared__ double2 Vbuf_[MPB * NN];
  __shared__ double2 LUb_[MPB * 2 * NN];
  __shared__ double2 bc_[MPB * 2 * N];
  __shared__ double2 rinv_[MPB * 2 * N];
  __shared__ int piv[MPB * 2 * N];
  __shared__ double Ks[3 * NN];
  zd* rinv = reinterpret_cast<zd*>(rinv_);
  zd* Abuf = reinterpret_cast<zd*>(Abuf_);
  zd* Vbuf = reinterpret_cast<zd*>(Vbuf_);
  zd* LUb = reinterpret_cast<zd*>(LUb_);
  zd* bc = reinterpret_cast<zd*>(bc_);
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int mw = lane / Cf::LPM;  // matrix within the wavefront
  const int lm = lane % Cf::LPM;
  const int mb = wave * Cf::MPW + mw;  // matrix within the block
  const bool inmat = mw < Cf::MPW;
  const double* Sg = Ks;
  const double* Cg = Ks + NN;
  const double* Wg = Ks + 2 * NN;  // S + C
  for (int e = tid; e < NN; e += Cf::THREADS) { Ks[e] = S[e]; Ks[NN + e] = C[e]; Ks[2 * NN + e] = S[e] + C[e]; }
  __syncthreads();
  const double a1 = mode == 0 ? 1.5 : mode == 1 ? 1.5 : -0.5;  // coefficient of S X_s
  const double a2 = mode == 0 ? 0.5 : -0.5;                     // coefficient of C X_c
  for (long b0 = (long)blockIdx.x * MPB; b0 < batch; b0 += (long)gridDim.x * MPB) {
    for (int idx = tid; idx < MPB * NN; idx += Cf::THREADS) {
      int m = idx % MPB, e = idx / MPB;
      long b = b0 + m;
      Abuf[m * NN + e] = b < batch ? chi[b + (long)e * es] : zd(0, 0);
    }
    __syncthreads();
    // static part, and for V_sigma -0.25 (S+C) chi0 (S+C), block-wide
    if (mode == 0) {
      for (int idx = tid; idx < MPB * NN; idx += Cf::THREADS) {
        int m = idx / NN, e = idx % NN, p = e % N, q = e / N;
        zd acc(0, 0);
        for (int r = 0; r < N; ++r) acc += Abuf[m * NN + p + N * r] * Wg[r + N * q];
        LUb[m * 2 * NN + e] = acc;
      }
      __syncthreads();
    }
    for (int idx = tid; idx < MPB * NN; idx += Cf::THREADS) {
      int m = idx / NN, e = idx % NN, p = e % N, q = e / N;
      double st = mode == 0 ? 1.5 * Sg[e] - 0.5 * Cg[e] : mode == 1 ? 0.5 * (Sg[e] + Cg[e]) : 0.5 * (Sg[e] - Cg[e]);
      zd v(st, 0);
      if (mode == 0) {
        zd acc(0, 0);
        for (int r = 0; r < N; ++r) acc += LUb[m * 2 * NN + r + N * q] * Wg[p + N * r];
        v -= acc * 0.25;
      }
      Vbuf[idx] = v;
    }
    __syncthreads();
    const zd* A = Abuf + mb * NN;
    for (int pass = 0; pass < (Cf::PAR ? 1 : 2); ++pass) {
      const int sv = Cf::PAR ? lm / (2 * N) : pass;
      const int c = Cf::PAR ? lm % (2 * N) : lm;
      const bool act = inmat && (Cf::PAR || lm < 2 * N);
      const int slot = mb * 2 + sv;
      const double* K = sv == 0 ? Sg : Cg;
      const int j = c % N;
      zd col[N];
#pragma unroll
      for (int p = 0; p < N; ++p) col[p] = zd(0, 0);
      if (act) {
        for (int r = 0; r < N; ++r) {
          const double kr = K[r + N * j];
#pragma unroll
          for (int p = 0; p < N; ++p) col[p] += A[p + N * r] * kr;
        }
        if (c < N) {
#pragma unroll
          for (int p = 0; p < N; ++p) {
            if (sv == 0) col[p] = zd(0, 0) - col[p];
            if (p == j) col[p] += zd(1.0, 0);
          }
        }
      }
      // ---- LU with partial pivoting; the RHS lanes are eliminated alongside
#pragma unroll
      for (int k = 0; k < N; ++k) {
        if (act && c == k) {
          int p = k;
          double best = abs1(col[k]);
#pragma unroll
          for (int r = k + 1; r < N; ++r) {
            double v = abs1(col[r]);
            if (v > best) { best = v; p = r; }
          }
          if (best == 0.0) atomicOr(err, 1);
          piv[slot * N + k] = p;
        }
        __syncthreads();
        if (act && c >= k) {
          const int p = piv[slot * N + k];
#pragma unroll
          for (int r = k + 1; r < N; ++r)
            if (r == p) { zd t = col[k]; col[k] = col[r]; col[r] = t; }
          if (c == k) {
            const zd d = col[k];
            const zd inv = (d.re == 0.0 && d.im == 0.0) ? zd(0, 0) : zd(1.0, 0) / d;
            rinv[slot * N + k] = inv;
#pragma unroll
            for (int r = k + 1; r < N; ++r) {
              col[r] = col[r] * inv;
              bc[slot * N + r] = col[r];
            }
          }
        }
        __syncthreads();
        if (act && c > k) {
#pragma unroll
          for (int r = k + 1; r < N; ++r) col[r] -= bc[slot * N + r] * col[k];
        }
      }
      if (act && c < N) {
#pragma unroll
        for (int r = 0; r < N; ++r)
          if (r <= c) LUb[slot * NN + r + N * c] = col[r];
      }
      __syncthreads();
      // ---- RHS lanes: back substitution with U (reciprocal pivots from LDS)
      if (act && c >= N) {
        const zd* LU = LUb + slot * NN;
        const zd* ri = rinv + slot * N;
#pragma unroll
        for (int k = N - 1; k >= 0; --k) {
          col[k] = col[k] * ri[k];
#pragma unroll
          for (int r = 0; r < k; ++r) col[r] -= LU[r + N * k] * col[k];
        }
      }
      // ---- add this solve's share K X to column q of the vertex (solves in turn)
      for (int turn = 0; turn < 2; ++turn) {
        if (act && c >= N && sv == turn) {
          const int q = c - N;
          const double ax = sv == 0 ? a1 : a2;
          zd* vb = Vbuf + mb * NN + N * q;
          for (int p = 0; p < N; ++p) {
            zd acc(0, 0);
#pragma unroll
            for (int r = 0; r < N; ++r) acc += col[r] * K[p + N * r];
            vb[p] += acc * ax;
          }
        }
        if (Cf::PAR) __syncthreads();
      }
      if (!Cf::PAR) __syncthreads();
    }
    for (int idx = tid; idx < MPB * NN; idx += Cf::THREADS) {
      int m = idx % MPB, e = idx / MPB;
      long b = b0 + m;
      if (b < batch) chi[b + (long)e * es] = Vbuf[m * NN + e];
    }
    __syncthreads();
  }
}

template <int N>
static void launch_vertex_reg(zd* chi, long batch, long es, const double* S, const double* C, int mode, int* err,
                              hipStream_t s) {
  long nb = (batch + VCfg<N>::MPB - 1) / VCfg<N>::MPB;
  unsigned grid = (unsigned)std::min<long>(nb, 256L * 32);
  hipLaunchKernelGGL((k_vertex_reg<N>), dim3(grid), dim3(VCfg<N>::THREADS), 0, s, chi, batch, es, S, C, mode, err);
  GMOD_CHECK_LAUNCH();
}

// ---------------------------------------------------------------------- sparse vertex matrices
// Kanamori S and C have at most VW_NZ nonzeros per row / column; the
// Gauss-Jordan vertex kernel below forms its products A S, A C and W A W
// from the sparse columns.
constexpr int VW_NZ = 4;      // nonzeros per row / column of S, C, S + C

struct VSparse {              // column j of K: rows idx[j][t], values val[j][t] (zero padded)
  int idx[3][16][VW_NZ];      // [S, C, W]
  double val[3][16][VW_NZ];
};

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Gauss-Jordan variant, one matrix row per lane: a 16-lane group solves one
// of the two systems of a (q, i nu) matrix ((I - A S) X = A S or
// (I + A C) X = A C), so a wavefront carries two matrices.  The solution
// comes from the inverse: X = (I - A S)^-1 - I and X = I - (I + A C)^-1,
// the inverse formed in place (no separate right-hand side).  Per elimination
// step the group finds the pivot by a maximum over its 16 lanes (DPP
// rotations within the row) and exchanges one thing through its LDS slot,
// the unnormalised pivot row: one LDS round trip per column instead of one
// cross-lane transfer per matrix element.  Pivots: the largest |Re|+|Im|
// among the rows not used yet (LAPACK's izamax criterion).
// a -= f b with four fused multiply-adds
__device__ __forceinline__ void csub_mul(zd& a, zd f, zd b) {
  a.re = fma(-f.re, b.re, fma(f.im, b.im, a.re));
  a.im = fma(-f.re, b.im, fma(-f.im, b.re, a.im));
}
// 1 / d through one real division (|d| is far from the overflow range here); 0 for d = 0
__device__ __forceinline__ zd crecip(zd d) {
  const double n = d.re * d.re + d.im * d.im;
  const double q = n == 0.0 ? 0.0 : 1.0 / n;
  return zd(d.re * q, -d.im * q);
}

// Pivot candidate of row r as one double whose maximum names the pivot:
// |.|_1 (non-negative; -1 for a used row; inf and NaN clamped to the
// largest double) with its low 4 mantissa bits replaced by 15 - r.  The
// bit patterns of non-negative doubles order like the values, so the
// maximum key carries the largest candidate and, among candidates equal
// in their upper 60 bits, the lowest row; a maximum over the group's lanes
// replaces a compare-and-select scan.
__device__ __forceinline__ long long gj_key(double v, int r) {
  const long long b = __double_as_longlong(fmin(v, DBL_MAX));
  return (b & ~0xFLL) | (long long)(15 - r);
}
// One rotation step of the keys within the DPP row of 16 lanes (the group of
// one system): the larger of this lane's key and that of the lane ROT
// places further round.  Keys of used rows are negative, every other key
// non-negative, so signed 64-bit order is the order of the candidates.
template <int ROT>
__device__ __forceinline__ long long gj_row_max_step(long long k) {
  constexpr int ctrl = 0x120 + ROT;  // DPP row_ror:ROT
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(k & 0xffffffffLL), ctrl, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(k >> 32), ctrl, 0xF, 0xF, false);
  const long long o = (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
  return o > k ? o : k;
}
// the maximum key of the lane's 16-lane row, in every lane of the row
__device__ __forceinline__ long long gj_row_max(long long k) {
  k = gj_row_max_step<8>(k);
  k = gj_row_max_step<4>(k);
  k = gj_row_max_step<2>(k);
  return gj_row_max_step<1>(k);
}

// Gauss-Jordan inversion of M (row r in lane r of a 16-lane group, in
// mrow) in place with partial pivoting and no row exchanges.  Column k of the
// row storage turns into a column of the inverse once column k is
// eliminated.  Rows stay unnormalised: the pivot lane publishes its row as is
// (prow, the group's LDS slot) and keeps 1 in the pivot column, the others
// form the multiplier from the published pivot, and each lane divides its row
// by its own pivot once at the end.  A lane's candidate for column k + 1 is
// formed right after column k + 1 of its row is updated.  At the end the lane
// that was pivot of column k holds row k of M^-1 (xrow = k), entry j in
// column piv[j] (the pivot row of column j).  Returns true if a pivot was
// zero or not finite.
template <int N>
__device__ __forceinline__ bool gj_invert(zd (&mrow)[N], int r, zd* prow, int* piv, int& xrow) {
  bool used = false;
  bool singular = false;
  zd dpiv(1.0, 0);
  long long key = gj_key(abs1(mrow[0]), r);  // my candidate for the next column
#pragma unroll
  for (int k = 0; k < N; ++k) {
    wave_sync();
    const long long kb = gj_row_max(key);
    const int pr = 15 - (int)(kb & 0xF);
    const double best = __longlong_as_double(kb & ~0xFLL);
    singular = singular || !(best > 0.0 && best < 1e300);  // zero, or a clamped inf / NaN
    if (r == pr) {
#pragma unroll
      for (int j = 0; j < N; ++j) prow[j] = mrow[j];
      piv[k] = r;
      dpiv = mrow[k];
      mrow[k] = zd(1.0, 0);
      used = true;
      xrow = k;
    }
    wave_sync();
    if (r != pr) {
      const zd f = mrow[k] * crecip(prow[k]);
      if (k + 1 < N) {
        csub_mul(mrow[k + 1], f, prow[k + 1]);
        key = gj_key(used ? -1.0 : abs1(mrow[k + 1]), r);
      }
#pragma unroll
      for (int j = 0; j < N; ++j)
        if (j != k && j != k + 1) csub_mul(mrow[j], f, prow[j]);
      mrow[k] = zd(0, 0) - f;
    } else if (k + 1 < N) {
      key = gj_key(-1.0, r);
    }
  }
  {
    const zd inv = crecip(dpiv);
#pragma unroll
    for (int j = 0; j < N; ++j) mrow[j] = mrow[j] * inv;
  }
  wave_sync();
  return singular;
}

constexpr int GJ_WAVES = 4;
constexpr int GJ_MAT = 2 * GJ_WAVES;  // matrices per block

// chi_in may equal chi_out (in place: a work-group reads all its matrices
// before it writes any).  thr > 0 zeroes real and imaginary parts of the
// input below thr as they are read; dv (with vprev) receives V - vprev in FP32.
// chi_out == nullptr (with dv): the vertex itself is not stored.
template <int N>
__global__ void __launch_bounds__(64 * GJ_WAVES) __attribute__((amdgpu_waves_per_eu(2)))
k_vertex_gj(const zd* chi_in, zd* chi_out, long batch, long es, const VSparse* __restrict__ spg,
            const double* __restrict__ Sd, const double* __restrict__ Cd, int mode, double thr,
            const zd* __restrict__ vprev, cpx<float>* __restrict__ dv, int* err) {
  static_assert(N == 16, "a 16-lane group per system");
  constexpr int NN = N * N;
  // matrix stride in the LDS buffers, padded by one entry: the block-wide
  // phases give consecutive lanes consecutive matrices (coalesced global
  // access), whose entries then fall in different LDS banks
  constexpr int MS = NN + 1;
  constexpr int PER = GJ_MAT * NN / (64 * GJ_WAVES);  // elements per thread in block-wide phases
  static_assert(64 * GJ_WAVES == NN, "one matrix entry per thread in the block-wide phases");
  __shared__ VSparse sp;
  __shared__ double2 Abuf_[GJ_MAT * MS];  // chi0, later X_s then X_c
  __shared__ double2 Vbuf_[GJ_MAT * MS];  // A W (mode 0), then the vertex
  // per-group slots padded so that the four groups of a wavefront start 4
  // dwords apart in the LDS banks (their broadcast reads do not conflict)
  constexpr int PROW = 2 * N + 1;  // complex entries per group slot
  __shared__ double2 Prow_[GJ_WAVES * 4 * PROW];
  __shared__ int Piv_[GJ_WAVES * 4 * N];  // per group: the pivot row of each column
  zd* Abuf = reinterpret_cast<zd*>(Abuf_);
  zd* Vbuf = reinterpret_cast<zd*>(Vbuf_);
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  {
    const int* src = reinterpret_cast<const int*>(spg);
    int* dst = reinterpret_cast<int*>(&sp);
    for (int i = tid; i < (int)(sizeof(VSparse) / sizeof(int)); i += 64 * GJ_WAVES) dst[i] = src[i];
  }
  const int g = lane >> 4, r = lane & 15;
  const int mloc = wave * 2 + (g >> 1);  // matrix of my group within the block
  const int sys = g & 1;                 // 0: S system, 1: C system
  zd* A = Abuf + mloc * MS;
  zd* prow = reinterpret_cast<zd*>(Prow_) + (wave * 4 + g) * PROW;
  int* piv = Piv_ + (wave * 4 + g) * N;
  const double a1 = mode == 2 ? -0.5 : 1.5;  // coefficient of S X_s
  const double a2 = mode == 0 ? 0.5 : -0.5;  // coefficient of C X_c
  const long bstep = (long)gridDim.x * GJ_MAT;
  // the work-group's next matrices are read into registers while the
  // current ones are solved (element i = tid + u 256 of the batch: matrix
  // i % GJ_MAT, entry i / GJ_MAT); a matrix past the batch reads entry 0 of
  // matrix 0 and is dropped
  zd nxt[PER];
  auto fetch = [&](long b0) {
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int i = tid + u * 64 * GJ_WAVES;
      const int m = i % GJ_MAT, e = i / GJ_MAT;
      const long b = b0 + m;
      nxt[u] = chi_in[b < batch ? b + (long)e * es : 0];
    }
  };
  __syncthreads();
  if ((long)blockIdx.x * GJ_MAT < batch) fetch((long)blockIdx.x * GJ_MAT);
  for (long b0 = (long)blockIdx.x * GJ_MAT; b0 < batch; b0 += bstep) {
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int i = tid + u * 64 * GJ_WAVES;
      const int m = i % GJ_MAT, e = i / GJ_MAT;
      zd v = b0 + m < batch ? nxt[u] : zd(0, 0);
      if (thr > 0.0) {
        if (fabs(v.re) < thr) v.re = 0.0;
        if (fabs(v.im) < thr) v.im = 0.0;
      }
      Abuf[m * MS + e] = v;
    }
    __syncthreads();
    if (b0 + bstep < batch) fetch(b0 + bstep);
    // ---- V = st (- 0.25 W A W for mode 0), block-wide
    if (mode == 0) {
      for (int i = tid; i < GJ_MAT * NN; i += 64 * GJ_WAVES) {
        const int m = i / NN, e = i % NN, p = e % N, q = e / N;
        zd acc(0, 0);
#pragma unroll
        for (int t = 0; t < VW_NZ; ++t) acc += Abuf[m * MS + p + N * sp.idx[2][q][t]] * sp.val[2][q][t];
        Vbuf[m * MS + e] = acc;  // (A W)[p][q]
      }
      __syncthreads();
    }
    zd vinit[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int i = tid + u * 64 * GJ_WAVES;
      const int m = i / NN, e = i % NN, p = e % N, q = e / N;
      const double sv = Sd[e], cv = Cd[e];
      const double st = mode == 0 ? 1.5 * sv - 0.5 * cv : mode == 1 ? 0.5 * (sv + cv) : 0.5 * (sv - cv);
      zd v(st, 0);
      if (mode == 0) {
        zd acc(0, 0);
#pragma unroll
        for (int t = 0; t < VW_NZ; ++t) acc += Vbuf[m * MS + sp.idx[2][p][t] + N * q] * sp.val[2][p][t];
        v -= acc * 0.25;
      }
      vinit[u] = v;
    }
    __syncthreads();  // every read of A W is complete
#pragma unroll
    for (int u = 0; u < PER; ++u) Vbuf[u * MS + tid] = vinit[u];  // entry tid of matrix u
    // ---- my row of M = I -+ A K (K = S or C)
    zd mrow[N];
#pragma unroll
    for (int j = 0; j < N; ++j) {
      zd acc(0, 0);
#pragma unroll
      for (int t = 0; t < VW_NZ; ++t) acc += A[r + N * sp.idx[sys][j][t]] * sp.val[sys][j][t];
      mrow[j] = sys == 0 ? zd(0, 0) - acc : acc;
      if (j == r) mrow[j] += zd(1.0, 0);
    }
    __syncthreads();  // all reads of A are complete before X overwrites it
    // ---- Gauss-Jordan inversion (the steps of gj_invert, kept inline here:
    // the headline kernel's register allocation is tuned on this form)
    bool used = false;
    int xrow = 0;  // the row of M^-1 this lane holds at the end
    bool singular = false;
    zd dpiv(1.0, 0);
    long long key = gj_key(abs1(mrow[0]), r);  // my candidate for the next column
#pragma unroll
    for (int k = 0; k < N; ++k) {
      wave_sync();
      const long long kb = gj_row_max(key);
      const int pr = 15 - (int)(kb & 0xF);
      const double best = __longlong_as_double(kb & ~0xFLL);
      singular = singular || !(best > 0.0 && best < 1e300);  // zero, or a clamped inf / NaN
      if (r == pr) {
#pragma unroll
        for (int j = 0; j < N; ++j) prow[j] = mrow[j];
        piv[k] = r;
        dpiv = mrow[k];
        mrow[k] = zd(1.0, 0);
        used = true;
        xrow = k;
      }
      wave_sync();
      if (r != pr) {
        const zd f = mrow[k] * crecip(prow[k]);
        if (k + 1 < N) {
          csub_mul(mrow[k + 1], f, prow[k + 1]);
          key = gj_key(used ? -1.0 : abs1(mrow[k + 1]), r);
        }
#pragma unroll
        for (int j = 0; j < N; ++j)
          if (j != k && j != k + 1) csub_mul(mrow[j], f, prow[j]);
        mrow[k] = zd(0, 0) - f;
      } else if (k + 1 < N) {
        key = gj_key(-1.0, r);
      }
    }
    {
      const zd inv = crecip(dpiv);
#pragma unroll
      for (int j = 0; j < N; ++j) mrow[j] = mrow[j] * inv;
    }
    wave_sync();
    if (singular && r == 0 && b0 + mloc < batch) atomicOr(err, 1);
    // ---- V += a1 S X_s, then V += a2 C X_c (X through my matrix's A slot)
#pragma unroll
    for (int pass = 0; pass < 2; ++pass) {
      if (sys == pass) {
        // X = (M^-1 - I) for the S system, I - M^-1 for the C system
#pragma unroll
        for (int j = 0; j < N; ++j) {
          const int q = piv[j];
          zd x = mrow[j];
          if (q == xrow) x -= zd(1.0, 0);
          A[xrow + N * q] = sys == 0 ? x : zd(0, 0) - x;
        }
      }
      __syncthreads();
      const double coef = pass == 0 ? a1 : a2;
#pragma unroll
      for (int u = 0; u < PER; ++u) {
        const int i = tid + u * 64 * GJ_WAVES;
        const int m = i / NN, e = i % NN, p = e % N, q = e / N;
        zd acc(0, 0);
#pragma unroll
        for (int t = 0; t < VW_NZ; ++t)
          acc += Abuf[m * MS + sp.idx[pass][p][t] + N * q] * sp.val[pass][p][t];
        Vbuf[m * MS + e] += acc * coef;
      }
      __syncthreads();
    }
    for (int i = tid; i < GJ_MAT * NN; i += 64 * GJ_WAVES) {
      const int m = i % GJ_MAT, e = i / GJ_MAT;
      const long b = b0 + m;
      if (b < batch) {
        const zd v = Vbuf[m * MS + e];
        const long o = b + (long)e * es;
        if (chi_out) chi_out[o] = v;
        if (dv) dv[o] = cpx<float>(v - vprev[o]);
      }
    }
    __syncthreads();
  }
}

// The vertices of the superconducting loop (ffeliash.f90:386-457) for
// Nchi = 16 with sparse S, C: the Gauss-Jordan scheme of k_vertex_gj with the
// S system solved on chi0_S and the C system on chi0_C (read from vs and vd),
// V_sigma = 3/2 S - 1/2 C + 3/2 S chi_s + 1/2 C chi_c - 1/4 W (chi0_S S +
// chi0_C C) written to vs and V_Delta (singlet: (S+C)/2 + 3/2 S chi_s - 1/2 C
// chi_c; triplet: (S-C)/2 - 1/2 S chi_s - 1/2 C chi_c) to vd.  Two waves and
// four matrices per work-group: chi0_S (later X), chi0_C (later V_Delta) and
// V_sigma take three 16 KB LDS buffers.
constexpr int SCGJ_WAVES = 2;
constexpr int SCGJ_MAT = 2 * SCGJ_WAVES;

template <int N>
__global__ void __launch_bounds__(64 * SCGJ_WAVES)
k_vertex_gj_sc(zd* vs, zd* vd, long batch, long es, const VSparse* __restrict__ spg, const double* __restrict__ Sd,
               const double* __restrict__ Cd, int singlet, int* err) {
  static_assert(N == 16, "a 16-lane group per system");
  constexpr int NN = N * N;
  constexpr int NT = 64 * SCGJ_WAVES;
  constexpr int PER = SCGJ_MAT * NN / NT;
  __shared__ VSparse sp;
  __shared__ double2 Abuf_[SCGJ_MAT * NN];  // chi0_S, later X_s then X_c
  __shared__ double2 Cbuf_[SCGJ_MAT * NN];  // chi0_C, then chi0_S S + chi0_C C, then V_Delta
  __shared__ double2 Vbuf_[SCGJ_MAT * NN];  // V_sigma
  constexpr int PROW = 2 * N + 1;
  __shared__ double2 Prow_[SCGJ_WAVES * 4 * PROW];
  __shared__ int Piv_[SCGJ_WAVES * 4 * N];
  zd* Abuf = reinterpret_cast<zd*>(Abuf_);
  zd* Cbuf = reinterpret_cast<zd*>(Cbuf_);
  zd* Vbuf = reinterpret_cast<zd*>(Vbuf_);
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  {
    const int* src = reinterpret_cast<const int*>(spg);
    int* dst = reinterpret_cast<int*>(&sp);
    for (int i = tid; i < (int)(sizeof(VSparse) / sizeof(int)); i += NT) dst[i] = src[i];
  }
  const int g = lane >> 4, r = lane & 15;
  const int mloc = wave * 2 + (g >> 1);
  const int sys = g & 1;  // 0: S system on chi0_S, 1: C system on chi0_C
  zd* Am = (sys == 0 ? Abuf : Cbuf) + mloc * NN;
  zd* prow = reinterpret_cast<zd*>(Prow_) + (wave * 4 + g) * PROW;
  int* piv = Piv_ + (wave * 4 + g) * N;
  const double d1 = singlet ? 1.5 : -0.5;  // V_Delta: coefficient of S X_s
  const double d2 = -0.5;                  // and of C X_c
  const long bstep = (long)gridDim.x * SCGJ_MAT;
  zd nxs[PER], nxc[PER];
  auto fetch = [&](long b0) {
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int i = tid + u * NT;
      const int m = i % SCGJ_MAT, e = i / SCGJ_MAT;
      const long b = b0 + m;
      const long o = b < batch ? b + (long)e * es : 0;
      nxs[u] = vs[o];
      nxc[u] = vd[o];
    }
  };
  __syncthreads();
  // (a prefetch of the next batch across the loop, as k_vertex_gj does, puts
  // this kernel's state in scratch memory: the loads are issued per batch)
  for (long b0 = (long)blockIdx.x * SCGJ_MAT; b0 < batch; b0 += bstep) {
    fetch(b0);
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int i = tid + u * NT;
      const int m = i % SCGJ_MAT, e = i / SCGJ_MAT;
      const bool in = b0 + m < batch;
      Abuf[m * NN + e] = in ? nxs[u] : zd(0, 0);
      Cbuf[m * NN + e] = in ? nxc[u] : zd(0, 0);
    }
    __syncthreads();
    // ---- my row of M = I - chi0_S S (S system) or I + chi0_C C (C system)
    zd mrow[N];
#pragma unroll
    for (int j = 0; j < N; ++j) {
      zd acc(0, 0);
#pragma unroll
      for (int t = 0; t < VW_NZ; ++t) acc += Am[r + N * sp.idx[sys][j][t]] * sp.val[sys][j][t];
      mrow[j] = sys == 0 ? zd(0, 0) - acc : acc;
      if (j == r) mrow[j] += zd(1.0, 0);
    }
    // ---- T = chi0_S S + chi0_C C, block-wide
    zd tt[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int i = tid + u * NT;
      const int m = i / NN, e = i % NN, p = e % N, q = e / N;
      zd acc(0, 0);
#pragma unroll
      for (int t = 0; t < VW_NZ; ++t) {
        acc += Abuf[m * NN + p + N * sp.idx[0][q][t]] * sp.val[0][q][t];
        acc += Cbuf[m * NN + p + N * sp.idx[1][q][t]] * sp.val[1][q][t];
      }
      tt[u] = acc;
    }
    __syncthreads();  // chi0_S and chi0_C are read
#pragma unroll
    for (int u = 0; u < PER; ++u) Cbuf[tid + u * NT] = tt[u];
    __syncthreads();
    // ---- V_sigma = st - 1/4 W T; V_Delta = its static part
    zd vsig[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int i = tid + u * NT;
      const int m = i / NN, e = i % NN, p = e % N, q = e / N;
      zd acc(0, 0);
#pragma unroll
      for (int t = 0; t < VW_NZ; ++t) acc += Cbuf[m * NN + sp.idx[2][p][t] + N * q] * sp.val[2][p][t];
      vsig[u] = zd(1.5 * Sd[e] - 0.5 * Cd[e], 0) - acc * 0.25;
    }
    __syncthreads();  // every read of T is complete
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int i = tid + u * NT;
      const int e = i % NN;
      Vbuf[i] = vsig[u];
      Cbuf[i] = singlet ? zd(0.5 * (Sd[e] + Cd[e]), 0) : zd(0.5 * (Sd[e] - Cd[e]), 0);
    }
    int xrow = 0;
    const bool singular = gj_invert<N>(mrow, r, prow, piv, xrow);
    if (singular && r == 0 && b0 + mloc < batch) atomicOr(err, 1);
    // ---- X_s then X_c through my matrix's A slot; both vertices accumulate
#pragma unroll
    for (int pass = 0; pass < 2; ++pass) {
      if (sys == pass) {
#pragma unroll
        for (int j = 0; j < N; ++j) {
          const int q = piv[j];
          zd x = mrow[j];
          if (q == xrow) x -= zd(1.0, 0);
          Abuf[mloc * NN + xrow + N * q] = sys == 0 ? x : zd(0, 0) - x;
        }
      }
      __syncthreads();
      const double cs = pass == 0 ? 1.5 : 0.5;
      const double cd = pass == 0 ? d1 : d2;
#pragma unroll
      for (int u = 0; u < PER; ++u) {
        const int i = tid + u * NT;
        const int m = i / NN, e = i % NN, p = e % N, q = e / N;
        zd acc(0, 0);
#pragma unroll
        for (int t = 0; t < VW_NZ; ++t) acc += Abuf[m * NN + sp.idx[pass][p][t] + N * q] * sp.val[pass][p][t];
        Vbuf[i] += acc * cs;
        Cbuf[i] += acc * cd;
      }
      __syncthreads();
    }
    for (int i = tid; i < SCGJ_MAT * NN; i += NT) {
      const int m = i % SCGJ_MAT, e = i / SCGJ_MAT;
      const long b = b0 + m;
      if (b < batch) {
        const long o = b + (long)e * es;
        vs[o] = Vbuf[m * NN + e];
        vd[o] = Cbuf[m * NN + e];
      }
    }
    __syncthreads();
  }
}

// sparse columns of S, C and S + C when every column has at most VW_NZ
// nonzeros and the matrices are symmetric (rows = columns); false otherwise
static bool vertex_sparse(const double* S_d, const double* C_d, int n, VSparse& out, hipStream_t s) {
  if (n > 16) return false;
  std::vector<double> S(n * n), C(n * n);
  GMOD_HIP(hipMemcpyAsync(S.data(), S_d, n * n * sizeof(double), hipMemcpyDeviceToHost, s));
  GMOD_HIP(hipMemcpyAsync(C.data(), C_d, n * n * sizeof(double), hipMemcpyDeviceToHost, s));
  GMOD_HIP(hipStreamSynchronize(s));
  memset(&out, 0, sizeof(out));
  for (int m = 0; m < 3; ++m)
    for (int j = 0; j < n; ++j) {
      int cnt = 0;
      for (int r = 0; r < n; ++r) {
        const double v = m == 0 ? S[r + n * j] : m == 1 ? C[r + n * j] : S[r + n * j] + C[r + n * j];
        const double vt = m == 0 ? S[j + n * r] : m == 1 ? C[j + n * r] : S[j + n * r] + C[j + n * r];
        if (v != vt) return false;
        if (v == 0.0) continue;
        if (cnt == VW_NZ) return false;
        out.idx[m][j][cnt] = r;
        out.val[m][j][cnt] = v;
        ++cnt;
      }
    }
  return true;
}

static void check_lds(size_t bytes) {
  if (bytes > 160 * 1024) throw Error("Nchi too large for the LDS-resident solver");
}

// the Gauss-Jordan kernel for Nchi = 16 with sparse symmetric S, C; false
// when it does not apply (the caller uses the general kernels)
static bool vertex_gj(const zd* in, zd* out, long batch, BatchLayout L, int n, const double* S, const double* C,
                      int mode, double thr, const zd* vprev, cpx<float>* dv, int* err, hipStream_t s) {
  if (!(L.bs == 1 && n == 16)) return false;
  if (!out && !dv) throw Error("vertex: neither the vertex nor its difference is stored");
  VSparse sp;
  if (!vertex_sparse(S, C, n, sp, s)) return false;
  DBuf<VSparse> spd(1);
  spd.upload(&sp, 1, s);
  const long nb = (batch + GJ_MAT - 1) / GJ_MAT;
  const unsigned grid = (unsigned)std::min<long>(nb, 256L * 32);
  hipLaunchKernelGGL(k_vertex_gj<16>, dim3(grid), dim3(64 * GJ_WAVES), 0, s, in, out, batch, L.es, spd.p, S, C,
                     mode, thr, vprev, dv, err);
  GMOD_CHECK_LAUNCH();
  GMOD_HIP(hipStreamSynchronize(s));  // spd is released at return
  return true;
}

void rpa_vertex_fused(const zd* chi_in, zd* chi_out, long batch, BatchLayout L, int n, const double* S,
                      const double* C, double thr, const zd* vprev, cpx<float>* dv, int* err, hipStream_t s,
                      bool store_v) {
  if (batch <= 0) return;
  if (vertex_gj(chi_in, store_v || !dv ? chi_out : nullptr, batch, L, n, S, C, 0, thr, vprev, dv, err, s)) return;
  const long nel = (long)n * n * L.es;  // the batch's extent (L.bs == 1 layouts; general kernels otherwise)
  if (L.bs != 1) throw Error("rpa_vertex_fused: matrix-major layouts only");
  if (thr > 0.0) threshold_copy(chi_out, chi_in, nel, thr, s);
  else if (chi_out != chi_in) GMOD_HIP(hipMemcpyAsync(chi_out, chi_in, nel * sizeof(zd), hipMemcpyDeviceToDevice, s));
  rpa_vertex(chi_out, batch, L, n, S, C, 0, err, s);
  if (dv) diff_to_f32(dv, chi_out, vprev, nel, s);
}

void rpa_vertex(zd* chi, long batch, BatchLayout L, int n, const double* S, const double* C, int mode, int* err,
                hipStream_t s) {
  if (batch <= 0) return;
  if (vertex_gj(chi, chi, batch, L, n, S, C, mode, 0.0, nullptr, nullptr, err, s)) return;
  if (L.bs == 1) {
    switch (n) {
      case 1: launch_vertex_reg<1>(chi, batch, L.es, S, C, mode, err, s); return;
      case 4: launch_vertex_reg<4>(chi, batch, L.es, S, C, mode, err, s); return;
      case 9: launch_vertex_reg<9>(chi, batch, L.es, S, C, mode, err, s); return;
      case 16: launch_vertex_reg<16>(chi, batch, L.es, S, C, mode, err, s); return;
      default: break;
    }
  }
  size_t shm = 2 * n * n * sizeof(double) + 5 * n * n * sizeof(zd);
  check_lds(shm);
  unsigned grid = (unsigned)std::min<long>(batch, 256L * 64);
  hipLaunchKernelGGL(k_rpa_vertex, dim3(grid), dim3(64), shm, s, chi, L, batch, n, S, C, mode, err);
  GMOD_CHECK_LAUNCH();
}

void sc_vertex(zd* vs, zd* vd, long batch, BatchLayout L, int n, const double* S, const double* C, bool singlet,
               int* err, hipStream_t s) {
  if (batch <= 0) return;
  VSparse sp;
  if (L.bs == 1 && n == 16 && vertex_sparse(S, C, n, sp, s)) {
    DBuf<VSparse> spd(1);
    spd.upload(&sp, 1, s);
    const long nb = (batch + SCGJ_MAT - 1) / SCGJ_MAT;
    const unsigned grid = (unsigned)std::min<long>(nb, 256L * 48);
    hipLaunchKernelGGL(k_vertex_gj_sc<16>, dim3(grid), dim3(64 * SCGJ_WAVES), 0, s, vs, vd, batch, L.es, spd.p, S, C,
                       singlet ? 1 : 0, err);
    GMOD_CHECK_LAUNCH();
    GMOD_HIP(hipStreamSynchronize(s));  // spd is released at return
    return;
  }
  size_t shm = 2 * n * n * sizeof(double) + 5 * n * n * sizeof(zd);
  check_lds(shm);
  unsigned grid = (unsigned)std::min<long>(batch, 256L * 64);
  hipLaunchKernelGGL(k_sc_vertex, dim3(grid), dim3(64), shm, s, vs, vd, L, batch, n, S, C, singlet ? 1 : 0, err);
  GMOD_CHECK_LAUNCH();
}

// ---------------------------------------------------------------- spinful vertex
// V_sigma = -V + V X with X = (I + chi0 V)^-1 chi0 V per (q, i nu) (fUJ.f90:367-399),
// chi0 read with the 1e-9 threshold of the loop applied (thr > 0).  One
// wavefront per matrix; chi0, I + chi0 V and the right-hand side live in LDS,
// or, when three n x n complex matrices exceed it, in a per-work-group slice
// of the global workspace gws.  out may alias chi_in.
__global__ void __launch_bounds__(64)
k_soc_vertex(const zd* chi_in, zd* out, BatchLayout L, long batch, int n, const double* __restrict__ V, double thr,
             zd* gws, int* err) {
  extern __shared__ unsigned char sm[];
  __shared__ int shi[2];
  const int nn = n * n;
  zd* A = gws ? gws + (size_t)blockIdx.x * 3 * nn : reinterpret_cast<zd*>(sm);
  zd* M = A + nn;
  zd* X = M + nn;
  const int tid = threadIdx.x, nt = blockDim.x;
  for (long b = blockIdx.x; b < batch; b += gridDim.x) {
    for (int e = tid; e < nn; e += nt) {
      zd a = chi_in[b * L.bs + e * L.es];
      if (thr > 0.0) {
        if (fabs(a.re) < thr) a.re = 0.0;
        if (fabs(a.im) < thr) a.im = 0.0;
      }
      A[e] = a;
    }
    __syncthreads();
    for (int e = tid; e < nn; e += nt) {
      const int p = e % n, q = e / n;
      zd av(0, 0);
      for (int r = 0; r < n; ++r) av += A[p + n * r] * V[r + n * q];
      M[e] = zd(p == q ? 1.0 : 0.0, 0) + av;
      X[e] = av;
    }
    __syncthreads();
    if (!lds_lu_solve(M, X, n, n, shi)) {
      if (tid == 0) atomicOr(err, 1);
      __syncthreads();
      continue;
    }
    for (int e = tid; e < nn; e += nt) {
      const int p = e % n, q = e / n;
      zd v(-V[e], 0);
      for (int r = 0; r < n; ++r) v += X[r + n * q] * V[p + n * r];
      out[b * L.bs + e * L.es] = v;
    }
    __syncthreads();
  }
}

void soc_vertex(const zd* chi_in, zd* out, long batch, BatchLayout L, int n, const double* V, double thr, int* err,
                hipStream_t s) {
  if (batch <= 0) return;
  const size_t shm = 3 * (size_t)n * n * sizeof(zd);
  if (shm <= 160 * 1024) {
    const unsigned grid = (unsigned)std::min<long>(batch, 256L * 64);
    hipLaunchKernelGGL(k_soc_vertex, dim3(grid), dim3(64), shm, s, chi_in, out, L, batch, n, V, thr, (zd*)nullptr,
                       err);
    GMOD_CHECK_LAUNCH();
    return;
  }
  const long grid = std::min<long>(batch, 2048L);
  DBuf<zd> ws((size_t)grid * 3 * n * n);
  hipLaunchKernelGGL(k_soc_vertex, dim3((unsigned)grid), dim3(64), 0, s, chi_in, out, L, batch, n, V, thr, ws.p,
                     err);
  GMOD_CHECK_LAUNCH();
  GMOD_HIP(hipStreamSynchronize(s));  // ws is released at return
}

__global__ void __launch_bounds__(64)
k_rpa_solve(zd* out, BatchLayout Lo, const zd* Ain, BatchLayout La, long batch, int n, const double* W,
            double alpha, int* err) {
  extern __shared__ unsigned char sm[];
  const int nn = n * n;
  double* Ws = reinterpret_cast<double*>(sm);
  zd* A = reinterpret_cast<zd*>(Ws + nn);
  zd* Mx = A + nn;
  zd* X = Mx + nn;
  __shared__ int shi[2];
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int e = tid; e < nn; e += nt) Ws[e] = W[e];
  for (long b = blockIdx.x; b < batch; b += gridDim.x) {
    const zd* base = Ain + b * La.bs;
    for (int e = tid; e < nn; e += nt) A[e] = base[e * La.es];
    __syncthreads();
    for (int e = tid; e < nn; e += nt) {
      int p = e % n, q = e / n;
      zd aw(0, 0);
      for (int r = 0; r < n; ++r) aw += A[p + n * r] * Ws[r + n * q];
      Mx[e] = zd(p == q ? 1.0 : 0.0, 0) + aw * alpha;
      X[e] = A[e];
    }
    __syncthreads();
    if (!lds_lu_solve(Mx, X, n, n, shi)) {
      if (tid == 0) atomicOr(err, 1);
      __syncthreads();
      continue;
    }
    zd* ob = out + b * Lo.bs;
    for (int e = tid; e < nn; e += nt) ob[e * Lo.es] = X[e];
    __syncthreads();
  }
}

void rpa_solve(zd* out, BatchLayout Lo, const zd* A, BatchLayout La, long batch, int n, const double* W,
               double alpha, int* err, hipStream_t s) {
  if (batch <= 0) return;
  size_t shm = n * n * sizeof(double) + 3 * n * n * sizeof(zd);
  check_lds(shm);
  unsigned grid = (unsigned)std::min<long>(batch, 256L * 64);
  hipLaunchKernelGGL(k_rpa_solve, dim3(grid), dim3(64), shm, s, out, Lo, A, La, batch, n, W, alpha, err);
  GMOD_CHECK_LAUNCH();
}

// ====================================================================== eigenvalues
// Complex general eigenvalues: Gaussian reduction to Hessenberg form with
// partial pivoting, then single-shift QR (Wilkinson shift, exceptional shift
// every 10 sweeps) with Givens rotations restricted to the active block.
// Column-major H (ld = n); element (i, j) at H[(i + n j) st] and eigenvalue i
// at w[i st], so a batch of matrices can be interleaved (st = batch size)
// and a wavefront's lanes read adjacent elements.  Returns 0 on success.
__host__ __device__ inline zd zsqrt_(zd a) {
  double r = hypot(a.re, a.im);
  if (r == 0.0) return zd(0, 0);
  double x = sqrt(0.5 * (r + fabs(a.re)));
  double y = 0.5 * a.im / x;
  if (a.re >= 0) return zd(x, y);
  return zd(fabs(y), a.im >= 0 ? x : -x);
}

__host__ __device__ int zhess_eig(zd* H, int n, zd* w, long st = 1) {
#define HH(i, j) H[((i) + n * (j)) * st]
#define WW(i) w[(i) * st]
  // Hessenberg reduction (elementary stabilised transformations)
  for (int m = 1; m < n - 1; ++m) {
    int piv = m;
    double best = abs1(HH(m, m - 1));
    for (int i = m + 1; i < n; ++i) {
      double v = abs1(HH(i, m - 1));
      if (v > best) { best = v; piv = i; }
    }
    if (piv != m) {
      for (int j = m - 1; j < n; ++j) { zd t = HH(piv, j); HH(piv, j) = HH(m, j); HH(m, j) = t; }
      for (int i = 0; i < n; ++i) { zd t = HH(i, piv); HH(i, piv) = HH(i, m); HH(i, m) = t; }
    }
    zd x = HH(m, m - 1);
    if (x.re != 0.0 || x.im != 0.0) {
      for (int i = m + 1; i < n; ++i) {
        zd y = HH(i, m - 1);
        if (y.re == 0.0 && y.im == 0.0) continue;
        y = y / x;
        HH(i, m - 1) = zd(0, 0);
        for (int j = m; j < n; ++j) HH(i, j) -= y * HH(m, j);
        for (int j = 0; j < n; ++j) HH(j, m) += y * HH(j, i);
      }
    }
  }
  double anorm = 0.0;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n && i <= j + 1; ++i) anorm = fmax(anorm, abs1(HH(i, j)));
  const double eps = 2.220446049250313e-16;
  int hi = n - 1;
  int iter = 0, total = 0;
  while (hi >= 0) {
    if (hi == 0) { WW(0) = HH(0, 0); break; }
    int l = hi;
    for (; l > 0; --l) {
      double sref = abs1(HH(l - 1, l - 1)) + abs1(HH(l, l));
      if (sref == 0.0) sref = anorm;
      if (abs1(HH(l, l - 1)) <= eps * sref) break;
    }
    if (l == hi) { WW(hi) = HH(hi, hi); --hi; iter = 0; continue; }
    ++iter;
    if (++total > 60 * n) return 1;
    zd sigma;
    if (iter % 10 == 0) {
      double ex = fabs(HH(hi, hi - 1).re) + (hi >= 2 ? fabs(HH(hi - 1, hi - 2).re) : 0.0);
      sigma = HH(hi, hi) + zd(ex, 0);
    } else {
      zd a = HH(hi - 1, hi - 1), b = HH(hi - 1, hi), c = HH(hi, hi - 1), d = HH(hi, hi);
      zd half = (a - d) * 0.5;
      zd disc = zsqrt_(half * half + b * c);
      zd m = (a + d) * 0.5;
      zd e1 = m + disc, e2 = m - disc;
      sigma = abs1(e1 - d) < abs1(e2 - d) ? e1 : e2;
    }
    for (int k = l; k <= hi; ++k) HH(k, k) -= sigma;
    // QR by Givens on rows (store rotations in the subdiagonal workspace)
    double cs[64];
    zd sn[64];
    const int nrot = hi - l;
    if (nrot > 64) return 2;
    for (int k = l; k < hi; ++k) {
      zd x = HH(k, k), y = HH(k + 1, k);
      double ax = hypot(x.re, x.im), ay = hypot(y.re, y.im);
      double r = hypot(ax, ay);
      double c;
      zd sv;
      if (r == 0.0) { c = 1.0; sv = zd(0, 0); }
      else if (ax == 0.0) { c = 0.0; sv = conj(y) * (1.0 / ay); }
      else { zd ph = x * (1.0 / ax); c = ax / r; sv = ph * conj(y) * (1.0 / r); }
      cs[k - l] = c;
      sn[k - l] = sv;
      for (int j = k; j <= hi; ++j) {
        zd h1 = HH(k, j), h2 = HH(k + 1, j);
        HH(k, j) = h1 * c + sv * h2;
        HH(k + 1, j) = h2 * c - conj(sv) * h1;
      }
    }
    for (int k = l; k < hi; ++k) {
      double c = cs[k - l];
      zd sv = sn[k - l];
      int rmax = k + 1 < hi ? k + 1 : hi;
      for (int i = l; i <= rmax; ++i) {
        zd h1 = HH(i, k), h2 = HH(i, k + 1);
        HH(i, k) = h1 * c + h2 * conj(sv);
        HH(i, k + 1) = h2 * c - h1 * sv;
      }
    }
    for (int k = l; k <= hi; ++k) HH(k, k) += sigma;
  }
  return 0;
#undef WW
#undef HH
}

__global__ void k_max_re_eig(double* out, const zd* A, BatchLayout La, long batch, int n, const double* W,
                             double alpha, zd* scratch, int* err) {
  // the matrices are interleaved in scratch (matrix b's element e at
  // scratch[e batch + b]): the lanes of a wavefront touch adjacent elements
  for (long b = blockIdx.x * (long)blockDim.x + threadIdx.x; b < batch; b += (long)gridDim.x * blockDim.x) {
    zd* H = scratch + b;
    zd* w = scratch + (long)n * n * batch + b;
    const zd* base = A + b * La.bs;
    for (int q = 0; q < n; ++q)
      for (int p = 0; p < n; ++p) {
        zd v(0, 0);
        if (W) {
          for (int r = 0; r < n; ++r) v += base[(p + n * r) * La.es] * W[r + n * q];
        } else {
          v = base[(p + n * q) * La.es];
        }
        H[(p + n * q) * batch] = v * alpha;
      }
    if (zhess_eig(H, n, w, batch)) atomicOr(err, 2);
    double m = -HUGE_VAL;
    for (int i = 0; i < n; ++i) m = fmax(m, w[i * batch].re);
    out[b] = m;
  }
}

void max_re_eig_ws(double* out, const zd* A, BatchLayout La, long batch, int n, const double* W, double alpha,
                   zd* work, int* err, hipStream_t s) {
  if (batch <= 0) return;
  hipLaunchKernelGGL(k_max_re_eig, dim3(nblk(batch, 64)), dim3(64), 0, s, out, A, La, batch, n, W, alpha, work,
                     err);
  GMOD_CHECK_LAUNCH();
}

void max_re_eig(double* out, const zd* A, BatchLayout La, long batch, int n, const double* W, double alpha,
                int* err, hipStream_t s) {
  if (batch <= 0) return;
  DBuf<zd> scratch(batch * (long)(n * n + n));
  max_re_eig_ws(out, A, La, batch, n, W, alpha, scratch.p, err, s);
  GMOD_HIP(hipStreamSynchronize(s));
}

// ====================================================================== register kernels
template <int NO>
__device__ inline bool gj_inverse(zd (&a)[NO][NO]) {
  zd inv[NO][NO];
#pragma unroll
  for (int p = 0; p < NO; ++p)
#pragma unroll
    for (int q = 0; q < NO; ++q) inv[p][q] = zd(p == q ? 1.0 : 0.0, 0);
#pragma unroll
  for (int k = 0; k < NO; ++k) {
    int piv = k;
    double best = abs1(a[k][k]);
#pragma unroll
    for (int r = k + 1; r < NO; ++r) {
      double v = abs1(a[r][k]);
      if (v > best) { best = v; piv = r; }
    }
    if (best == 0.0) return false;
#pragma unroll
    for (int r = k + 1; r < NO; ++r) {
      if (r == piv) {
#pragma unroll
        for (int c = 0; c < NO; ++c) {
          zd t = a[k][c]; a[k][c] = a[r][c]; a[r][c] = t;
          t = inv[k][c]; inv[k][c] = inv[r][c]; inv[r][c] = t;
        }
      }
    }
    zd d = zd(1.0, 0) / a[k][k];
#pragma unroll
    for (int c = 0; c < NO; ++c) { a[k][c] = a[k][c] * d; inv[k][c] = inv[k][c] * d; }
#pragma unroll
    for (int r = 0; r < NO; ++r) {
      if (r == k) continue;
      zd f = a[r][k];
#pragma unroll
      for (int c = 0; c < NO; ++c) { a[r][c] -= f * a[k][c]; inv[r][c] -= f * inv[k][c]; }
    }
  }
#pragma unroll
  for (int p = 0; p < NO; ++p)
#pragma unroll
    for (int q = 0; q < NO; ++q) a[p][q] = inv[p][q];
  return true;
}

// The per-(k, w) register kernels below declare the block size they are
// launched with (128 threads; 256 for the filling reductions): the bound
// lets the compiler keep their No x No complex matrices in registers
// (without it, it plans for 1024-thread blocks, caps a thread at 128
// registers and spills the matrices to scratch memory).
template <int NO>
__global__ void __launch_bounds__(128) k_dyson(zd* G, const zd* sig, const zd* H, double mu, double temp, long Nk, int Nw, int* err) {
  const long tot = Nk * (long)Nw;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
    long k = e % Nk;
    long j = e / Nk;
    zd z(mu, (double)(2 * j + 1) * PI_D * temp);
    zd a[NO][NO];
#pragma unroll
    for (int p = 0; p < NO; ++p)
#pragma unroll
      for (int q = 0; q < NO; ++q) {
        zd v = zd(0, 0) - (H[p + NO * (q + NO * k)] + sig[e + tot * (p + NO * q)]);
        if (p == q) v += z;
        a[p][q] = v;
      }
    if (!gj_inverse<NO>(a)) atomicOr(err, 4);
#pragma unroll
    for (int p = 0; p < NO; ++p)
#pragma unroll
      for (int q = 0; q < NO; ++q) G[e + tot * (p + NO * q)] = a[p][q];
  }
}

template <int NO>
__global__ void __launch_bounds__(256) k_filling(const zd* sig, const zd* H, const double* eig, double rmu, double temp, long Nk, int Nw,
                          double* part, int* err) {
  __shared__ double sh[2][256];
  const long tot = Nk * (long)Nw;
  double fsum = 0.0, dsum = 0.0;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
    long k = e % Nk;
    long j = e / Nk;
    zd z(rmu, (double)(2 * j + 1) * PI_D * temp);
    zd a[NO][NO];
#pragma unroll
    for (int p = 0; p < NO; ++p)
#pragma unroll
      for (int q = 0; q < NO; ++q) {
        zd v = zd(0, 0) - (H[p + NO * (q + NO * k)] + sig[e + tot * (p + NO * q)]);
        if (p == q) v += z;
        a[p][q] = v;
      }
    if (!gj_inverse<NO>(a)) atomicOr(err, 8);
    double d = 0.0;
#pragma unroll
    for (int n = 0; n < NO; ++n) {
      d += a[n][n].re;
      d -= (zd(1.0, 0) / (z - zd(eig[n + NO * k], 0))).re;
    }
    dsum += d;
    if (j == 0) {
#pragma unroll
      for (int n = 0; n < NO; ++n) fsum += 0.5 * (1.0 - tanh(0.5 * (eig[n + NO * k] - rmu) / temp));
    }
  }
  sh[0][threadIdx.x] = fsum;
  sh[1][threadIdx.x] = dsum;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + o];
      sh[1][threadIdx.x] += sh[1][threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) { part[2 * blockIdx.x] = sh[0][0]; part[2 * blockIdx.x + 1] = sh[1][0]; }
}

// Characteristic polynomial of M = H(k) + Sigma(k, i w_j) per (k, j):
// det(z - M) = z^NO + a_1 z^(NO-1) + ... + a_NO by Faddeev-LeVerrier
// (B_0 = I, A_i = M B_(i-1), a_i = -Tr(A_i) / i, B_i = A_i + a_i I).
// coef(e, i-1) = a_i.  Tr (z - M)^-1 = p'(z) / p(z) then serves every trial
// mu of the chemical-potential search at the cost of a Horner evaluation.
template <int NO>
__global__ void __launch_bounds__(128) k_charpoly(zd* coef, const zd* sig, const zd* H, long Nk, int Nw) {
  const long tot = Nk * (long)Nw;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
    const long k = e % Nk;
    zd M[NO][NO], B[NO][NO];
#pragma unroll
    for (int p = 0; p < NO; ++p)
#pragma unroll
      for (int q = 0; q < NO; ++q) {
        M[p][q] = H[p + NO * (q + NO * k)] + sig[e + tot * (p + NO * q)];
        B[p][q] = M[p][q];  // A_1 = M
      }
#pragma unroll
    for (int i = 1; i <= NO; ++i) {
      if (i > 1) {  // A_i = M B_(i-1), into B
        zd A[NO][NO];
#pragma unroll
        for (int p = 0; p < NO; ++p)
#pragma unroll
          for (int q = 0; q < NO; ++q) {
            zd acc(0, 0);
#pragma unroll
            for (int r = 0; r < NO; ++r) acc += M[p][r] * B[r][q];
            A[p][q] = acc;
          }
#pragma unroll
        for (int p = 0; p < NO; ++p)
#pragma unroll
          for (int q = 0; q < NO; ++q) B[p][q] = A[p][q];
      }
      zd tr(0, 0);
#pragma unroll
      for (int p = 0; p < NO; ++p) tr += B[p][p];
      const zd a = tr * (-1.0 / (double)i);
      coef[e + tot * (i - 1)] = a;
#pragma unroll
      for (int p = 0; p < NO; ++p) B[p][p] += a;  // B_i = A_i + a_i I
    }
  }
}

// n(mu) pieces from the characteristic polynomials (see k_charpoly)
template <int NO>
__global__ void __launch_bounds__(256) k_filling_cp(const zd* coef, const double* eig, double rmu, double temp, long Nk, int Nw,
                             double* part) {
  __shared__ double sh[2][256];
  const long tot = Nk * (long)Nw;
  double fsum = 0.0, dsum = 0.0;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
    const long k = e % Nk;
    const long j = e / Nk;
    const zd z(rmu, (double)(2 * j + 1) * PI_D * temp);
    zd p(1.0, 0), dp(0, 0);  // Horner for p and p'
#pragma unroll
    for (int i = 0; i < NO; ++i) {
      dp = dp * z + p;
      p = p * z + coef[e + tot * i];
    }
    double d = (dp / p).re;
#pragma unroll
    for (int n = 0; n < NO; ++n) d -= (zd(1.0, 0) / (z - zd(eig[n + NO * k], 0))).re;
    dsum += d;
    if (j == 0) {
#pragma unroll
      for (int n = 0; n < NO; ++n) fsum += 0.5 * (1.0 - tanh(0.5 * (eig[n + NO * k] - rmu) / temp));
    }
  }
  sh[0][threadIdx.x] = fsum;
  sh[1][threadIdx.x] = dsum;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + o];
      sh[1][threadIdx.x] += sh[1][threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) { part[2 * blockIdx.x] = sh[0][0]; part[2 * blockIdx.x + 1] = sh[1][0]; }
}

template <int NO>
__global__ void k_green0(zd* G, const double* eig, const zd* U, double mu, double temp, long Nk, int Nw) {
  const long tot = Nk * (long)Nw;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
    long k = e % Nk;
    long j = e / Nk;
    zd z(mu, (double)(2 * j + 1) * PI_D * temp);
    zd r[NO];
#pragma unroll
    for (int n = 0; n < NO; ++n) r[n] = zd(1.0, 0) / (z - zd(eig[n + NO * k], 0));
#pragma unroll
    for (int l = 0; l < NO; ++l)
#pragma unroll
      for (int m = 0; m < NO; ++m) {
        zd acc(0, 0);
#pragma unroll
        for (int n = 0; n < NO; ++n) acc += U[m + NO * (n + NO * k)] * conj(U[l + NO * (n + NO * k)]) * r[n];
        G[e + tot * (m + NO * l)] = acc;
      }
  }
}

template <int NO>
__global__ void __launch_bounds__(128) k_anom_f(zd* F, const zd* G, const zd* D, long Nk, int Nw) {
  const long tot = Nk * (long)Nw;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
    zd g[NO][NO], gd[NO][NO];
#pragma unroll
    for (int p = 0; p < NO; ++p)
#pragma unroll
      for (int q = 0; q < NO; ++q) g[p][q] = G[e + tot * (p + NO * q)];
#pragma unroll
    for (int p = 0; p < NO; ++p)
#pragma unroll
      for (int q = 0; q < NO; ++q) {
        zd acc(0, 0);
#pragma unroll
        for (int r = 0; r < NO; ++r) acc += g[p][r] * D[e + tot * (r + NO * q)];
        gd[p][q] = acc;
      }
    zd f[NO][NO];
#pragma unroll
    for (int p = 0; p < NO; ++p)
#pragma unroll
      for (int q = 0; q < NO; ++q) {
        zd acc(0, 0);
#pragma unroll
        for (int r = 0; r < NO; ++r) acc += gd[p][r] * conj(g[q][r]);
        f[p][q] = zd(0, 0) - acc;
      }
#pragma unroll
    for (int l = 0; l < NO; ++l) {
      f[l][l].im = 0.0;
#pragma unroll
      for (int m = l + 1; m < NO; ++m) f[l][m] = conj(f[m][l]);
    }
#pragma unroll
    for (int p = 0; p < NO; ++p)
#pragma unroll
      for (int q = 0; q < NO; ++q) F[e + tot * (p + NO * q)] = f[p][q];
  }
}

// one step of the superconducting Dyson equations per (k, i w)
// (ffeliash.f90:315-384): G <- G0 (I + D F^+) with the previous F, then
// F <- -G0 D G^+ with the new G, diagonal real and upper = conj(lower)
template <int NO>
__global__ void __launch_bounds__(128) k_sc_green(zd* G, zd* F, const zd* G0, const zd* D, long Nk, int Nw) {
  const long tot = Nk * (long)Nw;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
    zd g0[NO][NO], d[NO][NO], t[NO][NO];
#pragma unroll
    for (int p = 0; p < NO; ++p)
#pragma unroll
      for (int q = 0; q < NO; ++q) {
        g0[p][q] = G0[e + tot * (p + NO * q)];
        d[p][q] = D[e + tot * (p + NO * q)];
      }
#pragma unroll
    for (int p = 0; p < NO; ++p)
#pragma unroll
      for (int q = 0; q < NO; ++q) {
        zd acc(p == q ? 1.0 : 0.0, 0);
#pragma unroll
        for (int r = 0; r < NO; ++r) acc += d[p][r] * conj(F[e + tot * (q + NO * r)]);
        t[p][q] = acc;  // I + D F^+
      }
    zd g[NO][NO];
#pragma unroll
    for (int p = 0; p < NO; ++p)
#pragma unroll
      for (int q = 0; q < NO; ++q) {
        zd acc(0, 0);
#pragma unroll
        for (int r = 0; r < NO; ++r) acc += g0[p][r] * t[r][q];
        g[p][q] = acc;
      }
#pragma unroll
    for (int p = 0; p < NO; ++p)
#pragma unroll
      for (int q = 0; q < NO; ++q) {
        zd acc(0, 0);
#pragma unroll
        for (int r = 0; r < NO; ++r) acc += d[p][r] * conj(g[q][r]);
        t[p][q] = acc;  // D G^+
      }
    zd f[NO][NO];
#pragma unroll
    for (int p = 0; p < NO; ++p)
#pragma unroll
      for (int q = 0; q < NO; ++q) {
        zd acc(0, 0);
#pragma unroll
        for (int r = 0; r < NO; ++r) acc += g0[p][r] * t[r][q];
        f[p][q] = zd(0, 0) - acc;
      }
#pragma unroll
    for (int l = 0; l < NO; ++l) {
      f[l][l].im = 0.0;
#pragma unroll
      for (int m = l + 1; m < NO; ++m) f[l][m] = conj(f[m][l]);
    }
#pragma unroll
    for (int p = 0; p < NO; ++p)
#pragma unroll
      for (int q = 0; q < NO; ++q) {
        G[e + tot * (p + NO * q)] = g[p][q];
        F[e + tot * (p + NO * q)] = f[p][q];
      }
  }
}

__global__ void k_hermitize_lower(zd* D, long tot, int No) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
    for (int l = 0; l < No; ++l) {
      D[e + tot * (l + No * l)].im = 0.0;
      for (int m = l + 1; m < No; ++m) D[e + tot * (l + No * m)] = conj(D[e + tot * (m + No * l)]);
    }
  }
}

// static part per (k, l, m) (fself.f90:384-417): mode 1 the Hermitian part
// of Sigma(k, i w_0), mode 2 the frequency average of the Hermitian part
__global__ void k_static_part(zd* sub, const zd* sig, long Nk, int Nw, int No, int mode) {
  const long tot = Nk * (long)Nw;
  const long n = Nk * No * No;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const long k = e % Nk;
    const int lm = (int)(e / Nk), l = lm % No, m = lm / No;
    zd v(0, 0);
    if (mode == 1) {
      v = (sig[k + tot * (l + No * m)] + conj(sig[k + tot * (m + No * l)])) * 0.5;
    } else {
      for (int j = 0; j < Nw; ++j)
        v += sig[k + Nk * j + tot * (l + No * m)] + conj(sig[k + Nk * j + tot * (m + No * l)]);
      v = v * (0.5 / (double)Nw);
    }
    sub[e] = v;  // sub[k + Nk (l + No m)]
  }
}
__global__ void k_sub_static(zd* sig, const zd* sub, long Nk, int Nw, int No) {
  const long tot = Nk * (long)Nw;
  const long n = tot * No * No;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const long k = e % Nk, lm = e / tot;
    sig[e] -= sub[k + Nk * lm];
  }
}

#define LAUNCH_NO(No, KERNEL, GRID, BLOCK, ...)                                            \
  switch (No) {                                                                             \
    case 1: hipLaunchKernelGGL((KERNEL<1>), GRID, BLOCK, 0, s, __VA_ARGS__); break;         \
    case 2: hipLaunchKernelGGL((KERNEL<2>), GRID, BLOCK, 0, s, __VA_ARGS__); break;         \
    case 3: hipLaunchKernelGGL((KERNEL<3>), GRID, BLOCK, 0, s, __VA_ARGS__); break;         \
    case 4: hipLaunchKernelGGL((KERNEL<4>), GRID, BLOCK, 0, s, __VA_ARGS__); break;         \
    case 5: hipLaunchKernelGGL((KERNEL<5>), GRID, BLOCK, 0, s, __VA_ARGS__); break;         \
    case 6: hipLaunchKernelGGL((KERNEL<6>), GRID, BLOCK, 0, s, __VA_ARGS__); break;         \
    case 7: hipLaunchKernelGGL((KERNEL<7>), GRID, BLOCK, 0, s, __VA_ARGS__); break;         \
    case 8: hipLaunchKernelGGL((KERNEL<8>), GRID, BLOCK, 0, s, __VA_ARGS__); break;         \
    default: throw Error("Norb > 8 is not supported by the register kernels");             \
  }

void dyson(zd* G, const zd* sigma, const zd* hamk, double mu, double temp, long Nk, int Nw, int No, int* err,
           hipStream_t s) {
  long tot = Nk * (long)Nw;
  LAUNCH_NO(No, k_dyson, dim3(nblk(tot, 128)), dim3(128), G, sigma, hamk, mu, temp, Nk, Nw, err);
  GMOD_CHECK_LAUNCH();
}

void filling_parts(double out[2], const zd* sigma, const zd* hamk, const double* eig, double rmu, double temp,
                   long Nk, int Nw, int No, int* err, hipStream_t s) {
  constexpr int NB = 1024;
  DBuf<double> part(2 * NB);
  long tot = Nk * (long)Nw;
  LAUNCH_NO(No, k_filling, dim3(NB), dim3(256), sigma, hamk, eig, rmu, temp, Nk, Nw, part.p, err);
  GMOD_CHECK_LAUNCH();
  (void)tot;
  std::vector<double> h(2 * NB);
  GMOD_HIP(hipMemcpyAsync(h.data(), part.p, 2 * NB * sizeof(double), hipMemcpyDeviceToHost, s));
  GMOD_HIP(hipStreamSynchronize(s));
  out[0] = out[1] = 0.0;
  for (int b = 0; b < NB; ++b) { out[0] += h[2 * b]; out[1] += h[2 * b + 1]; }
}

void filling_charpoly(zd* coef, const zd* sigma, const zd* hamk, long Nk, int Nw, int No, hipStream_t s) {
  const long tot = Nk * (long)Nw;
  LAUNCH_NO(No, k_charpoly, dim3(nblk(tot, 128)), dim3(128), coef, sigma, hamk, Nk, Nw);
  GMOD_CHECK_LAUNCH();
}

void filling_parts_cp(double out[2], const zd* coef, const double* eig, double rmu, double temp, long Nk, int Nw,
                      int No, hipStream_t s) {
  constexpr int NB = 1024;
  DBuf<double> part(2 * NB);
  LAUNCH_NO(No, k_filling_cp, dim3(NB), dim3(256), coef, eig, rmu, temp, Nk, Nw, part.p);
  GMOD_CHECK_LAUNCH();
  std::vector<double> h(2 * NB);
  GMOD_HIP(hipMemcpyAsync(h.data(), part.p, 2 * NB * sizeof(double), hipMemcpyDeviceToHost, s));
  GMOD_HIP(hipStreamSynchronize(s));
  out[0] = out[1] = 0.0;
  for (int b = 0; b < NB; ++b) { out[0] += h[2 * b]; out[1] += h[2 * b + 1]; }
}

void green0(zd* G, const double* eig, const zd* uni, double mu, double temp, long Nk, int Nw, int No,
            hipStream_t s) {
  long tot = Nk * (long)Nw;
  LAUNCH_NO(No, k_green0, dim3(nblk(tot, 128)), dim3(128), G, eig, uni, mu, temp, Nk, Nw);
  GMOD_CHECK_LAUNCH();
}

void anomalous_f(zd* F, const zd* G, const zd* D, long Nk, int Nw, int No, hipStream_t s) {
  long tot = Nk * (long)Nw;
  LAUNCH_NO(No, k_anom_f, dim3(nblk(tot, 128)), dim3(128), F, G, D, Nk, Nw);
  GMOD_CHECK_LAUNCH();
}

void sc_green_update(zd* G, zd* F, const zd* G0, const zd* D, long Nk, int Nw, int No, hipStream_t s) {
  long tot = Nk * (long)Nw;
  LAUNCH_NO(No, k_sc_green, dim3(nblk(tot, 128)), dim3(128), G, F, G0, D, Nk, Nw);
  GMOD_CHECK_LAUNCH();
}

void hermitize_lower(zd* D, long Nk, int Nw, int No, hipStream_t s) {
  long tot = Nk * (long)Nw;
  hipLaunchKernelGGL(k_hermitize_lower, dim3(nblk(tot, 256)), dim3(256), 0, s, D, tot, No);
  GMOD_CHECK_LAUNCH();
}

void subtract_static(zd* sigma, long Nk, int Nw, int No, int mode, hipStream_t s) {
  if (No > 8) throw Error("Norb > 8 not supported");
  DBuf<zd> sub((size_t)Nk * No * No);
  hipLaunchKernelGGL(k_static_part, dim3(nblk((long)Nk * No * No, 256)), dim3(256), 0, s, sub.p, sigma, Nk, Nw, No,
                     mode);
  GMOD_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_sub_static, dim3(nblk((long)Nk * Nw * No * No, 256)), dim3(256), 0, s, sigma, sub.p, Nk, Nw,
                     No);
  GMOD_CHECK_LAUNCH();
}

template <int NO>
__global__ void __launch_bounds__(128) k_inv(zd* A, long tot, int* err) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
    zd a[NO][NO];
#pragma unroll
    for (int p = 0; p < NO; ++p)
#pragma unroll
      for (int q = 0; q < NO; ++q) a[p][q] = A[e + tot * (p + NO * q)];
    if (!gj_inverse<NO>(a)) atomicOr(err, 4);
#pragma unroll
    for (int p = 0; p < NO; ++p)
#pragma unroll
      for (int q = 0; q < NO; ++q) A[e + tot * (p + NO * q)] = a[p][q];
  }
}

void batched_inv(zd* A, long tot, int No, int* err, hipStream_t s) {
  LAUNCH_NO(No, k_inv, dim3(nblk(tot, 128)), dim3(128), A, tot, err);
  GMOD_CHECK_LAUNCH();
}

// G^-1 = (i w_j + mu) I - H(k) - Sigma  (fself.f90:43-83)
__global__ void k_green_inv(zd* G, const zd* sig, const zd* H, double mu, double temp, long Nk, int Nw, int No) {
  const long tot = Nk * (long)Nw;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
    long k = e % Nk, j = e / Nk;
    zd z(mu, (double)(2 * j + 1) * PI_D * temp);
    for (int p = 0; p < No; ++p)
      for (int q = 0; q < No; ++q) {
        zd v = zd(0, 0) - (H[p + No * (q + (long)No * k)] + sig[e + tot * (p + No * q)]);
        if (p == q) v += z;
        G[e + tot * (p + No * q)] = v;
      }
  }
}

void green_inv(zd* G, const zd* sigma, const zd* hamk, double mu, double temp, long Nk, int Nw, int No,
               hipStream_t s) {
  long tot = Nk * (long)Nw;
  hipLaunchKernelGGL(k_green_inv, dim3(nblk(tot, 128)), dim3(128), 0, s, G, sigma, hamk, mu, temp, Nk, Nw, No);
  GMOD_CHECK_LAUNCH();
}

// ====================================================================== host helpers
int host_eigvals(std::vector<zd>& A, int n, std::vector<zd>& w) {
  w.assign(n, zd(0, 0));
  std::vector<zd> H = A;
  return zhess_eig(H.data(), n, w.data());
}

static bool host_lu_solve(std::vector<zd> M, int n, std::vector<zd>& b) {
  std::vector<int> piv(n);
  for (int k = 0; k < n; ++k) {
    int p = k;
    double best = abs1(M[k + n * k]);
    for (int r = k + 1; r < n; ++r)
      if (abs1(M[r + n * k]) > best) { best = abs1(M[r + n * k]); p = r; }
    if (best == 0.0) return false;
    if (p != k) {
      for (int c = 0; c < n; ++c) std::swap(M[k + n * c], M[p + n * c]);
      std::swap(b[k], b[p]);
    }
    for (int r = k + 1; r < n; ++r) {
      zd l = M[r + n * k] / M[k + n * k];
      for (int c = k + 1; c < n; ++c) M[r + n * c] -= l * M[k + n * c];
      b[r] -= l * b[k];
    }
  }
  for (int k = n - 1; k >= 0; --k) {
    zd x = b[k];
    for (int j = k + 1; j < n; ++j) x -= M[k + n * j] * b[j];
    b[k] = x / M[k + n * k];
  }
  return true;
}

std::vector<zd> host_eigvec(const std::vector<zd>& A, int n, zd lam) {
  // inverse iteration on (A - lam I) with a tiny relative perturbation
  double scale = 0.0;
  for (auto v : A) scale = std::max(scale, abs1(v));
  if (scale == 0.0) scale = 1.0;
  std::vector<zd> M = A;
  zd shift = lam + zd(scale * 1e-13, scale * 1e-13);
  for (int i = 0; i < n; ++i) M[i + n * i] -= shift;
  std::vector<zd> v(n);
  for (int i = 0; i < n; ++i) v[i] = zd(1.0 / std::sqrt((double)n), 1e-3 * (i + 1));
  for (int it = 0; it < 4; ++it) {
    std::vector<zd> b = v;
    if (!host_lu_solve(M, n, b)) {
      // exact eigenvalue hit: perturb more
      for (int i = 0; i < n; ++i) M[i + n * i] -= zd(scale * 1e-10, 0);
      continue;
    }
    double nr = 0;
    for (auto x : b) nr += x.re * x.re + x.im * x.im;
    nr = std::sqrt(nr);
    for (int i = 0; i < n; ++i) v[i] = b[i] * (1.0 / nr);
  }
  // LAPACK zgeev normalisation
  int kmax = 0;
  double best = -1;
  for (int i = 0; i < n; ++i) {
    double m2 = v[i].re * v[i].re + v[i].im * v[i].im;
    if (m2 > best) { best = m2; kmax = i; }
  }
  zd ph = conj(v[kmax]) * (1.0 / std::sqrt(best));
  for (int i = 0; i < n; ++i) v[i] = v[i] * ph;
  v[kmax].im = 0.0;
  return v;
}

bool host_dgesv(std::vector<double>& A, int n, std::vector<double>& b) {
  for (int k = 0; k < n; ++k) {
    int p = k;
    double best = std::fabs(A[k + n * k]);
    for (int r = k + 1; r < n; ++r)
      if (std::fabs(A[r + n * k]) > best) { best = std::fabs(A[r + n * k]); p = r; }
    if (best == 0.0) return false;
    if (p != k) {
      for (int c = 0; c < n; ++c) std::swap(A[k + n * c], A[p + n * c]);
      std::swap(b[k], b[p]);
    }
    for (int r = k + 1; r < n; ++r) {
      double l = A[r + n * k] / A[k + n * k];
      for (int c = k + 1; c < n; ++c) A[r + n * c] -= l * A[k + n * c];
      b[r] -= l * b[k];
    }
  }
  for (int k = n - 1; k >= 0; --k) {
    double x = b[k];
    for (int j = k + 1; j < n; ++j) x -= A[k + n * j] * b[j];
    b[k] = x / A[k + n * k];
  }
  return true;
}

// ====================================================================== spinful Eliashberg
// Arrays on the full zone: element (i, j, a, b) of X(Nkall, Nw, No, No) at
// i + Nkall (j + Nw (a + No b)); G(Nk, Nw, No, No) on the irreducible set.
// sg[a + No b] = s_a s_b, inv[a] the 0-based spin-flipped orbital.

// initial gap (feliashwithsoc.f90:665-748) before its normalisation: one
// thread per (i, a, b), the same value at every frequency
__global__ void k_initial_delta_soc(zd* D, const double* init, const zd* U, const int64_t* invk, const int* sl,
                                    const int* inv, long Nkall, long Nk, int Nw, int No, long gap_sym) {
  const long tot = Nkall * No * No;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
    const long i = e % Nkall;
    const int a = (int)((e / Nkall) % No), b = (int)(e / (Nkall * No));
    const long k = invk[3 * i] - 1;
    const long f = invk[3 * i + 1];
    zd v(0, 0);
    if (f == 0 || f == 1) {
      for (int n = 0; n < No; ++n) {
        const double d = init[k + Nk * n];
        if (f == 0)  // -U diag(init) (s_b conj U(-b, :))^T
          v -= U[a + No * (n + (long)No * k)] * conj(U[inv[b] + No * (n + (long)No * k)]) * (d * sl[b]);
        else  // (s_a conj U(-a, :)) diag(init) U^T
          v += conj(U[inv[a] + No * (n + (long)No * k)]) * U[b + No * (n + (long)No * k)] * (d * sl[a]);
      }
    }
    if (gap_sym > 0 && sl[a] * sl[b] >= 0) v = zd(0, 0);
    for (int j = 0; j < Nw; ++j) D[i + Nkall * (j + (long)Nw * (a + (long)No * b))] = v;
  }
}

// F = -G(k) Delta(k) G(-k)^dagger with time reversal (feliashwithsoc.f90:445-470):
// one thread per (i, j, m, l)
__global__ void k_anom_f_soc(zd* F, const zd* G, const zd* D, const int64_t* invk, const double* sg, const int* inv,
                             long Nkall, long Nk, int Nw, int No) {
  const long NW = Nkall * Nw, NkNw = Nk * (long)Nw;
  const long tot = NW * No * No;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
    const long ij = e % NW;
    const long i = ij % Nkall, j = ij / Nkall;
    const int m = (int)((e / NW) % No), l = (int)(e / (NW * No));
    const long g0 = (invk[3 * i] - 1) + Nk * j;
    const bool dir = invk[3 * i + 1] == 0;
    auto Gat = [&](int p, int q) { return G[g0 + NkNw * (p + (long)No * q)]; };
    zd acc(0, 0);
    for (int b = 0; b < No; ++b) {
      zd t(0, 0);  // (G1 Delta)(m, b)
      for (int a = 0; a < No; ++a) {
        const zd g1 = dir ? Gat(m, a) : Gat(inv[a], inv[m]) * sg[a + No * m];
        t += g1 * D[ij + NW * (a + (long)No * b)];
      }
      const zd g2 = dir ? Gat(inv[l], inv[b]) * sg[b + No * l] : Gat(b, l);
      acc += t * conj(g2);
    }
    F[e] = zd(-acc.re, -acc.im);
  }
}

// spin structure of F (feliashwithsoc.f90:472-517), in the reference's order,
// one thread per (i, j)
__global__ void k_fk_sym_soc(zd* F, const int* sl, const int* inv, long NW, int No, long gap_sym) {
  for (long ij = blockIdx.x * (long)blockDim.x + threadIdx.x; ij < NW; ij += (long)gridDim.x * blockDim.x) {
    auto at = [&](int p, int q) -> zd& { return F[ij + NW * (p + (long)No * q)]; };
    for (int l = 0; l < No; ++l) {
      if (sl[l] != 1) continue;
      for (int m = 0; m < No; ++m) {
        if (sl[m] == 1) {
          if (gap_sym < 0) {
            const zd v = at(m, l);
            at(inv[l], inv[m]) = zd(-v.re, v.im);
          } else {
            at(m, l) = zd(0, 0);
            at(inv[l], inv[m]) = zd(0, 0);
          }
        } else if (inv[m] == l) {
          at(inv[l], inv[m]) = zd(at(m, l).re, 0);
        }
      }
    }
    for (int l = 0; l < No; ++l) {
      if (sl[l] != 1) continue;
      for (int m = 0; m < No; ++m) {
        if (sl[m] != -1 || gap_sym == -10) continue;
        const zd v = at(m, l);
        at(inv[m], inv[l]) = gap_sym < 0 ? v : zd(-v.re, -v.im);
      }
    }
  }
}

// spin structure of the new gap (feliashwithsoc.f90:614-662), one thread per (i, j)
__global__ void k_delta_sym_soc(zd* D, const int* sl, const int* inv, long NW, int No, long gap_sym) {
  for (long ij = blockIdx.x * (long)blockDim.x + threadIdx.x; ij < NW; ij += (long)gridDim.x * blockDim.x) {
    auto at = [&](int p, int q) -> zd& { return D[ij + NW * (p + (long)No * q)]; };
    for (int l = 0; l < No; ++l) {
      if (sl[l] != 1) continue;
      for (int m = 0; m < No; ++m) {
        if (sl[m] == 1) {
          const zd v = at(m, l);
          at(inv[l], inv[m]) = zd(-v.re, v.im);
        } else if (inv[m] == l) {
          at(l, m) = zd(at(l, m).re, 0);
        } else if (l > inv[m]) {
          at(inv[m], inv[l]) = conj(at(l, m));
        }
      }
    }
    for (int l = 0; l < No; ++l) {
      if (sl[l] != 1) continue;
      for (int m = 0; m < No; ++m) {
        if (sl[m] != -1 || gap_sym == -10) continue;
        const zd v = at(l, m);
        at(inv[l], inv[m]) = gap_sym > 0 ? zd(-v.re, -v.im) : v;
      }
    }
  }
}

void initial_delta_soc(zd* D, const double* init, const zd* U, const int64_t* invk, const int* sl, const int* inv,
                       long Nkall, long Nk, int Nw, int No, long gap_sym, hipStream_t s) {
  const long tot = Nkall * No * No;
  hipLaunchKernelGGL(k_initial_delta_soc, dim3(nblk(tot, 256)), dim3(256), 0, s, D, init, U, invk, sl, inv, Nkall, Nk,
                     Nw, No, gap_sym);
  GMOD_CHECK_LAUNCH();
}

void anomalous_f_soc(zd* F, const zd* G, const zd* D, const int64_t* invk, const double* sg, const int* sl,
                     const int* inv, long Nkall, long Nk, int Nw, int No, long gap_sym, hipStream_t s) {
  const long NW = Nkall * Nw;
  hipLaunchKernelGGL(k_anom_f_soc, dim3(nblk(NW * No * No, 256)), dim3(256), 0, s, F, G, D, invk, sg, inv, Nkall, Nk,
                     Nw, No);
  GMOD_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_fk_sym_soc, dim3(nblk(NW, 256)), dim3(256), 0, s, F, sl, inv, NW, No, gap_sym);
  GMOD_CHECK_LAUNCH();
}

void delta_sym_soc(zd* D, const int* sl, const int* inv, long NW, int No, long gap_sym, hipStream_t s) {
  hipLaunchKernelGGL(k_delta_sym_soc, dim3(nblk(NW, 256)), dim3(256), 0, s, D, sl, inv, NW, No, gap_sym);
  GMOD_CHECK_LAUNCH();
}

}  // namespace gmod
