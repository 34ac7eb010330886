// Wannier-R representation of Sigma(k, i w_n) or Delta(k, i w_n): the
// k -> R transform behind the reference's output_self_wannier /
// output_gap_wannier (libs/plibs/_wannier_io.py:45-143, 145-216) on the
// device.  The irreducible data are unfolded onto the full k-grid (a point
// reached through -k takes the complex conjugate, _wannier_io.py:60-66) and
// the first Ncut frequencies of every orbital pair are transformed by the
// library's box FFT over the spatial axes (normalised backward transform,
// numpy's ifftn).  Zeroing, R-vector selection and the file formats stay on
// the host (pyrpa_amd/wannier_io.py).
#include "../../include/gmod.h"
#include "fft.h"

namespace gmod {

// out (Nx, Ny, Nz, M), x fastest, slot m = (i No + j) Ncut + n for the
// Python element data[i, j, n, k] = in[k + Nk (n + Nw (j + No i))].
// full != 0: the data already cover the full grid (k = full index).
__global__ void k_wannier_unfold(zd* out, const zd* __restrict__ in, const int64_t* __restrict__ kmap,
                                 const int64_t* __restrict__ invk, long Nkall, long Nk, int Nw, int No, int Ncut,
                                 int Nx, int Ny, int full) {
  const long M = (long)No * No * Ncut;
  const long tot = Nkall * M;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
    const long p = e % Nkall, m = e / Nkall;
    const int n = (int)(m % Ncut), j = (int)((m / Ncut) % No), i = (int)(m / ((long)Ncut * No));
    long k;
    bool tr = false;
    if (full) {
      k = p;
    } else {
      k = invk[3 * p] - 1;
      tr = invk[3 * p + 1] != 0;
    }
    zd v = in[k + (full ? Nkall : Nk) * (n + (long)Nw * (j + (long)No * i))];
    if (tr) v = conj(v);
    const long sp = kmap[3 * p] + (long)Nx * (kmap[3 * p + 1] + (long)Ny * kmap[3 * p + 2]);
    out[sp + Nkall * m] = v;
  }
}

}  // namespace gmod

using namespace gmod;

// data: Python [No, No, Nw, Nk] (Nk irreducible, or Nkall when full != 0);
// out: Python [No, No, Ncut, Nz, Ny, Nx] = (1/Nkall) sum_k e^{+2 pi i k R} data(k)
extern "C" void gmod_wannier_r(gmod_z* out, const gmod_z* data, const int64_t* kmap, const int64_t* invk,
                               int64_t Nx, int64_t Ny, int64_t Nz, int64_t Nw, int64_t Nk, int64_t Norb,
                               int64_t Ncut, int full) {
  try {
    std::lock_guard<std::mutex> g(ctx().lock);
    hipStream_t s = ctx().stream;
    const long Nkall = Nx * Ny * Nz;
    if (Nkall < 1 || Norb < 1 || Ncut < 1 || Ncut > Nw) throw Error("gmod_wannier_r: bad sizes");
    if (full && Nk != Nkall) throw Error("gmod_wannier_r: full-grid data need Nk = Nx Ny Nz");
    // every index the unfold kernel follows stays inside its buffers
    for (long p = 0; p < Nkall; ++p) {
      if (kmap[3 * p] < 0 || kmap[3 * p] >= Nx || kmap[3 * p + 1] < 0 || kmap[3 * p + 1] >= Ny ||
          kmap[3 * p + 2] < 0 || kmap[3 * p + 2] >= Nz)
        throw Error("gmod_wannier_r: kmap entry outside the grid");
      if (!full && (invk[3 * p] < 1 || invk[3 * p] > Nk)) throw Error("gmod_wannier_r: invk entry outside 1..Nk");
    }
    const long M = Norb * Norb * Ncut;
    const long nin = (full ? Nkall : Nk) * Nw * Norb * Norb;
    DBuf<zd> in(nin), buf(Nkall * M), tmp(Nkall * M);
    DBuf<int64_t> km(3 * Nkall), iv(3 * Nkall);
    in.upload((const zd*)data, nin, s);
    km.upload(kmap, 3 * Nkall, s);
    iv.upload(invk, 3 * Nkall, s);
    GMOD_HIP(hipMemsetAsync(buf.p, 0, buf.bytes(), s));
    const long tot = Nkall * M;
    const unsigned blocks = (unsigned)std::min<long>((tot + 255) / 256, 1L << 20);
    hipLaunchKernelGGL(k_wannier_unfold, dim3(blocks), dim3(256), 0, s, buf.p, in.p, km.p, iv.p, Nkall, (long)Nk,
                       (int)Nw, (int)Norb, (int)Ncut, (int)Nx, (int)Ny, full);
    GMOD_CHECK_LAUNCH();
    const FFTPlan<double>& p = fft_plan<double>((int)Nx, (int)Ny, (int)Nz, (int)M, 7);
    p.exec(buf.p, tmp.p, +1, 1.0 / (double)Nkall, s);
    buf.download((zd*)out, Nkall * M, s);
    GMOD_HIP(hipStreamSynchronize(s));
  } catch (std::exception& e) {
    set_error(e.what());
  }
}
