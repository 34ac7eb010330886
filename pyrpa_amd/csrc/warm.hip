// Warm-start re-gridding of a FLEX self-energy: Sigma of a solve on a coarse
// mesh (k-mesh and Matsubara count) onto a finer mesh at the same
// temperature, on the device.
//
//  * k: Fourier interpolation over the full Brillouin zone.  The irreducible
//    coarse Sigma is unfolded with the time-reversal rule of G
//    (Sigma(-k) = Sigma(k)^T, libs/src/fchi.f90:573-579), transformed to R
//    space, zero-padded onto the fine R grid (a Nyquist plane is split in
//    halves between +R and -R) and transformed back; the fine irreducible
//    points are gathered from their direct full-BZ points.
//  * frequency: the fine mesh shares the Matsubara frequencies w_n of the
//    coarse one (same T).  Coarse values are kept for n < Nwc - 1 (the
//    outermost stored frequency carries the wrap-around of the circular
//    convolution; it is kept too when both meshes have the same cutoff);
//    beyond, the tail c0 + c1/(i w) is fitted by least squares
//    on the window [0.55, 0.85] Nwc and diagonal Im Sigma is clipped to <= 0
//    (causality) -- the frequency treatment of the reference's Python
//    re-gridder, libs/plibs/_calc.py:52-182.
#include <cmath>

#include "../../include/gmod.h"
#include "engine.h"

namespace gmod {

// fine R-space grid <- weighted coarse R-space grid (zero padding)
__global__ void k_pad_rspace(zd* __restrict__ fine, const zd* __restrict__ coarse, int cx, int cy, int cz,
                             int fx, int fy, int fz, int nslot, double scale) {
  const long Sf = (long)fx * fy * fz, Sc = (long)cx * cy * cz;
  const long tot = Sf * nslot;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
    const long sp = e % Sf, t = e / Sf;
    const int x = (int)(sp % fx), y = (int)((sp / fx) % fy), z = (int)(sp / ((long)fx * fy));
    double w = scale;
    int c[3];
    const int f[3] = {x, y, z}, nf[3] = {fx, fy, fz}, nc[3] = {cx, cy, cz};
    for (int d = 0; d < 3; ++d) {
      const int h = nc[d] / 2;
      if (nc[d] == nf[d]) { c[d] = f[d]; continue; }
      if (f[d] < (nc[d] + 1) / 2 && !((nc[d] % 2 == 0) && f[d] == h)) c[d] = f[d];
      else if (f[d] > nf[d] - (nc[d] + 1) / 2) c[d] = f[d] - (nf[d] - nc[d]);
      else if ((nc[d] % 2 == 0) && (f[d] == h || f[d] == nf[d] - h)) { c[d] = h; w *= 0.5; }
      else { w = 0.0; c[d] = 0; }
    }
    zd v(0, 0);
    if (w != 0.0) v = coarse[(c[0] + (long)cx * (c[1] + (long)cy * c[2])) + Sc * t] * w;
    fine[e] = v;
  }
}

// tail fit per (k, a, b): Sigma(w_n) ~ c0 + c1 / (i w_n) on [i0, i1), filled
// for n >= nkeep; diagonal Im Sigma clipped to <= 0 for all n
__global__ void k_sigma_tail(zd* __restrict__ sig, long Nk, int Nw, int No, int nkeep, int i0, int i1, double temp) {
  const long tot = Nk * No * No;
  const double PI = 3.141592653589793238462643383279;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
    const long k = e % Nk;
    const long ab = e / Nk;
    const int a = (int)(ab % No), b = (int)(ab / No);
    zd* s = sig + k + Nk * (long)Nw * ab;
    if (nkeep < Nw) {
      // normal equations with basis (1, phi), phi = 1/(i w) = -i/w
      double n0 = 0, s11 = 0;
      zd s01(0, 0), r0(0, 0), r1(0, 0);
      for (int n = i0; n < i1; ++n) {
        const double w = (2.0 * n + 1.0) * PI * temp;
        const zd phi(0.0, -1.0 / w);
        const zd y = s[Nk * (long)n];
        n0 += 1.0;
        s01 += phi;
        s11 += 1.0 / (w * w);
        r0 += y;
        r1 += conj(phi) * y;
      }
      // [n0  s01; conj(s01) s11] [c0; c1] = [r0; r1]
      const zd det = zd(n0 * s11, 0) - s01 * conj(s01);
      const zd c0 = (r0 * s11 - s01 * r1) / det;
      const zd c1 = (r1 * n0 - conj(s01) * r0) / det;
      for (int n = nkeep; n < Nw; ++n) {
        const double w = (2.0 * n + 1.0) * PI * temp;
        s[Nk * (long)n] = c0 + c1 * zd(0.0, -1.0 / w);
      }
    }
    if (a == b)
      for (int n = 0; n < Nw; ++n) {
        zd v = s[Nk * (long)n];
        if (v.im > 0.0) { v.im = 0.0; s[Nk * (long)n] = v; }
      }
  }
}

}  // namespace gmod

using namespace gmod;

extern "C" void gmod_regrid_sigma(gmod_z* sig_f, const gmod_z* sig_c, const int64_t* kmap_c, const int64_t* invk_c,
                                  int64_t Nk_c, int64_t Nx_c, int64_t Ny_c, int64_t Nz_c, int64_t Nw_c,
                                  const int64_t* kmap_f, const int64_t* invk_f, int64_t Nk_f, int64_t Nx_f,
                                  int64_t Ny_f, int64_t Nz_f, int64_t Nw_f, int64_t Norb, double temp) {
  try {
    std::lock_guard<std::mutex> g(ctx().lock);
    hipStream_t s = ctx().stream;
    const int No = (int)Norb;
    if (Nx_f < Nx_c || Ny_f < Ny_c || Nz_f < Nz_c) throw Error("regrid: the fine k-mesh must not be coarser");
    if (Nw_c < 8) throw Error("regrid: the coarse Matsubara count must be at least 8");
    const long Nkall_c = Nx_c * Ny_c * Nz_c, Nkall_f = Nx_f * Ny_f * Nz_f;
    Mesh mc, mf;
    mc.build(kmap_c, invk_c, Nkall_c, Nk_c, (int)Nx_c, (int)Ny_c, (int)Nz_c, (int)Nw_c, s);
    mf.build(kmap_f, invk_f, Nkall_f, Nk_f, (int)Nx_f, (int)Ny_f, (int)Nz_f, (int)Nw_f, s);
    // frequencies carried through the k interpolation: the coarse ones the fine mesh has
    const int nslot = (int)std::min<int64_t>(Nw_c, Nw_f);
    const size_t nc = (size_t)Nk_c * Nw_c * No * No, nf = (size_t)Nk_f * Nw_f * No * No;
    DBuf<zd> dc(nc), df(nf), gc((size_t)Nkall_c * nslot), gf((size_t)Nkall_f * nslot), tmp;
    tmp.alloc((size_t)std::max(Nkall_c, Nkall_f) * nslot);
    dc.upload((const zd*)sig_c, nc, s);
    GMOD_HIP(hipMemsetAsync(df.p, 0, nf * sizeof(zd), s));
    const FFTPlan<double>& pc = fft_plan<double>((int)Nx_c, (int)Ny_c, (int)Nz_c, nslot, 7);
    const FFTPlan<double>& pf = fft_plan<double>((int)Nx_f, (int)Ny_f, (int)Nz_f, nslot, 7);
    const MeshDev mdc = mesh_dev(mc), mdf = mesh_dev(mf);
    const long NkNw_f = Nk_f * Nw_f;
    for (int b = 0; b < No; ++b)
      for (int a = 0; a < No; ++a) {
        pc.exec_fused(GreenLoad<double>{dc.p, a, b, No, mdc}, GridStore<double>{gc.p, (unsigned)Nkall_c}, gc.p,
                      tmp.p, -1, 1.0, s);
        const long tot = Nkall_f * (long)nslot;
        const unsigned grid = (unsigned)std::min<long>((tot + 255) / 256, 1L << 20);
        hipLaunchKernelGGL(k_pad_rspace, dim3(grid), dim3(256), 0, s, gf.p, gc.p, (int)Nx_c, (int)Ny_c, (int)Nz_c,
                           (int)Nx_f, (int)Ny_f, (int)Nz_f, nslot, (double)Nkall_f / (double)Nkall_c);
        GMOD_CHECK_LAUNCH();
        pf.exec_fused(GridLoad<double>{gf.p, (unsigned)Nkall_f},
                      GatherStore<double>{df.p + NkNw_f * (a + (long)No * b), 1.0, zd(0, 0), nullptr, 0, mdf}, gf.p,
                      tmp.p, +1, 1.0 / (double)Nkall_f, s);
      }
    // the outermost coarse frequency is replaced by the tail unless the fine
    // mesh has the same cutoff (then it carries the same wrap-around)
    const int nkeep = (int)(Nw_c == Nw_f ? Nw_f : std::min<int64_t>(Nw_c - 1, Nw_f));
    const int i0 = (int)(0.55 * Nw_c), i1 = std::max((int)(0.85 * Nw_c), i0 + 4);
    const long tot = Nk_f * No * No;
    hipLaunchKernelGGL(k_sigma_tail, dim3((unsigned)std::min<long>((tot + 255) / 256, 1L << 20)), dim3(256), 0, s,
                       df.p, Nk_f, (int)Nw_f, No, nkeep, i0, std::min(i1, (int)Nw_c - 1), temp);
    GMOD_CHECK_LAUNCH();
    df.download((zd*)sig_f, nf, s);
    GMOD_HIP(hipStreamSynchronize(s));
  } catch (std::exception& e) {
    set_error(e.what());
  }
}
