// Grid kernels of the convolution path: irreducible -> full-grid scatter with
// the reference's symmetry unfolding, pointwise products, full -> irreducible
// gather, thresholds and reductions.
//
// k-space arrays keep the reference's Fortran layouts (column-major):
//   Gk(Nk,Nw,No,No), chi(Nk,Nw,Nchi,Nchi), hamk(No,No,Nk), eig(No,Nk),
//   uni(No,No,Nk).
// Full grids are (Nx,Ny,Nz,M=2Nw), x fastest; the spatial index of full-BZ
// point i is sp(i) = kmap(1,i) + Nx*kmap(2,i) + Nx*Ny*kmap(3,i).
#pragma once
#include "gmod_common.h"

namespace gmod {

// Device copy of the k-mesh bookkeeping (finv.f90 outputs).
struct Mesh {
  int Nx = 0, Ny = 0, Nz = 0, Nw = 0, M = 0;
  long Nkall = 0, Nk = 0;
  DBuf<int> irr;      // invk(1,i)-1
  DBuf<int> flag;     // invk(2,i)
  DBuf<int> sp;       // spatial grid index of full point i
  DBuf<int> rev;      // spatial index of -r for spatial index s
  DBuf<int> irr_sp;   // spatial index of the flag-0 full point of irreducible k (-1: none)
  DBuf<double> kw;    // full-BZ multiplicity of irreducible k
  std::vector<double> kw_host;
  long N() const { return Nkall * (long)M; }
  void build(const int64_t* kmap, const int64_t* invk, long Nkall, long Nk, int Nx, int Ny, int Nz,
             int Nw, hipStream_t s);
};

// ---- scatters (write the whole full grid)
template <typename R>
void scatter_green(cpx<R>* grid, const zd* G, int a, int b, int No, const Mesh& m, hipStream_t s);
template <typename R>
void scatter_boson(cpx<R>* grid, const zd* V, int mi, int li, int Nchi, zd stat, const Mesh& m,
                   hipStream_t s);
template <typename R>
void scatter_anom(cpx<R>* grid, const zd* F, int a, int b, int No, double pp, double sgn, double dprt,
                  const Mesh& m, hipStream_t s);
// analytic band kernels on the tau grid (fchi.f90:703-729): which=0 -> g(+tau), 1 -> g(-tau)
template <typename R>
void scatter_tau_ref(cpx<R>* grid, const double* eig, const zd* uni, int a, int b, int No, double mu,
                     double temp, int which, const Mesh& m, hipStream_t s);

// ---- products
// out = coef * A(r,t) * B(-r,-t)          (rev_t = true)   [fchi.f90:593]
// out = coef * A(r,t) * B(-r, t)          (rev_t = false)  [fchi.f90:824]
template <typename R>
void product_rev(cpx<R>* out, const cpx<R>* A, const cpx<R>* B, R coef, bool rev_t, bool accumulate,
                 const Mesh& m, hipStream_t s);
// acc (+)= A * B
template <typename R>
void mul_acc(cpx<R>* acc, const cpx<R>* A, const cpx<R>* B, bool first, long n, hipStream_t s);

// ---- gathers
// dst(k, j) (+)= weight * grid(sp(k), j) + J * sawc(j)   for j < Nw; dst is a (Nk,Nw) slice
template <typename R>
void gather(zd* dst, const cpx<R>* grid, double weight, bool accumulate, zd J, const zd* sawc,
            const Mesh& m, hipStream_t s);

// ---- elementwise
void conj_copy(zd* dst, const zd* src, long n, hipStream_t s);
void threshold(zd* a, long n, double eps, hipStream_t s);
void scale(zd* a, long n, double f, hipStream_t s);
void axpy(zd* y, const zd* x, zd alpha, long n, hipStream_t s);      // y += alpha x
void copy(zd* dst, const zd* src, long n, hipStream_t s);
void lincomb(zd* y, const zd* a, double ca, const zd* b, double cb, long n, hipStream_t s);  // y = ca a + cb b

// ---- deterministic reductions (two-stage, fixed order)
// sum over k-space arrays laid out (Nk, rest): weights kw[k] (or 1)
zd dot_w(const zd* u, const zd* v, const double* kw, long Nk, long rest, hipStream_t s);  // sum kw conj(u) v
double max_rel_err(const zd* s0, const zd* sd, long n, double floor, hipStream_t s, long* where);
double max_real(const zd* a, long n, hipStream_t s);
double min_real(const zd* a, long n, hipStream_t s);

}  // namespace gmod
