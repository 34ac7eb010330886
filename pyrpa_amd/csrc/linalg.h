// Batched small dense complex linear algebra on the device (FP64).
#pragma once
#include "gmod_common.h"

namespace gmod {

// Batched matrix addressing: matrix b, element (p,q) (column-major, p + n*q)
// lives at base[b*bs + (p + n*q)*es].
struct BatchLayout {
  long bs, es;
};

// RPA vertex per (q,i nu), in place (fUJ.f90:253-365):
//  mode 0: V_sigma = (1.5S-0.5C) + 1.5 S X_s + 0.5 C X_c - 0.25 (S+C)(chi0 S + chi0 C)
//  mode 1: V_Delta singlet = (S+C)/2 + 1.5 S X_s - 0.5 C X_c
//  mode 2: V_Delta triplet = (S-C)/2 - 0.5 S X_s - 0.5 C X_c
//  X_s = (I - chi0 S)^-1 chi0 S,  X_c = (I + chi0 C)^-1 chi0 C  (LU with partial pivoting)
void rpa_vertex(zd* chi, long batch, BatchLayout L, int n, const double* S, const double* C, int mode,
                int* err, hipStream_t s);
// mode 0 out of place on a matrix-major batch (L.bs = 1): reads chi_in with
// the 1e-9 threshold of the FLEX loop applied as it is read (thr > 0),
// writes V_sigma to chi_out (may equal chi_in) and, when vprev is given,
// dv = V_sigma - vprev in FP32 (the increment the mixed-precision loop
// transforms); with dv and store_v false, V_sigma itself may be left unstored
// (chi_out then keeps what it held, or receives V_sigma where the
// Gauss-Jordan kernel does not apply)
void rpa_vertex_fused(const zd* chi_in, zd* chi_out, long batch, BatchLayout L, int n, const double* S,
                      const double* C, double thr, const zd* vprev, cpx<float>* dv, int* err, hipStream_t s,
                      bool store_v = true);
// spinful FLEX vertex in a batch (fUJ.f90:367-399): out = -V + V (I + chi0 V)^-1 chi0 V,
// chi0 thresholded at thr (> 0) as it is read; out may equal chi_in
void soc_vertex(const zd* chi_in, zd* out, long batch, BatchLayout L, int n, const double* V, double thr, int* err,
                hipStream_t s);
// X = (I + alpha A W)^-1 A  (fchi.f90:319-350, fself.f90:783-816); out may alias nothing
void rpa_solve(zd* out, BatchLayout Lo, const zd* A, BatchLayout La, long batch, int n, const double* W,
               double alpha, int* err, hipStream_t s);
// max Re eig(alpha * A * W) per matrix (W = nullptr: identity)  (fchi.f90:373-389)
void max_re_eig(double* out, const zd* A, BatchLayout La, long batch, int n, const double* W, double alpha,
                int* err, hipStream_t s);
// the same with a caller-provided workspace of batch (n^2 + n) elements; no
// allocation, no synchronisation
void max_re_eig_ws(double* out, const zd* A, BatchLayout La, long batch, int n, const double* W, double alpha,
                   zd* work, int* err, hipStream_t s);

// Dyson: G = [(i w_j + mu) I - H(k) - Sigma(k, i w_j)]^-1   (fself.f90:43-83, 131-157)
void dyson(zd* G, const zd* sigma, const zd* hamk, double mu, double temp, long Nk, int Nw, int No, int* err,
           hipStream_t s);
// n(mu) pieces (fself.f90:687-714): returns {sum_k,n f(e), sum_{k,j} Re[Tr(z-H-Sigma)^-1 - sum_n 1/(z-e_n)]}
void filling_parts(double out[2], const zd* sigma, const zd* hamk, const double* eig, double rmu, double temp,
                   long Nk, int Nw, int No, int* err, hipStream_t s);
// The same n(mu) pieces through the characteristic polynomials of H + Sigma:
// filling_charpoly writes coef(NkNw, No) once per Sigma, filling_parts_cp
// evaluates Tr (z - H - Sigma)^-1 = p'(z)/p(z) at any mu
void filling_charpoly(zd* coef, const zd* sigma, const zd* hamk, long Nk, int Nw, int No, hipStream_t s);
void filling_parts_cp(double out[2], const zd* coef, const double* eig, double rmu, double temp, long Nk, int Nw,
                      int No, hipStream_t s);
// G0 (fself.f90:1-41)
void green0(zd* G, const double* eig, const zd* uni, double mu, double temp, long Nk, int Nw, int No,
            hipStream_t s);
// in-place inverse of a (tot, No, No) batch (zgetrf + zgetri, fself.f90:131-157)
void batched_inv(zd* A, long tot, int No, int* err, hipStream_t s);
// G^-1 = (i w + mu) I - H - Sigma (fself.f90:43-83)
void green_inv(zd* G, const zd* sigma, const zd* hamk, double mu, double temp, long Nk, int Nw, int No,
               hipStream_t s);
// F = -G Delta G^dagger, Hermitised (feliash.f90:445-484)
void anomalous_f(zd* F, const zd* G, const zd* D, long Nk, int Nw, int No, hipStream_t s);
// superconducting Dyson step (ffeliash.f90:315-384): G <- G0 (I + D F^+) with
// the F given, then F <- -G0 D G^+ (Hermitised), in place on G and F
void sc_green_update(zd* G, zd* F, const zd* G0, const zd* D, long Nk, int Nw, int No, hipStream_t s);
// vertices of the superconducting loop (ffeliash.f90:386-457): vs holds chi0_S
// and vd chi0_C on entry, V_sigma and V_Delta on exit; err set on a singular system
void sc_vertex(zd* vs, zd* vd, long batch, BatchLayout L, int n, const double* S, const double* C, bool singlet,
               int* err, hipStream_t s);
// Hermitise: upper = conj(lower), diagonal real  (feliash.f90:618-634)
void hermitize_lower(zd* D, long Nk, int Nw, int No, hipStream_t s);
// spinful Eliashberg on the full zone, X(Nkall, Nw, No, No); sl = slist,
// inv = 0-based spin-flipped orbital, sg = s_a s_b (device arrays):
// the initial gap before normalisation (feliashwithsoc.f90:665-730)
void initial_delta_soc(zd* D, const double* init, const zd* U, const int64_t* invk, const int* sl, const int* inv,
                       long Nkall, long Nk, int Nw, int No, long gap_sym, hipStream_t s);
// F = -G Delta G(-k)^dagger with its spin structure (feliashwithsoc.f90:426-518)
void anomalous_f_soc(zd* F, const zd* G, const zd* D, const int64_t* invk, const double* sg, const int* sl,
                     const int* inv, long Nkall, long Nk, int Nw, int No, long gap_sym, hipStream_t s);
// spin structure of the new gap (feliashwithsoc.f90:614-662), NW = Nkall Nw
void delta_sym_soc(zd* D, const int* sl, const int* inv, long NW, int No, long gap_sym, hipStream_t s);
// sub_sigma = 1 / 2 static subtraction (fself.f90:384-417)
void subtract_static(zd* sigma, long Nk, int Nw, int No, int mode, hipStream_t s);

// Host-side complex Hessenberg QR eigen-solver (also compiled for the device).
int host_eigvals(std::vector<zd>& A, int n, std::vector<zd>& w);
// Right eigenvector of A (n x n) for eigenvalue lam, LAPACK zgeev normalisation
// (unit 2-norm, largest-modulus component real and positive).
std::vector<zd> host_eigvec(const std::vector<zd>& A, int n, zd lam);
// Real linear solve with partial pivoting (dgesv); returns false if singular.
bool host_dgesv(std::vector<double>& A, int n, std::vector<double>& b);

}  // namespace gmod
