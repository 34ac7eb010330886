/* libgmod.so -- MI355X-native drop-in for the chi0/RPA, FLEX self-energy and
 * linearized-Eliashberg entry points of pyrpa's libfmod.so.
 *
 * Calling convention (identical to the reference's bind(C) Fortran
 * subroutines): every argument by pointer, void return, arrays column-major
 * with the reference's Fortran shapes, caller-allocated outputs.
 *   int64_t*  : integer(c_int64_t)      double* : real(c_double)
 *   bool*     : logical(1) / ctypes c_bool (1 byte)
 *   gmod_z*   : complex(c_double), interleaved (re, im)
 *   int32_t*  : integer(c_int32_t) (irr_chi, chi_map only)
 * Orbital indices inside olist are 1-based like the reference.
 *
 * Citations are to the reference source (ktszk/pyrpa) that each entry point
 * replaces; the ctypes wrapper that binds it in the reference is given in [].
 */
#ifndef GMOD_H
#define GMOD_H
#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double re, im; } gmod_z;

/* ---------------- FLEX self-energy loop ---------------------------------- */
/* libs/src/fself.f90:287-448 (mkself) [libs/flibs/_selfenergy.py:76-127, argtypes :14-37]
 * sigmak(Nk,Nw,Norb,Norb) out; mu out.  sw_in reads sigma.bin from the CWD
 * (fself.f90:742-780) unless an in-memory seed was given with
 * gmod_set_sigma_seed; sw_out writes it. */
void mkself(gmod_z* sigmak, double* mu, const double* Smat, const double* Cmat, const int64_t* kmap,
            const int64_t* invk, const int64_t* olist, const gmod_z* hamk, const double* eig, const gmod_z* uni,
            const double* mu_init, const double* rfill, const double* temp, const int64_t* scf_loop,
            const double* pp, const double* eps, const int64_t* Nkall, const int64_t* Nk, const int64_t* Nw,
            const int64_t* Norb, const int64_t* Nchi, const int64_t* Nx, const int64_t* Ny, const int64_t* Nz,
            const int64_t* sub_sigma, const bool* sw_out, const bool* sw_in, const int64_t* m_diis,
            const bool* sw_rescale, const bool* sw_tail, const double* sigma_scale);

/* ---------------- chi0 ---------------------------------------------------- */
/* libs/src/fchi.f90:401-419 (get_chi0 = get_chi_map + get_chi0_conv + ckchi_impl)
 * [libs/flibs/_susceptibility.py:432-458] */
void get_chi0(gmod_z* chi, const double* Smat, const double* Cmat, const gmod_z* Gk, const int64_t* kmap,
              const int64_t* invk, const int64_t* olist, const double* temp, const int64_t* Nx,
              const int64_t* Ny, const int64_t* Nz, const int64_t* Nw, const int64_t* Nk, const int64_t* Nkall,
              const int64_t* Norb, const int64_t* Nchi, double* maxchi0s_out);
/* libs/src/fchi.f90:894-918 (tail-corrected chi0 = conv[G]-conv[G0]+ref[G0], fchi.f90:852-892)
 * [libs/flibs/_susceptibility.py:460-494] */
void get_chi0_tail(gmod_z* chi, const double* Smat, const double* Cmat, const gmod_z* Gk, const double* eig,
                   const gmod_z* uni, const int64_t* kmap, const int64_t* invk, const int64_t* olist,
                   const double* mu, const double* temp, const int64_t* Nx, const int64_t* Ny, const int64_t* Nz,
                   const int64_t* Nw, const int64_t* Nk, const int64_t* Nkall, const int64_t* Norb,
                   const int64_t* Nchi, double* maxchi0s_out);
/* libs/src/fchi.f90:615-635 (bind name get_chi0_conv_) [libs/flibs/_susceptibility.py:244-269] */
void get_chi0_conv_(gmod_z* chi, const gmod_z* Gk, const int64_t* kmap, const int64_t* invk,
                    const int32_t* irr_chi, const int32_t* chi_map, const int64_t* olist, const double* temp,
                    const int64_t* Nx, const int64_t* Ny, const int64_t* Nz, const int64_t* Nw, const int64_t* Nk,
                    const int64_t* Nkall, const int64_t* Norb, const int64_t* Nchi);

/* ---------------- RPA vertices / susceptibilities ------------------------- */
/* libs/src/fUJ.f90:318-365 (bind name get_vsigma_flex_nosoc_) [libs/flibs/_selfenergy.py:62-74] */
void get_vsigma_flex_nosoc_(gmod_z* chi, const double* Smat, const double* Cmat, const int64_t* Nk,
                            const int64_t* Nw, const int64_t* Nchi);
/* libs/src/fchi.f90:319-350 [libs/flibs/_susceptibility.py:419-430] */
void get_chis(gmod_z* chis, const gmod_z* chi0, const double* Smat, const int64_t* Nchi, const int64_t* Nw);
/* libs/src/fself.f90:783-816 [libs/flibs/_susceptibility.py:568-583] */
void get_chis_chic(gmod_z* chis, gmod_z* chic, const gmod_z* chi, const double* Smat, const double* Cmat,
                   const int64_t* Nk, const int64_t* Nw, const int64_t* Nchi);
/* libs/src/fself.f90:818-850 [libs/flibs/_susceptibility.py:615-629] */
void get_eig_or_tr_chi(gmod_z* chiqout, const gmod_z* chi, const int64_t* invk, const int64_t* Nkall,
                       const int64_t* Nk, const int64_t* Nchi, const bool* sw_eig);

/* Addition (no single reference entry point; the reference offers chi_s at
 * one q for all nu, get_chis, and at nu = 0 for all q, get_chis_chic): chi_s =
 * (I - chi0 S)^-1 chi0 and chi_c = (I + chi0 C)^-1 chi0 at every (q, i nu),
 * the algebra of libs/src/fchi.f90:319-350 and fself.f90:783-816 applied to
 * the whole batch.  chi, chis, chic: (Nk,Nw,Nchi,Nchi) like get_chi0's
 * output; chis or chic may be NULL to skip that susceptibility. */
void gmod_get_chis_chic_all(gmod_z* chis, gmod_z* chic, const gmod_z* chi, const double* Smat, const double* Cmat,
                            const int64_t* Nk, const int64_t* Nw, const int64_t* Nchi);

/* No counterpart in the reference: an allocation-free variant of
 * get_eig_or_tr_chi for callers that cannot allocate device memory (code
 * under graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  chiqout (Nkall), chi (Nk,Nchi,Nchi) and invk (3,Nkall) are
 * device pointers; sw_eig nonzero: leading real eigenvalue per k, else the
 * trace.  The workspace (error flag, per-k maxima, interleaved Hessenberg
 * matrices: about Nk (Nchi^2 + Nchi) 16 bytes) is a fixed 64 MiB device
 * array of the library, one per GPU, held until the process ends; a batch
 * needing more is refused.  Calls that overlap on one GPU race.  Returns 0,
 * or 1 on an error (gmod_last_error). */
int gmod_eig_or_tr_chi_noalloc(gmod_z* chiqout, const gmod_z* chi, const int64_t* invk, int64_t Nkall, int64_t Nk,
                               int64_t Nchi, int sw_eig);

/* ---------------- real-frequency susceptibility, Pade continuation ------- */
/* libs/src/fchi.f90:1-92 (bind name get_qshift_) [libs/flibs/_hamiltonian.py:330-341]:
 * qshift(Nk) = 1-based index of k + q in klist(3,Nk) (periodic in [0,1)). */
void get_qshift_(const double* qpoint, const double* klist, int64_t* qshift, const int64_t* Nk);
/* libs/src/fchi.f90:206-240 (calc_chi fchi.f90:98-157) [libs/flibs/_susceptibility.py:271-293]:
 * chi(Nchi,Nchi,Nw) = chi0(q, wl + i idelta) by the band sum over k and band pairs. */
void get_chi_irr(gmod_z* chi, const gmod_z* uni, const double* eig, const double* ffermi, const int64_t* qshift,
                 const int64_t* ol, const double* wl, const int64_t* Nchi, const int64_t* Norb, const int64_t* Nk,
                 const int64_t* Nw, const double* idelta, const double* eps, const double* temp);
/* libs/src/fchi.f90:242-317 [libs/flibs/_susceptibility.py:295-324]: traces of chi_s and chi0
 * (orbital-diagonal pairs) at w = ecut on the q plane; trchis, trchi (Ny,Nx). */
void chiq_map(gmod_z* trchis, gmod_z* trchi, const gmod_z* uni, const double* eig, const double* ffermi,
              const double* klist, const double* Smat, const int64_t* ol, const double* temp, const double* ecut,
              const double* idelta, const double* eps, const int64_t* Nx, const int64_t* Ny, const int64_t* Nk,
              const int64_t* Norb, const int64_t* Nchi);
/* libs/src/fchi.f90:160-204 [libs/flibs/_susceptibility.py:326-344] */
void get_tr_chi(gmod_z* trchis, gmod_z* trchi0, gmod_z* chis_orb, const gmod_z* chis, const gmod_z* chi0,
                const int64_t* olist, const int64_t* Nw, const int64_t* Nchi, const int64_t* Norb);
/* libs/src/fpade.f90:1-42 (bind name get_a_) [libs/flibs/_green.py:177-188]: continued-fraction
 * coefficients a(Np) of the data inp_data(Np) at the points xn(Np). */
void get_a_(gmod_z* a, const gmod_z* xn, const gmod_z* inp_data, const int64_t* Np);
/* libs/src/fpade.f90:44-85 (bind name get_qp_) [libs/flibs/_green.py:190-204]: numerator P(Nw)
 * and denominator Q(Nw) of the continued fraction at wlist(Nw). */
void get_qp_(gmod_z* P, gmod_z* Q, const gmod_z* a, const gmod_z* xn, const gmod_z* wlist, const int64_t* Nw,
             const int64_t* Np);
/* libs/src/fpade.f90:87-106 (bind name pade_analytic_continuation_arrays_):
 * arrayout(Nk,Nw) = Pade continuation of every row of arrayin(Nk,Niw). */
void pade_analytic_continuation_arrays_(const gmod_z* arrayin, gmod_z* arrayout, const gmod_z* iwlist,
                                        const gmod_z* wlist, const int64_t* Nk, const int64_t* Niw,
                                        const int64_t* Nw);
/* libs/src/fpade.f90:108-129 [libs/flibs/_green.py:206-220]: arrayout(Nw,Nk) = sum over the
 * orbital-diagonal blocks of arrayin(Nk,Niw,Norb,Norb) of their continuations. */
void pade_with_trace(const gmod_z* arrayin, gmod_z* arrayout, const gmod_z* iwlist, const gmod_z* wlist,
                     const int64_t* Nk, const int64_t* Niw, const int64_t* Nw, const int64_t* Norb);

/* ---------------- Wannier-R output --------------------------------------- */
/* libs/plibs/_wannier_io.py:48-99 (_irr_to_full_kgrid, the ifftn of _wannier_all_nonzero),
 * the transform behind output_self_wannier / output_gap_wannier (:145-216):
 * out(Nx,Ny,Nz,Ncut,Norb,Norb) = (1/Nkall) sum_k e^{+2 pi i k.R} data(k) for the first Ncut
 * frequencies of data(Nk,Nw,Norb,Norb).  kmap, invk (3,Nkall) as gen_irr_k_TRS returns them
 * (0-based grid indices; 1-based irreducible index and time-reversal flag).  full != 0: data
 * hold every grid point (Nk = Nkall, placed through kmap alone). */
void gmod_wannier_r(gmod_z* out, const gmod_z* data, const int64_t* kmap, const int64_t* invk, int64_t Nx,
                    int64_t Ny, int64_t Nz, int64_t Nw, int64_t Nk, int64_t Norb, int64_t Ncut, int full);

/* ---------------- Green functions ---------------------------------------- */
/* libs/src/fself.f90:1-41 (bind name gen_green0_) [libs/flibs/_green.py:80-97] */
void gen_green0_(gmod_z* Gk, const double* eig, const gmod_z* uni, const double* mu, const double* temp,
                 const int64_t* Nk, const int64_t* Nw, const int64_t* Norb);
/* libs/src/fself.f90:43-83 (bind name gen_green_inv_) [libs/flibs/_green.py:99-117] */
void gen_green_inv_(gmod_z* Gk, const gmod_z* self, const gmod_z* hamk, const double* mu, const double* temp,
                    const int64_t* Nk, const int64_t* Nw, const int64_t* Norb);
/* libs/src/fself.f90:131-157 (bind name getinv_) [libs/flibs/_green.py:116] */
void getinv_(gmod_z* Gk, const int64_t* Nk, const int64_t* Nw, const int64_t* Norb);

/* ---------------- linearized Eliashberg ---------------------------------- */
/* libs/src/feliash.f90:1-443 [libs/flibs/_eliashberg.py:145-181].  chi is
 * overwritten with the pairing vertex V_Delta, as in the reference. */
void lin_eliash(gmod_z* delta, gmod_z* chi, const gmod_z* Gk, const gmod_z* uni, const double* init_delta,
                const double* Smat, const double* Cmat, const int64_t* olist, const double* prt,
                const int64_t* kmap, const int64_t* invk, const double* temp, const double* eps,
                const int64_t* Nkall, const int64_t* Nk, const int64_t* Nw, const int64_t* Nchi,
                const int64_t* Norb, const int64_t* Nx, const int64_t* Ny, const int64_t* Nz,
                const int64_t* itemax, const int64_t* gap_sym, const int64_t* arnoldi_m, double* lambda_out);
/* libs/src/feliash.f90:445-484 (bind name mkfk_trs_nsoc_) [libs/flibs/_eliashberg.py:343-355] */
void mkfk_trs_nsoc_(gmod_z* fk, const gmod_z* Gk, const gmod_z* delta, const int64_t* Nk, const int64_t* Nw,
                    const int64_t* Norb);
/* libs/src/feliash.f90:637-731 (bind name get_initial_delta_) [libs/flibs/_eliashberg.py:378-399] */
void get_initial_delta_(gmod_z* delta, const double* init_delta, const gmod_z* uni, const int64_t* kmap,
                        const int64_t* invk, const int64_t* Nkall, const int64_t* Nk, const int64_t* Nw,
                        const int64_t* Norb, const int64_t* gap_sym);
/* libs/src/feliash.f90:733-785 [libs/flibs/_eliashberg.py:306-323] */
void conv_delta_orb_to_band(gmod_z* deltab, const gmod_z* delta, const gmod_z* uni, const double* prt,
                            const int64_t* invk, const int64_t* Norb, const int64_t* Nkall, const int64_t* Nk,
                            const int64_t* Nw, const int64_t* gap_sym);
/* libs/src/feliash.f90:787-821 (bind name remap_delta_) [libs/flibs/_eliashberg.py:426-441] */
void remap_delta_(gmod_z* delta, const gmod_z* delta0, const double* prt, const int64_t* invk,
                  const int64_t* Nkall, const int64_t* Nk, const int64_t* Nw, const int64_t* Norb,
                  const int64_t* gap_sym);

/* ---------------- nonlinear FLEX-Eliashberg ------------------------------- */
/* libs/src/ffeliash.f90:2-313 [libs/flibs/_eliashberg.py:230-304].  delta (in:
 * the seed, out: the gap), sigmak (out; zero without sw_sigma) and Gk (in: the
 * normal G the loop starts from, out: the last normal G, as the reference's
 * intent(inout) leaves it) are (Nk,Nw,Norb,Norb).  Precision 0 FP64, 1 and 2
 * FP32 convolutions. */
void eliashberg(gmod_z* delta, gmod_z* sigmak, gmod_z* Gk, const gmod_z* hamk, const double* Smat,
                const double* Cmat, const int64_t* olist, const double* prt, const int64_t* kmap, const int64_t* invk,
                const double* mu, const double* temp, const double* eps, const int64_t* Nkall, const int64_t* Nk,
                const int64_t* Nw, const int64_t* Nchi, const int64_t* Norb, const int64_t* Nx, const int64_t* Ny,
                const int64_t* Nz, const int64_t* itemax, const int64_t* gap_sym, const bool* sw_sigma,
                const bool* sw_Vconst, const int64_t* m_diis, const double* gap_min, const bool* sw_amp_newton);
/* No counterpart in the reference (its loop prints these): of the last
 * eliashberg call, [iterations, status (0 itemax reached, 1 converged,
 * 2 gap collapsed), re-anchors, last max|D_new - D| / max|D|]. */
int gmod_eliashberg_stats(double* out, int n);
/* libs/src/ffeliash.f90:459-489 (get_chi0sc_, internal to eliashberg):
 * Vsigma <- chi^GG + chi^FF, Vdelta <- chi^GG - chi^FF, (Nk,Nw,Nchi,Nchi) */
void get_chi0sc_(gmod_z* Vsigma, gmod_z* Vdelta, const gmod_z* Gk, const gmod_z* fk, const int64_t* kmap,
                 const int64_t* invk, const int32_t* irr_chi, const int32_t* chi_map, const int64_t* olist,
                 const double* prt, const double* temp, const int64_t* Nx, const int64_t* Ny, const int64_t* Nz,
                 const int64_t* Nw, const int64_t* Nk, const int64_t* Nkall, const int64_t* Norb,
                 const int64_t* Nchi, const bool* sw_pair);
/* libs/src/ffeliash.f90:386-457 (mkv_flex_nosoc_, internal to eliashberg): in
 * place, Vsigma chi0_S -> V_sigma and Vdelta chi0_C -> V_Delta */
void mkv_flex_nosoc_(gmod_z* Vdelta, gmod_z* Vsigma, const double* Smat, const double* Cmat, const int64_t* Nk,
                     const int64_t* Nw, const int64_t* Nchi, const bool* sw_pair);

/* ---------------- spinful basis (spin-orbit coupling) -------------------- */
/* Orbitals carry spin: slist(Norb) = +-1, invs(Norb) the 1-based spin-flipped
 * partner, olist over all Norb^2 pairs, Nchi = Norb^2. */
/* libs/src/fselfwithsoc.f90:1-539 [libs/flibs/_selfenergy.py:129-174]; precision 0 FP64,
 * 1 and 2 FP32 transforms of the full fields */
void mkself_soc(gmod_z* sigmak, double* mu, const double* Vmat, const int64_t* kmap, const int64_t* invk,
                const int64_t* invs, const int64_t* olist, const int64_t* slist, const gmod_z* hamk, const double* eig,
                const gmod_z* uni, const double* mu_init, const double* rfill, const double* temp,
                const int64_t* scf_loop, const double* pp, const double* eps, const int64_t* Nkall, const int64_t* Nk,
                const int64_t* Nw, const int64_t* Norb, const int64_t* Nchi, const int64_t* Nx, const int64_t* Ny,
                const int64_t* Nz, const int64_t* sub_sigma, const bool* sw_out, const bool* sw_in,
                const int64_t* m_diis, const bool* sw_rescale, const double* sigma_scale);
/* libs/src/fchi0withsoc.f90:111-214 [libs/flibs/_susceptibility.py:496-528] */
void get_chi0_soc(gmod_z* chi, double* sgnsig, double* sgnsig2, int64_t* invschi, const double* Vmat,
                  const gmod_z* Gk, const int64_t* kmap, const int64_t* invk, const int64_t* invs,
                  const int64_t* olist, const int64_t* slist, const double* temp, const int64_t* Nx,
                  const int64_t* Ny, const int64_t* Nz, const int64_t* Nw, const int64_t* Nk, const int64_t* Nkall,
                  const int64_t* Nchi, const int64_t* Norb);
/* libs/src/fchi0withsoc.f90:281-326 [libs/flibs/_susceptibility.py:530-566]: chic, chiszz,
 * chispm (Nk, Nchi/4, Nchi/4), added to the arrays passed in and then scaled, as the reference does */
void get_chis_chic_soc(gmod_z* chic, gmod_z* chiszz, gmod_z* chispm, const gmod_z* chi, const double* Vmat,
                       const int64_t* orb_list, const int64_t* olist, const int64_t* slist, const int64_t* invs,
                       const int64_t* Nk, const int64_t* Nw, const int64_t* Nchi, const int64_t* Norb);
/* libs/src/fUJ.f90:367-399 (get_V_soc_flex, internal to mkself_soc): in place on
 * chi(Nk,Nw,Nchi,Nchi), -V + V (I + chi0 V)^-1 chi0 V */
void gmod_v_soc_flex(gmod_z* chi, const double* Vmat, const int64_t* Nk, const int64_t* Nw, const int64_t* Nchi);
/* libs/src/feliashwithsoc.f90:1-424 [libs/flibs/_eliashberg.py:183-223]: delta(Nkall,Nw,Norb,Norb)
 * on the full zone; chi leaves holding V_Delta; the eigenvalue through gmod_last_lambda */
void lin_eliash_soc(gmod_z* delta, gmod_z* chi, const gmod_z* Gk, const gmod_z* uni, const double* init_delta,
                    const double* Vmat, const double* sgnsig, const double* sgnsig2, const double* prt,
                    const int64_t* olist, const int64_t* slist, const int64_t* kmap, const int64_t* invk,
                    const int64_t* invs, const int64_t* invschi, const double* temp, const double* eps,
                    const int64_t* Nkall, const int64_t* Nk, const int64_t* Nw, const int64_t* Nchi,
                    const int64_t* Norb, const int64_t* Nx, const int64_t* Ny, const int64_t* Nz,
                    const int64_t* itemax, const int64_t* gap_sym, const int64_t* arnoldi_m);
/* No counterpart in the reference (which prints it): the eigenvalue of the last lin_eliash_soc call */
double gmod_last_lambda(void);
/* libs/src/feliashwithsoc.f90:426-518 (bind name mkfk_trs_soc_) [libs/flibs/_eliashberg.py:357-376] */
void mkfk_trs_soc_(gmod_z* fk, const gmod_z* Gk, const gmod_z* delta, const double* sgnsig, const int64_t* slist,
                   const int64_t* invk, const int64_t* invs, const int64_t* Nkall, const int64_t* Nk,
                   const int64_t* Nw, const int64_t* Norb, const int64_t* gap_sym);
/* libs/src/feliashwithsoc.f90:750-794 [libs/flibs/_eliashberg.py:325-343]: deltab(Nkall,Norb,Norb) */
void conv_delta_orb_to_band_soc(gmod_z* deltab, const gmod_z* delta, const gmod_z* uni, const int64_t* invk,
                                const int64_t* invs, const int64_t* slist, const int64_t* Norb,
                                const int64_t* Nkall, const int64_t* Nk, const int64_t* Nw);
/* libs/src/fUJ.f90:109-180 [libs/flibs/_interaction.py:83-100] */
void get_vmat_soc(double* Vmat, const int64_t* ol, const int64_t* sl, const int64_t* site, const int64_t* invs,
                  const double* Uval, const double* Jval, const int64_t* Nchi, const int64_t* Norb);
/* libs/src/fUJ.f90:182-251 [libs/flibs/_interaction.py:102-119] */
void get_vmat_soc_orb(double* Vmat, const int64_t* ol, const int64_t* sl, const int64_t* site,
                      const int64_t* invs, const double* Umat, const double* Jmat, const int64_t* Nchi,
                      const int64_t* Norb);

/* ---------------- interaction vertices and setup ------------------------- */
/* libs/src/fUJ.f90:1-57 [libs/flibs/_interaction.py:47-62] */
void get_scmat(double* Smat, double* Cmat, const int64_t* ol, const int64_t* site, const double* Uval,
               const double* Jval, const int64_t* Nchi);
/* libs/src/fUJ.f90:59-107 [libs/flibs/_interaction.py:64-81] */
void get_scmat_orb(double* Smat, double* Cmat, const int64_t* ol, const int64_t* site, const double* Umat,
                   const double* Jmat, const int64_t* Nchi, const int64_t* Norb);
/* libs/src/finv.f90:1-290 [libs/flibs/_hamiltonian.py:356-382] */
void generate_irr_kpoint_inv(double* klist, int64_t* kmap, int64_t* invk, const int64_t* Nk, const int64_t* Nx,
                             const int64_t* Ny, const int64_t* Nz);
/* libs/src/finv.f90:292-305 [libs/flibs/_hamiltonian.py:384-395] */
void get_kweight(double* weight, const int64_t* invk, const int64_t* Nk, const int64_t* Nkall);
/* libs/src/fmod.f90:162-207 [libs/flibs/_hamiltonian.py:134-158]; formed on the device */
void gen_ham(gmod_z* ham_k, const double* klist, const gmod_z* ham_r, const double* rvec, const int64_t* Nk,
             const int64_t* Nr, const int64_t* Norb);
/* libs/src/fmod.f90:209-248 [libs/flibs/_hamiltonian.py:160-184]; batched Hermitian
 * Jacobi on the device (eigenvalues ascending as zheev; eigenvector columns up to a
 * phase), Norb <= 64 */
void get_eig(double* eig, gmod_z* uni, const gmod_z* ham_k, const int64_t* Nk, const int64_t* Norb);
/* libs/src/fmod.f90:312-340 [libs/flibs/_hamiltonian.py:186-200] */
void get_ffermi(double* ffermi, const double* eig, const double* mu, const double* temp, const int64_t* Nk,
                const int64_t* Norb);
/* libs/src/fmod.f90:7-20 [libs/flibs/_hamiltonian.py:122-132]: reports the HIP device count */
void openmp_params(int64_t* omp_num, bool* omp_check);

/* ---------------- additions (no reference counterpart) ------------------- */
/* Sticky error channel replacing the reference's print+STOP: returns the
 * error count since the last call and copies the last message. */
int gmod_last_error(char* buf, int len);
/* 0: FP64 everywhere (default).  1: FP32 FFT convolutions with FP64 LU,
 * Dyson, mu search, tail and eigenvalue; the FLEX loop then stores the
 * convolutions as FP64 reference + FP32 increments so that it still reaches
 * esterr < 1e-8.  Also read from the environment variable GMOD_PRECISION. */
void gmod_set_precision(int mode);
int gmod_get_precision(void);
/* In-memory warm start for mkself(sw_in=true): Sigma(Nk,Nw,Norb,Norb), mu, mu_OLD.
 * Consumed by the next mkself call instead of sigma.bin. */
void gmod_set_sigma_seed(const gmod_z* sigma, double mu, double mu_old, int64_t count);
/* Statistics of the last mkself call: [iterations, esterr, stoner, mu,
 * seconds, seconds in chi0, seconds in vertex, seconds in sigma, seconds in
 * mu search + Dyson, FP32-increment iterations, device-resident solve seconds
 * (inputs in HBM to converged Sigma and mu, no PCIe copies), moves of the
 * increments' reference iterate]; returns the number of values written. */
int gmod_flex_stats(double* out, int n);
/* Kernel profiling with hipEvents on the library stream.  gmod_profile(1)
 * resets and enables; gmod_profile_read returns, per kernel class
 * [fft_pass (FP64), rpa_vertex, dyson, scatter, product, gather,
 * fft_pass_fp32]: ms, launches, algorithmic bytes; 3 values per class. */
void gmod_profile(int on);
int gmod_profile_read(double* out, int n);
void gmod_set_verbose(int level);
/* Warm-start re-gridding (no C counterpart in the reference; its Python
 * re-gridder libs/plibs/_calc.py:52-182 covers the frequency part): Sigma
 * (Nk_c,Nw_c,Norb,Norb) of a solve on a coarse mesh onto Sigma
 * (Nk_f,Nw_f,Norb,Norb) of a finer mesh at the same temperature.  k: Fourier
 * interpolation over the full BZ with the TRS unfolding Sigma(-k) =
 * Sigma(k)^T; frequency: shared Matsubara frequencies n < Nw_c - 1 kept (all
 * of them when Nw_c == Nw_f), the tail c0 + c1/(i w) fitted on
 * [0.55, 0.85] Nw_c beyond, diagonal Im Sigma
 * clipped to <= 0.  kmap/invk are the (3, Nkall) tables of each mesh. */
void gmod_regrid_sigma(gmod_z* sig_f, const gmod_z* sig_c, const int64_t* kmap_c, const int64_t* invk_c,
                       int64_t Nk_c, int64_t Nx_c, int64_t Ny_c, int64_t Nz_c, int64_t Nw_c,
                       const int64_t* kmap_f, const int64_t* invk_f, int64_t Nk_f, int64_t Nx_f, int64_t Ny_f,
                       int64_t Nz_f, int64_t Nw_f, int64_t Norb, double temp);
/* Host test hook (no reference counterpart; used by the CPU test suite only):
 * runs the box-pass FFT plan of an (nx,ny,nz,m) grid (x fastest, interleaved
 * complex, in place) on the host with the kernel's own load / sub-pass /
 * store code and bounds checks.  mask selects the axes (bit 0: x .. bit 3:
 * t), sign -1 forward / +1 normalised backward (2: the same backward
 * transform with the passes in the forward order, spatial axes first, as
 * the chi0 inverse transforms run it), precision 0 FP64 / 1 FP32,
 * lds_bytes the per-work-group LDS limit to plan for; npasses receives the
 * number of box passes (data == NULL: plan only).  Returns 0, 1 if the box
 * engine does not cover the shape, 2 on an indexing error. */
int gmod_box_fft_host(double* data, int nx, int ny, int nz, int m, int mask, int sign, int precision,
                      long lds_bytes, int* npasses);
/* Library extension (no counterpart in the reference; a test and benchmark
 * hook of the transform engine every convolution above runs on): `reps`
 * in-place transforms of an (nx,ny,nz,m) grid (x fastest, interleaved
 * complex) in device memory `buf`, with `tmp` of the same size as the
 * intermediate buffer, then a synchronisation.  mask, sign and precision as
 * for gmod_box_fft_host.  Returns 0, or 1 on an error (gmod_last_error). */
int gmod_fft_device(void* buf, void* tmp, int nx, int ny, int nz, int m, int mask, int sign, int precision,
                    int reps);
/* ---------------- learned warm start: the surrogate network --------------
 * No reference counterpart.  A fully connected network (SiLU hidden layers,
 * linear output) whose GEMMs run on the matrix cores with BF16 operands and
 * FP32 accumulation; FP32 master weights and Adam moments.  Handles are
 * small non-negative integers; functions returning a value return -1 on an
 * error (gmod_last_error), the others record it only.
 * Flat parameter / gradient order: W0 (dims[0] x dims[1], row-major), b0,
 * W1, b1, ...  X: (B, dims[0]) row-major FP32; Y, T: (B, dims[nlayers]). */
int64_t gmod_nn_create(int64_t nlayers, const int64_t* dims, uint64_t seed);
void gmod_nn_destroy(int64_t h);
int64_t gmod_nn_num_params(int64_t h);
void gmod_nn_get_params(int64_t h, float* out);
void gmod_nn_set_params(int64_t h, const float* in);
void gmod_nn_forward(int64_t h, const float* X, int64_t B, float* Y);
/* mean-squared-error loss of one batch; grad (may be NULL) receives the flat gradient */
double gmod_nn_grad(int64_t h, const float* X, const float* T, int64_t B, float* grad);
/* training set resident on the device: n rows of X and T */
void gmod_nn_set_data(int64_t h, const float* X, const float* T, int64_t n);
/* one pass over order[0..n_used) (row indices into the resident set) in
 * mini-batches of `batch`, one Adam step per batch; returns the mean batch loss */
double gmod_nn_train_epoch(int64_t h, const int64_t* order, int64_t n_used, int64_t batch, double lr);

/* Device buffers for the device-pointer entry points (gmod_fft_device and
 * the allocation-free variant below), so that a caller needs no second HIP
 * runtime in its process.  gmod_memcpy kind: 1 host to device, 2 device to
 * host, 3 device to device; synchronous; returns 0, or 1 on an error. */
void* gmod_device_alloc(size_t bytes);
void gmod_device_free(void* p);
int gmod_memcpy(void* dst, const void* src, size_t bytes, int kind);

/* Device memory held by the library's caching pool, in bytes. */
double gmod_pool_bytes(void);
void gmod_release_memory(void);

#ifdef __cplusplus
}
#endif
#endif /* GMOD_H */
