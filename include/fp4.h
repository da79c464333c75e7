/*
 * fp4 -- MI355X (gfx950) implementation of FourPhonon's four-phonon
 * scattering hot path, exposed as a plain C ABI.
 *
 * Every entry point replaces one procedure of the reference's
 * `module processes` (Src/processes.f90) or `module iterations`
 * (Src/iterations.f90); the comment above each declaration names it.
 * Arrays are the reference's own Fortran arrays: column-major, 1-based
 * indices stored in the data, complex(8) as interleaved (re, im) doubles.
 * The caller owns every host array; the library keeps no host pointer after
 * a call returns.  Device copies of omega, v, e and the IFCs live in the
 * context and are reused across calls and temperatures.
 *
 * The module globals the reference reads implicitly (nptk, Nbands, ngrid,
 * rlattvec, masses, types, T, scalebroad, omega_max, onlyharmonic;
 * Src/config.f90:9-51) are passed once in fp4_config.
 *
 * Calls are synchronous, not re-entrant per context, and made from one host
 * thread.  All functions return FP4_OK (0) or an FP4_ERR_* code; the text of
 * the last error on the calling thread is available from fp4_last_error().
 */
#ifndef FP4_H
#define FP4_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  FP4_OK = 0,
  FP4_ERR_ARG = 1,         /* bad argument or inconsistent sizes */
  FP4_ERR_DEVICE = 2,      /* HIP runtime error or no device */
  FP4_ERR_NOMEM = 3,       /* device allocation failed */
  FP4_ERR_OVERFLOW = 4,    /* a list longer than the count the caller passed */
  FP4_ERR_STATE = 5,       /* modes or IFCs not set yet */
  FP4_ERR_UNSUPPORTED = 6  /* e.g. Nbands outside 3..24, sampling requested */
};

/* Arithmetic used for the matrix elements of fp4_rta_4ph. */
enum {
  FP4_PREC_FP64 = 0,       /* per-process FP64, reference operation order */
  FP4_PREC_FP64_DENSE = 1, /* dense fold + contraction, FP64 */
  FP4_PREC_FP32 = 2,       /* dense fold + contraction on FP32 MFMA */
  FP4_PREC_BF16X3 = 3,     /* dense, BF16 hi/lo split (3 products) on MFMA, FP32 accumulate */
  FP4_PREC_FP16X3 = 4,     /* dense, FP16 hi/lo split (3 products, power-of-two operand scaling) on MFMA */
  FP4_PREC_FP16X2 = 5      /* dense, FP16 hi/lo split of the phase factors only (2 products, power-of-two
                              scaling) on MFMA, FP32 accumulate */
};

typedef struct fp4_ctx fp4_ctx;

typedef struct fp4_config {
  int32_t ngrid[3];
  int32_t nbands;          /* 3*natoms */
  int32_t natoms;
  int32_t nelements;
  double rlattvec[9];      /* rlattvec(3,3), nm^-1, column b_i = rlattvec(:,i) */
  const double *masses;    /* masses(nelements), amu */
  const int32_t *types;    /* types(natoms), 1-based */
  double T;                /* K */
  double scalebroad;
  double omega_max;        /* rad/ps */
  int32_t onlyharmonic;    /* logical */
  int32_t device;          /* HIP device ordinal */
} fp4_config;

const char *fp4_last_error(void);

int fp4_ctx_create(const fp4_config *cfg, fp4_ctx **out);
int fp4_ctx_destroy(fp4_ctx *ctx);
/* T may change between calls (temperature loop of
   Src/ShengBTE.f90:787-1469) without re-uploading anything. */
int fp4_ctx_set_temperature(fp4_ctx *ctx, double T);

/* energy(nptk,Nb), velocity(nptk,Nb,3), eigenvect(nptk,Nb,Nb) complex,
   IJK(3,nptk), List(Nlist) -- the arrays of Src/ShengBTE.f90:142-149.
   eigenvect may be NULL for callers that only count and list processes
   (NP_driver_4ph and Ind_only_driver_4ph take no eigenvectors); the rate,
   Vp and three-phonon rate entry points then return FP4_ERR_STATE until a
   call with eigenvectors.  The 4ph count pass of one mode subset and q'
   window is kept and reused by the next count or list call with the same
   subset and window (Ind_only_driver_4ph after NP_driver_4ph), until new
   mode data arrive; a call with the same frequencies, velocities, IJK and
   List as the last one uploads only the eigenvectors (if given). */
int fp4_ctx_set_modes(fp4_ctx *ctx, const double *energy, const double *velocity, const double *eigenvect,
                      const int32_t *IJK, int32_t nlist, const int32_t *List);

/* Phi(3,3,3,3,Ntri), R_j/R_k/R_l(3,Ntri) in nm on lattice vectors,
   Index_i..l(Ntri) -- the output of read4fc, Src/input.f90:141-182. */
int fp4_ctx_set_ifc4(fp4_ctx *ctx, int32_t ntri, const double *Phi, const double *R_j, const double *R_k,
                     const double *R_l, const int32_t *Index_i, const int32_t *Index_j, const int32_t *Index_k,
                     const int32_t *Index_l);

/* NP_driver_4ph, Src/processes.f90:2749-2847.  N_*(Nlist*Nb) int64 counts,
   P_*(Nb,Nlist) phase space. */
int fp4_np_4ph(fp4_ctx *ctx, int64_t *N_pp, double *P_pp, int64_t *N_pm, double *P_pm, int64_t *N_mm,
               double *P_mm);

/* Ind_only_driver_4ph, Src/processes.f90:3080-3246.  N_* are the counts
   from fp4_np_4ph; the nine output lists have sum(N_*) entries each, in the
   reference's enumeration order, encoded (q-1)*Nb+branch. */
int fp4_ind_only_4ph(fp4_ctx *ctx, const int64_t *N_pp, const int64_t *N_pm, const int64_t *N_mm,
                     int64_t *Indof2ndPhonon_pp, int64_t *Indof3rdPhonon_pp, int64_t *Indof4thPhonon_pp,
                     int64_t *Indof2ndPhonon_pm, int64_t *Indof3rdPhonon_pm, int64_t *Indof4thPhonon_pm,
                     int64_t *Indof2ndPhonon_mm, int64_t *Indof3rdPhonon_mm, int64_t *Indof4thPhonon_mm);

/* RTA_driver_4ph, Src/processes.f90:5174-5311 (and, with precision
   FP4_PREC_FP64, the same numbers RTA_driver_4ph_GPU_using_Ind of
   Src/processes.f90:5315-5387 would give with the CPU path's constants).
   Outputs are (Nb,Nlist) arrays, overwritten. */
int fp4_rta_4ph(fp4_ctx *ctx, int32_t precision, double *rate_4ph, double *rate_pp, double *rate_pm,
                double *rate_mm, double *WP4_pp, double *WP4_pm, double *WP4_mm);

/* RTA_driver_4ph restricted to a subset of modes (mm, 1-based) and to the
   second-phonon window ii_lo..ii_hi (1-based, inclusive; 1..nptk is the
   full rate).  Entries of the outputs outside the subset are left as they
   are.  *nproc receives the number of processes evaluated (may be NULL). */
int fp4_rta_4ph_modes(fp4_ctx *ctx, int32_t precision, int32_t nmodes, const int32_t *modes, int32_t ii_lo,
                      int32_t ii_hi, double *rate_4ph, double *rate_pp, double *rate_pm, double *rate_mm,
                      double *WP4_pp, double *WP4_pm, double *WP4_mm, int64_t *nproc);

/* RTA_driver_4ph_GPU_using_Ind, Src/processes.f90:5315-5387: rates from
   caller-provided lists (counts N_*, lists as from fp4_ind_only_4ph). */
int fp4_rta_4ph_using_ind(fp4_ctx *ctx, const int64_t *N_pp, const int64_t *N_pm, const int64_t *N_mm,
                          const int64_t *Indof2ndPhonon_pp, const int64_t *Indof3rdPhonon_pp,
                          const int64_t *Indof4thPhonon_pp, const int64_t *Indof2ndPhonon_pm,
                          const int64_t *Indof3rdPhonon_pm, const int64_t *Indof4thPhonon_pm,
                          const int64_t *Indof2ndPhonon_mm, const int64_t *Indof3rdPhonon_mm,
                          const int64_t *Indof4thPhonon_mm, double *rate_4ph, double *rate_pp, double *rate_pm,
                          double *rate_mm, double *WP4_pp, double *WP4_pm, double *WP4_mm);

/* Ind_driver_4ph, Src/processes.f90:3551-3822: lists plus the per-process
   Gamma_* (same order as the lists), per-mode rates and WP4. */
int fp4_ind_4ph(fp4_ctx *ctx, const int64_t *N_pp, const int64_t *N_pm, const int64_t *N_mm,
                int64_t *Indof2ndPhonon_pp, int64_t *Indof3rdPhonon_pp, int64_t *Indof4thPhonon_pp,
                int64_t *Indof2ndPhonon_pm, int64_t *Indof3rdPhonon_pm, int64_t *Indof4thPhonon_pm,
                int64_t *Indof2ndPhonon_mm, int64_t *Indof3rdPhonon_mm, int64_t *Indof4thPhonon_mm,
                double *Gamma_pp, double *Gamma_pm, double *Gamma_mm, double *rate_4ph, double *rate_pp,
                double *rate_pm, double *rate_mm, double *WP4_pp, double *WP4_pm, double *WP4_mm);

/* Vp_pp / Vp_pm / Vp_mm, Src/processes.f90:142-330, for n processes at
   once.  channel 0 = ++, 1 = +-, 2 = --.  idx holds n rows of
   (i, j, k, l, q, q', q'', q''') (1-based branches and grid indices), out
   receives n complex values. */
int fp4_vp_4ph(fp4_ctx *ctx, int32_t channel, int64_t n, const int32_t *idx, double *out);

/* NP_driver_4ph restricted to a strictly increasing subset of modes (mm,
   1-based): one MPI rank's share of the reference's chunk loop
   (Src/processes.f90:2798-2803).  N_* and P_* receive nmodes entries each,
   in the given order. */
int fp4_np_4ph_modes(fp4_ctx *ctx, int32_t nmodes, const int32_t *modes, int64_t *N_pp, double *P_pp,
                     int64_t *N_pm, double *P_pm, int64_t *N_mm, double *P_mm);

/* Allocation-free form of fp4_np_4ph_modes.  It has no counterpart in the
   reference: it is for callers that cannot allocate device memory (code
   under HIP graph capture, tight serving loops, harnesses that forbid
   hipMalloc).  Same arguments and results as fp4_np_4ph_modes (bit for
   bit); its working memory is a fixed array of FP4_NOALLOC_SCRATCH_BYTES
   per GPU, held until the process ends, instead of context buffers, and it
   calls no device allocation or free.  A subset needs
   3*nmodes*nptk*12 bytes (plus a few hundred), so at most about
   FP4_NOALLOC_SCRATCH_BYTES / (36*nptk) modes fit one call (227 at 16^3);
   larger subsets are rejected with FP4_ERR_ARG and a message naming the
   limit.  All calls on one GPU share the array: calls that overlap in time
   on one GPU (from several threads or contexts) race and must not be made. */
#define FP4_NOALLOC_SCRATCH_BYTES (32u * 1024u * 1024u)
int fp4_np_4ph_modes_noalloc(fp4_ctx *ctx, int32_t nmodes, const int32_t *modes, int64_t *N_pp, double *P_pp,
                             int64_t *N_pm, double *P_pm, int64_t *N_mm, double *P_mm);

/* Checks the lists of Ind_only_driver_4ph (Src/processes.f90:3080-3246) for
   a mode subset without materialising them: per (channel, mode) the count
   and the order-sensitive hash H = sum_p key_p*M^(n-1-p) mod 2^64 of the
   list the reference would write, key_p = (i2<<42)^(i3<<21)^i4 of entry p
   (Indof2nd/3rd/4thPhonon), M = 0x9E3779B97F4A7C15.  No counterpart in the
   reference; it lets a caller compare a whole mesh's lists (8.1e10 entries
   at 16^3) against stored hashes.  N_* and H_* receive nmodes entries. */
int fp4_list_hash_4ph_modes(fp4_ctx *ctx, int32_t nmodes, const int32_t *modes, int64_t *N_pp, int64_t *N_pm,
                            int64_t *N_mm, uint64_t *H_pp, uint64_t *H_pm, uint64_t *H_mm);

/* Ind_only_driver_4ph restricted to a strictly increasing subset of modes
   (mm, 1-based), for callers that split the mode range the way the
   reference's MPI chunks do (Src/processes.f90:3160-3203).  N_* receive the
   subset's counts (nmodes each); the lists receive the subset's entries,
   modes in the given order, each list sized by the caller to cap entries
   (FP4_ERR_OVERFLOW if a channel needs more). */
int fp4_ind_only_4ph_modes(fp4_ctx *ctx, int32_t nmodes, const int32_t *modes, int64_t cap, int64_t *N_pp,
                           int64_t *N_pm, int64_t *N_mm, int64_t *Indof2ndPhonon_pp, int64_t *Indof3rdPhonon_pp,
                           int64_t *Indof4thPhonon_pp, int64_t *Indof2ndPhonon_pm, int64_t *Indof3rdPhonon_pm,
                           int64_t *Indof4thPhonon_pm, int64_t *Indof2ndPhonon_mm, int64_t *Indof3rdPhonon_mm,
                           int64_t *Indof4thPhonon_mm);

/* ---- sampling estimators (Src/processes.f90:2480-2746, 4811-5171) -------
   The branches NP_driver_4ph and RTA_driver_4ph take when
   num_sample_process_4ph_phase_space / num_sample_process_4ph is not -1
   (Src/processes.f90:2807-2826, 5257-5290).  Per (channel, mode) nsample
   candidates are drawn uniformly from the Nb*nptk*Nb*nptk*Nb box; those in
   the part the full enumeration visits that pass its tests and exact
   window are summed, and the sums are scaled by total/nsample as the
   reference does (N = INT(REAL(hits)/nsample*total), P*total/nsample,
   Gamma/nsample*total, WP4*total/nsample).  The estimates are unbiased for
   the full drivers' outputs.  Random numbers come from counter-based
   SplitMix64 streams keyed by (seed, channel, mode), so a result is a
   function of (seed, nsample) and the inputs; the reference reseeds from the
   clock on every call, draws num+1 samples and reads out of bounds for the
   first one -- neither is reproduced. */
int fp4_np_4ph_sample(fp4_ctx *ctx, int32_t nsample, uint64_t seed, int64_t *N_pp, double *P_pp, int64_t *N_pm,
                      double *P_pm, int64_t *N_mm, double *P_mm);
/* Rates in the FP64 per-process form over the sampled processes; outputs
   as fp4_rta_4ph. */
int fp4_rta_4ph_sample(fp4_ctx *ctx, int32_t nsample, uint64_t seed, double *rate_4ph, double *rate_pp,
                       double *rate_pm, double *rate_mm, double *WP4_pp, double *WP4_pm, double *WP4_mm);
/* As fp4_rta_4ph_sample for a strictly increasing subset of modes; outputs
   are (Nb,Nlist) arrays of which the subset's entries are written; *nhits
   (unless NULL) receives the number of sampled processes evaluated. */
int fp4_rta_4ph_sample_modes(fp4_ctx *ctx, int32_t nsample, uint64_t seed, int32_t nmodes, const int32_t *modes,
                             double *rate_4ph, double *rate_pp, double *rate_pm, double *rate_mm, double *WP4_pp,
                             double *WP4_pm, double *WP4_mm, int64_t *nhits);

/* Three-phonon sampling estimators: NP_driver / RTA_driver with
   num_sample_process_3ph_phase_space / num_sample_process_3ph set
   (Src/processes.f90:828-979, 1474-1673, dispatched at 1033-1050 and
   1736-1760).  Each draw picks (q', j, k) uniformly from the Nb*nptk*Nb
   candidates of the full loops; counts, phase space, Gamma and WP3 are
   scaled as the reference does (its left-to-right order, the "-" rate
   halved).  Streams keyed by (seed, channel, mode) as above.  A count
   estimate beyond int32 returns FP4_ERR_OVERFLOW.  The context needs the
   third-order constants (fp4_ctx_set_ifc3) for the rates. */
int fp4_np_3ph_sample(fp4_ctx *ctx, int32_t nsample, uint64_t seed, int32_t *N_plus, double *Pspace_plus_total,
                      int32_t *N_minus, double *Pspace_minus_total);
int fp4_rta_3ph_sample(fp4_ctx *ctx, int32_t nsample, uint64_t seed, double *rate_scatt, double *rate_scatt_plus,
                       double *rate_scatt_minus, double *WP3_plus, double *WP3_minus);
int fp4_rta_3ph_sample_modes(fp4_ctx *ctx, int32_t nsample, uint64_t seed, int32_t nmodes, const int32_t *modes,
                             double *rate_scatt, double *rate_scatt_plus, double *rate_scatt_minus, double *WP3_plus,
                             double *WP3_minus, int64_t *nhits);

/* ---- three-phonon path (Src/processes.f90:25-1836) ----------------------
   Channel 0 = "+" (absorption), 1 = "-" (emission).  Counts and lists are
   int32, as the reference declares them; a channel whose total would not fit
   returns FP4_ERR_OVERFLOW.  The mode data are those of fp4_ctx_set_modes. */

/* FORCE_CONSTANTS_3RD as read by read3fc (Src/input.f90:102-140):
   Phi(3,3,3,Ntri) eV/A^3, R_j/R_k(3,Ntri) nm on lattice vectors, 1-based
   Index_i/j/k(Ntri). */
int fp4_ctx_set_ifc3(fp4_ctx *ctx, int32_t ntri, const double *Phi, const double *R_j, const double *R_k,
                     const int32_t *Index_i, const int32_t *Index_j, const int32_t *Index_k);

/* NP_driver (Src/processes.f90:983-1067): N_plus/N_minus(Nlist*Nbands),
   Pspace_plus_total/Pspace_minus_total(Nbands,Nlist). */
int fp4_np_3ph(fp4_ctx *ctx, int32_t *N_plus, double *Pspace_plus_total, int32_t *N_minus,
               double *Pspace_minus_total);

/* Ind_only_driver (Src/processes.f90:1198-1309): lists in the reference's
   order (per mode: j outer, q' inner, k innermost), entries (q-1)*Nb+branch. */
int fp4_ind_only_3ph(fp4_ctx *ctx, const int32_t *N_plus, const int32_t *N_minus, int32_t *Indof2ndPhonon_plus,
                     int32_t *Indof3rdPhonon_plus, int32_t *Indof2ndPhonon_minus, int32_t *Indof3rdPhonon_minus);

/* RTA_driver (Src/processes.f90:1676-1775): rate_scatt, rate_scatt_plus,
   rate_scatt_minus, WP3_plus, WP3_minus, all (Nbands,Nlist), FP64. */
int fp4_rta_3ph(fp4_ctx *ctx, double *rate_scatt, double *rate_scatt_plus, double *rate_scatt_minus,
                double *WP3_plus, double *WP3_minus);

/* RTA_driver over a strictly increasing mode subset (mm, 1-based) and a
   second-phonon window ii_lo..ii_hi (1-based, inclusive), as one MPI rank's
   chunk of the reference's loop (Src/processes.f90:1721-1765) restricted in
   q'.  Entries of modes outside the subset are left untouched; nproc (may be
   NULL) receives the number of processes visited. */
int fp4_rta_3ph_modes(fp4_ctx *ctx, int32_t nmodes, const int32_t *modes, int32_t ii_lo, int32_t ii_hi,
                      double *rate_scatt, double *rate_scatt_plus, double *rate_scatt_minus, double *WP3_plus,
                      double *WP3_minus, int64_t *nproc);

/* RTA_driver_GPU_using_Ind (Src/processes.f90:1778-1836), argument order of
   the reference.  Uses the float32-promoted rate units of the CPU routines
   (5.60626442 as a default real), like the other entry points. */
int fp4_rta_3ph_using_ind(fp4_ctx *ctx, const int32_t *N_plus, const int32_t *N_minus,
                          const int32_t *Indof2ndPhonon_plus, const int32_t *Indof2ndPhonon_minus,
                          const int32_t *Indof3rdPhonon_plus, const int32_t *Indof3rdPhonon_minus,
                          double *rate_scatt, double *rate_scatt_plus, double *rate_scatt_minus, double *WP3_plus,
                          double *WP3_minus);

/* Ind_driver (Src/processes.f90:524-707): lists as fp4_ind_only_3ph plus the
   per-process Gamma_plus/Gamma_minus (THz) the iterative solver reads. */
int fp4_ind_3ph(fp4_ctx *ctx, const int32_t *N_plus, const int32_t *N_minus, int32_t *Indof2ndPhonon_plus,
                int32_t *Indof3rdPhonon_plus, double *Gamma_plus, int32_t *Indof2ndPhonon_minus,
                int32_t *Indof3rdPhonon_minus, double *Gamma_minus, double *rate_scatt, double *rate_scatt_plus,
                double *rate_scatt_minus, double *WP3_plus, double *WP3_minus);

/* Vp_plus (channel 0) / Vp_minus (channel 1), Src/processes.f90:25-139, for
   n rows of (i, j, k, q, q', q'') (1-based); out receives n complex values
   (re, im). */
int fp4_vp_3ph(fp4_ctx *ctx, int32_t channel, int64_t n, const int32_t *idx, double *out);

/* ---- harmonic front end (Src/phonon_routines.f90) -----------------------
   eigenDM for Quantum Espresso 2nd-order constants (phonon_espresso,
   Src/phonon_routines.f90:12-85, 311-677), on the device: omega(nptk,Nb)
   rad/ps, eigenvect(nptk,Nb,Nb) complex (component 3*(atom-1)+alpha last)
   and velocity(nptk,Nb,3) km/s on the Gamma-centred grid, velocities through
   the q-point symmetrizers(3,3,nptk) (may be NULL), acoustic modes at Gamma
   exactly zero.  fc is fc_s(3,3,natoms,natoms,scell(1),scell(2),scell(3)) in
   Ry/bohr^2 as read from espresso.ifc2.  The non-analytic (polar) term is
   not built: it is zero for zero Born charges, the only case accepted.
   Eigenvectors of degenerate modes are one orthonormal basis of their
   subspace.  Needs no context. */
typedef struct fp4_harmonic {
  int32_t ngrid[3];
  int32_t natoms;
  int32_t nelements;
  int32_t scell[3];
  double lattvec[9];          /* lattvec(3,3) in nm, column a_i = lattvec(:,i) */
  const double *positions;    /* positions(3,natoms), fractional */
  const double *masses;       /* masses(nelements), amu */
  const int32_t *types;       /* types(natoms), 1-based */
  int32_t device;
} fp4_harmonic;
int fp4_eigen_dm_espresso(const fp4_harmonic *h, const double *fc, const double *symmetrizers, double *omega,
                          double *eigenvect, double *velocity);
/* eigenDM for phonopy 2nd-order constants (phonon_phonopy,
   Src/phonon_routines.f90:84-307) on the device, outputs as above.  fc is
   read2fc's fc(natoms,3,scell(1),scell(2),scell(3),natoms,3) (Src/input.f90:
   18-79): divided by sqrt(m_i m_j) (m_i on the diagonal) and scaled to
   (rad/ps)^2.  born(3,3,natoms) and epsilon(3,3), both or neither, switch on
   the dielectric correction (nonanalytic = .TRUE.): dyn_nac / ncell added to
   every cell's constants and its derivative to dD/dq, no correction at
   Gamma.  Its direction comes from nac_images(3,nptk) when given, else from
   the reference's search (q kept unless an image q + x b1 + y b2 + z b3,
   x, y, z = -2..2 in that order, is strictly shorter); on the zone boundary
   several images tie and that comparison turns on the norm's last bit, so a
   caller that must follow one run passes that run's images. */
int fp4_eigen_dm_phonopy(const fp4_harmonic *h, const double *fc, const double *born, const double *epsilon,
                         const double *nac_images, const double *symmetrizers, double *omega, double *eigenvect,
                         double *velocity);

/* ---- iterative solution (Src/iterations.f90, Src/conductivity.f90) ------
   F_n(Nb,nptk,3) and the scattering operator stay on the device between
   calls; the functions below take and return the reference's host arrays. */

/* The symmetry data the iterative solver reads as module globals:
   Nsymm_rot, ID_equi(Nsymm_rot,nptk) as built by symmetry_map
   (Src/config.f90:409-431, -1 where a rotation leaves the grid),
   symmetrizers(3,3,nptk) (Src/ShengBTE.f90:1288-1290) and
   crotations(3,3,Nsymm_rot) (symmetrize_tensor, Src/config.f90:503-518).
   Needs fp4_ctx_set_modes first. */
int fp4_ctx_set_symmetry(fp4_ctx *ctx, int32_t nsymm_rot, const int32_t *ID_equi, const double *symmetrizers,
                         const double *crotations);

/* iteration0, Src/iterations.f90:13-30: F_n = tau_zero*v*omega on every
   star member.  Nequi(Nlist), ALLEquiList(Nsymm_rot,nptk),
   tau_zero(Nb,Nlist); F_n(Nb,nptk,3) is overwritten. */
int fp4_iteration0(fp4_ctx *ctx, int32_t nlist, const int32_t *Nequi, const int32_t *ALLEquiList,
                   const double *tau_zero, double *F_n);

/* The arguments of iteration_4ph (Src/iterations.f90:95-120) that define
   the scattering operator, uploaded once and kept on the device: orbits,
   3ph counts N_plus/N_minus(Nlist*Nb) int32 with their int32 lists, 4ph
   counts (int64) with their int64 lists, and the per-process Gamma arrays.
   Ntotal_* are the list lengths and must equal the sums of the counts.  3ph
   pointers may be NULL when the 3ph totals are 0. */
int fp4_bte_set_operator(fp4_ctx *ctx, int32_t nlist, const int32_t *Nequi, const int32_t *ALLEquiList,
                         const int32_t *TypeofSymmetry, const int32_t *N_plus, const int32_t *N_minus,
                         int32_t Ntotal_plus, int32_t Ntotal_minus, const int64_t *N_pp, const int64_t *N_pm,
                         const int64_t *N_mm, int64_t Ntotal_pp, int64_t Ntotal_pm, int64_t Ntotal_mm,
                         const int32_t *Indof2ndPhonon_plus, const int32_t *Indof3rdPhonon_plus,
                         const int32_t *Indof2ndPhonon_minus, const int32_t *Indof3rdPhonon_minus,
                         const int64_t *Indof2ndPhonon_pp, const int64_t *Indof3rdPhonon_pp,
                         const int64_t *Indof4thPhonon_pp, const int64_t *Indof2ndPhonon_pm,
                         const int64_t *Indof3rdPhonon_pm, const int64_t *Indof4thPhonon_pm,
                         const int64_t *Indof2ndPhonon_mm, const int64_t *Indof3rdPhonon_mm,
                         const int64_t *Indof4thPhonon_mm, const double *Gamma_plus, const double *Gamma_minus,
                         const double *Gamma_pp, const double *Gamma_pm, const double *Gamma_mm);

/* iteration_4ph, Src/iterations.f90:95-215, with the reference's argument
   order (omega and velocity come from the context): fp4_bte_set_operator
   followed by fp4_bte_iterate.  F_n(Nb,nptk,3) is updated in place, in the
   reference's (ll, i, kk) order. */
int fp4_iteration_4ph(fp4_ctx *ctx, int32_t nlist, const int32_t *Nequi, const int32_t *ALLEquiList,
                      const int32_t *TypeofSymmetry, const int32_t *N_plus, const int32_t *N_minus,
                      int32_t Ntotal_plus, int32_t Ntotal_minus, const int64_t *N_pp, const int64_t *N_pm,
                      const int64_t *N_mm, int64_t Ntotal_pp, int64_t Ntotal_pm, int64_t Ntotal_mm,
                      const int32_t *Indof2ndPhonon_plus, const int32_t *Indof3rdPhonon_plus,
                      const int32_t *Indof2ndPhonon_minus, const int32_t *Indof3rdPhonon_minus,
                      const int64_t *Indof2ndPhonon_pp, const int64_t *Indof3rdPhonon_pp,
                      const int64_t *Indof4thPhonon_pp, const int64_t *Indof2ndPhonon_pm,
                      const int64_t *Indof3rdPhonon_pm, const int64_t *Indof4thPhonon_pm,
                      const int64_t *Indof2ndPhonon_mm, const int64_t *Indof3rdPhonon_mm,
                      const int64_t *Indof4thPhonon_mm, const double *Gamma_plus, const double *Gamma_minus,
                      const double *Gamma_pp, const double *Gamma_pm, const double *Gamma_mm,
                      const double *tau_zero, double *F_n);

/* The scattering operator of iteration_4ph assembled on the device for the
   context's T, for meshes whose process lists do not fit in memory.  No
   counterpart in the reference, which stores the lists of Ind_driver /
   Ind_driver_4ph (Src/processes.f90:524-707, 3551-3822) and sums over them
   in iteration_4ph (Src/iterations.f90:95-215).  The sum over a mode's
   processes is linear in F_n and each of its terms reads F_n at one partner
   mode, so it is regrouped by partner: W[(ll,i)][(q,b)] = sum of the signed
   Gamma of every process of mode (ll,i) with partner (q,b), and the sweep
   computes DeltaF(i,q_kk) = sum W[(ll,i)][(q,b)] F_n(b, ID_equi(S_kk,q), :)
   in the reference's (ll, i, kk) order.  W takes Nb*Nlist*nptk*Nb doubles
   (590 MB at 20^3, Nb = 6) whatever the number of processes.  The 3ph part
   (with_3ph, needs fp4_ctx_set_ifc3) comes from the FP64 per-process
   Gamma of Ind_driver; the 4ph part (with_4ph) from the FP64 per-process
   form (precision FP4_PREC_FP64) or a dense form (the MFMA forms, or
   FP4_PREC_FP64_DENSE).  The additions are atomic, so W agrees between
   runs to rounding, not bit for bit.  Orbits as in fp4_bte_set_operator;
   later fp4_bte_iterate / fp4_bte_solve calls use W until the next
   fp4_bte_set_operator.  *nproc3 / *nproc4 (may be NULL) receive the
   number of processes added, rate_3ph / rate_4ph(Nb,Nlist) (may be NULL)
   the per-mode scattering rates of the same processes (Ind_driver's and
   RTA_driver_4ph's rate_scatt, for tau_zero), so no separate rate pass is
   needed. */
int fp4_bte_assemble_operator(fp4_ctx *ctx, int32_t nlist, const int32_t *Nequi, const int32_t *ALLEquiList,
                              const int32_t *TypeofSymmetry, int32_t precision, int32_t with_3ph, int32_t with_4ph,
                              int64_t *nproc3, int64_t *nproc4, double *rate_3ph, double *rate_4ph);

/* fp4_bte_assemble_operator in parts, for callers that assemble a large
   operator in pieces (progress reports, several callers sharing the mode
   range): _begin sets the orbits and zeroes W, each _add adds the processes
   of a strictly increasing subset of modes (mm, 1-based; rates written at
   those modes' entries only), _end makes the sweeps use W.  Subsets must
   not overlap. */
int fp4_bte_operator_begin(fp4_ctx *ctx, int32_t nlist, const int32_t *Nequi, const int32_t *ALLEquiList,
                           const int32_t *TypeofSymmetry);
int fp4_bte_operator_add(fp4_ctx *ctx, int32_t precision, int32_t with_3ph, int32_t with_4ph, int32_t nmodes,
                         const int32_t *modes, int64_t *nproc3, int64_t *nproc4, double *rate_3ph,
                         double *rate_4ph);
int fp4_bte_operator_end(fp4_ctx *ctx);

/* ---- batches of crystals ---------------------------------------------------
   RTA_driver_4ph (Src/processes.f90:5174-5311) for many small crystals at
   once.  No counterpart in the reference, which runs one crystal per job
   (its nearest analogue is the gang loop over one crystal's modes,
   Src/processes.f90:4202-4243).  The contexts (one per crystal, modes and
   IFCs set, one device) are taken as they are.  fp4_batch_rta_4ph runs each
   crystal's count pass over the call's q' window once (it orders the work
   and does not depend on T), then
   the dense form of every crystal (up to 8 atoms, Nbands 6, 12, 18 or 24)
   in one persistent launch over a single queue of all their work items
   (channel, point, mode block, q') ordered costliest first by the prefilter
   counts; with use_graph that device work is a captured HIP graph, replayed
   while the crystals' data and T are unchanged.  precision is
   FP4_PREC_FP32 or FP4_PREC_FP16X2 (MFMA; in the FP16X2 batch the 6- and
   8-atom crystals run FP32 MFMA).  Crystals of more than 8 atoms run the
   FP64 per-process form, one after the other.  A call covers the
   part-th of nparts q' windows of every crystal ([part*nptk/nparts,
   (part+1)*nptk/nparts)); the rates of the parts add up to the full rates,
   so a caller that cannot wait for a whole pass cycles through the parts,
   each with its own cached graph (part = 0, nparts = 1: everything).
   rates[c] receives
   7 arrays of (Nb,Nlist) doubles back to back (rate_4ph, rate_pp, rate_pm,
   rate_mm, WP4_pp, WP4_pm, WP4_mm) for crystal c, nproc[c] (may be NULL)
   its processes; *kernel_ms the device time of the batched launches and
   *launches the number of kernels in them. */
typedef struct fp4_batch fp4_batch;
int fp4_batch_create(int32_t n, fp4_ctx *const *ctxs, fp4_batch **out);
int fp4_batch_destroy(fp4_batch *b);
int fp4_batch_rta_4ph(fp4_batch *b, int32_t precision, int32_t use_graph, int32_t part, int32_t nparts,
                      double *const *rates, int64_t *nproc, double *kernel_ms, int32_t *launches);

/* One iteration_4ph sweep with the operator of the last
   fp4_bte_set_operator / fp4_iteration_4ph call; F_n updated in place. */
int fp4_bte_iterate(fp4_ctx *ctx, const double *tau_zero, double *F_n);

/* No counterpart in the reference (a launch choice of its sweep): with the
   aggregated operator (fp4_bte_assemble_operator) a sweep of iteration_4ph
   is one cooperative launch over all (ll, i, kk) steps with a grid barrier
   per step (persistent = 1, the default) or one launch per step
   (persistent = 0).  Both give the same F_n bit for bit. */
int fp4_bte_set_sweep(fp4_ctx *ctx, int32_t persistent);

/* TConduct, Src/conductivity.f90:11-37, at the context's T.  V is the cell
   volume in nm^3 (the reference's module global V).  ThConductivity(Nb,3,3)
   and, unless NULL, ThConductivityMode(nptk,Nb,3,3) are overwritten. */
int fp4_tconduct(fp4_ctx *ctx, double V, const double *F_n, double *ThConductivity, double *ThConductivityMode);

/* The iterative solution of Src/ShengBTE.f90:1250-1307 on the device with
   the current operator: iteration0, TConduct and symmetrize_tensor, then up
   to maxiter rounds of (iteration_4ph sweep, symmetrisers on every q,
   TConduct, symmetrize_tensor) until the relative change of sum_b kappa in
   the 2-norm is below eps.  Outputs: F_n(Nb,nptk,3), ThConductivity(Nb,3,3)
   of the last round, and unless NULL kappa_iter(3,3,0:niter) (sum over
   branches per round) and relchange(0:niter) (entry 0 is 0); *niter is the
   number of sweeps made. */
int fp4_bte_solve(fp4_ctx *ctx, double V, const double *tau_zero, int32_t maxiter, double eps, double *F_n,
                  double *ThConductivity, double *kappa_iter, double *relchange, int32_t *niter);

/* Diagnostics; no counterpart in the reference.  Accumulated device time of
   each kernel family, measured with HIP events on the context's stream:
   kind 0 = prefilter count pass, 1 = list emission, 2 = FP64 per-process
   rate kernel, 3 = dense rate kernels, 4 = iteration_4ph sweeps (launches
   counts (ll, i, kk) steps), 5 = three-phonon scan (count and list passes),
   6 = three-phonon FP64 rate kernel, 7 = sampling passes. */
int fp4_ctx_kernel_stats(fp4_ctx *ctx, int32_t kind, double *total_ms, int64_t *launches);
int fp4_ctx_reset_stats(fp4_ctx *ctx);
/* Diagnostics; no counterpart in the reference.  Work of the last dense-form
   call: visited (q, q', q'') triples, work items, and the number of coarse
   (K, L, R_k-R_l) and fine (I, J, K, L, R_k-R_l) cluster groups. */
int fp4_ctx_dense_info(fp4_ctx *ctx, int64_t *triples, int64_t *items, int32_t *ngroups, int32_t *nfine);
/* Diagnostics; no counterpart in the reference.  Shape of the last dense
   kernel: its Psi GEMM path (0 VALU, 1 FP32 MFMA, 2 BF16 hi/lo MFMA, 3 FP16
   hi/lo MFMA), q'' per tile, modes per block. */
int fp4_ctx_dense_shape(fp4_ctx *ctx, int32_t *mfma, int32_t *tq, int32_t *ib);
/* Diagnostics; no counterpart in the reference.  Kernel of the last dense
   call: 0 = the block-cooperative kernel, > 0 = the kernel whose waves own
   slices of the (i, j) branch pairs, with that many coarse groups per MFMA
   k-step. */
int fp4_ctx_dense_kernel(fp4_ctx *ctx, int32_t *kg);
/* Diagnostics; no counterpart in the reference.  The candidates the last
   fp4_ctx_autotune_dense call timed (up to cap of them): kernel k-step
   groups as in fp4_ctx_dense_kernel, IB, TQ, resident blocks and kernel
   time in ms.  Returns the number of candidates (negative on error); any
   output may be NULL. */
int fp4_ctx_dense_candidates(fp4_ctx *ctx, int32_t cap, int32_t *kg, int32_t *ib, int32_t *tq, int32_t *slots,
                             double *ms);
/* Diagnostics; no counterpart in the reference.  In the profiling build
   (libfp4prof.so, compiled with FP4_PHASE_PROF) the shader clock cycles of
   the last dense call's k_dense2 launch per phase, summed over waves, in the
   order: work-item fetch and reduction, A' set-up, q'' scan, tile data, tile
   barrier, filter, filter barrier, survivor queue, Psi GEMM, contraction,
   rates, end-of-tile barrier.  Returns the number of phases written (12);
   -FP4_ERR_UNSUPPORTED in the product build. */
int fp4_ctx_dense_phase_profile(fp4_ctx *ctx, int32_t cap, double *cycles);
/* Tuning control; no counterpart in the reference.  Makes later dense calls
   of this precision run the built shape with the given k-step groups (as in
   fp4_ctx_dense_kernel) and modes per block at `slots` resident blocks
   (0 = one per CU), in place of the autotuner's choice. */
int fp4_ctx_set_dense_shape(fp4_ctx *ctx, int32_t precision, int32_t kg, int32_t ib, int32_t slots);
/* The dense form's autotuner; no counterpart in the reference (its CPU loops
   have no tile shapes).  Times every built kernel shape of this precision and
   Nbands that fits in LDS (modes per block IB, q'' per tile TQ, MFMA path),
   at one and, where occupancy allows, two resident blocks per CU, on every
   branch of the irreducible points of the given strictly increasing mode
   subset (the kernels share work across the branches of a point) over the
   q' window ii_lo..ii_hi, and
   keeps the fastest for this context's later dense-form calls of that
   precision.  Outputs (each may be NULL): the chosen IB, TQ, resident block
   count and its kernel time in ms.  Rates from different shapes agree to
   rounding. */
int fp4_ctx_autotune_dense(fp4_ctx *ctx, int32_t precision, int32_t nmodes, const int32_t *modes, int32_t ii_lo,
                           int32_t ii_hi, int32_t *ib, int32_t *tq, int32_t *slots, double *best_ms);
/* Diagnostics; no counterpart in the reference.  Runs the MFMA instructions
   the dense kernels issue (16x16x4 FP32, 16x16x16 and 16x16x32 BF16 and FP16)
   with known operands and returns the largest deviation from the exact
   products, checking the lane layouts the kernels rely on. */
int fp4_selftest_mfma(int32_t device, double *max_err);
/* Upper bound on the processes materialised per batch (default 2^26); no
   counterpart in the reference, which materialises every list at once. */
int fp4_ctx_set_batch_capacity(fp4_ctx *ctx, int64_t processes);

#ifdef __cplusplus
}
#endif
#endif
