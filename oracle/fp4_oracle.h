/*
 * CPU restatement of FourPhonon's four-phonon hot path and its three-phonon
 * companion -- TEST INFRASTRUCTURE.
 *
 * This library is the checker for the HIP implementation.  Only the test
 * suite, the smoke check and the benchmark's CPU-baseline leg may load it.
 * Nothing in the product path links it or calls it.
 *
 * Every routine states the reference lines it follows.  Arithmetic follows
 * the reference's expression order with floating-point contraction off
 * (build with -ffp-contract=off), and the reference's unsuffixed real
 * literals are reproduced as float32 values promoted to double.
 *
 * Channels: 0 = "++" (plusplus), 1 = "+-" (plusminus), 2 = "--" (minusminus).
 * All arrays are the reference's Fortran column-major arrays, 1-based
 * indices stored in the data; complex arrays are interleaved (re, im).
 */
#ifndef FP4_ORACLE_H
#define FP4_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_case {
  int32_t ngrid[3];
  int32_t nptk;
  int32_t nbands;
  int32_t nlist;
  double rlattvec[9];          /* rlattvec(3,3), nm^-1 */
  double T;
  double scalebroad;
  double omega_max;
  int32_t onlyharmonic;
  const double *masses;        /* masses(nelements) */
  const int32_t *types;        /* types(natoms) */
  const double *energy;        /* energy(nptk,Nb) rad/ps */
  const double *velocity;      /* velocity(nptk,Nb,3) km/s */
  const double *eigenvect;     /* eigenvect(nptk,Nb,Nb) complex */
  const int32_t *IJK;          /* IJK(3,nptk) */
  const int32_t *List;         /* List(Nlist) */
  int32_t ntri;
  const double *Phi;           /* Phi(3,3,3,3,Ntri) eV/A^4 */
  const double *R_j, *R_k, *R_l; /* (3,Ntri) nm */
  const int32_t *Index_i, *Index_j, *Index_k, *Index_l;
  /* third-order clusters (FORCE_CONSTANTS_3RD as read by read3fc) */
  int32_t ntri3;
  const double *Phi3;          /* Phi(3,3,3,Ntri) eV/A^3 */
  const double *R3_j, *R3_k;   /* (3,Ntri) nm */
  const int32_t *Index3_i, *Index3_j, *Index3_k;
} orc_case;

double orc_base_sigma(const orc_case *c, const double v[3]);

/* Per-mode kernels, mm is 1-based as in the reference. */
void orc_np_mode(const orc_case *c, int ch, int mm, int64_t *N, double *P);
int64_t orc_ind_only_mode(const orc_case *c, int ch, int mm, int64_t cap,
                          int64_t *i2, int64_t *i3, int64_t *i4);
/* Streamed order-sensitive hash of the list orc_ind_only_mode would write
   (sum_p key_p M^(n-1-p) mod 2^64), with its count and phase space. */
uint64_t orc_list_hash_mode(const orc_case *c, int ch, int mm, int64_t *N, double *P);
void orc_rta_mode(const orc_case *c, int ch, int mm, int32_t ii_lo, int32_t ii_hi,
                  double *gamma, double *wp4);
int64_t orc_ind_mode(const orc_case *c, int ch, int mm, int64_t cap,
                     int64_t *i2, int64_t *i3, int64_t *i4, double *gamma, double *wp4);

void orc_vp(const orc_case *c, int ch, int i, int j, int k, int l,
            int q, int qp, int qdp, int qtp,
            const double rqp[3], const double rqdp[3], const double rqtp[3], double out[2]);
/* Vp from grid indices only: builds the q-vectors the way the RTA loops do. */
void orc_vp_idx(const orc_case *c, int ch, int i, int j, int k, int l,
                int q, int qp, int qdp, int qtp, double out[2]);

/* Drivers (OpenMP dynamic,1 over modes; nthreads<=0 means the runtime default). */
void orc_np_driver(const orc_case *c, int nthreads, int64_t *N_pp, double *P_pp,
                   int64_t *N_pm, double *P_pm, int64_t *N_mm, double *P_mm);
int orc_ind_only_driver(const orc_case *c, int nthreads, const int64_t *N_pp,
                        const int64_t *N_pm, const int64_t *N_mm,
                        int64_t *i2_pp, int64_t *i3_pp, int64_t *i4_pp,
                        int64_t *i2_pm, int64_t *i3_pm, int64_t *i4_pm,
                        int64_t *i2_mm, int64_t *i3_mm, int64_t *i4_mm);
/* RTA rates for a subset of modes (mm 1-based); outputs are (Nb,Nlist) arrays
   that the caller zeroes.  ii_lo..ii_hi (1-based, inclusive) restricts the
   second-phonon q index; pass 1..nptk for the full reference result. */
void orc_rta_driver(const orc_case *c, int nthreads, int nmodes, const int32_t *modes,
                    int32_t ii_lo, int32_t ii_hi,
                    double *rate_4ph, double *rate_pp, double *rate_pm, double *rate_mm,
                    double *wp4_pp, double *wp4_pm, double *wp4_mm);
void orc_sample_mode(const orc_case *c, int ch, int mm, int32_t nsample, uint64_t seed, int want_rates,
                     int64_t *hits, double *P, double *gamma, double *wp4);
void orc_np_sample_driver(const orc_case *c, int nthreads, int32_t nsample, uint64_t seed,
                          int64_t *N_pp, double *P_pp, int64_t *N_pm, double *P_pm,
                          int64_t *N_mm, double *P_mm);
void orc_rta_sample_driver(const orc_case *c, int nthreads, int32_t nsample, uint64_t seed,
                           int nmodes, const int32_t *modes, double *rate_4ph, double *rate_pp,
                           double *rate_pm, double *rate_mm, double *wp4_pp, double *wp4_pm,
                           double *wp4_mm, int64_t *hits);
int orc_ind_driver(const orc_case *c, int nthreads, const int64_t *N_pp,
                   const int64_t *N_pm, const int64_t *N_mm,
                   int64_t *i2_pp, int64_t *i3_pp, int64_t *i4_pp,
                   int64_t *i2_pm, int64_t *i3_pm, int64_t *i4_pm,
                   int64_t *i2_mm, int64_t *i3_mm, int64_t *i4_mm,
                   double *g_pp, double *g_pm, double *g_mm,
                   double *rate_4ph, double *rate_pp, double *rate_pm, double *rate_mm,
                   double *wp4_pp, double *wp4_pm, double *wp4_mm);

/* iterations.f90 / conductivity.f90 */
void orc_iteration0(int nptk, int nbands, int nlist, int nsymm_rot, const int32_t *Nequi,
                    const int32_t *ALLEquiList, const double *omega, const double *velocity,
                    const double *tau_zero, double *F_n);
void orc_iteration_4ph(int nptk, int nbands, int nlist, int nsymm_rot,
                       const int32_t *Nequi, const int32_t *ALLEquiList,
                       const int32_t *TypeofSymmetry, const int32_t *ID_equi,
                       const int32_t *N_plus, const int32_t *N_minus, const int32_t *i2_plus,
                       const int32_t *i3_plus, const int32_t *i2_minus, const int32_t *i3_minus,
                       const double *g_plus, const double *g_minus,
                       const int64_t *N_pp, const int64_t *N_pm, const int64_t *N_mm,
                       const int64_t *i2_pp, const int64_t *i3_pp, const int64_t *i4_pp,
                       const int64_t *i2_pm, const int64_t *i3_pm, const int64_t *i4_pm,
                       const int64_t *i2_mm, const int64_t *i3_mm, const int64_t *i4_mm,
                       const double *omega, const double *velocity,
                       const double *g_pp, const double *g_pm, const double *g_mm,
                       const double *tau_zero, const double *symmetrizers, double *F_n);
void orc_tconduct(int nptk, int nbands, double T, double V, const double *omega,
                  const double *velocity, const double *F_n, double *kappa);
void orc_symmetrize_F(int nptk, int nbands, const double *symmetrizers, double *F_n);
void orc_symmetrize_tensor(int nsymm_rot, const double *crotations, double *tensor);

/* Three-phonon path.  Channels: 0 = "+" (absorption), 1 = "-" (emission).
   Counts and lists are int32 as in the reference. */
void orc3_np_mode(const orc_case *c, int ch, int mm, int64_t *N, double *P);
int64_t orc3_ind_only_mode(const orc_case *c, int ch, int mm, int64_t cap, int32_t *i2, int32_t *i3);
void orc3_vp(const orc_case *c, int ch, int i, int j, int k, int q, int qp, int qdp, const double rqp[3],
             const double rqdp[3], double out[2]);
void orc3_vp_idx(const orc_case *c, int ch, int i, int j, int k, int q, int qp, int qdp, double out[2]);
void orc3_rta_mode(const orc_case *c, int ch, int mm, int32_t ii_lo, int32_t ii_hi, double *gamma, double *wp3);
int64_t orc3_ind_mode(const orc_case *c, int ch, int mm, int64_t cap, int32_t *i2, int32_t *i3, double *gamma,
                      double *wp3);
void orc3_np_driver(const orc_case *c, int nthreads, int32_t *N_plus, double *P_plus, int32_t *N_minus,
                    double *P_minus);
int orc3_ind_only_driver(const orc_case *c, int nthreads, const int32_t *N_plus, const int32_t *N_minus,
                         int32_t *i2_plus, int32_t *i3_plus, int32_t *i2_minus, int32_t *i3_minus);
void orc3_rta_driver(const orc_case *c, int nthreads, int nmodes, const int32_t *modes, int32_t ii_lo, int32_t ii_hi,
                     double *rate, double *rate_plus, double *rate_minus, double *wp3_plus, double *wp3_minus);
void orc3_sample_mode(const orc_case *c, int ch, int mm, int32_t nsample, uint64_t seed, int want_rates,
                      int64_t *hits, double *P, double *gamma, double *wp3);
int orc3_np_sample_driver(const orc_case *c, int nthreads, int32_t nsample, uint64_t seed, int32_t *N_plus,
                          double *P_plus, int32_t *N_minus, double *P_minus);
void orc3_rta_sample_driver(const orc_case *c, int nthreads, int32_t nsample, uint64_t seed, int nmodes,
                            const int32_t *modes, double *rate, double *rate_plus, double *rate_minus,
                            double *wp3_plus, double *wp3_minus, int64_t *hits);
int orc3_ind_driver(const orc_case *c, int nthreads, const int32_t *N_plus, const int32_t *N_minus, int32_t *i2_plus,
                    int32_t *i3_plus, double *g_plus, int32_t *i2_minus, int32_t *i3_minus, double *g_minus,
                    double *rate, double *rate_plus, double *rate_minus, double *wp3_plus, double *wp3_minus);

#ifdef __cplusplus
}
#endif
#endif
