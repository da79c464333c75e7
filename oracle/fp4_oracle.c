/*
 * CPU restatement of FourPhonon's four-phonon hot path and its three-phonon
 * companion -- TEST INFRASTRUCTURE.
 *
 * Only the tests, the smoke check and the benchmark's CPU-baseline leg use
 * this file; it is the checker, never the thing measured on the GPU or
 * shipped.  Build with -O2 -ffp-contract=off -fopenmp (oracle/Makefile).
 *
 * Reference lines followed (FourPhonon v1.3, Src/):
 *   constants            data.f90:5-8, processes.f90:20 (hbarp = hbar*1e22,
 *                        the literal 1e22 is a default-real (float32) value)
 *   phexp                misc.f90:70-77
 *   base_sigma           config.f90:481-499
 *   Index_N / IJK        processes.f90:2277-2283, config.f90:458-478
 *   Vp_pp/pm/mm          processes.f90:142-330
 *   NP_plusplus/..       processes.f90:2263-2476, NP_driver_4ph 2749-2847
 *   Ind_only_*           processes.f90:2854-3074, Ind_only_driver_4ph 3080-3246
 *   Ind_* (iterative)    processes.f90:3254-3545, Ind_driver_4ph 3551-3822
 *   RTA_*                processes.f90:3828-4129, RTA_driver_4ph 5174-5311
 *   *_sample             processes.f90:2480-2746, 4811-5171 (sampling
 *                        estimators; the random stream is this file's own)
 *   3ph *_sample         processes.f90:828-979, 1474-1673
 *   iteration0           iterations.f90:13-30
 *   iteration_4ph        iterations.f90:95-215 (3ph and 4ph terms)
 *   TConduct             conductivity.f90:11-37
 *   symmetrisers on F_n  ShengBTE.f90:1288-1290
 *   symmetrize_tensor    config.f90:503-518
 *   3ph Vp_plus/minus    processes.f90:25-139
 *   3ph NP_plus/minus    processes.f90:712-823, NP_driver 983-1067
 *   3ph Ind_only_*       processes.f90:1074-1192, Ind_only_driver 1198-1309
 *   3ph RTA_plus/minus   processes.f90:1313-1471, RTA_driver 1676-1775
 *   3ph Ind_plus/minus   processes.f90:361-521, Ind_driver 524-707
 *
 * Float32-promoted literals of the reference (unsuffixed reals are default
 * REAL): 0.001 (sigma clamp), 3.37617087 (4ph rate units), 5.60626442 (3ph
 * rate units), 1e22 (hbarp), 1e21 (kappa prefactor).  1.25, 0.5, 2. and 8.d0
 * are exact.
 *
 * Pinning.  The reference cannot be compiled in this image, so every
 * routine here is checked by an independent route (brute-force enumeration,
 * closed forms, invariants, dense linear algebra; tests/test_oracle*.py),
 * and the three-phonon drivers also against the reference's own Si example
 * run (counts equal to its totals, phase space and rates to its ten printed
 * digits; tests/test_reference_example.py).  The four-phonon rate assembly
 * (prefactor, unit constant, 1/nptk^2, occupation factors, the call to Vp_*)
 * is rebuilt in numpy from the definition, with an independent numpy V and
 * the unit constants derived from CODATA (tests/test_oracle_rates4.py).
 * Parity unpinned: which candidates exactly at the window edge the
 * reference's compiled code keeps (this file follows its expression order and
 * rounding, but only a run of the reference could confirm ties), and the
 * four-phonon matrix elements and rates against a reference run (the
 * examples ship no fourth-order constants).
 */
#include "fp4_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PI_D 3.141592653589793238
#define KB_D 1.380648813e-23
#define HBAR_D 1.05457172647e-22

static const float F_CLAMP = 0.001f;
static const float F_GAMMA_UNITS = 3.37617087f;
static const float F_1E22 = 1e22f;
static const float F_1E21 = 1e21f;

typedef struct { double re, im; } cplx;

static inline cplx cmul(cplx a, cplx b) {
  cplx r;
  r.re = a.re * b.re - a.im * b.im;
  r.im = a.re * b.im + a.im * b.re;
  return r;
}
static inline cplx rmul(double a, cplx b) {
  cplx r;
  r.re = a * b.re;
  r.im = a * b.im;
  return r;
}
static inline cplx phexp(double x) {
  cplx r;
  r.re = cos(x);
  r.im = sin(x);
  return r;
}
static inline cplx conj_(cplx a) {
  a.im = -a.im;
  return a;
}

#define E(c, q, b) ((c)->energy[(size_t)((q)-1) + (size_t)(c)->nptk * ((b)-1)])
#define VEL(c, q, b, a) \
  ((c)->velocity[(size_t)((q)-1) + (size_t)(c)->nptk * ((size_t)((b)-1) + (size_t)(c)->nbands * ((a)-1))])

static inline cplx EV(const orc_case *c, int q, int b, int comp) {
  size_t idx = (size_t)(q - 1) + (size_t)c->nptk * ((size_t)(b - 1) + (size_t)c->nbands * (size_t)(comp - 1));
  cplx r;
  r.re = c->eigenvect[2 * idx];
  r.im = c->eigenvect[2 * idx + 1];
  return r;
}

static inline int fmodulo(int a, int n) {
  int r = a % n;
  return r < 0 ? r + n : r;
}

/* config.f90:481-499 */
double orc_base_sigma(const orc_case *c, const double v[3]) {
  double s[3];
  for (int nu = 0; nu < 3; nu++) {
    const double *b = c->rlattvec + 3 * nu;
    double d = 0.0;
    d = d + b[0] * v[0];
    d = d + b[1] * v[1];
    d = d + b[2] * v[2];
    double t = d / (double)c->ngrid[nu];
    s[nu] = t * t;
  }
  double m = s[0];
  if (s[1] > m) m = s[1];
  if (s[2] > m) m = s[2];
  return sqrt(m / 2.0);
}

/* matmul(rlattvec, q/dble(ngrid)) */
static void realq_of(const orc_case *c, int qidx, double out[3]) {
  const int32_t *g = c->IJK + 3 * (size_t)(qidx - 1);
  double x[3];
  for (int k = 0; k < 3; k++) x[k] = (double)g[k] / (double)c->ngrid[k];
  for (int i = 0; i < 3; i++) {
    double s = 0.0;
    for (int k = 0; k < 3; k++) s = s + c->rlattvec[i + 3 * k] * x[k];
    out[i] = s;
  }
}

static inline double bose(const orc_case *c, double w) {
  return 1.0 / (exp(HBAR_D * w / KB_D / c->T) - 1.0);
}

/* --------------------------------------------------------------------------
 * Candidate enumeration shared by the count, list, RTA and iterative loops.
 * Order and tests follow processes.f90:2291-2330 (++), 2382-2409 (+-),
 * 2456-2474 (--).
 * ------------------------------------------------------------------------ */
typedef struct {
  int ii, jj, ss, j, k, l;
  double omega, omegap, omegadp, omegatp;
  double delta;  /* the energy mismatch tested against sigma */
  double sigma;  /* after the 0.001 clamp */
} orc_hit;

typedef void (*visit_fn)(const orc_case *c, const orc_hit *h, void *ud);

/* sigma of the pair (jj, k; ss, l) in channel ch: scalebroad*base_sigma of
   -v(jj,k)+v(ss,l) (++) or v(jj,k)-v(ss,l) (+-, --), before the clamp. */
static double pair_sigma4(const orc_case *c, int ch, int jj, int k, int ss, int l) {
  double dv[3];
  for (int a = 1; a <= 3; a++) {
    if (ch == 0)
      dv[a - 1] = -VEL(c, jj, k, a) + VEL(c, ss, l, a);
    else
      dv[a - 1] = VEL(c, jj, k, a) - VEL(c, ss, l, a);
  }
  return c->scalebroad * orc_base_sigma(c, dv);
}

/* sig[(k-1)*nb + (l-1)] = pair_sigma4(c, ch, jj, k, ss, l) for all (k, l):
   the same operations per entry, with the velocities gathered first. */
static void pair_sigma4_table(const orc_case *c, int ch, int jj, int ss, double *sig) {
  const int nb = c->nbands;
  double vj[24][3], vs[24][3];
  for (int b = 1; b <= nb; b++)
    for (int a = 1; a <= 3; a++) {
      vj[b - 1][a - 1] = VEL(c, jj, b, a);
      vs[b - 1][a - 1] = VEL(c, ss, b, a);
    }
  const double *bv = c->rlattvec;
  const double n0 = (double)c->ngrid[0], n1 = (double)c->ngrid[1], n2 = (double)c->ngrid[2];
  const double sb = c->scalebroad;
  for (int k = 0; k < nb; k++)
    for (int l = 0; l < nb; l++) {
      double dv0, dv1, dv2;
      if (ch == 0) {
        dv0 = -vj[k][0] + vs[l][0]; dv1 = -vj[k][1] + vs[l][1]; dv2 = -vj[k][2] + vs[l][2];
      } else {
        dv0 = vj[k][0] - vs[l][0]; dv1 = vj[k][1] - vs[l][1]; dv2 = vj[k][2] - vs[l][2];
      }
      /* orc_base_sigma's operations, in its order */
      double d0 = 0.0, d1 = 0.0, d2 = 0.0;
      d0 = d0 + bv[0] * dv0; d0 = d0 + bv[1] * dv1; d0 = d0 + bv[2] * dv2;
      d1 = d1 + bv[3] * dv0; d1 = d1 + bv[4] * dv1; d1 = d1 + bv[5] * dv2;
      d2 = d2 + bv[6] * dv0; d2 = d2 + bv[7] * dv1; d2 = d2 + bv[8] * dv2;
      const double t0 = d0 / n0, t1 = d1 / n1, t2 = d2 / n2;
      const double s0 = t0 * t0, s1 = t1 * t1, s2 = t2 * t2;
      double m = s0;
      if (s1 > m) m = s1;
      if (s2 > m) m = s2;
      sig[k * nb + l] = sb * sqrt(m / 2.0);
    }
}

/* The tests of one candidate (ii, jj, ss; j, k, l) of mode (i, qi) with
   frequency omega in channel ch: nonzero frequencies, the +- pair
   exclusions and the window |delta| <= sigma, then the 0.001 clamp.  op,
   odp, otp are E(ii,j), E(jj,k), E(ss,l); sigma is pair_sigma4(ch, jj, k,
   ss, l), which does not depend on j, so enumerate() evaluates it once per
   (jj, ss, k, l).  Fills h and returns 1 for a listed process. */
static inline int test4(const orc_case *c, int ch, int i, int qi, double omega, int ii, int jj, int ss, int j, int k,
                        int l, double op, double odp, double otp, double sigma, orc_hit *h) {
  (void)c;
  if (!((op != 0.0) && (odp != 0.0) && (otp != 0.0))) return 0;
  if (ch == 1) {
    if (!((qi != jj || i != k) && (qi != ss || i != l) && (ii != jj || j != k) && (ii != ss || j != l))) return 0;
  }
  double delta;
  if (ch == 0)
    delta = omega + op + odp - otp;
  else if (ch == 1)
    delta = omega + op - odp - otp;
  else
    delta = omega - op - odp - otp;
  if (!(fabs(delta) <= sigma)) return 0;
  if (sigma <= (double)F_CLAMP) sigma = 1.0 / sqrt(PI_D);
  h->omega = omega;
  h->ii = ii; h->jj = jj; h->ss = ss; h->j = j; h->k = k; h->l = l;
  h->omegap = op; h->omegadp = odp; h->omegatp = otp;
  h->delta = delta;
  h->sigma = sigma;
  return 1;
}

/* ss of channel ch for grid points qi, ii, jj (1-based), Index_N of
   processes.f90:2277-2298 */
static int third_partner(const orc_case *c, int ch, int qi, int ii, int jj) {
  const int32_t *q = c->IJK + 3 * (size_t)(qi - 1);
  const int32_t *qp = c->IJK + 3 * (size_t)(ii - 1);
  const int32_t *qdp = c->IJK + 3 * (size_t)(jj - 1);
  const int n1 = c->ngrid[0], n2 = c->ngrid[1], n3 = c->ngrid[2];
  int t0, t1, t2;
  if (ch == 0) {
    t0 = q[0] + qp[0] + qdp[0]; t1 = q[1] + qp[1] + qdp[1]; t2 = q[2] + qp[2] + qdp[2];
  } else if (ch == 1) {
    t0 = q[0] + qp[0] - qdp[0]; t1 = q[1] + qp[1] - qdp[1]; t2 = q[2] + qp[2] - qdp[2];
  } else {
    t0 = q[0] - qp[0] - qdp[0]; t1 = q[1] - qp[1] - qdp[1]; t2 = q[2] - qp[2] - qdp[2];
  }
  t0 = fmodulo(t0, n1); t1 = fmodulo(t1, n2); t2 = fmodulo(t2, n3);
  return (t2 * n2 + t1) * n1 + t0 + 1;
}

static void enumerate(const orc_case *c, int ch, int mm, int ii_lo, int ii_hi, visit_fn f, void *ud) {
  const int nb = c->nbands, nptk = c->nptk;
  const int i = fmodulo(mm - 1, nb) + 1;
  const int ll = (mm - 1) / nb + 1;
  const int qi = c->List[ll - 1];
  const double omega = E(c, qi, i);
  orc_hit h;
  if (!(omega != 0.0)) return;
  if (ii_lo < 1) ii_lo = 1;
  if (ii_hi > nptk) ii_hi = nptk;
  /* per (jj, ss): the sigma table and the three phonons' frequencies */
  double *sig = (double *)malloc(sizeof(double) * ((size_t)nb * nb + 3 * (size_t)nb));
  if (!sig) abort();
  double *Eii = sig + (size_t)nb * nb, *Ejj = Eii + nb, *Ess = Ejj + nb;
  for (int ii = ii_lo; ii <= ii_hi; ii++) {
    for (int b = 1; b <= nb; b++) Eii[b - 1] = E(c, ii, b);
    for (int jj = (ch == 1 ? 1 : ii); jj <= nptk; jj++) {
      const int ss = third_partner(c, ch, qi, ii, jj);
      if (ch == 0 ? !(ss >= 1) : !(ss >= jj)) continue;
      for (int b = 1; b <= nb; b++) {
        Ejj[b - 1] = E(c, jj, b);
        Ess[b - 1] = E(c, ss, b);
      }
      pair_sigma4_table(c, ch, jj, ss, sig);
      for (int j = 1; j <= nb; j++) {
        const int kstart = (ch != 1 && jj == ii) ? j : 1;
        for (int k = kstart; k <= nb; k++) {
          const int lstart = (ch != 0 && ss == jj) ? k : 1;
          for (int l = lstart; l <= nb; l++)
            if (test4(c, ch, i, qi, omega, ii, jj, ss, j, k, l, Eii[j - 1], Ejj[k - 1], Ess[l - 1],
                      sig[(k - 1) * nb + (l - 1)], &h))
              f(c, &h, ud);
        }
      }
    }
  }
  free(sig);
}

/* ---------------------------------------------------------------- counts */
typedef struct { int64_t n; double p; double dn3, dnb4; } np_acc;

/* P += exp(-d**2/sigma**2)/(sigma*sqrt(Pi)*dble(nptk)**3*nbands**4) */
static void visit_np(const orc_case *c, const orc_hit *h, void *ud) {
  np_acc *a = (np_acc *)ud;
  (void)c;
  a->n = a->n + 1;
  const double num = exp(-(h->delta * h->delta) / (h->sigma * h->sigma));
  const double den = h->sigma * sqrt(PI_D) * a->dn3 * a->dnb4;
  a->p = a->p + num / den;
}

void orc_np_mode(const orc_case *c, int ch, int mm, int64_t *N, double *P) {
  np_acc a;
  const double n = (double)c->nptk;
  a.n = 0;
  a.p = 0.0;
  a.dn3 = n * n * n;
  a.dnb4 = (double)(c->nbands * c->nbands * c->nbands * c->nbands);
  enumerate(c, ch, mm, 1, c->nptk, visit_np, &a);
  *N = a.n;
  *P = a.p;
}

/* ----------------------------------------------------------------- lists */
typedef struct { int64_t n, cap; int64_t *i2, *i3, *i4; int nb; } list_acc;

static void visit_list(const orc_case *c, const orc_hit *h, void *ud) {
  list_acc *a = (list_acc *)ud;
  (void)c;
  if (a->n < a->cap) {
    a->i2[a->n] = (int64_t)((h->ii - 1) * a->nb + h->j);
    a->i3[a->n] = (int64_t)((h->jj - 1) * a->nb + h->k);
    a->i4[a->n] = (int64_t)((h->ss - 1) * a->nb + h->l);
  }
  a->n = a->n + 1;
}

int64_t orc_ind_only_mode(const orc_case *c, int ch, int mm, int64_t cap, int64_t *i2, int64_t *i3, int64_t *i4) {
  list_acc a;
  a.n = 0; a.cap = cap; a.i2 = i2; a.i3 = i3; a.i4 = i4; a.nb = c->nbands;
  enumerate(c, ch, mm, 1, c->nptk, visit_list, &a);
  return a.n;
}

/* Order-sensitive hash of one mode's list, streamed (no list is kept):
   h = sum_p key_p * M^(n-1-p) mod 2^64 with key = (i2 << 42) ^ (i3 << 21) ^ i4
   and M = 0x9E3779B97F4A7C15, evaluated by Horner's rule h = h*M + key in
   list order.  Also returns the count and phase space of the same pass. */
typedef struct { uint64_t h; np_acc np; int nb; } hash_acc;

static void visit_hash(const orc_case *c, const orc_hit *h, void *ud) {
  hash_acc *a = (hash_acc *)ud;
  const uint64_t i2 = (uint64_t)((h->ii - 1) * a->nb + h->j);
  const uint64_t i3 = (uint64_t)((h->jj - 1) * a->nb + h->k);
  const uint64_t i4 = (uint64_t)((h->ss - 1) * a->nb + h->l);
  a->h = a->h * 0x9E3779B97F4A7C15ULL + ((i2 << 42) ^ (i3 << 21) ^ i4);
  visit_np(c, h, &a->np);
}

uint64_t orc_list_hash_mode(const orc_case *c, int ch, int mm, int64_t *N, double *P) {
  hash_acc a;
  const double n = (double)c->nptk;
  a.h = 0;
  a.nb = c->nbands;
  a.np.n = 0;
  a.np.p = 0.0;
  a.np.dn3 = n * n * n;
  a.np.dnb4 = (double)(c->nbands * c->nbands * c->nbands * c->nbands);
  enumerate(c, ch, mm, 1, c->nptk, visit_hash, &a);
  *N = a.np.n;
  *P = a.np.p;
  return a.h;
}

/* ------------------------------------------------------ matrix elements */
/* processes.f90:142-330.  Phase signs (phonons 2,3,4): ++ (+,+,-), +- (+,-,-),
   -- (-,-,-); conj on phonon 4 (++), 3 and 4 (+-), 2, 3 and 4 (--). */
void orc_vp(const orc_case *c, int ch, int i, int j, int k, int l, int q, int qp, int qdp, int qtp,
            const double rqp[3], const double rqdp[3], const double rqtp[3], double out[2]) {
  cplx V = {0.0, 0.0};
  const int nt = c->ntri;
  for (int t = 0; t < nt; t++) {
    const int Ii = c->Index_i[t], Ij = c->Index_j[t], Ik = c->Index_k[t], Il = c->Index_l[t];
    double mprod = c->masses[c->types[Ii - 1] - 1] * c->masses[c->types[Ij - 1] - 1];
    mprod = mprod * c->masses[c->types[Ik - 1] - 1];
    mprod = mprod * c->masses[c->types[Il - 1] - 1];
    const double *rj = c->R_j + 3 * (size_t)t, *rk = c->R_k + 3 * (size_t)t, *rl = c->R_l + 3 * (size_t)t;
    double dj = 0.0, dk = 0.0, dl = 0.0;
    for (int a = 0; a < 3; a++) dj = dj + rqp[a] * rj[a];
    for (int a = 0; a < 3; a++) dk = dk + rqdp[a] * rk[a];
    for (int a = 0; a < 3; a++) dl = dl + rqtp[a] * rl[a];
    cplx p2 = phexp(ch == 2 ? -dj : dj);
    cplx p3 = phexp(ch == 0 ? dk : -dk);
    cplx p4 = phexp(-dl);
    cplx pre = rmul(1.0 / sqrt(mprod), p2);
    pre = cmul(pre, p3);
    pre = cmul(pre, p4);
    cplx V0 = {0.0, 0.0};
    const double *phi = c->Phi + 81 * (size_t)t;
    for (int rr = 1; rr <= 3; rr++) {
      for (int ss = 1; ss <= 3; ss++) {
        for (int tt = 1; tt <= 3; tt++) {
          for (int uu = 1; uu <= 3; uu++) {
            const double f = phi[(uu - 1) + 3 * (tt - 1) + 9 * (ss - 1) + 27 * (rr - 1)];
            cplx e1 = EV(c, q, i, uu + 3 * (Ii - 1));
            cplx e2 = EV(c, qp, j, tt + 3 * (Ij - 1));
            cplx e3 = EV(c, qdp, k, ss + 3 * (Ik - 1));
            cplx e4 = EV(c, qtp, l, rr + 3 * (Il - 1));
            if (ch == 2) e2 = conj_(e2);
            if (ch != 0) e3 = conj_(e3);
            e4 = conj_(e4);
            cplx x = rmul(f, e1);
            x = cmul(x, e2);
            x = cmul(x, e3);
            x = cmul(x, e4);
            V0.re = V0.re + x.re;
            V0.im = V0.im + x.im;
          }
        }
      }
    }
    cplx y = cmul(pre, V0);
    V.re = V.re + y.re;
    V.im = V.im + y.im;
  }
  out[0] = V.re;
  out[1] = V.im;
}

void orc_vp_idx(const orc_case *c, int ch, int i, int j, int k, int l, int q, int qp, int qdp, int qtp,
                double out[2]) {
  double a[3], b[3], d[3];
  realq_of(c, qp, a);
  realq_of(c, qdp, b);
  realq_of(c, qtp, d);
  orc_vp(c, ch, i, j, k, l, q, qp, qdp, qtp, a, b, d, out);
}

/* ------------------------------------------------------------------- RTA */
typedef struct {
  int i, qi, ch, nb;
  double gamma, wp4;
  /* iterative variant */
  int store;
  int64_t n, cap;
  int64_t *i2, *i3, *i4;
  double *g;
} rta_acc;

/* WP4 of one process, processes.f90:3895-3901 (++), 3988-3994 (+-), 4095-4101 (--) */
static double wp4_of(const orc_case *c, int ch, const orc_hit *h) {
  const double fp = bose(c, h->omegap), fdp = bose(c, h->omegadp), ftp = bose(c, h->omegatp);
  double occ;
  if (ch == 0)
    occ = fp * fdp * (1 + ftp) - (1 + fp) * (1 + fdp) * ftp;
  else if (ch == 1)
    occ = fp * (1 + fdp) * (1 + ftp) - (1 + fp) * fdp * ftp;
  else
    occ = (1 + fp) * (1 + fdp) * (1 + ftp) - fp * fdp * ftp;
  double w = occ * exp(-(h->delta * h->delta) / (h->sigma * h->sigma)) / h->sigma / sqrt(PI_D) /
             (h->omega * h->omegap * h->omegadp * h->omegatp);
  if (h->omegap <= 1.25 || h->omegadp <= 1.25 || h->omegatp <= 1.25) w = 0;
  return w;
}

static void visit_rta(const orc_case *c, const orc_hit *h, void *ud) {
  rta_acc *a = (rta_acc *)ud;
  const double w = wp4_of(c, a->ch, h);
  a->wp4 = a->wp4 + w;
  const double hbarp = HBAR_D * (double)F_1E22;
  if (a->store) {
    double V[2];
    orc_vp_idx(c, a->ch, a->i, h->j, h->k, h->l, a->qi, h->ii, h->jj, h->ss, V);
    double av = hypot(V[0], V[1]);
    double g = hbarp * hbarp * PI_D / 8.0 * w * (av * av);
    g = g * (double)F_GAMMA_UNITS * 1.e9 / c->nptk / c->nptk;
    if (a->n < a->cap) {
      a->i2[a->n] = (int64_t)((h->ii - 1) * a->nb + h->j);
      a->i3[a->n] = (int64_t)((h->jj - 1) * a->nb + h->k);
      a->i4[a->n] = (int64_t)((h->ss - 1) * a->nb + h->l);
      a->g[a->n] = g;
    }
    a->n = a->n + 1;
  } else if (!c->onlyharmonic) {
    double V[2];
    orc_vp_idx(c, a->ch, a->i, h->j, h->k, h->l, a->qi, h->ii, h->jj, h->ss, V);
    double av = hypot(V[0], V[1]);
    a->gamma = a->gamma + hbarp * hbarp * PI_D / 8.0 * w * (av * av);
  }
}

void orc_rta_mode(const orc_case *c, int ch, int mm, int32_t ii_lo, int32_t ii_hi, double *gamma, double *wp4) {
  rta_acc a;
  memset(&a, 0, sizeof(a));
  a.i = fmodulo(mm - 1, c->nbands) + 1;
  a.qi = c->List[(mm - 1) / c->nbands];
  a.ch = ch;
  a.nb = c->nbands;
  const double omega = E(c, a.qi, a.i);
  enumerate(c, ch, mm, ii_lo, ii_hi, visit_rta, &a);
  if (omega != 0.0) a.wp4 = a.wp4 / c->nptk / c->nptk;
  *gamma = a.gamma * (double)F_GAMMA_UNITS * 1.e9 / c->nptk / c->nptk;
  *wp4 = a.wp4;
}

int64_t orc_ind_mode(const orc_case *c, int ch, int mm, int64_t cap, int64_t *i2, int64_t *i3, int64_t *i4,
                     double *gamma, double *wp4) {
  rta_acc a;
  memset(&a, 0, sizeof(a));
  a.i = fmodulo(mm - 1, c->nbands) + 1;
  a.qi = c->List[(mm - 1) / c->nbands];
  a.ch = ch;
  a.nb = c->nbands;
  a.store = 1;
  a.cap = cap;
  a.i2 = i2; a.i3 = i3; a.i4 = i4; a.g = gamma;
  const double omega = E(c, a.qi, a.i);
  enumerate(c, ch, mm, 1, c->nptk, visit_rta, &a);
  if (omega != 0.0) a.wp4 = a.wp4 / c->nptk / c->nptk;
  *wp4 = a.wp4;
  return a.n;
}

/* --------------------------------------------------------------- drivers */
static void set_threads(int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
}

/* processes.f90:2749-2847 */
void orc_np_driver(const orc_case *c, int nthreads, int64_t *N_pp, double *P_pp, int64_t *N_pm, double *P_pm,
                   int64_t *N_mm, double *P_mm) {
  const int nm = c->nbands * c->nlist;
  set_threads(nthreads);
  for (int m = 0; m < nm; m++) {
    N_pp[m] = N_pm[m] = N_mm[m] = 0;
    P_pp[m] = P_pm[m] = P_mm[m] = 0.0;
  }
#pragma omp parallel for schedule(dynamic, 1)
  for (int mm = 1; mm <= nm; mm++) {
    const int i = fmodulo(mm - 1, c->nbands) + 1;
    const int ll = (mm - 1) / c->nbands + 1;
    if (E(c, c->List[ll - 1], i) <= c->omega_max) {
      orc_np_mode(c, 0, mm, &N_pp[mm - 1], &P_pp[mm - 1]);
      orc_np_mode(c, 1, mm, &N_pm[mm - 1], &P_pm[mm - 1]);
      orc_np_mode(c, 2, mm, &N_mm[mm - 1], &P_mm[mm - 1]);
    }
  }
}

/* processes.f90:3080-3246.  Returns 0, or -1 when a list is longer than its count. */
int orc_ind_only_driver(const orc_case *c, int nthreads, const int64_t *N_pp, const int64_t *N_pm,
                        const int64_t *N_mm, int64_t *i2_pp, int64_t *i3_pp, int64_t *i4_pp, int64_t *i2_pm,
                        int64_t *i3_pm, int64_t *i4_pm, int64_t *i2_mm, int64_t *i3_mm, int64_t *i4_mm) {
  const int nm = c->nbands * c->nlist;
  const int64_t *Ns[3] = {N_pp, N_pm, N_mm};
  int64_t *out[3][3] = {{i2_pp, i3_pp, i4_pp}, {i2_pm, i3_pm, i4_pm}, {i2_mm, i3_mm, i4_mm}};
  int64_t *acc[3];
  int bad = 0;
  for (int ch = 0; ch < 3; ch++) {
    acc[ch] = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nm + 1));
    acc[ch][0] = 0;
    for (int m = 0; m < nm; m++) acc[ch][m + 1] = acc[ch][m] + Ns[ch][m];
    for (int64_t t = 0; t < acc[ch][nm]; t++) out[ch][0][t] = out[ch][1][t] = out[ch][2][t] = 0;
  }
  set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) reduction(| : bad)
  for (int mm = 1; mm <= nm; mm++) {
    const int i = fmodulo(mm - 1, c->nbands) + 1;
    const int ll = (mm - 1) / c->nbands + 1;
    if (!(E(c, c->List[ll - 1], i) <= c->omega_max)) continue;
    for (int ch = 0; ch < 3; ch++) {
      const int64_t n = Ns[ch][mm - 1];
      if (n == 0) continue;
      const int64_t off = acc[ch][mm - 1];
      int64_t got = orc_ind_only_mode(c, ch, mm, n, out[ch][0] + off, out[ch][1] + off, out[ch][2] + off);
      if (got != n) bad |= 1;
    }
  }
  for (int ch = 0; ch < 3; ch++) free(acc[ch]);
  return bad ? -1 : 0;
}

/* processes.f90:5174-5311, restricted to a list of modes and a q' window. */
void orc_rta_driver(const orc_case *c, int nthreads, int nmodes, const int32_t *modes, int32_t ii_lo, int32_t ii_hi,
                    double *rate_4ph, double *rate_pp, double *rate_pm, double *rate_mm, double *wp4_pp,
                    double *wp4_pm, double *wp4_mm) {
  set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
  for (int t = 0; t < nmodes; t++) {
    const int mm = modes[t];
    const int i = fmodulo(mm - 1, c->nbands) + 1;
    const int ll = (mm - 1) / c->nbands + 1;
    if (!(E(c, c->List[ll - 1], i) <= c->omega_max)) continue;
    double g, w;
    orc_rta_mode(c, 0, mm, ii_lo, ii_hi, &g, &w);
    rate_pp[mm - 1] = g; wp4_pp[mm - 1] = w;
    orc_rta_mode(c, 1, mm, ii_lo, ii_hi, &g, &w);
    rate_pm[mm - 1] = g; wp4_pm[mm - 1] = w;
    orc_rta_mode(c, 2, mm, ii_lo, ii_hi, &g, &w);
    rate_mm[mm - 1] = g; wp4_mm[mm - 1] = w;
  }
  for (int t = 0; t < nmodes; t++) {
    const int m = modes[t] - 1;
    rate_4ph[m] = rate_pp[m] + rate_pm[m] + rate_mm[m];
  }
}

/* ---------------------------------------------------- sampling estimators
 * processes.f90:2480-2746 (NP_*_sample) and 4811-5171 (RTA_*_sample),
 * dispatched at 2807-2826 and 5257-5290.  Candidates (ii, j, jj, k, l) are
 * drawn uniformly from the Nb*nptk*Nb*nptk*Nb box and kept when they lie in
 * the part the full enumeration visits (jj >= ii, k >= j on the diagonal,
 * for ++ and --; ss >= jj, l >= k on the diagonal, for +- and --) and pass
 * test4; sums are scaled by total/nsample in the reference's order.
 * Random numbers: SplitMix64 (Steele, Lea & Flood, OOPSLA 2014) evaluated
 * at counter positions: segment key mix64(seed + G*(3*(mm-1)+ch+1)), draw t
 * of sample s mix64(key + G*(5*s+t+1)), an index below n as the high 64 bits
 * of draw*n.  Exactly nsample samples are drawn (the reference draws num+1,
 * the first from out-of-bounds entries of its random array).
 * ------------------------------------------------------------------------ */
#define ORC_GOLDEN 0x9E3779B97F4A7C15ull

static uint64_t orc_mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static int orc_below(uint64_t u, int n) { return (int)(((unsigned __int128)u * (uint64_t)n) >> 64); }

/* Raw sums of one (channel, mode): hits, phase space, Gamma and WP4 (before
   the unit and sampling factors); rates only when want_rates. */
void orc_sample_mode(const orc_case *c, int ch, int mm, int32_t nsample, uint64_t seed, int want_rates,
                     int64_t *hits, double *P, double *gamma, double *wp4) {
  const int nb = c->nbands, nptk = c->nptk;
  const int i = fmodulo(mm - 1, nb) + 1;
  const int qi = c->List[(mm - 1) / nb];
  const double omega = E(c, qi, i);
  np_acc na;
  const double n = (double)nptk;
  na.n = 0;
  na.p = 0.0;
  na.dn3 = n * n * n;
  na.dnb4 = (double)(nb * nb * nb * nb);
  rta_acc ra;
  memset(&ra, 0, sizeof(ra));
  ra.i = i;
  ra.qi = qi;
  ra.ch = ch;
  ra.nb = nb;
  if (omega != 0.0) {
    const uint64_t key = orc_mix64(seed + ORC_GOLDEN * (uint64_t)(3 * (mm - 1) + ch + 1));
    for (int64_t s = 0; s < nsample; s++) {
      const uint64_t base = key + ORC_GOLDEN * (uint64_t)(5 * s + 1);
      const int ii = 1 + orc_below(orc_mix64(base), nptk);
      const int j = 1 + orc_below(orc_mix64(base + ORC_GOLDEN), nb);
      const int jj = 1 + orc_below(orc_mix64(base + 2 * ORC_GOLDEN), nptk);
      const int k = 1 + orc_below(orc_mix64(base + 3 * ORC_GOLDEN), nb);
      const int l = 1 + orc_below(orc_mix64(base + 4 * ORC_GOLDEN), nb);
      const int ss = third_partner(c, ch, qi, ii, jj);
      if (ch != 1 && !(jj >= ii && (jj != ii || k >= j))) continue;
      if (ch != 0 && !(ss >= jj && (ss != jj || l >= k))) continue;
      orc_hit h;
      if (!test4(c, ch, i, qi, omega, ii, jj, ss, j, k, l, E(c, ii, j), E(c, jj, k), E(c, ss, l),
                 pair_sigma4(c, ch, jj, k, ss, l), &h))
        continue;
      visit_np(c, &h, &na);
      if (want_rates) visit_rta(c, &h, &ra);
    }
  }
  *hits = na.n;
  *P = na.p;
  if (gamma) *gamma = ra.gamma;
  if (wp4) *wp4 = ra.wp4;
}

static double orc_sample_total(const orc_case *c) {
  return (double)((int64_t)c->nbands * c->nptk * c->nbands * c->nptk * c->nbands);
}

/* NP_driver_4ph with num_sample_process_4ph_phase_space = nsample */
void orc_np_sample_driver(const orc_case *c, int nthreads, int32_t nsample, uint64_t seed, int64_t *N_pp,
                          double *P_pp, int64_t *N_pm, double *P_pm, int64_t *N_mm, double *P_mm) {
  const int nm = c->nbands * c->nlist;
  int64_t *N[3] = {N_pp, N_pm, N_mm};
  double *P[3] = {P_pp, P_pm, P_mm};
  const double total = orc_sample_total(c), num = (double)nsample;
  set_threads(nthreads);
  for (int m = 0; m < nm; m++)
    for (int ch = 0; ch < 3; ch++) {
      N[ch][m] = 0;
      P[ch][m] = 0.0;
    }
#pragma omp parallel for schedule(dynamic, 1)
  for (int mm = 1; mm <= nm; mm++) {
    const int i = fmodulo(mm - 1, c->nbands) + 1;
    const int ll = (mm - 1) / c->nbands + 1;
    if (!(E(c, c->List[ll - 1], i) <= c->omega_max)) continue;
    for (int ch = 0; ch < 3; ch++) {
      int64_t hits;
      double p;
      orc_sample_mode(c, ch, mm, nsample, seed, 0, &hits, &p, NULL, NULL);
      /* processes.f90:2562-2564 */
      N[ch][mm - 1] = (int64_t)((double)hits / num * total);
      P[ch][mm - 1] = p * total / num;
    }
  }
}

/* RTA_driver_4ph with num_sample_process_4ph = nsample, over a mode list;
   hits (unless NULL) receives the number of sampled processes evaluated. */
void orc_rta_sample_driver(const orc_case *c, int nthreads, int32_t nsample, uint64_t seed, int nmodes,
                           const int32_t *modes, double *rate_4ph, double *rate_pp, double *rate_pm, double *rate_mm,
                           double *wp4_pp, double *wp4_pm, double *wp4_mm, int64_t *hits) {
  double *R[3] = {rate_pp, rate_pm, rate_mm};
  double *W[3] = {wp4_pp, wp4_pm, wp4_mm};
  const double total = orc_sample_total(c), num = (double)nsample;
  int64_t nh = 0;
  set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : nh)
  for (int t = 0; t < nmodes; t++) {
    const int mm = modes[t];
    const int i = fmodulo(mm - 1, c->nbands) + 1;
    const int ll = (mm - 1) / c->nbands + 1;
    const double omega = E(c, c->List[ll - 1], i);
    if (!(omega <= c->omega_max)) continue;
    for (int ch = 0; ch < 3; ch++) {
      int64_t h;
      double p, g, w;
      orc_sample_mode(c, ch, mm, nsample, seed, 1, &h, &p, &g, &w);
      nh += h;
      /* processes.f90:4918-4923 */
      if (omega != 0.0) w = w / c->nptk / c->nptk;
      g = g * (double)F_GAMMA_UNITS * 1.e9 / c->nptk / c->nptk;
      R[ch][mm - 1] = g / num * total;
      W[ch][mm - 1] = w * total / num;
    }
  }
  for (int t = 0; t < nmodes; t++) {
    const int m = modes[t] - 1;
    rate_4ph[m] = rate_pp[m] + rate_pm[m] + rate_mm[m];
  }
  if (hits) *hits = nh;
}

/* processes.f90:3551-3822 (the momentum test of Ind_* selects the same ss as
   the direct Index_N lookup, so the enumeration is shared). */
int orc_ind_driver(const orc_case *c, int nthreads, const int64_t *N_pp, const int64_t *N_pm, const int64_t *N_mm,
                   int64_t *i2_pp, int64_t *i3_pp, int64_t *i4_pp, int64_t *i2_pm, int64_t *i3_pm, int64_t *i4_pm,
                   int64_t *i2_mm, int64_t *i3_mm, int64_t *i4_mm, double *g_pp, double *g_pm, double *g_mm,
                   double *rate_4ph, double *rate_pp, double *rate_pm, double *rate_mm, double *wp4_pp,
                   double *wp4_pm, double *wp4_mm) {
  const int nm = c->nbands * c->nlist;
  const int64_t *Ns[3] = {N_pp, N_pm, N_mm};
  int64_t *out[3][3] = {{i2_pp, i3_pp, i4_pp}, {i2_pm, i3_pm, i4_pm}, {i2_mm, i3_mm, i4_mm}};
  double *gs[3] = {g_pp, g_pm, g_mm};
  double *rates[3] = {rate_pp, rate_pm, rate_mm};
  double *wps[3] = {wp4_pp, wp4_pm, wp4_mm};
  int64_t *acc[3];
  int bad = 0;
  for (int ch = 0; ch < 3; ch++) {
    acc[ch] = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nm + 1));
    acc[ch][0] = 0;
    for (int m = 0; m < nm; m++) acc[ch][m + 1] = acc[ch][m] + Ns[ch][m];
    for (int m = 0; m < nm; m++) rates[ch][m] = wps[ch][m] = 0.0;
  }
  set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) reduction(| : bad)
  for (int mm = 1; mm <= nm; mm++) {
    for (int ch = 0; ch < 3; ch++) {
      const int64_t n = Ns[ch][mm - 1];
      if (n == 0) continue;
      const int64_t off = acc[ch][mm - 1];
      double w;
      int64_t got = orc_ind_mode(c, ch, mm, n, out[ch][0] + off, out[ch][1] + off, out[ch][2] + off, gs[ch] + off, &w);
      if (got != n) bad |= 1;
      wps[ch][mm - 1] = w;
      double s = 0.0;
      for (int64_t t = 0; t < n; t++) s = s + gs[ch][off + t];
      rates[ch][mm - 1] = s;
    }
  }
  for (int m = 0; m < nm; m++) rate_4ph[m] = rate_pp[m] + rate_pm[m] + rate_mm[m];
  for (int ch = 0; ch < 3; ch++) free(acc[ch]);
  return bad ? -1 : 0;
}

/* ------------------------------------------------------------ iterations */
#define FN(F, nb, nptk, b, q, a) (F)[(size_t)((b)-1) + (size_t)(nb) * ((size_t)((q)-1) + (size_t)(nptk) * ((a)-1))]

/* iterations.f90:13-30 */
void orc_iteration0(int nptk, int nbands, int nlist, int nsymm_rot, const int32_t *Nequi, const int32_t *ALLEquiList,
                    const double *omega, const double *velocity, const double *tau_zero, double *F_n) {
  for (int ll = 1; ll <= nlist; ll++)
    for (int ii = 1; ii <= nbands; ii++)
      for (int kk = 1; kk <= Nequi[ll - 1]; kk++) {
        const int q = ALLEquiList[(kk - 1) + (size_t)nsymm_rot * (ll - 1)];
        for (int a = 1; a <= 3; a++) {
          const double v = velocity[(size_t)(q - 1) + (size_t)nptk * ((size_t)(ii - 1) + (size_t)nbands * (a - 1))];
          const double w = omega[(size_t)(q - 1) + (size_t)nptk * (ii - 1)];
          FN(F_n, nbands, nptk, ii, q, a) = tau_zero[(ii - 1) + (size_t)nbands * (ll - 1)] * v * w;
        }
      }
}

/* iterations.f90:95-215.  The 3ph arguments (N_plus, N_minus and their
   int32 lists) may be NULL, which means empty 3ph lists. */
void orc_iteration_4ph(int nptk, int nbands, int nlist, int nsymm_rot, const int32_t *Nequi,
                       const int32_t *ALLEquiList, const int32_t *TypeofSymmetry, const int32_t *ID_equi,
                       const int32_t *N_plus, const int32_t *N_minus, const int32_t *i2_plus,
                       const int32_t *i3_plus, const int32_t *i2_minus, const int32_t *i3_minus,
                       const double *g_plus, const double *g_minus,
                       const int64_t *N_pp, const int64_t *N_pm, const int64_t *N_mm, const int64_t *i2_pp,
                       const int64_t *i3_pp, const int64_t *i4_pp, const int64_t *i2_pm, const int64_t *i3_pm,
                       const int64_t *i4_pm, const int64_t *i2_mm, const int64_t *i3_mm, const int64_t *i4_mm,
                       const double *omega, const double *velocity, const double *g_pp, const double *g_pm,
                       const double *g_mm, const double *tau_zero, const double *symmetrizers, double *F_n) {
  const int nb = nbands;
  double *dF = (double *)calloc((size_t)nb * nptk * 3, sizeof(double));
  int64_t acc[3] = {0, 0, 0};
  int64_t acc3[2] = {0, 0};
  const int32_t *N3[2] = {N_plus, N_minus};
  const int32_t *I2_3[2] = {i2_plus, i2_minus}, *I3_3[2] = {i3_plus, i3_minus};
  const double *G3[2] = {g_plus, g_minus};
  const int64_t *Ns[3] = {N_pp, N_pm, N_mm};
  const int64_t *I2[3] = {i2_pp, i2_pm, i2_mm};
  const int64_t *I3[3] = {i3_pp, i3_pm, i3_mm};
  const int64_t *I4[3] = {i4_pp, i4_pm, i4_mm};
  const double *G[3] = {g_pp, g_pm, g_mm};
#define IDEQ(s, q) ID_equi[(size_t)((s)-1) + (size_t)nsymm_rot * ((q)-1)]
  for (int ll = 1; ll <= nlist; ll++) {
    for (int i = 1; i <= nb; i++) {
      const int mm = (ll - 1) * nb + i;
      if (mm != 1) {
        for (int ch = 0; ch < 3; ch++) acc[ch] = acc[ch] + Ns[ch][mm - 2];
        for (int t = 0; t < 2; t++)
          if (N3[t]) acc3[t] = acc3[t] + N3[t][mm - 2];
      }
      for (int kk = 1; kk <= Nequi[ll - 1]; kk++) {
        const int qk = ALLEquiList[(kk - 1) + (size_t)nsymm_rot * (ll - 1)];
        const int sym = TypeofSymmetry[(kk - 1) + (size_t)nsymm_rot * (ll - 1)];
        /* 3ph terms, iterations.f90:146-167 */
        for (int t = 0; t < 2; t++) {
          const int64_t n = N3[t] ? N3[t][mm - 1] : 0;
          for (int64_t x = 0; x < n; x++) {
            const int64_t p = acc3[t] + x;
            const int j = (I2_3[t][p] - 1) % nb + 1, qm = (I2_3[t][p] - 1) / nb + 1;
            const int k = (I3_3[t][p] - 1) % nb + 1, qn = (I3_3[t][p] - 1) / nb + 1;
            const int m2 = IDEQ(sym, qm), n2 = IDEQ(sym, qn);
            for (int a = 1; a <= 3; a++) {
              const double Fj = FN(F_n, nb, nptk, j, m2, a);
              const double Fk = FN(F_n, nb, nptk, k, n2, a);
              const double term = (t == 0) ? G3[t][p] * (Fk - Fj) : G3[t][p] * (Fk + Fj) * 5.e-1;
              FN(dF, nb, nptk, i, qk, a) = FN(dF, nb, nptk, i, qk, a) + term;
            }
          }
        }
        for (int ch = 0; ch < 3; ch++) {
          const int64_t n = Ns[ch][mm - 1];
          for (int64_t t = 0; t < n; t++) {
            const int64_t p = acc[ch] + t;
            const int j = (int)((I2[ch][p] - 1) % nb) + 1, qm = (int)((I2[ch][p] - 1) / nb) + 1;
            const int k = (int)((I3[ch][p] - 1) % nb) + 1, qn = (int)((I3[ch][p] - 1) / nb) + 1;
            const int l = (int)((I4[ch][p] - 1) % nb) + 1, qs = (int)((I4[ch][p] - 1) / nb) + 1;
            const int m2 = IDEQ(sym, qm), n2 = IDEQ(sym, qn), s2 = IDEQ(sym, qs);
            for (int a = 1; a <= 3; a++) {
              const double Fj = FN(F_n, nb, nptk, j, m2, a);
              const double Fk = FN(F_n, nb, nptk, k, n2, a);
              const double Fl = FN(F_n, nb, nptk, l, s2, a);
              double term;
              if (ch == 0)
                term = Fl - Fj - Fk;
              else if (ch == 1)
                term = Fl - Fj + Fk;
              else
                term = Fl + Fj + Fk;
              FN(dF, nb, nptk, i, qk, a) = FN(dF, nb, nptk, i, qk, a) + G[ch][p] * term;
            }
          }
        }
        const double tz = tau_zero[(i - 1) + (size_t)nb * (ll - 1)];
        for (int a = 1; a <= 3; a++) {
          const double v = velocity[(size_t)(qk - 1) + (size_t)nptk * ((size_t)(i - 1) + (size_t)nb * (a - 1))];
          const double w = omega[(size_t)(qk - 1) + (size_t)nptk * (i - 1)];
          FN(F_n, nb, nptk, i, qk, a) = tz * v * w + tz * FN(dF, nb, nptk, i, qk, a);
        }
        const double *S = symmetrizers + 9 * (size_t)(qk - 1);
        for (int b = 1; b <= nb; b++) {
          double x[3], y[3];
          for (int a = 0; a < 3; a++) x[a] = FN(F_n, nb, nptk, b, qk, a + 1);
          for (int a = 0; a < 3; a++) {
            double s = 0.0;
            for (int c2 = 0; c2 < 3; c2++) s = s + S[a + 3 * c2] * x[c2];
            y[a] = s;
          }
          for (int a = 0; a < 3; a++) FN(F_n, nb, nptk, b, qk, a + 1) = y[a];
        }
      }
    }
  }
#undef IDEQ
  free(dF);
}

/* conductivity.f90:11-37; kappa is (Nbands,3,3) */
void orc_tconduct(int nptk, int nbands, double T, double V, const double *omega, const double *velocity,
                  const double *F_n, double *kappa) {
  for (int t = 0; t < nbands * 9; t++) kappa[t] = 0.0;
  for (int jj = 1; jj <= nbands; jj++) {
    for (int ii = 2; ii <= nptk; ii++) {
      double tmp[3][3];
      for (int d1 = 1; d1 <= 3; d1++)
        for (int d2 = 1; d2 <= 3; d2++)
          tmp[d1 - 1][d2 - 1] = velocity[(size_t)(ii - 1) + (size_t)nptk * ((size_t)(jj - 1) + (size_t)nbands * (d1 - 1))] *
                                FN(F_n, nbands, nptk, jj, ii, d2);
      const double w = omega[(size_t)(ii - 1) + (size_t)nptk * (jj - 1)];
      const double f = 1.0 / (exp(HBAR_D * w / KB_D / T) - 1.0);
      for (int d1 = 0; d1 < 3; d1++)
        for (int d2 = 0; d2 < 3; d2++) {
          const double mode = f * (f + 1) * w * tmp[d1][d2];
          kappa[(jj - 1) + nbands * (d1 + 3 * d2)] = kappa[(jj - 1) + nbands * (d1 + 3 * d2)] + mode;
        }
    }
  }
  const double pre = (double)F_1E21 * (HBAR_D * HBAR_D);
  const double den = KB_D * T * T * V * nptk;
  for (int t = 0; t < nbands * 9; t++) kappa[t] = pre * kappa[t] / den;
}

/* F_n(:,q,:) = transpose(matmul(symmetrizers(:,:,q), transpose(F_n(:,q,:)))) for
   every q, ShengBTE.f90:1288-1290 */
void orc_symmetrize_F(int nptk, int nbands, const double *symmetrizers, double *F_n) {
  for (int q = 1; q <= nptk; q++) {
    const double *S = symmetrizers + 9 * (size_t)(q - 1);
    for (int b = 1; b <= nbands; b++) {
      double x[3], y[3];
      for (int a = 0; a < 3; a++) x[a] = FN(F_n, nbands, nptk, b, q, a + 1);
      for (int a = 0; a < 3; a++) {
        double s = 0.0;
        for (int c2 = 0; c2 < 3; c2++) s = s + S[a + 3 * c2] * x[c2];
        y[a] = s;
      }
      for (int a = 0; a < 3; a++) FN(F_n, nbands, nptk, b, q, a + 1) = y[a];
    }
  }
}

/* symmetrize_tensor, config.f90:503-518: tensor(3,3) <- sum_isym C T C^T / Nsymm_rot */
void orc_symmetrize_tensor(int nsymm_rot, const double *crotations, double *tensor) {
  double tmp[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int is = 0; is < nsymm_rot; is++) {
    const double *Cr = crotations + 9 * (size_t)is;
    double M1[9];
    for (int a = 0; a < 3; a++)
      for (int c2 = 0; c2 < 3; c2++) {
        double s = 0.0;
        for (int k = 0; k < 3; k++) s = s + tensor[a + 3 * k] * Cr[c2 + 3 * k];
        M1[a + 3 * c2] = s;
      }
    for (int a = 0; a < 3; a++)
      for (int c2 = 0; c2 < 3; c2++) {
        double s = 0.0;
        for (int k = 0; k < 3; k++) s = s + Cr[a + 3 * k] * M1[k + 3 * c2];
        tmp[a + 3 * c2] = tmp[a + 3 * c2] + s;
      }
  }
  for (int t = 0; t < 9; t++) tensor[t] = tmp[t] / nsymm_rot;
}

/* ==========================================================================
 * Three-phonon path (processes.f90:25-139, 361-1775).  Channel 0 = "+"
 * (absorption, q'' = q + q'), 1 = "-" (emission, q'' = q - q').  The window
 * is |delta| <= 2 sigma with sigma from v(q',j) - v(q'',k) and no clamp;
 * NP_* and RTA_* loop ii outer, j inner (:740-741, :1352-1353), Ind_only_*
 * and Ind_* loop j outer, ii inner (:1105-1106, :398-399).
 * ======================================================================== */
typedef struct {
  int ii, ss, j, k;
  double omega, omegap, omegadp, delta, sigma;
} orc3_hit;

typedef void (*visit3_fn)(const orc_case *c, const orc3_hit *h, void *ud);

/* branches k_lo..k_hi of the third phonon for (ii, j) */
static void test3(const orc_case *c, int ch, const int32_t *q, double omega, int ii, int j, int k_lo, int k_hi,
                  visit3_fn f, void *ud) {
  const int n1 = c->ngrid[0], n2 = c->ngrid[1], n3 = c->ngrid[2];
  const int32_t *qp = c->IJK + 3 * (size_t)(ii - 1);
  const double omegap = E(c, ii, j);
  for (int k = k_lo; k <= k_hi; k++) {
    int t0, t1, t2;
    if (ch == 0) {
      t0 = q[0] + qp[0]; t1 = q[1] + qp[1]; t2 = q[2] + qp[2];
    } else {
      t0 = q[0] - qp[0]; t1 = q[1] - qp[1]; t2 = q[2] - qp[2];
    }
    t0 = fmodulo(t0, n1); t1 = fmodulo(t1, n2); t2 = fmodulo(t2, n3);
    const int ss = (t2 * n2 + t1) * n1 + t0 + 1;
    const double omegadp = E(c, ss, k);
    if (!((omegap != 0.0) && (omegadp != 0.0))) continue;
    double dv[3];
    for (int a = 1; a <= 3; a++) dv[a - 1] = VEL(c, ii, j, a) - VEL(c, ss, k, a);
    const double sigma = c->scalebroad * orc_base_sigma(c, dv);
    const double delta = (ch == 0) ? omega + omegap - omegadp : omega - omegap - omegadp;
    if (fabs(delta) <= 2.0 * sigma) {
      orc3_hit h;
      h.ii = ii; h.ss = ss; h.j = j; h.k = k;
      h.omega = omega; h.omegap = omegap; h.omegadp = omegadp;
      h.delta = delta; h.sigma = sigma;
      f(c, &h, ud);
    }
  }
}

/* order 0: ii outer, j inner; order 1: j outer, ii inner.  ii_lo..ii_hi
   restricts the second phonon's q index. */
static void enumerate3(const orc_case *c, int ch, int mm, int order, int ii_lo, int ii_hi, visit3_fn f, void *ud) {
  const int nb = c->nbands, nptk = c->nptk;
  const int i = fmodulo(mm - 1, nb) + 1;
  const int qi = c->List[(mm - 1) / nb];
  const int32_t *q = c->IJK + 3 * (size_t)(qi - 1);
  const double omega = E(c, qi, i);
  if (!(omega != 0.0)) return;
  if (ii_lo < 1) ii_lo = 1;
  if (ii_hi > nptk) ii_hi = nptk;
  if (order == 0) {
    for (int ii = ii_lo; ii <= ii_hi; ii++)
      for (int j = 1; j <= nb; j++) test3(c, ch, q, omega, ii, j, 1, nb, f, ud);
  } else {
    for (int j = 1; j <= nb; j++)
      for (int ii = ii_lo; ii <= ii_hi; ii++) test3(c, ch, q, omega, ii, j, 1, nb, f, ud);
  }
}

typedef struct { int64_t n; double p; double den; } np3_acc;

/* P += exp(-d**2/sigma**2)/(sigma*sqrt(Pi)*nptk**2*nbands**3), processes.f90:756-758
   (nptk**2 and nbands**3 are integer products, converted once each) */
static void visit_np3(const orc_case *c, const orc3_hit *h, void *ud) {
  np3_acc *a = (np3_acc *)ud;
  (void)c;
  a->n = a->n + 1;
  const double num = exp(-(h->delta * h->delta) / (h->sigma * h->sigma));
  const double nb3 = (double)(c->nbands * c->nbands * c->nbands);
  const double den = h->sigma * sqrt(PI_D) * a->den * nb3;
  a->p = a->p + num / den;
}

void orc3_np_mode(const orc_case *c, int ch, int mm, int64_t *N, double *P) {
  np3_acc a;
  a.n = 0;
  a.p = 0.0;
  a.den = (double)((int64_t)c->nptk * c->nptk);
  enumerate3(c, ch, mm, 0, 1, c->nptk, visit_np3, &a);
  *N = a.n;
  *P = a.p;
}

typedef struct { int64_t n, cap; int32_t *i2, *i3; int nb; } list3_acc;

static void visit_list3(const orc_case *c, const orc3_hit *h, void *ud) {
  list3_acc *a = (list3_acc *)ud;
  (void)c;
  if (a->n < a->cap) {
    a->i2[a->n] = (int32_t)((h->ii - 1) * a->nb + h->j);
    a->i3[a->n] = (int32_t)((h->ss - 1) * a->nb + h->k);
  }
  a->n = a->n + 1;
}

int64_t orc3_ind_only_mode(const orc_case *c, int ch, int mm, int64_t cap, int32_t *i2, int32_t *i3) {
  list3_acc a;
  a.n = 0; a.cap = cap; a.i2 = i2; a.i3 = i3; a.nb = c->nbands;
  enumerate3(c, ch, mm, 1, 1, c->nptk, visit_list3, &a);
  return a.n;
}

/* Vp_plus / Vp_minus, processes.f90:25-139: prefactor
   1/sqrt(m_i m_j m_k) * phexp(+-q'.R_j) * phexp(-q''.R_k); the 27-term sum
   runs rr (phonon 3) outer, ss, tt (phonon 1) inner as
   ((Phi e1) e2) conj(e3), with e2 conjugated for "-". */
void orc3_vp(const orc_case *c, int ch, int i, int j, int k, int q, int qp, int qdp, const double rqp[3],
             const double rqdp[3], double out[2]) {
  cplx V = {0.0, 0.0};
  for (int t = 0; t < c->ntri3; t++) {
    const int Ii = c->Index3_i[t], Ij = c->Index3_j[t], Ik = c->Index3_k[t];
    double mprod = c->masses[c->types[Ii - 1] - 1] * c->masses[c->types[Ij - 1] - 1];
    mprod = mprod * c->masses[c->types[Ik - 1] - 1];
    const double *rj = c->R3_j + 3 * (size_t)t, *rk = c->R3_k + 3 * (size_t)t;
    double dj = 0.0, dk = 0.0;
    for (int a = 0; a < 3; a++) dj = dj + rqp[a] * rj[a];
    for (int a = 0; a < 3; a++) dk = dk + rqdp[a] * rk[a];
    cplx pre = rmul(1.0 / sqrt(mprod), phexp(ch == 0 ? dj : -dj));
    pre = cmul(pre, phexp(-dk));
    cplx V0 = {0.0, 0.0};
    const double *phi = c->Phi3 + 27 * (size_t)t;
    for (int rr = 1; rr <= 3; rr++)
      for (int ss = 1; ss <= 3; ss++)
        for (int tt = 1; tt <= 3; tt++) {
          const double f = phi[(tt - 1) + 3 * (ss - 1) + 9 * (rr - 1)];
          cplx e1 = EV(c, q, i, tt + 3 * (Ii - 1));
          cplx e2 = EV(c, qp, j, ss + 3 * (Ij - 1));
          cplx e3 = conj_(EV(c, qdp, k, rr + 3 * (Ik - 1)));
          if (ch == 1) e2 = conj_(e2);
          cplx x = rmul(f, e1);
          x = cmul(x, e2);
          x = cmul(x, e3);
          V0.re = V0.re + x.re;
          V0.im = V0.im + x.im;
        }
    cplx y = cmul(pre, V0);
    V.re = V.re + y.re;
    V.im = V.im + y.im;
  }
  out[0] = V.re;
  out[1] = V.im;
}

void orc3_vp_idx(const orc_case *c, int ch, int i, int j, int k, int q, int qp, int qdp, double out[2]) {
  double a[3], b[3];
  realq_of(c, qp, a);
  realq_of(c, qdp, b);
  orc3_vp(c, ch, i, j, k, q, qp, qdp, a, b, out);
}

/* float32-promoted 3ph rate units, processes.f90:433, 512, 1389, 1469 */
static const float F_GAMMA3_UNITS = 5.60626442f;

typedef struct {
  int i, qi, ch, nb, store, onlyharmonic;
  double gamma, wp3;
  int64_t n, cap;
  int32_t *i2, *i3;
  double *g;
} rta3_acc;

/* WP3: "+" (f' - f'') and "-" (f' + f'' + 1), times
   exp(-d**2/sigma**2)/sigma/sqrt(Pi)/(omega*omega'*omega''), processes.f90:1365-1367, 1444-1446 */
static double wp3_of(const orc_case *c, int ch, const orc3_hit *h) {
  const double fp = bose(c, h->omegap), fdp = bose(c, h->omegadp);
  const double occ = (ch == 0) ? fp - fdp : fp + fdp + 1;
  return occ * exp(-(h->delta * h->delta) / (h->sigma * h->sigma)) / h->sigma / sqrt(PI_D) /
         (h->omega * h->omegap * h->omegadp);
}

static void visit_rta3(const orc_case *c, const orc3_hit *h, void *ud) {
  rta3_acc *a = (rta3_acc *)ud;
  const double w = wp3_of(c, a->ch, h);
  a->wp3 = a->wp3 + w;
  const double hbarp = HBAR_D * (double)F_1E22;
  if (a->store) {
    /* Ind_plus / Ind_minus: per-process Gamma, always with Vp (:418-433) */
    double V[2];
    orc3_vp_idx(c, a->ch, a->i, h->j, h->k, a->qi, h->ii, h->ss, V);
    const double av = hypot(V[0], V[1]);
    double g = hbarp * PI_D / 4.0 * w * (av * av);
    g = g * (double)F_GAMMA3_UNITS * 1.e8 / c->nptk;
    if (a->n < a->cap) {
      a->i2[a->n] = (int32_t)((h->ii - 1) * a->nb + h->j);
      a->i3[a->n] = (int32_t)((h->ss - 1) * a->nb + h->k);
      a->g[a->n] = g;
    }
    a->n = a->n + 1;
  } else if (!a->onlyharmonic) {
    double V[2];
    orc3_vp_idx(c, a->ch, a->i, h->j, h->k, a->qi, h->ii, h->ss, V);
    const double av = hypot(V[0], V[1]);
    a->gamma = a->gamma + hbarp * PI_D / 4.0 * w * (av * av);
  }
}

/* RTA_plus / RTA_minus (processes.f90:1313-1471): Gamma in THz before the
   driver's factor 1/2 for "-", WP3 after its 1/nptk (and 1/2 for "-"). */
void orc3_rta_mode(const orc_case *c, int ch, int mm, int32_t ii_lo, int32_t ii_hi, double *gamma, double *wp3) {
  rta3_acc a;
  memset(&a, 0, sizeof(a));
  a.i = fmodulo(mm - 1, c->nbands) + 1;
  a.qi = c->List[(mm - 1) / c->nbands];
  a.ch = ch;
  a.nb = c->nbands;
  a.onlyharmonic = c->onlyharmonic;
  const double omega = E(c, a.qi, a.i);
  enumerate3(c, ch, mm, 0, ii_lo, ii_hi, visit_rta3, &a);
  if (omega != 0.0) a.wp3 = (ch == 0) ? a.wp3 / c->nptk : a.wp3 * 5.e-1 / c->nptk;
  *gamma = a.gamma * (double)F_GAMMA3_UNITS * 1.e8 / c->nptk;
  *wp3 = a.wp3;
}

/* Ind_plus / Ind_minus (processes.f90:361-521): list in j-outer order and
   per-process Gamma (THz). */
int64_t orc3_ind_mode(const orc_case *c, int ch, int mm, int64_t cap, int32_t *i2, int32_t *i3, double *gamma,
                      double *wp3) {
  rta3_acc a;
  memset(&a, 0, sizeof(a));
  a.i = fmodulo(mm - 1, c->nbands) + 1;
  a.qi = c->List[(mm - 1) / c->nbands];
  a.ch = ch;
  a.nb = c->nbands;
  a.store = 1;
  a.cap = cap;
  a.i2 = i2; a.i3 = i3; a.g = gamma;
  const double omega = E(c, a.qi, a.i);
  enumerate3(c, ch, mm, 1, 1, c->nptk, visit_rta3, &a);
  if (omega != 0.0) a.wp3 = (ch == 0) ? a.wp3 / c->nptk : a.wp3 * 5.e-1 / c->nptk;
  *wp3 = a.wp3;
  return a.n;
}

/* NP_driver, processes.f90:983-1067 */
void orc3_np_driver(const orc_case *c, int nthreads, int32_t *N_plus, double *P_plus, int32_t *N_minus,
                    double *P_minus) {
  const int nm = c->nbands * c->nlist;
  set_threads(nthreads);
  for (int m = 0; m < nm; m++) {
    N_plus[m] = N_minus[m] = 0;
    P_plus[m] = P_minus[m] = 0.0;
  }
#pragma omp parallel for schedule(dynamic, 1)
  for (int mm = 1; mm <= nm; mm++) {
    const int i = fmodulo(mm - 1, c->nbands) + 1;
    const int ll = (mm - 1) / c->nbands + 1;
    if (E(c, c->List[ll - 1], i) <= c->omega_max) {
      int64_t n;
      orc3_np_mode(c, 0, mm, &n, &P_plus[mm - 1]);
      N_plus[mm - 1] = (int32_t)n;
      orc3_np_mode(c, 1, mm, &n, &P_minus[mm - 1]);
      N_minus[mm - 1] = (int32_t)n;
    }
  }
}

/* Ind_only_driver, processes.f90:1198-1309.  Returns 0, or -1 when a list is
   longer than its count. */
int orc3_ind_only_driver(const orc_case *c, int nthreads, const int32_t *N_plus, const int32_t *N_minus,
                         int32_t *i2_plus, int32_t *i3_plus, int32_t *i2_minus, int32_t *i3_minus) {
  const int nm = c->nbands * c->nlist;
  const int32_t *Ns[2] = {N_plus, N_minus};
  int32_t *out[2][2] = {{i2_plus, i3_plus}, {i2_minus, i3_minus}};
  int64_t *acc[2];
  int bad = 0;
  for (int ch = 0; ch < 2; ch++) {
    acc[ch] = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nm + 1));
    acc[ch][0] = 0;
    for (int m = 0; m < nm; m++) acc[ch][m + 1] = acc[ch][m] + Ns[ch][m];
    for (int64_t t = 0; t < acc[ch][nm]; t++) out[ch][0][t] = out[ch][1][t] = 0;
  }
  set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) reduction(| : bad)
  for (int mm = 1; mm <= nm; mm++) {
    const int i = fmodulo(mm - 1, c->nbands) + 1;
    const int ll = (mm - 1) / c->nbands + 1;
    if (!(E(c, c->List[ll - 1], i) <= c->omega_max)) continue;
    for (int ch = 0; ch < 2; ch++) {
      const int64_t n = Ns[ch][mm - 1];
      if (n == 0) continue;
      const int64_t off = acc[ch][mm - 1];
      if (orc3_ind_only_mode(c, ch, mm, n, out[ch][0] + off, out[ch][1] + off) != n) bad |= 1;
    }
  }
  for (int ch = 0; ch < 2; ch++) free(acc[ch]);
  return bad ? -1 : 0;
}

/* RTA_driver, processes.f90:1676-1775, restricted to a list of modes and a
   q' window; outputs are (Nb,Nlist) arrays that the caller zeroes. */
void orc3_rta_driver(const orc_case *c, int nthreads, int nmodes, const int32_t *modes, int32_t ii_lo, int32_t ii_hi,
                     double *rate, double *rate_plus, double *rate_minus, double *wp3_plus, double *wp3_minus) {
  set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
  for (int t = 0; t < nmodes; t++) {
    const int mm = modes[t];
    const int i = fmodulo(mm - 1, c->nbands) + 1;
    const int ll = (mm - 1) / c->nbands + 1;
    if (!(E(c, c->List[ll - 1], i) <= c->omega_max)) continue;
    double g, w;
    orc3_rta_mode(c, 0, mm, ii_lo, ii_hi, &g, &w);
    rate_plus[mm - 1] = g; wp3_plus[mm - 1] = w;
    orc3_rta_mode(c, 1, mm, ii_lo, ii_hi, &g, &w);
    rate_minus[mm - 1] = g * 5.e-1; wp3_minus[mm - 1] = w;
  }
  for (int t = 0; t < nmodes; t++) {
    const int m = modes[t] - 1;
    rate[m] = rate_plus[m] + rate_minus[m];
  }
}

/* Ind_driver, processes.f90:524-707 */
int orc3_ind_driver(const orc_case *c, int nthreads, const int32_t *N_plus, const int32_t *N_minus, int32_t *i2_plus,
                    int32_t *i3_plus, double *g_plus, int32_t *i2_minus, int32_t *i3_minus, double *g_minus,
                    double *rate, double *rate_plus, double *rate_minus, double *wp3_plus, double *wp3_minus) {
  const int nm = c->nbands * c->nlist;
  const int32_t *Ns[2] = {N_plus, N_minus};
  int32_t *out[2][2] = {{i2_plus, i3_plus}, {i2_minus, i3_minus}};
  double *gs[2] = {g_plus, g_minus};
  double *rates[2] = {rate_plus, rate_minus};
  double *wps[2] = {wp3_plus, wp3_minus};
  int64_t *acc[2];
  int bad = 0;
  for (int ch = 0; ch < 2; ch++) {
    acc[ch] = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nm + 1));
    acc[ch][0] = 0;
    for (int m = 0; m < nm; m++) acc[ch][m + 1] = acc[ch][m] + Ns[ch][m];
    for (int m = 0; m < nm; m++) rates[ch][m] = wps[ch][m] = 0.0;
  }
  set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) reduction(| : bad)
  for (int mm = 1; mm <= nm; mm++) {
    for (int ch = 0; ch < 2; ch++) {
      const int64_t n = Ns[ch][mm - 1];
      if (n == 0) continue;
      const int64_t off = acc[ch][mm - 1];
      double w;
      if (orc3_ind_mode(c, ch, mm, n, out[ch][0] + off, out[ch][1] + off, gs[ch] + off, &w) != n) bad |= 1;
      wps[ch][mm - 1] = w;
      double s = 0.0;
      for (int64_t t = 0; t < n; t++) s = s + gs[ch][off + t];
      rates[ch][mm - 1] = (ch == 0) ? s : s * 5.e-1;
    }
  }
  for (int m = 0; m < nm; m++) rate[m] = rate_plus[m] + rate_minus[m];
  for (int ch = 0; ch < 2; ch++) free(acc[ch]);
  return bad ? -1 : 0;
}

/* ------------------------------------------ three-phonon sampling estimators
 * processes.f90:828-979 (NP_plus/minus_sample) and 1474-1673
 * (RTA_plus/minus_sample), dispatched at 1033-1050 and 1736-1760.  One draw
 * per sample picks (ii, j, k) uniformly from the Nb*nptk*Nb candidates of the
 * full loops (index rand_num-1 = ((ii-1)*Nb + j-1)*Nb + k-1, as the
 * reference decodes it); every candidate is one the full drivers visit, so
 * the sums scale by Nb*nptk*Nb/nsample.  Draw s of segment (ch, mm):
 * mix64(key + G*(s+1)), key = mix64(seed + G*(2*(mm-1)+ch+1)).
 * ------------------------------------------------------------------------ */
void orc3_sample_mode(const orc_case *c, int ch, int mm, int32_t nsample, uint64_t seed, int want_rates,
                      int64_t *hits, double *P, double *gamma, double *wp3) {
  const int nb = c->nbands, nptk = c->nptk;
  const int i = fmodulo(mm - 1, nb) + 1;
  const int qi = c->List[(mm - 1) / nb];
  const int32_t *q = c->IJK + 3 * (size_t)(qi - 1);
  const double omega = E(c, qi, i);
  np3_acc na;
  na.n = 0;
  na.p = 0.0;
  na.den = (double)((int64_t)nptk * nptk);
  rta3_acc ra;
  memset(&ra, 0, sizeof(ra));
  ra.i = i;
  ra.qi = qi;
  ra.ch = ch;
  ra.nb = nb;
  ra.onlyharmonic = c->onlyharmonic;
  if (omega != 0.0) {
    const uint64_t key = orc_mix64(seed + ORC_GOLDEN * (uint64_t)(2 * (mm - 1) + ch + 1));
    for (int64_t s = 0; s < nsample; s++) {
      const int r = orc_below(orc_mix64(key + ORC_GOLDEN * (uint64_t)(s + 1)), nb * nptk * nb);
      const int k = r % nb + 1;
      const int j = (r / nb) % nb + 1;
      const int ii = r / (nb * nb) + 1;
      test3(c, ch, q, omega, ii, j, k, k, visit_np3, &na);
      if (want_rates) test3(c, ch, q, omega, ii, j, k, k, visit_rta3, &ra);
    }
  }
  *hits = na.n;
  *P = na.p;
  if (gamma) *gamma = ra.gamma;
  if (wp3) *wp3 = ra.wp3;
}

/* NP_driver with num_sample_process_3ph_phase_space = nsample; N and P
   scaled in the reference's left-to-right order (processes.f90:898-899).
   Returns -1 when a count estimate exceeds int32. */
int orc3_np_sample_driver(const orc_case *c, int nthreads, int32_t nsample, uint64_t seed, int32_t *N_plus,
                          double *P_plus, int32_t *N_minus, double *P_minus) {
  const int nm = c->nbands * c->nlist;
  const int nb = c->nbands, nptk = c->nptk;
  int32_t *N[2] = {N_plus, N_minus};
  double *P[2] = {P_plus, P_minus};
  int bad = 0;
  set_threads(nthreads);
  for (int m = 0; m < nm; m++)
    for (int ch = 0; ch < 2; ch++) {
      N[ch][m] = 0;
      P[ch][m] = 0.0;
    }
#pragma omp parallel for schedule(dynamic, 1) reduction(| : bad)
  for (int mm = 1; mm <= nm; mm++) {
    const int i = fmodulo(mm - 1, nb) + 1;
    const int ll = (mm - 1) / nb + 1;
    if (!(E(c, c->List[ll - 1], i) <= c->omega_max)) continue;
    for (int ch = 0; ch < 2; ch++) {
      int64_t hits;
      double p;
      orc3_sample_mode(c, ch, mm, nsample, seed, 0, &hits, &p, NULL, NULL);
      const double n = (double)hits / (double)nsample * nb * nptk * nb;
      if (n > 2147483647.0) bad |= 1;
      N[ch][mm - 1] = (int32_t)n;
      P[ch][mm - 1] = p * nb * nptk * nb / (double)nsample;
    }
  }
  return bad ? -1 : 0;
}

/* RTA_driver with num_sample_process_3ph = nsample over a mode list
   (processes.f90:1559-1570, 1660-1671, minus channel halved at :1759). */
void orc3_rta_sample_driver(const orc_case *c, int nthreads, int32_t nsample, uint64_t seed, int nmodes,
                            const int32_t *modes, double *rate, double *rate_plus, double *rate_minus,
                            double *wp3_plus, double *wp3_minus, int64_t *hits) {
  const int nb = c->nbands, nptk = c->nptk;
  double *R[2] = {rate_plus, rate_minus};
  double *W[2] = {wp3_plus, wp3_minus};
  int64_t nh = 0;
  set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : nh)
  for (int t = 0; t < nmodes; t++) {
    const int mm = modes[t];
    const int i = fmodulo(mm - 1, nb) + 1;
    const int ll = (mm - 1) / nb + 1;
    const double omega = E(c, c->List[ll - 1], i);
    if (!(omega <= c->omega_max)) continue;
    for (int ch = 0; ch < 2; ch++) {
      int64_t h;
      double p, g, w;
      orc3_sample_mode(c, ch, mm, nsample, seed, 1, &h, &p, &g, &w);
      nh += h;
      if (omega != 0.0) w = (ch == 0) ? w / nptk : w * 5.e-1 / nptk;
      g = g * nb * nptk * nb / (double)nsample;
      w = w * nb * nptk * nb / (double)nsample;
      g = g * (double)F_GAMMA3_UNITS * 1.e8 / nptk;
      R[ch][mm - 1] = (ch == 0) ? g : g * 5.e-1;
      W[ch][mm - 1] = w;
    }
  }
  for (int t = 0; t < nmodes; t++) {
    const int m = modes[t] - 1;
    rate[m] = rate_plus[m] + rate_minus[m];
  }
  if (hits) *hits = nh;
}
