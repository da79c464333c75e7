// Iterative BTE on the device: iteration0, iteration_4ph, TConduct and the
// convergence loop around them (Src/iterations.f90:13-30, 95-215;
// Src/conductivity.f90:11-37; Src/ShengBTE.f90:1250-1307).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace fp4 {

// Arguments of one (ll, i, kk) step of iteration_4ph.  List types:
// 0 = 3ph +, 1 = 3ph -, 2 = 4ph ++, 3 = 4ph +-, 4 = 4ph --.
struct IterStep {
  const int *e2[5], *e3[5], *e4[3];  // encoded (q-1)*Nb+branch, 1-based, int32 on the device
  const double *g[5];                // per-process Gamma
  long long off[5], n[5];            // this mode's slice of each list
  const int *ideq;                   // ID_equi(Nsymm_rot, nptk), 1-based
  const double *symm;                // symmetrizers(3,3,nptk)
  const double *omega, *vel;         // energy(nptk,Nb), velocity(nptk,Nb,3)
  double *F;                         // F_n on the device, [q][b][alpha]
  double *part;                      // 3 partial sums per block
  unsigned int *ticket;              // block-completion counter, 0 between launches
  int nb, nptk, nsymm;
  int i, qk, sym;                    // 0-based branch, star member, symmetry operation
  double tau;
  // aggregated operator instead of the lists when W is set: this mode's row
  // of W[Nb*Nlist][nptk*Nb] (partner columns q*Nb + b)
  const double *W;
  long long wcols;
  int wrow;
};

struct BteModes {
  const double *omega;  // device energy(nptk,Nb)
  const double *vel;    // device velocity(nptk,Nb,3)
  int nptk, nb, nlist;
  hipStream_t stream;
};

struct BteState;

BteState *bte_create();
void bte_destroy(BteState *s);
const char *bte_error();

// ID_equi(Nsymm_rot,nptk) 1-based (-1 where a rotation leaves the grid),
// symmetrizers(3,3,nptk), crotations(3,3,Nsymm_rot).
int bte_set_symmetry(BteState *s, const BteModes &m, int nsymm, const int32_t *ID_equi, const double *symmetrizers,
                     const double *crotations);
// Nequi(Nlist), ALLEquiList(Nsymm_rot,nptk), TypeofSymmetry(Nsymm_rot,nptk).
int bte_set_orbits(BteState *s, const BteModes &m, const int32_t *Nequi, const int32_t *ALLEquiList,
                   const int32_t *TypeofSymmetry);
// Scattering operator: per-mode counts and lists of the five list types
// (null counts = empty type).  3ph lists are int32, 4ph lists int64, as in
// the reference.
int bte_set_operator(BteState *s, const BteModes &m, const int32_t *const N3[2], const int32_t *const L3[2][2],
                     const double *const G3[2], const int64_t *const N4[3], const int64_t *const L4[3][3],
                     const double *const G4[3]);

// Aggregated operator: bte_w_begin allocates W[Nb*Nlist][nptk*Nb] on the
// device, zeroed (*W, *cols receive it; the caller adds the processes into
// it), bte_w_finish makes the sweeps use it instead of the lists (until the
// next bte_set_operator).  Needs the orbits (bte_set_orbits) first.
int bte_w_begin(BteState *s, const BteModes &m, double **W, long long *cols);
int bte_w_finish(BteState *s, const BteModes &m);
int bte_iteration0(BteState *s, const BteModes &m, const double *tau_zero);  // tau_zero(Nb,Nlist) host
int bte_sweep(BteState *s, const BteModes &m, const double *tau_zero);       // one iteration_4ph
// With the aggregated operator a sweep is one cooperative launch (on, the
// default) or one launch per (ll, i, kk) step (off); the results are the same
// bit for bit.
int bte_set_persistent(BteState *s, bool on);
int bte_symmetrize_all(BteState *s, const BteModes &m);                      // F_n(:,q,:) for every q
int bte_get_F(BteState *s, const BteModes &m, double *F_n);                  // F_n(Nb,nptk,3) host
int bte_put_F(BteState *s, const BteModes &m, const double *F_n);
// TConduct on the resident F_n; kappa(Nb,3,3) and (optional) mode(nptk,Nb,3,3).
int bte_tconduct(BteState *s, const BteModes &m, double T, double V, double *kappa, double *mode);
// Convergence loop; kappa_iter receives sum_b kappa (3,3) per iteration
// (0..niter), kappa the last per-branch tensor.
int bte_solve(BteState *s, const BteModes &m, double T, double V, const double *tau_zero, int maxiter, double eps,
              double *kappa, double *kappa_iter, double *relchange, int *niter);
double bte_last_ms(const BteState *s);
long long bte_last_launches(const BteState *s);

}  // namespace fp4
