// Dense (fold + contraction) form of the four-phonon rates; see fp4_dense.hip.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "fp4_device.hpp"

namespace fp4 {

// Atoms per cell the dense kernels take (8: Nbands up to 24), and the (K, L)
// atom pairs that follow from it.
constexpr int kDenseMaxAtoms = 8;
constexpr int kDenseMaxPairs = kDenseMaxAtoms * kDenseMaxAtoms;

// Coarse phase group: all clusters whose third and fourth atoms are (K, L)
// and whose R_k - R_l is the same lattice vector d (in phase units).
struct DGroup {
  int K, L;
  int d[3];
  int f_first, f_count;   // fine groups of this coarse group
};
// Fine group: clusters of one coarse group that also share atoms (I, J).
struct FGroup {
  int I, J;
  int c_first, c_count;   // clusters (sorted) of this fine group
};
// Cluster in dense order: q'-dependent phase multipliers and mass factor.
struct DCluster {
  int pjl[3];             // pj - pl
  int pl[3];
  double minv;
};

struct DenseTuning {
  int slots = 0;          // resident blocks of the persistent kernel; 0 = one per CU
  int shape = -1;         // kernel shape (IB, TQ, MFMA path) chosen by dense_autotune; -1 = first that fits
  int shape_prec = -1;    // the precision the shape and slots were tuned for
};

struct DenseState;        // device copies of the grouped clusters (opaque)

DenseState *dense_create(const std::vector<ClusterRec> &cl, const std::vector<double> &phi, int natoms, int nb,
                         int L, int *rc);
void dense_destroy(DenseState *s);
int dense_ngroups(const DenseState *s);
// Work of the last dense_rates call: visited triples and work items, and the
// coarse / fine group counts of the clusters.
void dense_info(const DenseState *s, long long *triples, long long *items, int *ngroups, int *nfine);
// Kernel shape of the last dense_rates call: MFMA path, q'' per tile, modes
// per block, and the coarse groups per MFMA k-step (0 = k_dense, > 0 =
// k_dense2).
void dense_shape(const DenseState *s, int *mfma, int *tq, int *ib, int *kg = nullptr);
// Checks the MFMA lane layout the dense kernel assumes; *err = max abs error.
int dense_selftest_mfma(double *err);

struct DenseRun {
  ModeView m;
  const double2 *roots;
  double T;
  int onlyharmonic;
  int natoms;
  int precision;          // FP4_PREC_FP32 / FP4_PREC_FP64_DENSE
  hipStream_t stream;
  DenseTuning *tuning;
  // aggregated scattering operator (iterative solution), or null: each
  // process also adds its Gamma (in THz) with its iteration_4ph sign to
  // wop[irreducible mode][partner's q*Nb + b], wcols = nptk*Nb columns
  double *wop = nullptr;
  long long wcols = 0;
};

// Rates of the given modes over the q' window [ii_lo0, ii_lo0+nii): g, w and
// n are per-(channel, mode-in-subset) sums of Gamma (unscaled), WP4 and the
// number of processes.
int dense_rates(DenseState *s, const DenseRun &r, const std::vector<int> &modes, int ii_lo0, int nii,
                std::vector<double> &g, std::vector<double> &w, std::vector<long long> &n, double *ms);
// The project's autotuner: runs every built kernel shape of r.precision and
// Nbands that fits in LDS, at one and, where occupancy allows, two resident
// blocks per CU, on the given modes and q' window, and stores the fastest
// (by the kernel's HIP-event time) in *out.  *ncand receives the number of
// candidates timed, best_ms the winner's kernel time.  Rates computed with
// different shapes agree to rounding (the per-item sums are regrouped).
int dense_autotune(DenseState *s, const DenseRun &r, const std::vector<int> &modes, int ii_lo0, int nii,
                   DenseTuning *out, double *best_ms, int *ncand);

// Profiling build only: shader cycles per k_dense2 phase (DensePhase order,
// summed over waves) of the last dense call; returns the number of phases,
// or -FP4_ERR_UNSUPPORTED in the product build.
int dense_phase_profile(const DenseState *s, int cap, double *cycles);
// Candidates timed by the last dense_autotune call: kernel k-step groups
// (0 = k_dense), IB, TQ, resident blocks and kernel time; returns the count.
int dense_candidates(const DenseState *s, int cap, int *kg, int *ib, int *tq, int *slots, double *ms);
// A tuning that runs the first shape of this precision and Nbands with the
// given IB and k-step groups (kg = 0: k_dense) at `slots` blocks (0 = one
// per CU); FP4_ERR_UNSUPPORTED when no such shape is built or fits in LDS.
int dense_pick(const DenseState *s, int precision, int nb, int kg, int ib, int slots, DenseTuning *out);

// Several crystals' dense rates in one persistent launch (k_dense_multi;
// FP32 or FP16X2 MFMA, cells of up to 8 atoms, Nbands 6, 12, 18, 24): every
// crystal's work items go into one queue, costliest first by the prefilter
// counts (counts:
// processes per (channel, mode-in-subset, q') in the count pass's layout, or
// empty), and with use_graph the device work of a call (the queue counter,
// the launch, the per-crystal reductions) is a captured HIP graph, replayed
// while the crystals' arguments are unchanged.
struct DenseBatchMember {
  DenseState *s = nullptr;
  DenseRun r;
  std::vector<int> modes;
  int ii_lo0 = 0, nii = 0;
  std::vector<int> counts;
  std::vector<double> g, w;              // results, as dense_rates
  std::vector<long long> n;
};
struct DenseBatchState;
DenseBatchState *dense_batch_create();
void dense_batch_destroy(DenseBatchState *b);
// part: the key of the cached device tables and graph (one per q' window
// the caller cycles through).
int dense_rates_batch(DenseBatchState *b, std::vector<DenseBatchMember> &mem, int precision, hipStream_t st,
                      bool use_graph, int part, double *kernel_ms, int *launches);
const char *dense_error();

}  // namespace fp4
