// Launch interfaces between the host runtime (fp4_api.hip) and the kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "fp4_device.hpp"

namespace fp4 {

// Items are (channel, mode-in-subset, ii) with ii fastest; item_lo..item_hi
// is the slice this launch covers.
struct PrefilterArgs {
  ModeView m;
  const int *modes;      // mm (1-based) of the subset, ascending
  int nsub;              // number of modes in the subset
  int nii;               // ii window length
  int ii_lo0;            // first ii (0-based)
  long item_lo, item_hi;
  // count pass
  int *cnt;              // [items]
  double *psum;          // [items]
  // emit pass
  const long long *offsets;  // exclusive prefix of cnt over all items
  long long out_base;        // offsets[item_lo]
  int *out32[3];
  long long *out64[3];
  // hash pass: per-(channel, mode) order-sensitive list hash, accumulated
  // with 64-bit atomic adds (zeroed by the caller)
  unsigned long long *seg_hash;
};

int launch_prefilter(const PrefilterArgs &a, int mode, hipStream_t st);
void launch_reduce_counts(const int *cnt, const double *psum, int nseg, int nii, long long *N, double *P,
                          hipStream_t st);
void launch_widen(const int *cnt, long long *out, long n, hipStream_t st);

// Sampling estimators: segments are (channel, mode-in-subset), ch-major.
struct SampleArgs {
  ModeView m;
  const int *modes;          // mm (1-based) of the subset
  int nsub;
  long long nsample;         // draws per segment
  unsigned long long seed;
  int seg_lo, seg_hi;        // segments of this launch
  // count pass
  int *cnt;                  // [3*nsub] hits
  double *psum;              // [3*nsub] phase-space sums
  // emit pass
  const long long *offsets;  // exclusive prefix of cnt, 3*nsub+1 entries
  long long out_base;        // offsets[seg_lo]
  int *out32[3];
};

int launch_sample4(const SampleArgs &a, int mode, hipStream_t st);
// three phonons: segments (channel, mode-in-subset) with 2 channels; the
// emit pass writes out32[0..1]
int launch_sample3(const SampleArgs &a, int mode, hipStream_t st);

// A block of at most 256 consecutive processes of one (channel, mode).
struct BlockDesc {
  long long start;   // first process, index into the batch's list arrays
  int count;
  int ch;
  int mm;            // 1-based mode
  int seg;           // output slot of the (channel, mode) segment
};

struct RateArgs {
  ModeView m;
  IfcView f;
  const int *lst[3];         // batch lists (reference encoding), int32
  const long long *lst64[3]; // or int64 (caller-provided lists); one of the two is set
  const BlockDesc *blocks;
  int nblocks;
  double T;
  int onlyharmonic;
  int store_gamma;           // Ind_driver_4ph: write per-process Gamma
  double *proc_gamma;        // [batch processes], when store_gamma
  double *blk_gamma;         // [nblocks]
  double *blk_wp4;           // [nblocks]
};

int launch_rate64(const RateArgs &a, hipStream_t st);

// Aggregated scattering operator of the iterative solution: the per-process
// Gamma of one batch (rate kernel with store_gamma) added to W[mode][col]
// with the signs of the F_n terms of iteration_4ph, col = the partner's
// (q-1)*Nb + branch - 1.  Block per BlockDesc, thread per process.
struct WScatterArgs {
  const BlockDesc *blocks;
  int nblocks;
  const int *lst[3];         // batch lists (reference encoding), int32
  const double *gamma;       // per-process Gamma of the batch, THz
  double *W;
  long long wcols;
  int order;                 // 3: lists +/- of three phonons; 4: ++/+-/--
};
int launch_w_scatter(const WScatterArgs &a, hipStream_t st);

// Per-segment sums of block partials, in block order.
void launch_seg_reduce(const double *blk_gamma, const double *blk_wp4, const int *seg_first, int nseg,
                       double *seg_gamma, double *seg_wp4, hipStream_t st);

struct VpArgs {
  ModeView m;
  IfcView f;
  int ch;
  long n;
  const int *idx;   // n rows of (i, j, k, l, q, q', q'', q''') 1-based
  double2 *out;
};
int launch_vp64(const VpArgs &a, hipStream_t st);

// Three-phonon scan: items (channel, mode-in-subset, j) with j fastest,
// second-phonon q window ii_lo0 .. ii_lo0 + nii - 1.
struct Scan3Args {
  ModeView m;
  const int *modes;
  int nsub;
  int nii, ii_lo0;
  long item_lo, item_hi;
  int *cnt;                  // count pass: [items]
  double *psum;              // count pass: [items]
  const long long *offsets;  // emit pass: exclusive prefix of cnt
  long long out_base;
  int *out[2];               // emit pass: Indof2ndPhonon, Indof3rdPhonon
};
int launch_scan3(const Scan3Args &a, int mode, hipStream_t st);

struct Rate3Args {
  ModeView m;
  Ifc3View f;
  const int *lst[2];
  const BlockDesc *blocks;
  int nblocks;
  double T;
  int onlyharmonic;
  int store_gamma;           // Ind_driver: per-process Gamma in THz
  double *proc_gamma;
  double *blk_gamma;
  double *blk_wp;
};
int launch_rate3(const Rate3Args &a, hipStream_t st);

struct Vp3Args {
  ModeView m;
  Ifc3View f;
  int ch;
  long n;
  const int *idx;   // n rows of (i, j, k, q, q', q'') 1-based
  double2 *out;
};
int launch_vp3(const Vp3Args &a, hipStream_t st);

// Harmonic front end (eigenDM, Quantum Espresso constants): one wavefront
// per grid point.
struct EigenDMArgs {
  int nptk, nb, nat;
  int n1, n2, n3;
  double b[9];              // rlattvec(3,3) column-major, nm^-1
  double bohr2nm, tothz;
  const int *img_start;     // [nat*nat + 1] image ranges per atom pair (iat*nat + jat)
  const double *img_t;      // [nimg][3] image vectors, bohr
  const double *img_fc;     // [nimg][9] Phi(alpha,beta) * Wigner-Seitz weight, alpha slow
  const double *mm;         // [nat*nat] sqrt(m_i m_j) (m_i on the diagonal), Ry mass units
  const double *symm;       // [nptk][9] symmetrizers(:,:,q) column-major, or null
  double *omega;            // omega(nptk,Nb)
  double2 *eig;             // eigenvect(nptk,Nb,Nb)
  double *vel;              // velocity(nptk,Nb,3)
  // phonopy's dielectric correction (phonon_phonopy); nac_q null: none
  const double *img_w;      // [nimg] weight of each image (1/neq)
  const double *nac_q;      // [nptk][3] the q image the correction's direction comes from, nm^-1
  const double *nac_mm;     // [nat*nat] m_i on the diagonal, sqrt(m_i m_j) off it, amu
  const double *born;       // born(3,3,nat) column-major
  double eps[9];            // epsilon(3,3) column-major
  double nac_scale;         // prefactor / V / (scell(1) scell(2) scell(3))
};
int launch_eigen_dm(const EigenDMArgs &a, hipStream_t st);

}  // namespace fp4
