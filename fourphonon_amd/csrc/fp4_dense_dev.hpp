// Device-side pieces shared by the dense-form kernels (fp4_dense.hip and
// fp4_dense2.hip): complex helpers, the kernel argument block, block scans,
// the sigma helpers of the tile filter and the BF16 / FP16 operand splits.
#pragma once
#include <hip/hip_runtime.h>

#include "fp4_dense.hpp"

namespace fp4 {

constexpr int NT_DENSE = 512;

// Phases of k_dense2 in the profiling build's cycle counts: work-item fetch
// and reduction, A' set-up, valid-q'' scan, tile data, wait at the tile
// barrier, filter (phases, screen, exact tests), wait at the filter barrier,
// per-wave survivor queue, Psi GEMM, contraction, rates, wait at the end of
// the tile.
enum DensePhase {
  kPhItem, kPhSetup, kPhScan, kPhTile, kPhBar1, kPhFilter, kPhBar2, kPhQueue, kPhGemm, kPhContract, kPhRate,
  kPhBar3, kDensePhases
};

template <class T>
struct C2 {
  T re, im;
};

struct DenseArgs {
  ModeView m;
  const double2 *roots;
  int L;
  const DGroup *g;
  const FGroup *f;
  const DCluster *c;
  const double *phi;
  int klf[kDenseMaxPairs + 1];  // first coarse group of each (K, L) pair, then ng
  const double *fbose;   // [Nb][nptk] Bose-Einstein factor of every mode at T
  int roots_lds;         // the root table is copied to LDS (L <= kRootsLds)
  int ng;
  const int *lls;        // selected irreducible points (0-based ll), ascending
  int nls;
  int nib;               // mode blocks per q
  int ii_lo0, nii;
  const unsigned char *sel;  // [Nb*Nlist] selected modes
  double T;
  int onlyharmonic;
  long nitems;
  int *counter;
  double *out_g;
  double *out_w;
  int *out_n;
  int *out_t;            // [nitems] (q, q', q'') triples with >= 1 surviving process
  int ntri;
  int smem_static;       // bytes of DenseSmem at the start of dynamic LDS
  float bf[9];           // rlattvec(:,nu)/ngrid(nu) as floats
  float sbf;             // scalebroad
  float a_scale;         // FP16 form: power of two with |A'| * a_scale <= 2^14
  // k_dense2 only
  int pair_off[17];      // first phase-table slot of each (K, L) pair (groups padded per pair)
  int gpt;               // phase-table slots per row
  const int *gslot;      // phase-table slot of each coarse group
  int lmask;             // L - 1 when L is a power of two, else -1
  unsigned long long *prof;  // profiling build: [kDensePhases] cycle sums
  // aggregated scattering operator (iterative solution): when wop is set,
  // every process also adds its Gamma * wscale, with the sign of its term in
  // iteration_4ph, to wop[row][col] (row = irreducible mode, col = the
  // partner's full-mesh mode q*Nb + b, wcols columns)
  double *wop;
  long long wcols;
  double wscale;
  // several crystals of one Nb in one launch: the per-crystal arguments in
  // device memory, and the launch's work items in queue order, each packed
  // as crystal << 40 | the crystal's own item index (nbatch = 0: one crystal,
  // these arguments)
  const DenseArgs *batch;
  const long long *order;
  int nbatch;
  // cells of more than 4 atoms: A' of every resident block in device memory,
  // ag_stride bytes per block (the table does not fit LDS beside the tile)
  void *ag;
  long long ag_stride;
};

// The arguments of the crystal a work item belongs to.
template <bool BATCH>
__device__ __forceinline__ const DenseArgs &dense_args_of(const DenseArgs &ka, int cr) {
  if constexpr (BATCH)
    return ka.batch[cr];
  else
    return ka;
}

// Adds one process's Gamma to the aggregated operator with the signs of the
// F_n terms of iteration_4ph (Src/iterations.f90:177-203): ++ -F(j,q2)
// -F(k,q3) +F(l,q4); +- -F(j,q2) +F(k,q3) +F(l,q4); -- all +.  c2, c3, c4
// are the partners' columns q*Nb + b.
__device__ __forceinline__ void wop_add(const DenseArgs &a, int row, int ch, int c2, int c3, int c4, double g) {
  double *r = a.wop + (size_t)row * (size_t)a.wcols;
  const double gs = g * a.wscale;
  unsafeAtomicAdd(r + c2, ch == 2 ? gs : -gs);
  unsafeAtomicAdd(r + c3, ch == 0 ? -gs : gs);
  unsafeAtomicAdd(r + c4, gs);
}

template <class T>
__device__ __forceinline__ C2<T> cm(C2<T> a, C2<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ double2 zm(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// block-wide exclusive scan of per-thread counts in thread order (NW waves)
template <int NW>
__device__ __forceinline__ int block_excl_scan(int v, int *s_wsum, int &tot) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int incl_w = wave_excl_scan(v, lane) + v;
  if (lane == 63) s_wsum[wv] = incl_w;
  __syncthreads();
  int base = 0;
  tot = 0;
#pragma unroll
  for (int w = 0; w < NW; w++) {
    if (w < wv) base += s_wsum[w];
    tot += s_wsum[w];
  }
  __syncthreads();
  return base + incl_w - v;
}

constexpr int kRootsLds = 256;

__host__ __device__ constexpr size_t align16(size_t x) { return (x + 15) & ~(size_t)15; }

// Bit mask with bit i*nb set for every i < ib.
constexpr unsigned long long rep_mask(int ib, int nb) {
  unsigned long long r = 0ull;
  for (int i = 0; i < ib; i++) r |= 1ull << (i * nb);
  return r;
}

// Velocities of the tile's third and fourth phonons live in LDS, so sigma is
// recomputed here with the same operations as pair_sigma().
__device__ __forceinline__ double tile_sigma(const ModeView &m, int ch, const double *v3, const double *v4) {
  double d[3];
#pragma unroll
  for (int a = 0; a < 3; a++) d[a] = (ch == 0) ? (-v3[a] + v4[a]) : (v3[a] - v4[a]);
  return m.scalebroad * base_sigma(m, d[0], d[1], d[2]);
}

// Float estimate of sigma for the cull: b_nu / N_nu and scalebroad as floats,
// velocity differences formed in FP64 and rounded once.  Its relative error
// is below 1e-6, far inside the 1e-4 band that is sent to the exact test.
__device__ __forceinline__ float sigma_estimate(const float *bf, int ch, const double *v3, const double *v4,
                                                float sbf) {
  float d[3];
#pragma unroll
  for (int a = 0; a < 3; a++) d[a] = (float)((ch == 0) ? (-v3[a] + v4[a]) : (v3[a] - v4[a]));
  float mx = 0.f;
#pragma unroll
  for (int nu = 0; nu < 3; nu++) {
    const float x = bf[3 * nu] * d[0] + bf[3 * nu + 1] * d[1] + bf[3 * nu + 2] * d[2];
    mx = fmaxf(mx, x * x);
  }
  return sbf * sqrtf(0.5f * mx);
}


typedef float floatx4 __attribute__((ext_vector_type(4)));

// BF16 operands: round-to-nearest-even to the upper 16 bits of a float.
typedef short bf16x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ unsigned bf16_rn(float f) {
  const unsigned u = __float_as_uint(f);
  return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
}
__device__ __forceinline__ float bf16_val(unsigned h) { return __uint_as_float(h << 16); }
// f = hi + lo + O(2^-16 |f|): the hi/lo split of one float
__device__ __forceinline__ void bf16_split(float f, unsigned &hi, unsigned &lo) {
  hi = bf16_rn(f);
  lo = bf16_rn(f - bf16_val(hi));
}
// two (lo16 = k, hi16 = k+1) words -> the 4 consecutive-k operand of one lane
__device__ __forceinline__ bf16x4 bf16_pack(unsigned w0, unsigned w1) {
  bf16x4 v;
  v[0] = (short)(w0 & 0xFFFFu);
  v[1] = (short)(w0 >> 16);
  v[2] = (short)(w1 & 0xFFFFu);
  v[3] = (short)(w1 >> 16);
  return v;
}

// FP16 operands (the FP16 hi/lo form): round-to-nearest-even binary16 bits.
typedef _Float16 f16x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ unsigned f16_rn(float f) {
  return (unsigned)__builtin_bit_cast(unsigned short, (_Float16)f);
}
__device__ __forceinline__ float f16_val(unsigned h) {
  return (float)__builtin_bit_cast(_Float16, (unsigned short)h);
}
// f = hi + lo + O(2^-22 |f|) for |f| in the normal binary16 range
__device__ __forceinline__ void f16_split(float f, unsigned &hi, unsigned &lo) {
  hi = f16_rn(f);
  lo = f16_rn(f - f16_val(hi));
}
__device__ __forceinline__ f16x4 f16_pack(unsigned w0, unsigned w1) {
  f16x4 v;
  v[0] = __builtin_bit_cast(_Float16, (unsigned short)(w0 & 0xFFFFu));
  v[1] = __builtin_bit_cast(_Float16, (unsigned short)(w0 >> 16));
  v[2] = __builtin_bit_cast(_Float16, (unsigned short)(w1 & 0xFFFFu));
  v[3] = __builtin_bit_cast(_Float16, (unsigned short)(w1 >> 16));
  return v;
}


// Scale of the phase operand in the FP16 form: |root| <= 1 becomes <= 2^14,
// well inside binary16's range, with the lo part normal for components above
// ~2^-10.
constexpr float kPhaseScaleF16 = 16384.0f;

// Geometry of one cluster set that a kernel shape's LDS size depends on.
struct DenseGeom {
  int ng, ntri, L;
  int npairs;            // (K, L) atom pairs
  int klf[kDenseMaxPairs + 1];  // first coarse group of each pair, then ng
};

// One kernel shape: precision, Nbands, modes per block (IB), q'' per tile
// (TQ), MFMA path and, for k_dense2, the groups per MFMA k-step.
struct DenseShape {
  int prec, nb, IB, TQ;
  int mfma;              // 0 VALU, 1 FP32 MFMA, 2 BF16 hi/lo MFMA, 3 FP16 hi/lo MFMA
  int kg;                // 0: k_dense; > 0: k_dense2 with kg coarse groups per k-step
  size_t (*bytes)(const DenseGeom &);
  size_t (*stat)();
  void (*launch)(const DenseArgs &, int, size_t, hipStream_t);
  int (*per_cu)(size_t);  // resident blocks per CU at this dynamic LDS size
};

// k_dense2 (fp4_dense2.hip): phase-table slots per row for kg groups per
// k-step, the first slot of every pair and the slot of every coarse group,
// its kernel shapes, and the check of the K=32 MFMA lane layout.
int dense2_gpt(const DenseGeom &g, int kg);
void dense2_slots(const DenseGeom &g, int kg, int *pair_off, int *gslot);
const DenseShape *dense2_shapes(int *n);
int dense2_selftest_mfma(double *err);

}  // namespace fp4
