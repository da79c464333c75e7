// Dense form of the four-phonon rates, MFMA kernel with per-wave ownership of
// the branch pairs (i, j).  Same algebra, work items, filter and outputs as
// k_dense in fp4_dense.hip (RTA_driver_4ph, Src/processes.f90:3828-4129 and
// 5174-5311; Vp_pp/pm/mm, Src/processes.f90:142-330); what changes is how a
// tile of 8 q'' is worked through on gfx950:
//
// * Ownership.  Wave w owns the (i, j) pairs [lo(w), hi(w)) of the item, i.e.
//   the Psi columns n = 9 (i, j) + (s, r) of that range.  It computes those
//   columns of every (K, L) pair's Psi GEMM, stores them in its own LDS
//   slice, and contracts exactly the surviving processes whose (i, j) it
//   owns.  The chain GEMM -> contraction -> rate therefore needs no block
//   barrier: one tile costs three barriers (tile data, filter, end of tile)
//   instead of two per (K, L) pair.
// * Screen.  The item's 36 partial sums x1 = w_i +- w'_j are sorted once.
//   The exact FP64 test |((x1 +- w'') - w''')| <= sigma is monotone in x1, so
//   for one (q'', k, l) the survivors are a contiguous run of the sorted x1.
//   A binary search brackets x1 within a conservative window around
//   w''' -+ w'' (the float sigma estimate widened by 1e-4 relative, plus an
//   absolute margin that bounds the FP64 rounding of the mismatch); only
//   candidates inside the bracket take the exact sigma and the exact test, so
//   the processes are exactly the reference's.  Most (q'', k, l) have an
//   empty bracket and cost one search.
// * MFMA operands.  A' (the B operand) sits in LDS as [group][NPAD] words,
//   (re | im << 16) for the 16-bit forms; NPAD is chosen so that the four
//   lane groups of a wave read disjoint bank halves.  The phase operand
//   (A operand) is padded per (K, L) pair to whole k-steps with zeros, so no
//   load is predicated: padded k pairs a zero phase with a finite A' row.
//   BF16 / FP16 hi/lo run on gfx950's v_mfma_f32_16x16x32 (8 groups x re/im
//   per lane group and k-step of 16 groups) or 16x16x16 (k-step of 8
//   groups); the FP32 form on v_mfma_f32_16x16x4f32.
//
// Results are per item and per mode in a fixed order (per-lane sums in queue
// order, then a fixed-tree block reduction), so they do not depend on which
// block ran which item.
#include <algorithm>
#include <cmath>

#include "../../include/fp4.h"
#include "fp4_dense.hpp"
#include "fp4_dense_dev.hpp"

namespace fp4 {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

// Lane layout of v_mfma_f32_16x16x32_bf16 / _f16 used below: A[i][k] in lane
// i + 16 (k / 8), element k % 8; B[k][j] in lane j + 16 (k / 8), element
// k % 8; D[i][j] in lane j + 16 (i / 4), register i % 4.  Checked exactly
// with small integers (all products and sums are exact in FP32).
__global__ void k_mfma_x32_selftest(double *err, int f16) {
  const int l = threadIdx.x;
  const int i = l & 15, kg = l >> 4;
  unsigned short av[8], bv[8];
  for (int q = 0; q < 8; q++) {
    const int k = 8 * kg + q;
    const float x = (float)(1 + i + 3 * k), y = (float)(2 + 5 * (l & 15) - k);
    if (f16) {
      av[q] = __builtin_bit_cast(unsigned short, (_Float16)x);
      bv[q] = __builtin_bit_cast(unsigned short, (_Float16)y);
    } else {
      av[q] = (unsigned short)bf16_rn(x);
      bv[q] = (unsigned short)bf16_rn(y);
    }
  }
  uint4 ua, ub;
  ua.x = av[0] | ((unsigned)av[1] << 16);
  ua.y = av[2] | ((unsigned)av[3] << 16);
  ua.z = av[4] | ((unsigned)av[5] << 16);
  ua.w = av[6] | ((unsigned)av[7] << 16);
  ub.x = bv[0] | ((unsigned)bv[1] << 16);
  ub.y = bv[2] | ((unsigned)bv[3] << 16);
  ub.z = bv[4] | ((unsigned)bv[5] << 16);
  ub.w = bv[6] | ((unsigned)bv[7] << 16);
  floatx4 acc = {0.f, 0.f, 0.f, 0.f};
  if (f16)
    acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, ua), __builtin_bit_cast(f16x8, ub), acc,
                                                 0, 0, 0);
  else
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, ua), __builtin_bit_cast(bf16x8, ub),
                                                  acc, 0, 0, 0);
  double e = 0.0;
  for (int r = 0; r < 4; r++) {
    const int ii = 4 * (l >> 4) + r, jj = l & 15;
    double want = 0.0;
    for (int k = 0; k < 32; k++) want += (double)(1 + ii + 3 * k) * (double)(2 + 5 * jj - k);
    e = fmax(e, fabs(want - (double)acc[r]));
  }
  for (int o = 32; o > 0; o >>= 1) e = fmax(e, __shfl_xor(e, o, 64));
  if (l == 0) *err = e;
}

int dense2_selftest_mfma(double *err) {
  double *d = nullptr;
  if (hipMalloc(&d, sizeof(double)) != hipSuccess) return FP4_ERR_NOMEM;
  double eb = 0.0, eh = 0.0;
  hipLaunchKernelGGL(k_mfma_x32_selftest, dim3(1), dim3(64), 0, 0, d, 0);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpy(&eb, d, sizeof(double), hipMemcpyDeviceToHost);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_mfma_x32_selftest, dim3(1), dim3(64), 0, 0, d, 1);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(&eh, d, sizeof(double), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  *err = std::max(eb, eh);
  return e == hipSuccess ? 0 : FP4_ERR_DEVICE;
}

// Which (i, j) pairs (index il*NB + j, 0..NIJ-1) each wave owns, how many
// 16-column MFMA tiles that takes, and the padded row length of A'.
template <int NIJ>
struct Own {
  static constexpr int NWV = NT_DENSE / 64;
  __host__ __device__ static constexpr int lo(int w) { return (w * NIJ) / NWV; }
  __host__ __device__ static constexpr int hi(int w) { return ((w + 1) * NIJ) / NWV; }
  static constexpr int ctw() {
    int c = 1;
    for (int w = 0; w < NWV; w++) {
      const int t = (9 * (hi(w) - lo(w)) + 15) / 16;
      if (t > c) c = t;
    }
    return c;
  }
  static constexpr int CTW = ctw();
  // every wave's tiles lie inside the row; NPAD = 4 mod 8 words puts the
  // rows of two consecutive 16-bit-form lane groups (4 groups apart) on
  // disjoint bank halves, and NPAD = 8 mod 16 does so for 2 groups apart
  static constexpr int npad(int mod, int rem) {
    int p = 9 * NIJ;
    for (int w = 0; w < NWV; w++) {
      const int e = 9 * lo(w) + 16 * CTW;
      if (e > p) p = e;
    }
    while (p % mod != rem) p++;
    return p;
  }
};

// One tile's data; the kernel keeps two, the current tile's and the next's.
template <int NB>
struct TileBuf {
  C2<float> e3[8][NB][NB];   // [t][comp][k], conj applied
  C2<float> e4[8][NB][NB];   // [t][comp][l], conj applied
  double Ejj[8][NB], Ess[8][NB], fb3[8][NB], fb4[8][NB];
  double inv3[8][NB], inv4[8][NB];      // 1 / w'' and 1 / w''', 0 at or below the 1.25 rad/ps cut
  double isg[8][NB][NB];                // 1 / clamped exact sigma where a candidate survived
  unsigned long long mask[8][NB][NB];   // surviving (i, j) of each (t, k, l)
  unsigned hitword;                     // tile triples with a survivor (bit t)
};

// The filter's inputs of one tile: frequencies and velocities of phonons 3
// and 4, staged two tiles ahead of the filter.
template <int NB>
struct FiltIn {
  double Ej[8][NB], Es[8][NB];
  double vj[8][NB][3], vs[8][NB][3];
};

template <int NB, int IB, int CTW>
struct Smem2 {
  static constexpr int NIJ = IB * NB;
  static constexpr int TQ = 8;
  static constexpr int NW = NT_DENSE / 64;
  static constexpr int QW = 128;    // survivor codes per wave and chunk
  TileBuf<NB> tb[2];
  FiltIn<NB> fi[2];
  double xs[NIJ];                       // x1 of the item, ascending
  int perm[NIJ];                        // (i, j) index of each sorted entry
  double x1[NIJ];
  double inv12[NIJ];                    // 1 / (w_i w'_j), 0 when w'_j is at or below the cut
  double fb1[NB], Eii[NB], om[IB];
  int act[IB];
  double2 e1[NB][IB];
  double2 e2[NB][NB];
  int vj[NT_DENSE], vs[NT_DENSE];
  int wsum[NW];
  int coop_ctr[2];   // chunks of the next tile's filter and phase work taken so far, by buffer
  long item;
  double red[NW][2][3];                 // per wave: sums of its (at most) two modes
  unsigned long long base_ok;           // (i, j) pairs allowed for the whole item
};

// LDS used only in the q'' loop: each wave's Psi slice for one (K, L) pair
// and its survivor codes (task | ij << 16)
template <int CTW>
struct Run2 {
  C2<float> psi[NT_DENSE / 64][8][16 * CTW];
  unsigned q[NT_DENSE / 64][128];
};
// LDS used only while folding A': one wave's sums over the clusters of a
// fine group and their contraction with the first eigenvectors
template <int IB>
struct SetupW {
  float2 F[81];
  double2 b1[IB][27];
};

// Dynamic LDS: U = the cluster coefficients and the per-wave set-up scratch,
// or Run2 in the q'' loop (never both); A' in the B layout; the phase operand;
// roots, group vectors and phase-table slots.
struct Dyn2 {
  size_t U, W, A, ph, rt, gd, gs, end;
};
template <int IB, int CTW>
__host__ __device__ inline Dyn2 dyn2_layout(size_t stat, int npad, int ng, int ntri, int gpt, int L_lds,
                                            int abytes) {
  Dyn2 d{};
  d.U = align16(stat);
  d.W = align16(d.U + sizeof(C2<float>) * (size_t)ntri);
  const size_t setup_end = d.W + sizeof(SetupW<IB>) * (NT_DENSE / 64);
  const size_t run_end = d.U + sizeof(Run2<CTW>);
  d.A = align16(setup_end > run_end ? setup_end : run_end);
  d.ph = align16(d.A + (size_t)abytes * ng * npad);         // A': word planes or float2 [ng][npad]
  d.rt = align16(d.ph + (size_t)256 * gpt);                 // phases: 2 tiles x 2 planes x 16 rows x gpt words
  d.gd = align16(d.rt + sizeof(double2) * (size_t)L_lds);
  d.gs = align16(d.gd + sizeof(int) * 3 * (size_t)ng);
  d.end = align16(d.gs + sizeof(int) * (size_t)ng);
  return d;
}

// acc += a * b with explicit FMAs (the library is built without contraction)
__device__ __forceinline__ void cmac(C2<float> &acc, C2<float> a, C2<float> b) {
  acc.re = fmaf(a.re, b.re, acc.re);
  acc.re = fmaf(-a.im, b.im, acc.re);
  acc.im = fmaf(a.re, b.im, acc.im);
  acc.im = fmaf(a.im, b.re, acc.im);
}

// e^{2 pi i k / L} table index of k: a mask when L is a power of two
__device__ __forceinline__ int root_index(int k, int L, int lmask) { return lmask >= 0 ? (k & lmask) : fmod_pos(k, L); }

// The (i, j) range of one wave spans at most two modes i.
template <int NIJ, int NB>
constexpr bool two_modes_per_wave() {
  for (int w = 0; w < NT_DENSE / 64; w++) {
    const int lo = (w * NIJ) / (NT_DENSE / 64), hi = ((w + 1) * NIJ) / (NT_DENSE / 64);
    if (hi > lo && (hi - 1) / NB - lo / NB > 1) return false;
  }
  return true;
}

constexpr double kInvSqrtPi = 0.56418958354775628695;   // 1 / sqrt(pi)

// Exclusive prefix sum over the wave of counts 0 <= v < 2^BITS and their
// total, by bit planes: a ballot and a lane-mask population count per bit,
// no cross-lane data movement.
template <int BITS>
__device__ __forceinline__ int wave_excl_scan_small(int v, int &total) {
  int pre = 0;
  total = 0;
#pragma unroll
  for (int b = 0; b < BITS; b++) {
    const unsigned long long bal = __ballot((v >> b) & 1);
    pre += (int)__builtin_amdgcn_mbcnt_hi((unsigned)(bal >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)bal, 0u)) << b;
    total += __popcll(bal) << b;
  }
  return pre;
}

// Phase profile of the profiling build (FP4_PHASE_PROF, libfp4prof.so): per
// wave, shader clock cycles (s_memtime) spent in each phase, summed over the
// launch into a.prof in the order of DensePhase.  Reading the clock waits for
// the wave's outstanding LDS and scalar loads, so the profiled kernel runs
// somewhat slower than the product build.
#ifdef FP4_PHASE_PROF
#define PROF_INIT                                    \
  unsigned long long prof_acc[kDensePhases] = {};    \
  unsigned long long prof_t = __builtin_readcyclecounter();
#define PROF(ph)                                                 \
  do {                                                           \
    const unsigned long long now_ = __builtin_readcyclecounter(); \
    prof_acc[ph] += now_ - prof_t;                               \
    prof_t = now_;                                               \
  } while (0)
#define PROF_FLUSH                                                                  \
  do {                                                                              \
    if (lane == 0)                                                                  \
      for (int p_ = 0; p_ < kDensePhases; p_++) atomicAdd(&a.prof[p_], prof_acc[p_]); \
  } while (0)
#else
#define PROF_INIT
#define PROF(ph) \
  do {           \
  } while (0)
#define PROF_FLUSH \
  do {             \
  } while (0)
#endif

__device__ __forceinline__ void lds_wave_sync() {
  // LDS traffic of one wave is processed in order; this waits for the
  // wave's own outstanding LDS operations and stops the compiler from
  // moving LDS accesses across the point.
  __asm__ volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
}

// A' of the coarse groups gi = wave, wave + NW, ...: each wave folds whole
// groups (fine groups in order) with wave-local synchronisation only, into
// the B layout.
template <int NB, int IB, int MF, int NPAD, class S>
__device__ __forceinline__ void fold_groups(const DGroup *g, const FGroup *fg, const double *phi, int ng, float a_scale,
                                         const S &s, SetupW<IB> &sw, const C2<float> *s_c, unsigned *Bh,
                                         unsigned *Bl, float *Bf, int wave, int lane) {
  constexpr int NOUT2 = IB * NB * 9, NW = NT_DENSE / 64;
  constexpr int NPL = (NOUT2 + 63) / 64;   // A' entries of a group per lane
  for (int gi = wave; gi < ng; gi += NW) {
    const DGroup G = g[gi];
    // FP32 partial sums (A' is rounded to FP32 or to BF16/FP16 pairs afterwards)
    float2 acc[NPL];
#pragma unroll
    for (int k = 0; k < NPL; k++) acc[k] = make_float2(0.f, 0.f);
    for (int fi = G.f_first; fi < G.f_first + G.f_count; fi++) {
      const FGroup F = fg[fi];
      for (int e = lane; e < 81; e += 64) {
        double sr = 0.0, si = 0.0;
        for (int ci = F.c_first; ci < F.c_first + F.c_count; ci++) {
          const double p0 = phi[(size_t)ci * 81 + e];
          const C2<float> c0 = s_c[ci];
          sr += p0 * (double)c0.re;
          si += p0 * (double)c0.im;
        }
        sw.F[e] = make_float2((float)sr, (float)si);
      }
      lds_wave_sync();
      for (int x = lane; x < IB * 27; x += 64) {
        const int il = x / 27, tsr = x % 27;
        double2 sacc = make_double2(0.0, 0.0);
#pragma unroll
        for (int uu = 0; uu < 3; uu++) {
          const float2 f = sw.F[uu + 3 * tsr];
          const double2 v = zm(make_double2((double)f.x, (double)f.y), s.e1[3 * F.I + uu][il]);
          sacc.x += v.x;
          sacc.y += v.y;
        }
        sw.b1[il][tsr] = sacc;
      }
      lds_wave_sync();
#pragma unroll
      for (int k = 0; k < NPL; k++) {
        const int n = lane + 64 * k;
        if (n < NOUT2) {
          const int il = n / (NB * 9), rem = n % (NB * 9), j = rem / 9, sr = rem % 9;
          double2 v3 = make_double2(0.0, 0.0);
#pragma unroll
          for (int t = 0; t < 3; t++) {
            const double2 v = zm(sw.b1[il][t + 3 * sr], s.e2[3 * F.J + t][j]);
            v3.x += v.x;
            v3.y += v.y;
          }
          acc[k].x += (float)v3.x;
          acc[k].y += (float)v3.y;
        }
      }
      lds_wave_sync();
    }
#pragma unroll
    for (int k = 0; k < NPL; k++) {
      const int n = lane + 64 * k;
      if (n < NOUT2) {
        if constexpr (MF == 1) {
          Bf[((size_t)gi * NPAD + n) * 2] = (float)acc[k].x;
          Bf[((size_t)gi * NPAD + n) * 2 + 1] = (float)acc[k].y;
        } else {
          unsigned rh, rl, ih, il;
          if constexpr (MF >= 3) {
            f16_split((float)acc[k].x * a_scale, rh, rl);
            f16_split((float)acc[k].y * a_scale, ih, il);
          } else {
            bf16_split((float)acc[k].x, rh, rl);
            bf16_split((float)acc[k].y, ih, il);
          }
          Bh[(size_t)gi * NPAD + n] = rh | (ih << 16);
          if constexpr (MF != 4) Bl[(size_t)gi * NPAD + n] = rl | (il << 16);
        }
      }
    }
  }
}

// MF: 1 = FP32 MFMA (KG = 2 groups per k-step), 2 = BF16 hi/lo, 3 = FP16
// hi/lo, 4 = FP16 hi/lo phases against one FP16 A' plane (KG = 16 on the
// K=32 instructions, 8 on the K=16 ones).
template <int NB, int IB, int MF, int KG>
__global__ void __launch_bounds__(NT_DENSE) k_dense2(DenseArgs a) {
  constexpr int NIJ = IB * NB;
  using O = Own<NIJ>;
  constexpr int CTW = O::CTW;
  constexpr int GPL = (MF == 1) ? 1 : KG / 4;        // groups per lane group and k-step (16-bit forms)
  constexpr int NPAD = (MF == 1) ? O::npad(1, 0) : (GPL == 4 ? O::npad(8, 4) : O::npad(16, 8));
  using S = Smem2<NB, IB, CTW>;
  constexpr int TQ = S::TQ, NW = S::NW, QW = S::QW, QPT = QW / 64;
  static_assert(QW == 128 && TQ == 8, "Run2 sizes");
  constexpr int NOUT2 = NIJ * 9;
  constexpr int NA = NB / 3;
  constexpr int NKLT = TQ * NB * NB;   // (t, k, l) tasks of a tile
  static_assert(NIJ <= 64, "candidate mask is 64 bits");
  static_assert(NKLT < 65536, "task index fits 16 bits");
  static_assert(MF == 1 ? KG == 2 : (KG == 16 || KG == 8), "k-step shape");
  static_assert(two_modes_per_wave<NIJ, NB>(), "each wave's (i, j) range spans at most two modes");
  constexpr unsigned long long kJall = (1ull << NB) - 1ull;
  constexpr unsigned long long kRep = rep_mask(IB, NB);
  const float psi_unscale = (MF >= 3) ? 1.0f / (kPhaseScaleF16 * a.a_scale) : 1.0f;
  constexpr int ABYTES = (MF == 4) ? 4 : 8;   // A' bytes per entry

  extern __shared__ __align__(16) unsigned char dyn[];
  S &s = *reinterpret_cast<S *>(dyn);
  const Dyn2 dl = dyn2_layout<IB, CTW>(a.smem_static, NPAD, a.ng, a.ntri, a.gpt, a.roots_lds ? a.L : 0, ABYTES);
  unsigned *Bh = reinterpret_cast<unsigned *>(dyn + dl.A);   // 16-bit forms: hi plane [ng][NPAD]
  unsigned *Bl = Bh + (size_t)a.ng * NPAD;                    // lo plane (not in the FP16X2 form)
  float *Bf = reinterpret_cast<float *>(dyn + dl.A);          // FP32: [ng][NPAD][re, im]
  C2<float> *s_c = reinterpret_cast<C2<float> *>(dyn + dl.U);
  Run2<CTW> &run = *reinterpret_cast<Run2<CTW> *>(dyn + dl.U);
  // phase operand of tile buffer b: 16-bit forms [16 rows][gpt] hi plane at
  // Ph + 32 b gpt, lo plane 16 gpt words later; FP32 [16 rows][gpt][qq] at
  // Pf + 32 b gpt
  unsigned *Ph = reinterpret_cast<unsigned *>(dyn + dl.ph);
  float *Pf = reinterpret_cast<float *>(dyn + dl.ph);
  double2 *s_rt = reinterpret_cast<double2 *>(dyn + dl.rt);
  int *s_gd = reinterpret_cast<int *>(dyn + dl.gd);
  int *s_gs = reinterpret_cast<int *>(dyn + dl.gs);

  const ModeView &m = a.m;
  const int tid = threadIdx.x;
  // the wave index is uniform within a wave; readfirstlane lets the compiler keep
  // everything derived from it (ownership range, slices) in scalar registers
  const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  SetupW<IB> &sw = reinterpret_cast<SetupW<IB> *>(dyn + dl.W)[wave];
  const double hbarp = kHbar * (double)k1e22;
  const double gpre = hbarp * hbarp * kPi / 8.0;
  const int ijl = O::lo(wave), ijh = O::hi(wave);
  const unsigned long long ownbits =
      ((ijh >= 64) ? ~0ull : ((1ull << ijh) - 1ull)) & ~((1ull << ijl) - 1ull);
  const int colbase = 9 * ijl;
  const int il_lo = ijl / NB;   // first mode of this wave's (i, j) range
  // bits of a lane's survivor count in the queue scan (at most ijh - ijl survivors per task)
  constexpr int QBITS = (NIJ / NW + 1 < 2) ? 1 : (NIJ / NW + 1 < 4) ? 2 : (NIJ / NW + 1 < 8) ? 3 : (NIJ / NW + 1 < 16) ? 4 : 7;

  // once per launch: group vectors, slots, roots; zero phase table (padded
  // slots stay zero) and the A' columns beyond NOUT2 (never written later)
  for (int x = tid; x < 3 * a.ng; x += NT_DENSE) s_gd[x] = a.g[x / 3].d[x % 3];
  for (int x = tid; x < a.ng; x += NT_DENSE) s_gs[x] = a.gslot[x];
  if (a.roots_lds)
    for (int x = tid; x < a.L; x += NT_DENSE) s_rt[x] = a.roots[x];
  for (int x = tid; x < 64 * a.gpt; x += NT_DENSE) Ph[x] = 0u;
  if (tid < 2) {
    s.coop_ctr[tid] = 0;
    s.tb[tid].hitword = 0u;
  }
  for (int x = tid; x < a.ng * (NPAD - NOUT2); x += NT_DENSE) {
    const int r = x / (NPAD - NOUT2), c = NOUT2 + x % (NPAD - NOUT2);
    if constexpr (MF == 1) {
      Bf[((size_t)r * NPAD + c) * 2] = 0.f;
      Bf[((size_t)r * NPAD + c) * 2 + 1] = 0.f;
    } else {
      Bh[(size_t)r * NPAD + c] = 0u;
      if constexpr (MF != 4) Bl[(size_t)r * NPAD + c] = 0u;
    }
  }
  __syncthreads();
  PROF_INIT

  for (;;) {
    if (tid == 0) s.item = atomicAdd(a.counter, 1);
    __syncthreads();
    const long item = s.item;
    __syncthreads();
    if (item >= a.nitems) break;
    const int ii0 = (int)(item % a.nii) + a.ii_lo0;
    long r_ = item / a.nii;
    const int ib = (int)(r_ % a.nib);
    r_ /= a.nib;
    const int ls = (int)(r_ % a.nls);
    const int ch = (int)(r_ / a.nls);
    const int ll0 = a.lls[ls];
    const int qi0 = m.list[ll0] - 1;
    const int s2 = (ch == 2) ? -1 : 1;
    const int s3 = (ch == 0) ? 1 : -1;
    int n1[3], n2[3];
#pragma unroll
    for (int t = 0; t < 3; t++) {
      n1[t] = m.ijk[3 * qi0 + t];
      n2[t] = m.ijk[3 * ii0 + t];
    }
    if (tid < IB) {
      const int i0 = ib * IB + tid;
      const double w = E(m, qi0, i0);
      s.om[tid] = w;
      s.act[tid] = (a.sel[ll0 * NB + i0] != 0) && (w != 0.0) && (w <= m.omega_max);
    }
    if (tid < NB) {
      s.Eii[tid] = E(m, ii0, tid);
      s.fb1[tid] = a.fbose[(size_t)tid * m.nptk + ii0];
    }
    __syncthreads();
    if (tid < NIJ) {
      const int il = tid / NB, j = tid % NB;
      s.x1[tid] = (ch == 2) ? (s.om[il] - s.Eii[j]) : (s.om[il] + s.Eii[j]);
      s.inv12[tid] = (s.Eii[j] <= 1.25 || s.om[il] == 0.0) ? 0.0 : 1.0 / (s.om[il] * s.Eii[j]);
    }
    int any = 0;
#pragma unroll
    for (int il = 0; il < IB; il++) any |= s.act[il];
    PROF(kPhItem);
    if (tid == 0) {
      unsigned long long bo = 0ull;
      for (int il = 0; il < IB; il++)
        for (int j = 0; j < NB; j++)
          if (s.act[il] && s.Eii[j] != 0.0) bo |= 1ull << (il * NB + j);
      s.base_ok = bo;
    }
    double acc_g[2] = {0.0, 0.0}, acc_w[2] = {0.0, 0.0};
    int acc_n[2] = {0, 0};
    int nhit = 0;

    if (any) {
      // ------------- setup 1: eigenvectors of phonons 1 and 2, cluster
      // coefficients, and the sorted partial sums x1
      for (int x = tid; x < NB * IB; x += NT_DENSE) {
        const int comp = x / IB, il = x % IB;
        s.e1[comp][il] = m.eig[((size_t)qi0 * NB + ib * IB + il) * NB + comp];
      }
      for (int x = tid; x < NB * NB; x += NT_DENSE) {
        const int comp = x / NB, j = x % NB;
        double2 e = m.eig[((size_t)ii0 * NB + j) * NB + comp];
        if (ch == 2) e.y = -e.y;
        s.e2[comp][j] = e;
      }
      for (int ci = tid; ci < a.ntri; ci += NT_DENSE) {
        const DCluster c = a.c[ci];
        const int k = s2 * (n2[0] * c.pjl[0] + n2[1] * c.pjl[1] + n2[2] * c.pjl[2]) -
                      (n1[0] * c.pl[0] + n1[1] * c.pl[1] + n1[2] * c.pl[2]);
        const double2 rt = a.roots[root_index(k, a.L, a.lmask)];
        s_c[ci] = C2<float>{(float)(c.minv * rt.x), (float)(c.minv * rt.y)};
      }
      if (tid < NIJ) {
        const double v = s.x1[tid];
        int rk = 0;
        for (int p = 0; p < NIJ; p++) {
          const double w = s.x1[p];
          rk += (w < v || (w == v && p < tid)) ? 1 : 0;
        }
        s.xs[rk] = v;
        s.perm[rk] = tid;
      }
      __syncthreads();
      // ------------- setup 2: A' of every coarse group, into the B layout
      fold_groups<NB, IB, MF, NPAD>(a.g, a.f, a.phi, a.ng, a.a_scale, s, sw, s_c, Bh, Bl, Bf, wave, lane);
      __syncthreads();
      PROF(kPhSetup);

      // Registers carrying a later tile's global data, loaded ahead and
      // written to LDS when that tile is staged.
      constexpr bool PF = (TQ * NB * NB <= NT_DENSE) && (4 * TQ * NB <= NT_DENSE);
      double2 pf_e3 = make_double2(0.0, 0.0), pf_e4 = make_double2(0.0, 0.0);
      double pf_s = 0.0;
      bool have_pf = false;
      auto fetch = [&](int tb, int tqb) {
        if (tid < TQ * NB * NB) {   // eigenvector components of phonons 3 and 4
          const int t = tid / (NB * NB), comp = (tid / NB) % NB, b = tid % NB;
          if (t < tqb) {
            pf_e3 = m.eig[((size_t)s.vj[tb + t] * NB + b) * NB + comp];
            pf_e4 = m.eig[((size_t)s.vs[tb + t] * NB + b) * NB + comp];
          }
        }
        if (tid < 4 * TQ * NB) {    // one of w'', w''', f'', f'''
          const int f = tid / (TQ * NB), y = tid % (TQ * NB), t = y / NB, b = y % NB;
          if (t < tqb) {
            const int q = (f == 0 || f == 2) ? s.vj[tb + t] : s.vs[tb + t];
            pf_s = (f < 2) ? E(m, q, b) : a.fbose[(size_t)b * m.nptk + q];
          }
        }
      };
      // filter inputs: one of w'', w''', v''(3), v'''(3) per thread
      constexpr bool PFV = (8 * TQ * NB <= NT_DENSE);
      double pf_v = 0.0;
      auto fetchv = [&](int tb, int tqb) {
        if (tid < 8 * TQ * NB) {
          const int f = tid / (TQ * NB), y = tid % (TQ * NB), t = y / NB, b = y % NB;
          if (t < tqb) {
            const int q = (f == 0 || (f >= 2 && f < 5)) ? s.vj[tb + t] : s.vs[tb + t];
            pf_v = (f < 2) ? E(m, q, b) : Vel(m, q, b, (f < 5) ? f - 2 : f - 5);
          }
        }
      };
      auto put_v = [&](FiltIn<NB> &F, int f, int t, int b, double v) {
        if (f == 0)
          F.Ej[t][b] = v;
        else if (f == 1)
          F.Es[t][b] = v;
        else if (f < 5)
          F.vj[t][b][f - 2] = v;
        else
          F.vs[t][b][f - 5] = v;
      };
      auto stagev = [&](int tqb, int fb) {   // from pf_v
        if (tid < 8 * TQ * NB) {
          const int f = tid / (TQ * NB), y = tid % (TQ * NB), t = y / NB, b = y % NB;
          if (t < tqb) put_v(s.fi[fb], f, t, b, pf_v);
        }
      };
      auto stagev_direct = [&](int tb, int tqb, int fb) {
        for (int x = tid; x < 8 * TQ * NB; x += NT_DENSE) {
          const int f = x / (TQ * NB), y = x % (TQ * NB), t = y / NB, b = y % NB;
          if (t < tqb) {
            const int q = (f == 0 || (f >= 2 && f < 5)) ? s.vj[tb + t] : s.vs[tb + t];
            put_v(s.fi[fb], f, t, b, (f < 2) ? E(m, q, b) : Vel(m, q, b, (f < 5) ? f - 2 : f - 5));
          }
        }
      };
      // tile data of (tb, tqb) into buffer bf: eigenvectors of phonons 3 and
      // 4, their frequencies, Bose-Einstein factors and reciprocals
      auto stage = [&](int tb, int tqb, int bf) {
        TileBuf<NB> &B = s.tb[bf];
        if constexpr (PF) {
          if (tid < TQ * NB * NB) {
            const int t = tid / (NB * NB), comp = (tid / NB) % NB, b = tid % NB;
            C2<float> v3 = {0.f, 0.f}, v4 = {0.f, 0.f};
            if (t < tqb) {
              v3 = C2<float>{(float)pf_e3.x, (float)((ch != 0) ? -pf_e3.y : pf_e3.y)};
              v4 = C2<float>{(float)pf_e4.x, (float)(-pf_e4.y)};
            }
            B.e3[t][comp][b] = v3;
            B.e4[t][comp][b] = v4;
          }
          if (tid < 4 * TQ * NB) {
            const int f = tid / (TQ * NB), y = tid % (TQ * NB), t = y / NB, b = y % NB;
            if (t < tqb) {
              const double iv = (pf_s > 1.25) ? 1.0 / pf_s : 0.0;
              if (f == 0) {
                B.Ejj[t][b] = pf_s;
                B.inv3[t][b] = iv;
              } else if (f == 1) {
                B.Ess[t][b] = pf_s;
                B.inv4[t][b] = iv;
              } else if (f == 2) {
                B.fb3[t][b] = pf_s;
              } else {
                B.fb4[t][b] = pf_s;
              }
            }
          }
        } else {
          for (int x = tid; x < TQ * NB * NB; x += NT_DENSE) {
            const int t = x / (NB * NB), comp = (x / NB) % NB, b = x % NB;
            C2<float> v3 = {0.f, 0.f}, v4 = {0.f, 0.f};
            if (t < tqb) {
              const double2 e3 = m.eig[((size_t)s.vj[tb + t] * NB + b) * NB + comp];
              const double2 e4 = m.eig[((size_t)s.vs[tb + t] * NB + b) * NB + comp];
              v3 = C2<float>{(float)e3.x, (float)((ch != 0) ? -e3.y : e3.y)};
              v4 = C2<float>{(float)e4.x, (float)(-e4.y)};
            }
            B.e3[t][comp][b] = v3;
            B.e4[t][comp][b] = v4;
          }
          for (int x = tid; x < TQ * NB; x += NT_DENSE) {
            const int t = x / NB, b = x % NB;
            if (t < tqb) {
              const int jv = s.vj[tb + t], sv = s.vs[tb + t];
              const double wj = E(m, jv, b), ws = E(m, sv, b);
              B.Ejj[t][b] = wj;
              B.Ess[t][b] = ws;
              B.inv3[t][b] = (wj > 1.25) ? 1.0 / wj : 0.0;
              B.inv4[t][b] = (ws > 1.25) ? 1.0 / ws : 0.0;
              B.fb3[t][b] = a.fbose[(size_t)b * m.nptk + jv];
              B.fb4[t][b] = a.fbose[(size_t)b * m.nptk + sv];
            }
          }
        }
      };
      // phase operand of tile (tb, tqb) into buffer bf, tasks x0 .. x0 + 63:
      // row (t, part), slot of group gi
      auto phases = [&](int tb, int tqb, int bf, int x0) {
        {
          const int x = x0 + lane;
          if (x >= TQ * a.ng) return;
          const int t = x / a.ng, gi = x % a.ng;
          double2 rt = make_double2(0.0, 0.0);
          if (t < tqb) {
            const int jv = s.vj[tb + t];
            const int k = s3 * (m.ijk[3 * jv] * s_gd[3 * gi] + m.ijk[3 * jv + 1] * s_gd[3 * gi + 1] +
                                m.ijk[3 * jv + 2] * s_gd[3 * gi + 2]);
            const int kk = root_index(k, a.L, a.lmask);
            rt = a.roots_lds ? s_rt[kk] : a.roots[kk];
          }
          const int slot = s_gs[gi];
          if constexpr (MF == 1) {
            // part 0: (re, -im), part 1: (im, re) over (qq = 0, 1)
            float *pb = Pf + (size_t)bf * 32 * a.gpt;
            float *r0 = pb + ((size_t)(2 * t) * a.gpt + slot) * 2;
            float *r1 = pb + ((size_t)(2 * t + 1) * a.gpt + slot) * 2;
            r0[0] = (float)rt.x;
            r0[1] = -(float)rt.y;
            r1[0] = (float)rt.y;
            r1[1] = (float)rt.x;
          } else {
            unsigned rh, rl, ih, il;
            if constexpr (MF >= 3) {
              f16_split((float)rt.x * kPhaseScaleF16, rh, rl);
              f16_split((float)rt.y * kPhaseScaleF16, ih, il);
            } else {
              bf16_split((float)rt.x, rh, rl);
              bf16_split((float)rt.y, ih, il);
            }
            // the sign flip is the top bit in both 16-bit formats
            unsigned *ph = Ph + (size_t)bf * 32 * a.gpt, *pl = ph + (size_t)16 * a.gpt;
            ph[(size_t)(2 * t) * a.gpt + slot] = rh | ((ih ^ 0x8000u) << 16);
            ph[(size_t)(2 * t + 1) * a.gpt + slot] = ih | (rh << 16);
            pl[(size_t)(2 * t) * a.gpt + slot] = rl | ((il ^ 0x8000u) << 16);
            pl[(size_t)(2 * t + 1) * a.gpt + slot] = il | (rl << 16);
          }
        }
      };
      // filter of tile (tb, tqb) into buffer bf from the staged inputs fi[fb]:
      // one (t, k, l) task per thread; screen by the sorted x1, exact sigma
      // and exact test only inside the bracket
      auto filter = [&](int tb, int tqb, int bf, int fb, int x0) {
        TileBuf<NB> &B = s.tb[bf];
        const FiltIn<NB> &F = s.fi[fb];
        unsigned hitbits = 0u;
        {
          const int x = x0 + lane;
          const int l = x % NB, k = (x / NB) % NB, t = x / (NB * NB);
          unsigned long long mask = 0ull;
          if (x < NKLT && t < tqb) {
            const int jj0 = s.vj[tb + t], ss0 = s.vs[tb + t];
            const double ej = F.Ej[t][k], es = F.Es[t][l];
            if (ej != 0.0 && es != 0.0 && !(ch != 0 && ss0 == jj0 && l < k)) {
              double v3[3], v4[3];
#pragma unroll
              for (int d = 0; d < 3; d++) {
                v3[d] = F.vj[t][k][d];
                v4[d] = F.vs[t][l][d];
              }
              const double odp = (ch == 0) ? ej : -ej;
              const double otp = es;
              const double c = otp - odp;
              const float sf = sigma_estimate(a.bf, ch, v3, v4, a.sbf);
              // sigma <= 1.0001 * estimate; 2^-40 (|x1| + |w''| + |w'''|) bounds the
              // rounding of the two FP64 additions of the mismatch and of c -+ W
              const double xsmax = fmax(fabs(s.xs[0]), fabs(s.xs[NIJ - 1]));
              const double W = (double)sf * 1.0001 + 0x1p-40 * (xsmax + fabs(odp) + fabs(otp));
              const double lo_v = c - W, hi_v = c + W;
              // first p with xs[p] >= lo_v: count the fences xs[4q + 3] below
              // lo_v (independent loads), then two binary steps in the block
              constexpr int NFEN = NIJ / 4;
              int p = 0;
#pragma unroll
              for (int q = 0; q < NFEN; q++) p += (s.xs[4 * q + 3] < lo_v) ? 4 : 0;
              if (p + 2 <= NIJ && s.xs[p + 1] < lo_v) p += 2;
              if (p + 1 <= NIJ && s.xs[p] < lo_v) p += 1;
              if (p < NIJ && s.xs[p] <= hi_v) {
                const double sg = tile_sigma(m, ch, v3, v4);
                unsigned long long hitm = 0ull;
                for (; p < NIJ; p++) {
                  const double xv = s.xs[p];
                  if (xv > hi_v) break;
                  const double delta = (xv + odp) - otp;
                  if (fabs(delta) <= sg) hitm |= 1ull << s.perm[p];
                }
                // exclusions of the reference's loops as bit masks over (i, j)
                unsigned long long excl = 0ull;
                if (ch != 1 && jj0 == ii0) excl |= kRep * (kJall & ~((2ull << k) - 1ull));   // j > k
                if (ch == 1) {
                  if (ii0 == jj0) excl |= kRep << k;                                         // j == k
                  if (ii0 == ss0) excl |= kRep << l;                                         // j == l
                  const int ilk = k - ib * IB, ill = l - ib * IB;
                  if (qi0 == jj0 && ilk >= 0 && ilk < IB) excl |= kJall << (ilk * NB);      // i == k
                  if (qi0 == ss0 && ill >= 0 && ill < IB) excl |= kJall << (ill * NB);      // i == l
                }
                mask = hitm & s.base_ok & ~excl;
                if (mask) B.isg[t][k][l] = 1.0 / ((sg <= (double)kClamp) ? 1.0 / sqrt(kPi) : sg);
              }
            }
          }
          if (x < NKLT) {
            B.mask[t][k][l] = mask;
            if (mask) hitbits |= 1u << t;
          }
        }
        unsigned hw = 0u;
#pragma unroll
        for (int t = 0; t < TQ; t++)
          if (__ballot((hitbits >> t) & 1u) != 0ull) hw |= 1u << t;
        if (lane == 0 && hw) atomicOr(&B.hitword, hw);
      };
      // the filter and phase work of one tile in chunks of 64 tasks, taken by
      // whichever wave is free first (each task writes only its own results)
      auto coop = [&](int tb, int tqb, int bf, int fb) {
        const int nfc = (NKLT + 63) / 64, npc = (TQ * a.ng + 63) / 64;
        for (;;) {
          int c = 0;
          if (lane == 0) c = atomicAdd(&s.coop_ctr[bf], 1);
          c = __shfl(c, 0, 64);
          if (c >= nfc + npc) break;
          if (c < nfc)
            filter(tb, tqb, bf, fb, 64 * c);
          else
            phases(tb, tqb, bf, 64 * (c - nfc));
        }
      };

      // ------------- q'' loop: windows of up to NT_DENSE valid q'', tiles of
      // TQ; tile T's per-wave work and the staging and filter of tile T + 1
      // run between the same two barriers (double-buffered tile data)
      const int jj_first = (ch == 1) ? 0 : ii0;
      int cur = 0;
      // range of the item's x1 for the triple screen
      const double xlo = s.xs[0], xhi = s.xs[NIJ - 1], xabs = fmax(fabs(xlo), fabs(xhi));
      for (int base = jj_first; base < m.nptk; base += NT_DENSE) {
        const int jj = base + tid;
        int ok = 0, ssv = 0;
        if (jj < m.nptk) {
          int n3[3];
#pragma unroll
          for (int t = 0; t < 3; t++) n3[t] = m.ijk[3 * jj + t];
          ssv = partner(m, ch, n1, n2, n3);
          ok = ((ch == 0) || (ssv >= jj)) && triple_may_survive<NB>(m, ch, jj, ssv, xlo, xhi, xabs);
        }
        int tot;
        const int pos = block_excl_scan<NW>(ok, s.wsum, tot);
        if (ok) {
          s.vj[pos] = jj;
          s.vs[pos] = ssv;
        }
        __syncthreads();
        PROF(kPhScan);
        if (tot == 0) continue;
        {
          const int tq0 = min(TQ, tot);
          // filter inputs of the window's first two tiles, then the third's into registers
          stagev_direct(0, tq0, 0);
          if (TQ < tot) stagev_direct(TQ, min(TQ, tot - TQ), 1);
          if constexpr (PFV)
            if (2 * TQ < tot) fetchv(2 * TQ, min(TQ, tot - 2 * TQ));
          if constexpr (PF) fetch(0, tq0);
          stage(0, tq0, cur);
          if constexpr (PF) {
            have_pf = TQ < tot;
            if (have_pf) fetch(TQ, min(TQ, tot - TQ));
          }
          PROF(kPhTile);
          __syncthreads();
          PROF(kPhBar1);
          coop(0, tq0, cur, 0);
          PROF(kPhFilter);
          __syncthreads();
          PROF(kPhBar1);
        }
        for (int t0 = 0; t0 < tot; t0 += TQ) {
          const TileBuf<NB> &B = s.tb[cur];
          const unsigned *Phc = Ph + (size_t)cur * 32 * a.gpt;
          const unsigned *Plc = Phc + (size_t)16 * a.gpt;
          const float *Pfc = Pf + (size_t)cur * 32 * a.gpt;
          // ---- per wave: survivors of the owned (i, j), Psi slice, contraction, rates
          if (ijh > ijl) {
            const unsigned long long *mflat = &B.mask[0][0][0];
            int total = 0;
            for (int cbase = 0; cbase == 0 || cbase < total; cbase += QW) {
              // queue the survivors with positions [cbase, cbase + QW) in
              // (t, k, l, i, j) order; the first pass also counts them all
              int cnt_all = 0;
              for (int x0 = 0; x0 < NKLT; x0 += 64) {
                const int x = x0 + lane;
                unsigned long long mk = (x < NKLT) ? (mflat[x] & ownbits) : 0ull;
                if (__ballot(mk != 0ull) == 0ull) continue;
                int tot_r;
                const int cnt = __popcll(mk);
                int pos = cnt_all + wave_excl_scan_small<QBITS>(cnt, tot_r);
                cnt_all += tot_r;
                if (pos + cnt > cbase && pos < cbase + QW) {
                  while (mk) {
                    const int b = __builtin_ctzll(mk);
                    mk &= mk - 1ull;
                    if (pos >= cbase && pos < cbase + QW) run.q[wave][pos - cbase] = (unsigned)x | ((unsigned)b << 16);
                    pos++;
                  }
                }
              }
              total = cnt_all;
              PROF(kPhQueue);
              if (total == 0) break;
              lds_wave_sync();
              const int qn = min(QW, total - cbase);
              int info[QPT], o_psi[QPT], o_e3[QPT], o_e4[QPT];
              C2<float> V[QPT];
#pragma unroll
              for (int u = 0; u < QPT; u++) {
                const int qi = lane + 64 * u;
                info[u] = (qi < qn) ? (int)run.q[wave][qi] : -1;
                V[u] = C2<float>{0.f, 0.f};
                const int x = info[u] & 0xFFFF, ij = (info[u] >> 16) & 63;
                const int l = x % NB, k = (x / NB) % NB, t = x / (NB * NB);
                o_psi[u] = t * 16 * CTW + 9 * (ij - ijl);
                o_e3[u] = t * NB * NB + k;
                o_e4[u] = t * NB * NB + l;
              }
              const C2<float> *psiw = &run.psi[wave][0][0];
              const C2<float> *e3f = &B.e3[0][0][0];
              const C2<float> *e4f = &B.e4[0][0][0];
              if (!a.onlyharmonic) {
                for (int kl = 0; kl < NA * NA; kl++) {
                  const int g0 = a.klf[kl], g1 = a.klf[kl + 1];
                  if (g1 == g0) continue;
                  floatx4 acc[CTW];
#pragma unroll
                  for (int c = 0; c < CTW; c++) acc[c] = floatx4{0.f, 0.f, 0.f, 0.f};
                  const int n = lane & 15, kg = lane >> 4, row = lane & 15;
                  const int nks = (g1 - g0 + KG - 1) / KG;
                  for (int ks = 0; ks < nks; ks++) {
                    if constexpr (MF == 1) {
                      const int g = g0 + 2 * ks + (kg >> 1), qq = kg & 1;
                      const float av = Pfc[((size_t)row * a.gpt + a.pair_off[kl] + 2 * ks + (kg >> 1)) * 2 + qq];
                      const int gr = (g < g1) ? g : g1 - 1;
                      const float *bp = Bf + ((size_t)gr * NPAD + colbase + n) * 2 + qq;
                      float bv[CTW];
#pragma unroll
                      for (int c = 0; c < CTW; c++) bv[c] = bp[32 * c];
#pragma unroll
                      for (int c = 0; c < CTW; c++) acc[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv[c], acc[c], 0, 0, 0);
                    } else {
                      const int slot = a.pair_off[kl] + KG * ks + GPL * kg;
                      int roff[GPL];
#pragma unroll
                      for (int e = 0; e < GPL; e++) {
                        const int g = g0 + KG * ks + GPL * kg + e;
                        roff[e] = ((g < g1) ? g : g1 - 1) * NPAD + colbase + n;
                      }
                      if constexpr (GPL == 4) {
                        const uint4 ah = *reinterpret_cast<const uint4 *>(Phc + (size_t)row * a.gpt + slot);
                        const uint4 al = *reinterpret_cast<const uint4 *>(Plc + (size_t)row * a.gpt + slot);
                        uint4 bh[CTW], bl[CTW];
#pragma unroll
                        for (int c = 0; c < CTW; c++) {
                          bh[c] = make_uint4(Bh[roff[0] + 16 * c], Bh[roff[1] + 16 * c], Bh[roff[2] + 16 * c],
                                             Bh[roff[3] + 16 * c]);
                          if constexpr (MF != 4)
                            bl[c] = make_uint4(Bl[roff[0] + 16 * c], Bl[roff[1] + 16 * c], Bl[roff[2] + 16 * c],
                                               Bl[roff[3] + 16 * c]);
                        }
                        if constexpr (MF == 4) {
                          // phases hi + lo against the single A' plane
                          const f16x8 xh = __builtin_bit_cast(f16x8, ah), xl = __builtin_bit_cast(f16x8, al);
#pragma unroll
                          for (int c = 0; c < CTW; c++)
                            acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(xh, __builtin_bit_cast(f16x8, bh[c]), acc[c], 0, 0, 0);
#pragma unroll
                          for (int c = 0; c < CTW; c++)
                            acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(xl, __builtin_bit_cast(f16x8, bh[c]), acc[c], 0, 0, 0);
                        } else if constexpr (MF == 3) {
                          const f16x8 xh = __builtin_bit_cast(f16x8, ah), xl = __builtin_bit_cast(f16x8, al);
#pragma unroll
                          for (int c = 0; c < CTW; c++)
                            acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(xh, __builtin_bit_cast(f16x8, bh[c]), acc[c], 0, 0, 0);
#pragma unroll
                          for (int c = 0; c < CTW; c++)
                            acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(xh, __builtin_bit_cast(f16x8, bl[c]), acc[c], 0, 0, 0);
#pragma unroll
                          for (int c = 0; c < CTW; c++)
                            acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(xl, __builtin_bit_cast(f16x8, bh[c]), acc[c], 0, 0, 0);
                        } else {
                          const bf16x8 xh = __builtin_bit_cast(bf16x8, ah), xl = __builtin_bit_cast(bf16x8, al);
#pragma unroll
                          for (int c = 0; c < CTW; c++)
                            acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh, __builtin_bit_cast(bf16x8, bh[c]), acc[c], 0, 0, 0);
#pragma unroll
                          for (int c = 0; c < CTW; c++)
                            acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh, __builtin_bit_cast(bf16x8, bl[c]), acc[c], 0, 0, 0);
#pragma unroll
                          for (int c = 0; c < CTW; c++)
                            acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xl, __builtin_bit_cast(bf16x8, bh[c]), acc[c], 0, 0, 0);
                        }
                      } else {
                        const uint2 ah = *reinterpret_cast<const uint2 *>(Phc + (size_t)row * a.gpt + slot);
                        const uint2 al = *reinterpret_cast<const uint2 *>(Plc + (size_t)row * a.gpt + slot);
                        uint2 bh[CTW], bl[CTW];
#pragma unroll
                        for (int c = 0; c < CTW; c++) {
                          bh[c] = make_uint2(Bh[roff[0] + 16 * c], Bh[roff[1] + 16 * c]);
                          if constexpr (MF != 4) bl[c] = make_uint2(Bl[roff[0] + 16 * c], Bl[roff[1] + 16 * c]);
                        }
                        if constexpr (MF == 4) {
                          const f16x4 xh = __builtin_bit_cast(f16x4, ah), xl = __builtin_bit_cast(f16x4, al);
#pragma unroll
                          for (int c = 0; c < CTW; c++)
                            acc[c] = __builtin_amdgcn_mfma_f32_16x16x16f16(xh, __builtin_bit_cast(f16x4, bh[c]), acc[c], 0, 0, 0);
#pragma unroll
                          for (int c = 0; c < CTW; c++)
                            acc[c] = __builtin_amdgcn_mfma_f32_16x16x16f16(xl, __builtin_bit_cast(f16x4, bh[c]), acc[c], 0, 0, 0);
                        } else if constexpr (MF == 3) {
                          const f16x4 xh = __builtin_bit_cast(f16x4, ah), xl = __builtin_bit_cast(f16x4, al);
#pragma unroll
                          for (int c = 0; c < CTW; c++)
                            acc[c] = __builtin_amdgcn_mfma_f32_16x16x16f16(xh, __builtin_bit_cast(f16x4, bh[c]), acc[c], 0, 0, 0);
#pragma unroll
                          for (int c = 0; c < CTW; c++)
                            acc[c] = __builtin_amdgcn_mfma_f32_16x16x16f16(xh, __builtin_bit_cast(f16x4, bl[c]), acc[c], 0, 0, 0);
#pragma unroll
                          for (int c = 0; c < CTW; c++)
                            acc[c] = __builtin_amdgcn_mfma_f32_16x16x16f16(xl, __builtin_bit_cast(f16x4, bh[c]), acc[c], 0, 0, 0);
                        } else {
                          const bf16x4 xh = __builtin_bit_cast(bf16x4, ah), xl = __builtin_bit_cast(bf16x4, al);
#pragma unroll
                          for (int c = 0; c < CTW; c++)
                            acc[c] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(xh, __builtin_bit_cast(bf16x4, bh[c]), acc[c], 0, 0, 0);
#pragma unroll
                          for (int c = 0; c < CTW; c++)
                            acc[c] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(xh, __builtin_bit_cast(bf16x4, bl[c]), acc[c], 0, 0, 0);
#pragma unroll
                          for (int c = 0; c < CTW; c++)
                            acc[c] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(xl, __builtin_bit_cast(bf16x4, bh[c]), acc[c], 0, 0, 0);
                        }
                      }
                    }
                  }
                  // this wave's Psi columns: rows 4 kg + r = (t, part)
                  {
                    const int t0r = 2 * kg;
#pragma unroll
                    for (int c = 0; c < CTW; c++) {
                      const int col = 16 * c + n;
                      run.psi[wave][t0r][col] = C2<float>{acc[c][0] * psi_unscale, acc[c][1] * psi_unscale};
                      run.psi[wave][t0r + 1][col] = C2<float>{acc[c][2] * psi_unscale, acc[c][3] * psi_unscale};
                    }
                  }
                  lds_wave_sync();
                  PROF(kPhGemm);
                  const int K = kl / NA, L = kl % NA;
#pragma unroll
                  for (int u = 0; u < QPT; u++) {
                    if (info[u] >= 0) {
                      const C2<float> *ps = psiw + o_psi[u];
                      const C2<float> *q3 = e3f + o_e3[u] + 3 * K * NB;
                      const C2<float> *q4 = e4f + o_e4[u] + 3 * L * NB;
                      const C2<float> f4[3] = {q4[0], q4[NB], q4[2 * NB]};
#pragma unroll
                      for (int sx = 0; sx < 3; sx++) {
                        C2<float> inner = {0.f, 0.f};
#pragma unroll
                        for (int rx = 0; rx < 3; rx++) cmac(inner, ps[sx + 3 * rx], f4[rx]);
                        cmac(V[u], inner, q3[sx * NB]);
                      }
                    }
                  }
                  lds_wave_sync();
                  PROF(kPhContract);
                }
              }
              // weights and rates in FP64
#pragma unroll
              for (int u = 0; u < QPT; u++) {
                if (info[u] >= 0) {
                  const int x = info[u] & 0xFFFF, ij = info[u] >> 16;
                  const int l = x % NB, k = (x / NB) % NB, t = x / (NB * NB);
                  const int il = ij / NB, j = ij % NB;
                  const double w = s.om[il], op = s.Eii[j], odp = B.Ejj[t][k], otp = B.Ess[t][l];
                  const double delta = mismatch(ch, w, op, odp, otp);
                  // WP4 of processes.f90:3895-3901 with the divisions as
                  // reciprocals (1 / sigma of the clamped sigma, 1 / (w w'),
                  // 1 / w'', 1 / w'''; the last three are 0 below the cut)
                  const double fp = s.fb1[j], fdp = B.fb3[t][k], ftp = B.fb4[t][l];
                  double occ;
                  if (ch == 0)
                    occ = fp * fdp * (1 + ftp) - (1 + fp) * (1 + fdp) * ftp;
                  else if (ch == 1)
                    occ = fp * (1 + fdp) * (1 + ftp) - (1 + fp) * fdp * ftp;
                  else
                    occ = (1 + fp) * (1 + fdp) * (1 + ftp) - fp * fdp * ftp;
                  const double is = B.isg[t][k][l];
                  const double wpv = occ * exp(-(delta * delta) * (is * is)) * (is * kInvSqrtPi) *
                                     (s.inv12[ij] * B.inv3[t][k] * B.inv4[t][l]);
                  double g = 0.0;
                  if (!a.onlyharmonic) {
                    const double vr = (double)V[u].re, vi = (double)V[u].im;
                    g = gpre * wpv * (vr * vr + vi * vi);
                  }
                  if (a.wop && g != 0.0)
                    wop_add(a, ll0 * NB + ib * IB + il, ch, ii0 * NB + j, s.vj[t0 + t] * NB + k,
                            s.vs[t0 + t] * NB + l, g);
                  const bool second = (il != il_lo);
                  acc_g[0] += second ? 0.0 : g;
                  acc_w[0] += second ? 0.0 : wpv;
                  acc_n[0] += second ? 0 : 1;
                  acc_g[1] += second ? g : 0.0;
                  acc_w[1] += second ? wpv : 0.0;
                  acc_n[1] += second ? 1 : 0;
                }
              }
              lds_wave_sync();
              PROF(kPhRate);
            }
          }
          if (tid == 0) {
            // this tile's hits; its buffer's word and chunk counter are next
            // used two tiles on, after the coming barrier
            nhit += __popc(s.tb[cur].hitword);
            s.tb[cur].hitword = 0u;
            s.coop_ctr[cur] = 0;
          }
          // staging, phase operand and filter of the next tile, into the other buffer
          const int tn = t0 + TQ;
          if (tn < tot) {
            const int tqn = min(TQ, tot - tn);
            const int tn2 = tn + TQ, nt = tn / TQ;   // nt: window tile index of tn
            if (tn2 < tot) {
              // filter inputs of tile tn2, then of the tile after it into registers
              if constexpr (PFV) {
                stagev(min(TQ, tot - tn2), (nt + 1) & 1);
                if (tn2 + TQ < tot) fetchv(tn2 + TQ, min(TQ, tot - tn2 - TQ));
              } else {
                stagev_direct(tn2, min(TQ, tot - tn2), (nt + 1) & 1);
              }
            }
            stage(tn, tqn, cur ^ 1);
            if constexpr (PF) {
              have_pf = tn2 < tot;
              if (have_pf) fetch(tn2, min(TQ, tot - tn2));
            }
            PROF(kPhTile);
            coop(tn, tqn, cur ^ 1, nt & 1);
            PROF(kPhFilter);
          }
          __syncthreads();
          PROF(kPhBar3);
          cur ^= 1;
        }
      }
    }
    // fixed-order block reduction of the per-thread partial sums
#pragma unroll
    for (int sl = 0; sl < 2; sl++) {
      const double g = wave_sum(acc_g[sl]);
      const double w = wave_sum(acc_w[sl]);
      const double n = wave_sum((double)acc_n[sl]);
      if (lane == 0) {
        s.red[wave][sl][0] = g;
        s.red[wave][sl][1] = w;
        s.red[wave][sl][2] = n;
      }
    }
    __syncthreads();
    if (tid < IB) {
      double g = 0.0, w = 0.0, n = 0.0;
#pragma unroll
      for (int wv = 0; wv < NW; wv++) {
        const int sl = tid - O::lo(wv) / NB;   // this wave's slot of mode tid, if any
        if (sl == 0 || sl == 1) {
          g = g + s.red[wv][sl][0];
          w = w + s.red[wv][sl][1];
          n = n + s.red[wv][sl][2];
        }
      }
      a.out_g[item * IB + tid] = g;
      a.out_w[item * IB + tid] = w;
      a.out_n[item * IB + tid] = (int)n;
    }
    const int nh = wave_sum_i(nhit);   // only threads of wave 0 count tile triples
    if (tid == 0) a.out_t[item] = nh;
    __syncthreads();
    PROF(kPhItem);
  }
  PROF_FLUSH;
}

// ------------------------------------------------------------------- host
namespace {

template <int NB, int IB, int MF, int KG>
constexpr int npad_of() {
  constexpr int GPL = (MF == 1) ? 1 : KG / 4;
  using O = Own<IB * NB>;
  return (MF == 1) ? O::npad(1, 0) : (GPL == 4 ? O::npad(8, 4) : O::npad(16, 8));
}

template <int NB, int IB, int MF, int KG>
size_t stat2() {
  return (sizeof(Smem2<NB, IB, Own<IB * NB>::CTW>) + 15) & ~size_t(15);
}

template <int NB, int IB, int MF, int KG>
size_t bytes2(const DenseGeom &g) {
  const int gpt = dense2_gpt(g, KG);
  return dyn2_layout<IB, Own<IB * NB>::CTW>(stat2<NB, IB, MF, KG>(), npad_of<NB, IB, MF, KG>(), g.ng, g.ntri, gpt,
                                             g.L <= kRootsLds ? g.L : 0, MF == 4 ? 4 : 8)
      .end;
}

template <int NB, int IB, int MF, int KG>
void launch2(const DenseArgs &a, int slots, size_t dyn, hipStream_t st) {
  (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_dense2<NB, IB, MF, KG>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
  hipLaunchKernelGGL((k_dense2<NB, IB, MF, KG>), dim3(slots), dim3(NT_DENSE), dyn, st, a);
}

template <int NB, int IB, int MF, int KG>
int per_cu2(size_t dyn) {
  (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_dense2<NB, IB, MF, KG>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, reinterpret_cast<const void *>(&k_dense2<NB, IB, MF, KG>),
                                                   NT_DENSE, dyn) != hipSuccess)
    return 1;
  return std::max(n, 1);
}

#define FP4_SHAPE2(P, NB, IB, MF, KG) \
  {P, NB, IB, 8, MF, KG, &bytes2<NB, IB, MF, KG>, &stat2<NB, IB, MF, KG>, &launch2<NB, IB, MF, KG>, &per_cu2<NB, IB, MF, KG>}

const DenseShape kShapes2[] = {
    FP4_SHAPE2(FP4_PREC_BF16X3, 6, 6, 2, 8),  FP4_SHAPE2(FP4_PREC_BF16X3, 6, 6, 2, 16),
    FP4_SHAPE2(FP4_PREC_BF16X3, 6, 3, 2, 8),  FP4_SHAPE2(FP4_PREC_FP16X3, 6, 3, 3, 8),
    FP4_SHAPE2(FP4_PREC_FP32, 6, 3, 1, 2),    FP4_SHAPE2(FP4_PREC_FP16X2, 6, 3, 4, 8),
    FP4_SHAPE2(FP4_PREC_BF16X3, 3, 3, 2, 8),  FP4_SHAPE2(FP4_PREC_BF16X3, 9, 3, 2, 8),
    FP4_SHAPE2(FP4_PREC_BF16X3, 12, 3, 2, 8), FP4_SHAPE2(FP4_PREC_BF16X3, 12, 2, 2, 8),
    FP4_SHAPE2(FP4_PREC_FP16X3, 6, 6, 3, 8),  FP4_SHAPE2(FP4_PREC_FP16X3, 6, 6, 3, 16),
    FP4_SHAPE2(FP4_PREC_FP16X3, 3, 3, 3, 8),  FP4_SHAPE2(FP4_PREC_FP16X3, 9, 3, 3, 8),
    FP4_SHAPE2(FP4_PREC_FP16X3, 12, 3, 3, 8), FP4_SHAPE2(FP4_PREC_FP16X3, 12, 2, 3, 8),
    FP4_SHAPE2(FP4_PREC_FP16X2, 6, 6, 4, 8),  FP4_SHAPE2(FP4_PREC_FP16X2, 6, 6, 4, 16),
    FP4_SHAPE2(FP4_PREC_FP16X2, 3, 3, 4, 8),  FP4_SHAPE2(FP4_PREC_FP16X2, 9, 3, 4, 8),
    FP4_SHAPE2(FP4_PREC_FP16X2, 12, 3, 4, 8), FP4_SHAPE2(FP4_PREC_FP16X2, 12, 2, 4, 8),
    FP4_SHAPE2(FP4_PREC_FP32, 6, 6, 1, 2),    FP4_SHAPE2(FP4_PREC_FP32, 3, 3, 1, 2),
    FP4_SHAPE2(FP4_PREC_FP32, 9, 3, 1, 2),    FP4_SHAPE2(FP4_PREC_FP32, 12, 3, 1, 2),
    FP4_SHAPE2(FP4_PREC_FP32, 12, 2, 1, 2),
};

}  // namespace

int dense2_gpt(const DenseGeom &g, int kg) {
  int t = 0;
  for (int p = 0; p < g.npairs; p++) {
    const int n = g.klf[p + 1] - g.klf[p];
    t += ((n + kg - 1) / kg) * kg;
  }
  return std::max(4, (t + 3) & ~3);
}

void dense2_slots(const DenseGeom &g, int kg, int *pair_off, int *gslot) {
  int t = 0;
  for (int p = 0; p < g.npairs; p++) {
    pair_off[p] = t;
    for (int gi = g.klf[p]; gi < g.klf[p + 1]; gi++) gslot[gi] = t + (gi - g.klf[p]);
    const int n = g.klf[p + 1] - g.klf[p];
    t += ((n + kg - 1) / kg) * kg;
  }
  pair_off[g.npairs] = t;
}

const DenseShape *dense2_shapes(int *n) {
  *n = (int)(sizeof(kShapes2) / sizeof(kShapes2[0]));
  return kShapes2;
}

}  // namespace fp4
