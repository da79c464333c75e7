// Dense form of the four-phonon rates (RTA_driver_4ph, Src/processes.f90:
// 3828-4129 and 5174-5311): the same processes and weights as the
// per-process form, with the matrix elements (Vp_pp/pm/mm,
// Src/processes.f90:142-330) regrouped so that most of the work is shared
// between processes.
//
// Algebra.  For channel signs (s2, s3) and q''' = q + s2 q' + s3 q'', the
// phase of cluster c is
//     s2 q'.(R_j - R_l) + s3 q''.(R_k - R_l) - q.R_l ,
// so with d = R_k - R_l every cluster factors into a (q, q')-dependent
// coefficient and a q''-dependent root of unity.  Clusters are grouped by
// gamma = (K, L, d) (atoms of phonons 3 and 4 and d): 1522 second-shell
// diamond clusters fall into 34 such groups.  Per work item (channel,
// irreducible q, q') the block builds
//     A'[gamma][i][j][s][r] = sum_{c in gamma} coef_c
//                             sum_{u,t} Phi_c[u,t,s,r] e1[(I,u),i] e2[(J,t),j]
// (the first two eigenvectors are fixed for the whole item), then for each
// tile of q'' values evaluates the skinny GEMM
//     Psi[q''][i][j][(K,s)][(L,r)] = sum_{gamma in (K,L)} root(q''.d) A'[gamma][i][j][s][r]
// and, for the processes that pass the reference's exact filter only,
//     V = sum_{C,D} Psi[q''][i][j][C][D] e3[C,k] e4[D,l].
// The filter is the prefilter's bit-exact FP64 test, so the processes are
// exactly the reference's; Bose-Einstein weights, WP4 and the Gamma sums
// stay in FP64.  T (float or double) is the arithmetic of A', Psi and V.
//
// Scheduling: one persistent 512-thread block per CU pulls work items from
// an atomic counter.  A', the per-cluster coefficients, the phase table and
// a tile of Psi live in LDS (up to ~150 KiB of gfx950's 160 KiB); the kernel
// shape (modes per block, q'' per tile) is the first one in the shape list whose
// LDS fits.  Results are written per item and reduced over q' in a fixed
// order, and each block's partial sums are reduced in a fixed tree order, so
// the output does not depend on which block ran which item.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <type_traits>

#include "../../include/fp4.h"
#include "fp4_dense.hpp"
#include "fp4_dense_dev.hpp"

namespace fp4 {


namespace {
thread_local std::string d_err;
int derr(int code, const char *msg) {
  d_err = msg;
  return code;
}
}  // namespace

const char *dense_error() { return d_err.c_str(); }

struct DenseState {
  int natoms = 0, nb = 0, L = 1, ng = 0, nf = 0, nc = 0;
  DGroup *g = nullptr;
  FGroup *f = nullptr;
  DCluster *c = nullptr;
  double *phi = nullptr;
  int klf[kDenseMaxPairs + 1] = {};   // first coarse group of each (K, L) pair, then ng
  void *scratch = nullptr;
  size_t scratch_bytes = 0;
  void *fbose = nullptr;   // [Nb][nptk] Bose-Einstein factors at the last T
  size_t fbose_bytes = 0;
  void *outb = nullptr;
  size_t out_bytes = 0;
  int *counter = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  long long last_triples = 0, last_items = 0;
  int last_mfma = 0, last_tq = 0, last_ib = 0, last_kg = 0;
  struct Cand {
    int shape, slots;
    double ms;
  };
  std::vector<Cand> cands;   // the last autotune's timings
  unsigned long long *prof = nullptr;   // profiling build: phase cycle sums of the last k_dense2 launch
  unsigned long long last_prof[kDensePhases] = {};
  float a_scale = 1.0f;    // FP16 form: 2^k with |A'| * a_scale <= 2^14 for every group
};


// Dynamic LDS behind DenseSmem: A' [ng][NOUT2], the cluster coefficients
// [ntri], the tile's phase operands, the root table (when L <= kRootsLds) and
// the coarse groups' d vectors; each region starts on a 16-byte boundary.
struct DynLayout {
  size_t A, c, ph, rt, gd, end;
};
// Cells of more than 4 atoms (Nbands > 12) keep A' in device memory (one
// region per resident block, DenseArgs::ag): a tile of 8 q'' of their
// eigenvectors already takes ~74 KiB of LDS at Nbands = 24.
template <int NB>
constexpr bool dense_big() {
  return NB > 12;
}

template <class T, int NB, int IB, int TQ, int MF>
__host__ __device__ DynLayout dyn_layout(size_t stat, int ng, int ntri, int L_lds) {
  DynLayout d{};
  d.A = align16(stat);
  d.c = align16(d.A + (dense_big<NB>() ? 0 : sizeof(C2<T>) * (size_t)ng * IB * NB * 9));
  d.ph = align16(d.c + sizeof(C2<T>) * (size_t)ntri);
  // MFMA path: the A operand of the Psi GEMM, [t][g][qq][part] floats
  d.rt = align16(d.ph + (MF != 0 ? sizeof(float) * 4 : sizeof(C2<T>)) * (size_t)TQ * ng);
  d.gd = align16(d.rt + sizeof(double2) * (size_t)L_lds);
  d.end = align16(d.gd + sizeof(int) * 3 * (size_t)ng);
  return d;
}


// Lane layout of v_mfma_f32_16x16x4f32 (one 16x16 block, K = 4):
//   A[i][k] is held by lane i + 16 k, B[k][j] by lane j + 16 k,
//   D[i][j] by lane j + 16 (i / 4), accumulator register i % 4.
// k_mfma_selftest checks exactly this mapping on the device.
__global__ void k_mfma_selftest(double *err) {
  const int l = threadIdx.x;
  const int i = l & 15, kk = l >> 4;
  const float av = (float)(1 + i + 3 * kk) * 0.25f;          // A[i][kk]
  const float bv = (float)(2 + (l & 15) * 5 - kk) * 0.125f;  // B[kk][j], j = l & 15
  floatx4 acc = {0.f, 0.f, 0.f, 0.f};
  acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc, 0, 0, 0);
  double e = 0.0;
  for (int r = 0; r < 4; r++) {
    const int ii = 4 * (l >> 4) + r, jj = l & 15;
    double want = 0.0;
    for (int k = 0; k < 4; k++) want += (double)((1 + ii + 3 * k) * 0.25f) * (double)((2 + jj * 5 - k) * 0.125f);
    e = fmax(e, fabs(want - (double)acc[r]));
  }
  for (int o = 32; o > 0; o >>= 1) e = fmax(e, __shfl_xor(e, o, 64));
  if (l == 0) *err = e;
}


// v_mfma_f32_16x16x16bf16_1k lane layout used by the BF16 form: A[i][k] in
// lane i + 16*(k/4), element k%4; B[k][j] in lane j + 16*(k/4), element k%4;
// D as for the FP32 instruction.  Checked exactly with small integers.
__global__ void k_mfma_bf16_selftest(double *err) {
  const int l = threadIdx.x;
  const int i = l & 15, kg = l >> 4;
  bf16x4 av, bv;
  for (int q = 0; q < 4; q++) {
    const int k = 4 * kg + q;
    av[q] = (short)bf16_rn((float)(1 + i + 3 * k));
    bv[q] = (short)bf16_rn((float)(2 + 5 * (l & 15) - k));
  }
  floatx4 acc = {0.f, 0.f, 0.f, 0.f};
  acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
  double e = 0.0;
  for (int r = 0; r < 4; r++) {
    const int ii = 4 * (l >> 4) + r, jj = l & 15;
    double want = 0.0;
    for (int k = 0; k < 16; k++) want += (double)(1 + ii + 3 * k) * (double)(2 + 5 * jj - k);
    e = fmax(e, fabs(want - (double)acc[r]));
  }
  for (int o = 32; o > 0; o >>= 1) e = fmax(e, __shfl_xor(e, o, 64));
  if (l == 0) *err = e;
}

// v_mfma_f32_16x16x16f16: the same lane layout as the BF16 instruction,
// checked the same way.
__global__ void k_mfma_f16_selftest(double *err) {
  const int l = threadIdx.x;
  const int i = l & 15, kg = l >> 4;
  f16x4 av, bv;
  for (int q = 0; q < 4; q++) {
    const int k = 4 * kg + q;
    av[q] = (_Float16)(float)(1 + i + 3 * k);
    bv[q] = (_Float16)(float)(2 + 5 * (l & 15) - k);
  }
  floatx4 acc = {0.f, 0.f, 0.f, 0.f};
  acc = __builtin_amdgcn_mfma_f32_16x16x16f16(av, bv, acc, 0, 0, 0);
  double e = 0.0;
  for (int r = 0; r < 4; r++) {
    const int ii = 4 * (l >> 4) + r, jj = l & 15;
    double want = 0.0;
    for (int k = 0; k < 16; k++) want += (double)(1 + ii + 3 * k) * (double)(2 + 5 * jj - k);
    e = fmax(e, fabs(want - (double)acc[r]));
  }
  for (int o = 32; o > 0; o >>= 1) e = fmax(e, __shfl_xor(e, o, 64));
  if (l == 0) *err = e;
}


template <class T, int NB, int IB, int TQ>
struct DenseSmem {
  static constexpr bool BIG = dense_big<NB>();
  static constexpr int NOUT2 = IB * NB * 9;   // A' entries per group = Psi entries per (K,L) and triple
  static constexpr int FMAX = 8;              // fine groups per setup chunk
  // survivors contracted together: at least one filter round (NB per thread)
  // for up to 12 branches; 8 per thread for the large cells, whose rounds
  // take several contractions
  static constexpr int QMAX = NT_DENSE * (BIG ? 8 : NB);
  // candidate codes (t, k, l, i, j): 16 bits where they fit
  using Code = std::conditional_t<(TQ * NB * NB * IB * NB <= 65536), unsigned short, unsigned int>;
  static constexpr int TQS = BIG ? 1 : TQ;    // large cells recompute sigma instead of tabling it
  union {
    C2<T> psi[TQ * NOUT2];                    // Psi of one (K,L) pair for the tile
    struct {
      double2 af[FMAX][81];
      double2 b1[FMAX][IB][27];
    } setup;
  } u;
  C2<T> e3[TQ][NB][NB];   // [t][comp][k]
  C2<T> e4[TQ][NB][NB];   // [t][comp][l]
  double Ejj[TQ][NB], Ess[TQ][NB];
  double vjj[TQ][NB][3], vss[TQ][NB][3];
  double sig[TQS][BIG ? 1 : NB][BIG ? 1 : NB];  // exact sigma of (q'',k) / (q''',l), -1 when excluded
  int jj[TQ], ss[TQ];
  int hit[TQ];             // tile triple t has at least one surviving process
  int qjj[TQ][3];          // grid coordinates of the tile's q''
  double fb1[NB];          // Bose-Einstein factors of phonon 2 (the item's q')
  double fb3[TQ][NB];      // ... of phonon 3 (the tile's q'')
  double fb4[TQ][NB];      // ... of phonon 4 (the tile's q''')
  double2 e1[NB][IB];
  double2 e2[NB][NB];
  int vj[NT_DENSE], vs[NT_DENSE];
  Code q[QMAX];            // candidate codes (t, k, l, i, j)
  int wsum[NT_DENSE / 64];
  double om[IB];
  int act[IB];
  double Eii[NB];
  double x1[IB][NB];       // w_i +- w'_j, the first partial sum of delta
  long item;
  double red[NT_DENSE / 64][3 * IB];
};

// MF: 0 = VALU Psi, 1 = FP32 MFMA Psi, 2 = BF16 hi/lo MFMA Psi, 3 = FP16 hi/lo MFMA Psi,
// 4 = FP16 Psi with hi/lo phases against one A' plane.
//
// BATCH: one launch over several crystals with the same Nb (the 24-crystal
// workload): the queue hands out items of every crystal, in the host's order
// (largest first), and a block reloads the crystal's group vectors and
// roots when its next item belongs to another crystal.  The LDS layout is
// the launch's (the largest group, cluster and root counts of the crystals).
//
// One work item of the dense form: item of crystal cr (cr = 0 outside a batch);
// key identifies the crystal across the launch, and the crystal's group vectors
// and roots are reloaded into LDS when it differs from cur_key.  ka holds the
// launch's LDS layout (and, in a batch, the per-crystal argument table).
template <class T, int NB, int IB, int TQ, int MF, bool BATCH>
__device__ __forceinline__ void dense_item(const DenseArgs &ka, int cr, long item, int key, int &cur_key) {
  using S = DenseSmem<T, NB, IB, TQ>;
  extern __shared__ __align__(16) unsigned char dyn[];
  constexpr int NA = NB / 3;
  constexpr int QMAX = S::QMAX;
  constexpr int FMAX = S::FMAX;
  constexpr int NOUT2 = S::NOUT2;
  constexpr int NW = NT_DENSE / 64;
  constexpr int QPT = QMAX / NT_DENSE;        // survivors per thread in a contraction
  static_assert(MF == 0 || (sizeof(T) == 4 && 2 * TQ == 16), "MFMA path: float, 8 q'' per tile");
  // FP16 form: A' is scaled by a_scale (a power of two that bounds it by 2^14)
  // and the phases by kPhaseScaleF16 before the split; Psi is unscaled after
  // the GEMM, exactly (powers of two)
  constexpr bool BIG = S::BIG;
  static_assert(!BIG || MF == 0 || MF == 1, "large cells: FP32 MFMA or VALU Psi");
  static_assert(NB <= 31 && TQ <= 15, "survivor info packs k, l in 5 bits and t in 4");
  constexpr unsigned long long kJall = (1ull << NB) - 1ull;   // bits of all j of one i
  constexpr unsigned long long kRep = rep_mask(IB, NB);       // bit j = 0 of every i
  S &s = *reinterpret_cast<S *>(dyn);
  const DynLayout dl = dyn_layout<T, NB, IB, TQ, MF>(ka.smem_static, ka.ng, ka.ntri, ka.roots_lds ? ka.L : 0);
  // [ng][NOUT2]: LDS, or this block's region of device memory for large cells
  C2<T> *s_A = BIG ? reinterpret_cast<C2<T> *>(reinterpret_cast<char *>(ka.ag) + (size_t)blockIdx.x * ka.ag_stride)
                   : reinterpret_cast<C2<T> *>(dyn + dl.A);
  C2<T> *s_c = reinterpret_cast<C2<T> *>(dyn + dl.c);     // [ntri]
  C2<T> *s_ph = reinterpret_cast<C2<T> *>(dyn + dl.ph);   // [TQ][ng] (VALU path)
  float *s_phA = reinterpret_cast<float *>(dyn + dl.ph);  // [TQ][ng][2][2] (MFMA path)
  double2 *s_rt = reinterpret_cast<double2 *>(dyn + dl.rt);
  int *s_gd = reinterpret_cast<int *>(dyn + dl.gd);       // [ng][3]

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const double hbarp = kHbar * (double)k1e22;
  const double gpre = hbarp * hbarp * kPi / 8.0;
  auto load_tables = [&](const DenseArgs &a) {
    for (int x = tid; x < 3 * a.ng; x += NT_DENSE) s_gd[x] = a.g[x / 3].d[x % 3];
    if (a.roots_lds)
      for (int x = tid; x < a.L; x += NT_DENSE) s_rt[x] = a.roots[x];
    __syncthreads();
  };
  const DenseArgs &a = dense_args_of<BATCH>(ka, cr);
  if (key != cur_key) {
    load_tables(a);
    cur_key = key;
  }
  const ModeView &m = a.m;
  const float psi_unscale = (MF >= 3) ? 1.0f / (kPhaseScaleF16 * a.a_scale) : 1.0f;
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
  if (tid < IB * NB) {
    const int il = tid / NB, j = tid % NB;
    s.x1[il][j] = (ch == 2) ? (s.om[il] - s.Eii[j]) : (s.om[il] + s.Eii[j]);
  }
  int any = 0;
#pragma unroll
  for (int il = 0; il < IB; il++) any |= s.act[il];
  // (i, j) pairs allowed for the whole item: active mode i, nonzero w'_j
  unsigned long long base_ok = 0ull;
#pragma unroll
  for (int il = 0; il < IB; il++)
#pragma unroll
    for (int j = 0; j < NB; j++)
      if (s.act[il] && s.Eii[j] != 0.0) base_ok |= 1ull << (il * NB + j);

  double acc_g[IB], acc_w[IB];
  int acc_n[IB];
  int nhit = 0;   // (q, q', q'') triples of this item with >= 1 surviving process
#pragma unroll
  for (int il = 0; il < IB; il++) {
    acc_g[il] = 0.0;
    acc_w[il] = 0.0;
    acc_n[il] = 0;
  }

  if (any) {
    // ------------- setup 1: eigenvectors of phonons 1, 2 and cluster coefficients
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
      const double2 rt = a.roots[fmod_pos(k, a.L)];
      s_c[ci] = C2<T>{(T)(c.minv * rt.x), (T)(c.minv * rt.y)};
    }
    __syncthreads();
    // ------------- setup 2: A'[gamma] for every coarse group, into LDS
    for (int gi = 0; gi < a.ng; gi++) {
      const DGroup G = a.g[gi];
      double2 acc = make_double2(0.0, 0.0);
      for (int f0 = 0; f0 < G.f_count; f0 += FMAX) {
        const int nf = min(FMAX, G.f_count - f0);
        for (int x = tid; x < nf * 81; x += NT_DENSE) {
          const int fl = x / 81, e = x % 81;
          const FGroup F = a.f[G.f_first + f0 + fl];
          double sr = 0.0, si = 0.0;
          int ci = F.c_first;
          const int ce = F.c_first + F.c_count;
          for (; ci + 1 < ce; ci += 2) {
            const double p0 = a.phi[(size_t)ci * 81 + e], p1 = a.phi[(size_t)(ci + 1) * 81 + e];
            const C2<T> c0 = s_c[ci], c1 = s_c[ci + 1];
            sr += p0 * (double)c0.re;
            si += p0 * (double)c0.im;
            sr += p1 * (double)c1.re;
            si += p1 * (double)c1.im;
          }
          if (ci < ce) {
            const double p0 = a.phi[(size_t)ci * 81 + e];
            const C2<T> c0 = s_c[ci];
            sr += p0 * (double)c0.re;
            si += p0 * (double)c0.im;
          }
          s.u.setup.af[fl][e] = make_double2(sr, si);
        }
        __syncthreads();
        for (int x = tid; x < nf * IB * 27; x += NT_DENSE) {
          const int fl = x / (IB * 27), il = (x / 27) % IB, tsr = x % 27;
          const int I = a.f[G.f_first + f0 + fl].I;
          double2 sacc = make_double2(0.0, 0.0);
#pragma unroll
          for (int uu = 0; uu < 3; uu++) {
            const double2 v = zm(s.u.setup.af[fl][uu + 3 * tsr], s.e1[3 * I + uu][il]);
            sacc.x += v.x;
            sacc.y += v.y;
          }
          s.u.setup.b1[fl][il][tsr] = sacc;
        }
        __syncthreads();
        if (tid < NOUT2) {
          const int il = tid / (NB * 9), rem = tid % (NB * 9), j = rem / 9, sr = rem % 9;
          for (int fl = 0; fl < nf; fl++) {
            const int J = a.f[G.f_first + f0 + fl].J;
#pragma unroll
            for (int t = 0; t < 3; t++) {
              const double2 v = zm(s.u.setup.b1[fl][il][t + 3 * sr], s.e2[3 * J + t][j]);
              acc.x += v.x;
              acc.y += v.y;
            }
          }
        }
        __syncthreads();
      }
      if (tid < NOUT2) {
        if constexpr (MF == 3 || MF == 4) {
          // hi plane and lo plane as in the BF16 form, binary16 values of A' * a_scale
          // (the FP16X2 form reads the hi plane only)
          unsigned rh, rl, ih, il;
          f16_split((float)acc.x * a.a_scale, rh, rl);
          f16_split((float)acc.y * a.a_scale, ih, il);
          unsigned *bh = reinterpret_cast<unsigned *>(s_A);
          bh[(size_t)gi * NOUT2 + tid] = rh | (ih << 16);
          bh[(size_t)a.ng * NOUT2 + (size_t)gi * NOUT2 + tid] = rl | (il << 16);
        } else if constexpr (MF == 2) {
          unsigned rh, rl, ih, il;
          bf16_split((float)acc.x, rh, rl);
          bf16_split((float)acc.y, ih, il);
          unsigned *bh = reinterpret_cast<unsigned *>(s_A);
          bh[(size_t)gi * NOUT2 + tid] = rh | (ih << 16);
          bh[(size_t)a.ng * NOUT2 + (size_t)gi * NOUT2 + tid] = rl | (il << 16);
        } else {
          s_A[(size_t)gi * NOUT2 + tid] = C2<T>{(T)acc.x, (T)acc.y};
        }
      }
    }
    __syncthreads();

    // Registers carrying the next tile's global data: loaded while the
    // current tile is filtered and contracted (barriers wait for LDS only),
    // written to LDS when that tile starts.
    constexpr bool PF = (MF != 0) && (TQ * NB * NB <= NT_DENSE) && (10 * TQ * NB <= NT_DENSE);
    double2 pf_e3 = make_double2(0.0, 0.0), pf_e4 = make_double2(0.0, 0.0);
    double pf_s = 0.0;
    int pf_q = 0;
    bool have_pf = false;
    auto fetch = [&](int tb, int tqb) {
      if (tid < TQ * NB * NB) {   // eigenvector components of phonons 3 and 4
        const int t = tid / (NB * NB), comp = (tid / NB) % NB, b = tid % NB;
        if (t < tqb) {
          pf_e3 = m.eig[((size_t)s.vj[tb + t] * NB + b) * NB + comp];
          pf_e4 = m.eig[((size_t)s.vs[tb + t] * NB + b) * NB + comp];
        }
      }
      if (tid < 10 * TQ * NB) {   // one of w'', w''', f'', f''', v''(3), v'''(3)
        const int f = tid / (TQ * NB), y = tid % (TQ * NB), t = y / NB, b = y % NB;
        if (t < tqb) {
          const int jv = s.vj[tb + t], sv = s.vs[tb + t];
          if (f == 0)
            pf_s = E(m, jv, b);
          else if (f == 1)
            pf_s = E(m, sv, b);
          else if (f == 2)
            pf_s = a.fbose[(size_t)b * m.nptk + jv];
          else if (f == 3)
            pf_s = a.fbose[(size_t)b * m.nptk + sv];
          else if (f < 7)
            pf_s = Vel(m, jv, b, f - 4);
          else
            pf_s = Vel(m, sv, b, f - 7);
        }
      }
      if (tid < 3 * TQ) {         // grid coordinates of q''
        const int t = tid / 3, c = tid % 3;
        pf_q = m.ijk[3 * ((t < tqb) ? s.vj[tb + t] : 0) + c];
      }
    };
    auto stash = [&](int tqb) {
      if (tid < TQ * NB * NB) {
        const int t = tid / (NB * NB), comp = (tid / NB) % NB, b = tid % NB;
        C2<T> v3 = {(T)0, (T)0}, v4 = {(T)0, (T)0};
        if (t < tqb) {
          v3 = C2<T>{(T)pf_e3.x, (T)((ch != 0) ? -pf_e3.y : pf_e3.y)};
          v4 = C2<T>{(T)pf_e4.x, (T)(-pf_e4.y)};
        }
        s.e3[t][comp][b] = v3;
        s.e4[t][comp][b] = v4;
      }
      if (tid < 10 * TQ * NB) {
        const int f = tid / (TQ * NB), y = tid % (TQ * NB), t = y / NB, b = y % NB;
        if (t < tqb) {
          if (f == 0)
            s.Ejj[t][b] = pf_s;
          else if (f == 1)
            s.Ess[t][b] = pf_s;
          else if (f == 2)
            s.fb3[t][b] = pf_s;
          else if (f == 3)
            s.fb4[t][b] = pf_s;
          else if (f < 7)
            s.vjj[t][b][f - 4] = pf_s;
          else
            s.vss[t][b][f - 7] = pf_s;
        }
      }
      if (tid < 3 * TQ) s.qjj[tid / 3][tid % 3] = pf_q;
    };

    // ------------- q'' loop
    const int jj_first = (ch == 1) ? 0 : ii0;
    // range of the item's x1 for the triple screen
    double xlo = s.x1[0][0], xhi = s.x1[0][0];
    for (int il = 0; il < IB; il++)
      for (int j = 0; j < NB; j++) {
        xlo = fmin(xlo, s.x1[il][j]);
        xhi = fmax(xhi, s.x1[il][j]);
      }
    const double xabs = fmax(fabs(xlo), fabs(xhi));
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
      for (int t0 = 0; t0 < tot; t0 += TQ) {
        const int tq = min(TQ, tot - t0);
        // tile data straight from the scan buffer (one round of global loads):
        // q'' and q''' of the tile, their grid coordinates, eigenvectors of
        // phonons 3 and 4, frequencies, Bose-Einstein factors and velocities
        if (tid < TQ) {
          s.jj[tid] = (tid < tq) ? s.vj[t0 + tid] : 0;
          s.ss[tid] = (tid < tq) ? s.vs[t0 + tid] : 0;
          s.hit[tid] = 0;
        }
        if constexpr (PF) {
          if (!have_pf) fetch(t0, tq);
          stash(tq);
          const int tn = t0 + TQ;
          have_pf = tn < tot;
          if (have_pf) fetch(tn, min(TQ, tot - tn));
        } else {
          if (tid < TQ) {
            const int jv = (tid < tq) ? s.vj[t0 + tid] : 0;
#pragma unroll
            for (int c = 0; c < 3; c++) s.qjj[tid][c] = m.ijk[3 * jv + c];
          }
          for (int x = tid; x < TQ * NB * NB; x += NT_DENSE) {
            const int t = x / (NB * NB), comp = (x / NB) % NB, b = x % NB;
            C2<T> v3 = {(T)0, (T)0}, v4 = {(T)0, (T)0};
            if (t < tq) {
              const double2 e3 = m.eig[((size_t)s.vj[t0 + t] * NB + b) * NB + comp];
              const double2 e4 = m.eig[((size_t)s.vs[t0 + t] * NB + b) * NB + comp];
              v3 = C2<T>{(T)e3.x, (T)((ch != 0) ? -e3.y : e3.y)};
              v4 = C2<T>{(T)e4.x, (T)(-e4.y)};
            }
            s.e3[t][comp][b] = v3;
            s.e4[t][comp][b] = v4;
          }
          for (int x = tid; x < TQ * NB; x += NT_DENSE) {
            const int t = x / NB, b = x % NB;
            if (t < tq) {
              const int jv = s.vj[t0 + t], sv = s.vs[t0 + t];
              s.Ejj[t][b] = E(m, jv, b);
              s.Ess[t][b] = E(m, sv, b);
              s.fb3[t][b] = a.fbose[(size_t)b * m.nptk + jv];
              s.fb4[t][b] = a.fbose[(size_t)b * m.nptk + sv];
#pragma unroll
              for (int d = 0; d < 3; d++) {
                s.vjj[t][b][d] = Vel(m, jv, b, d);
                s.vss[t][b][d] = Vel(m, sv, b, d);
              }
            }
          }
        }
        __syncthreads();
        for (int x = tid; x < TQ * a.ng; x += NT_DENSE) {
          const int t = x / a.ng, gi = x % a.ng;
          double2 rt = make_double2(0.0, 0.0);
          if (t < tq) {
            const int k = s3 * (s.qjj[t][0] * s_gd[3 * gi] + s.qjj[t][1] * s_gd[3 * gi + 1] +
                                s.qjj[t][2] * s_gd[3 * gi + 2]);
            const int kk = fmod_pos(k, a.L);
            rt = a.roots_lds ? s_rt[kk] : a.roots[kk];
          }
          if constexpr (MF == 2 || MF == 3 || MF == 4) {
            // A operand rows (t, part), k = (g, qq): part 0 (re, -im), part 1 (im, re), hi and lo
            // planes; the sign flip is the top bit in both 16-bit formats
            unsigned rh, rl, ih, il;
            if constexpr (MF >= 3) {
              f16_split((float)rt.x * kPhaseScaleF16, rh, rl);
              f16_split((float)rt.y * kPhaseScaleF16, ih, il);
            } else {
              bf16_split((float)rt.x, rh, rl);
              bf16_split((float)rt.y, ih, il);
            }
            unsigned *ph = reinterpret_cast<unsigned *>(s_phA);
            unsigned *pl = ph + (size_t)TQ * a.ng * 2;
            ph[2 * x] = rh | ((ih ^ 0x8000u) << 16);
            ph[2 * x + 1] = ih | (rh << 16);
            pl[2 * x] = rl | ((il ^ 0x8000u) << 16);
            pl[2 * x + 1] = il | (rl << 16);
          } else if constexpr (MF == 1) {
            // A operand rows (t, part), k = (g, qq): (re, im | -im, re)
            float *o = s_phA + (size_t)x * 4;
            o[0] = (float)rt.x;
            o[1] = (float)rt.y;
            o[2] = -(float)rt.y;
            o[3] = (float)rt.x;
          } else {
            s_ph[x] = C2<T>{(T)rt.x, (T)rt.y};
          }
        }
        __syncthreads();
        // contraction of the queued survivors: for each (K,L) pair, Psi of
        // the tile (skinny GEMM over that pair's phase groups), then the
        // survivors' partial sums; finally weights and rates in FP64.
        auto contract = [&](int qlen) {
          C2<T> V[QPT];
          // survivor u of this thread: packed (psi row offset, t, k, l), -1 when empty
          int info[QPT];
#pragma unroll
          for (int u = 0; u < QPT; u++) {
            V[u] = C2<T>{(T)0, (T)0};
            const int qi = tid + u * NT_DENSE;
            info[u] = -1;
            if (qi < qlen) {
              const int cc = s.q[qi];
              const int j = cc % NB, il = (cc / NB) % IB, l = (cc / (NB * IB)) % NB;
              const int k = (cc / (NB * IB * NB)) % NB, t = cc / (NB * IB * NB * NB);
              info[u] = ((t * NOUT2 + (il * NB + j) * 9) << 14) | (t << 10) | (k << 5) | l;
            }
          }
          for (int kl = 0; kl < NA * NA; kl++) {
            const int g0 = a.klf[kl], g1 = a.klf[kl + 1];
            if (g1 == g0) continue;
            if constexpr (MF == 1) {
              constexpr int NTILE = (NOUT2 + 15) / 16;
              const int ksteps = (2 * (g1 - g0) + 3) / 4;
              const int i = lane & 15, t = i >> 1, part = i & 1, kg = lane >> 4;
              const float *Af = reinterpret_cast<const float *>(s_A);   // [g][n][re, im]
              // each wave owns column tiles wave, wave + NW, ...; their MFMA
              // chains are independent and share the A operand, so they are
              // interleaved, with the operands of four k-steps loaded first
              constexpr int TPW = (NTILE + NW - 1) / NW;
              int ncol[TPW];
              bool nin[TPW];
              floatx4 acc[TPW];
#pragma unroll
              for (int c = 0; c < TPW; c++) {
                const int n = (wave + NW * c) * 16 + (lane & 15);
                nin[c] = n < NOUT2;
                ncol[c] = nin[c] ? n : NOUT2 - 1;
                acc[c] = floatx4{0.f, 0.f, 0.f, 0.f};
              }
              for (int ks0 = 0; ks0 < ksteps; ks0 += 4) {
                float av[4], bv[TPW][4];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                  const int kp = 4 * (ks0 + u) + kg;
                  const int g = g0 + (kp >> 1), qq = kp & 1;
                  const bool ok = g < g1;
                  const int gc = ok ? g : g0;
                  const float x = s_phA[((t * a.ng + gc) * 2 + qq) * 2 + part];
                  av[u] = ok ? x : 0.f;
#pragma unroll
                  for (int c = 0; c < TPW; c++) {
                    const float y = Af[((size_t)gc * NOUT2 + ncol[c]) * 2 + qq];
                    bv[c][u] = (ok && nin[c]) ? y : 0.f;
                  }
                }
#pragma unroll
                for (int u = 0; u < 4; u++)
                  if (ks0 + u < ksteps)
#pragma unroll
                    for (int c = 0; c < TPW; c++)
                      if (wave + NW * c < NTILE)
                        acc[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[u], bv[c][u], acc[c], 0, 0, 0);
              }
#pragma unroll
              for (int c = 0; c < TPW; c++) {
                const int n = (wave + NW * c) * 16 + (lane & 15);
                if (n < NOUT2) {
#pragma unroll
                  for (int r = 0; r < 4; r++) {
                    const int row = 4 * (lane >> 4) + r;
                    if ((row & 1) == 0)
                      s.u.psi[(row >> 1) * NOUT2 + n].re = (T)acc[c][r];
                    else
                      s.u.psi[(row >> 1) * NOUT2 + n].im = (T)acc[c][r];
                  }
                }
              }
            } else if constexpr (MF == 2 || MF == 3 || MF == 4) {
              // BF16 (FP16) hi/lo form: Psi = A_hi B_hi + A_hi B_lo + A_lo B_hi
              // on v_mfma_f32_16x16x16bf16_1k (v_mfma_f32_16x16x16f16), FP32
              // accumulation; a k-step is 8 groups, lane group kg holding
              // groups 2kg and 2kg+1
              constexpr int NTILE = (NOUT2 + 15) / 16;
              constexpr int TPW = (NTILE + NW - 1) / NW;
              const int i = lane & 15, t = i >> 1, part = i & 1, kg = lane >> 4;
              const unsigned *Bh = reinterpret_cast<const unsigned *>(s_A);       // [g][n] (re | im << 16)
              const unsigned *Bl = Bh + (size_t)a.ng * NOUT2;
              const unsigned *Ph = reinterpret_cast<const unsigned *>(s_phA);     // [t][g][part] (qq0 | qq1 << 16)
              const unsigned *Pl = Ph + (size_t)TQ * a.ng * 2;
              int ncol[TPW];
              bool nin[TPW];
              floatx4 acc[TPW];
#pragma unroll
              for (int c = 0; c < TPW; c++) {
                const int n = (wave + NW * c) * 16 + (lane & 15);
                nin[c] = n < NOUT2;
                ncol[c] = nin[c] ? n : NOUT2 - 1;
                acc[c] = floatx4{0.f, 0.f, 0.f, 0.f};
              }
              const int nks = (g1 - g0 + 7) / 8;
              for (int ks = 0; ks < nks; ks++) {
                const int g = g0 + 8 * ks + 2 * kg;
                const bool v0 = g < g1, v1 = g + 1 < g1;
                const int gc0 = v0 ? g : g0, gc1 = v1 ? g + 1 : g0;
                const unsigned pa = (t * a.ng + gc0) * 2 + part, pb = (t * a.ng + gc1) * 2 + part;
                const unsigned ph0 = v0 ? Ph[pa] : 0u, ph1 = v1 ? Ph[pb] : 0u;
                const unsigned pl0 = v0 ? Pl[pa] : 0u, pl1 = v1 ? Pl[pb] : 0u;
#pragma unroll
                for (int c = 0; c < TPW; c++) {
                  if (wave + NW * c >= NTILE) continue;
                  const size_t b0 = (size_t)gc0 * NOUT2 + ncol[c], b1 = (size_t)gc1 * NOUT2 + ncol[c];
                  const bool w0 = v0 && nin[c], w1 = v1 && nin[c];
                  const unsigned bh0 = w0 ? Bh[b0] : 0u, bh1 = w1 ? Bh[b1] : 0u;
                  const unsigned bl0 = w0 ? Bl[b0] : 0u, bl1 = w1 ? Bl[b1] : 0u;
                  if constexpr (MF == 4) {
                    // hi/lo phases against the single FP16 A' plane
                    const f16x4 ah = f16_pack(ph0, ph1), al = f16_pack(pl0, pl1);
                    const f16x4 bh = f16_pack(bh0, bh1);
                    (void)bl0;
                    (void)bl1;
                    acc[c] = __builtin_amdgcn_mfma_f32_16x16x16f16(ah, bh, acc[c], 0, 0, 0);
                    acc[c] = __builtin_amdgcn_mfma_f32_16x16x16f16(al, bh, acc[c], 0, 0, 0);
                  } else if constexpr (MF == 3) {
                    const f16x4 ah = f16_pack(ph0, ph1), al = f16_pack(pl0, pl1);
                    const f16x4 bh = f16_pack(bh0, bh1), bl = f16_pack(bl0, bl1);
                    acc[c] = __builtin_amdgcn_mfma_f32_16x16x16f16(ah, bh, acc[c], 0, 0, 0);
                    acc[c] = __builtin_amdgcn_mfma_f32_16x16x16f16(ah, bl, acc[c], 0, 0, 0);
                    acc[c] = __builtin_amdgcn_mfma_f32_16x16x16f16(al, bh, acc[c], 0, 0, 0);
                  } else {
                    const bf16x4 ah = bf16_pack(ph0, ph1), al = bf16_pack(pl0, pl1);
                    const bf16x4 bh = bf16_pack(bh0, bh1), bl = bf16_pack(bl0, bl1);
                    acc[c] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bh, acc[c], 0, 0, 0);
                    acc[c] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bl, acc[c], 0, 0, 0);
                    acc[c] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al, bh, acc[c], 0, 0, 0);
                  }
                }
              }
#pragma unroll
              for (int c = 0; c < TPW; c++) {
                const int n = (wave + NW * c) * 16 + (lane & 15);
                if (n < NOUT2) {
#pragma unroll
                  for (int r = 0; r < 4; r++) {
                    const int row = 4 * (lane >> 4) + r;
                    if ((row & 1) == 0)
                      s.u.psi[(row >> 1) * NOUT2 + n].re = (T)(acc[c][r] * psi_unscale);
                    else
                      s.u.psi[(row >> 1) * NOUT2 + n].im = (T)(acc[c][r] * psi_unscale);
                  }
                }
              }
            } else {
              for (int n = tid; n < NOUT2; n += NT_DENSE) {
                C2<T> acc[TQ];
#pragma unroll
                for (int t = 0; t < TQ; t++) acc[t] = C2<T>{(T)0, (T)0};
                for (int g = g0; g < g1; g++) {
                  const C2<T> av = s_A[(size_t)g * NOUT2 + n];
#pragma unroll
                  for (int t = 0; t < TQ; t++) {
                    const C2<T> p = s_ph[t * a.ng + g];
                    acc[t].re += p.re * av.re - p.im * av.im;
                    acc[t].im += p.re * av.im + p.im * av.re;
                  }
                }
#pragma unroll
                for (int t = 0; t < TQ; t++) s.u.psi[t * NOUT2 + n] = acc[t];
              }
            }
            __syncthreads();
            const int K = kl / NA, L = kl % NA;
#pragma unroll
            for (int u = 0; u < QPT; u++) {
              if (info[u] >= 0) {
                const int t = (info[u] >> 10) & 15, k = (info[u] >> 5) & 31, l = info[u] & 31;
                const C2<T> *psi = &s.u.psi[info[u] >> 14];
#pragma unroll
                for (int sx = 0; sx < 3; sx++) {
                  C2<T> inner = {(T)0, (T)0};
#pragma unroll
                  for (int rx = 0; rx < 3; rx++) {
                    const C2<T> xx = cm(psi[sx + 3 * rx], s.e4[t][3 * L + rx][l]);
                    inner.re += xx.re;
                    inner.im += xx.im;
                  }
                  const C2<T> y = cm(inner, s.e3[t][3 * K + sx][k]);
                  V[u].re += y.re;
                  V[u].im += y.im;
                }
              }
            }
            __syncthreads();
          }
#pragma unroll
          for (int u = 0; u < QPT; u++) {
            const int qi = tid + u * NT_DENSE;
            if (qi < qlen) {
              const int cc = s.q[qi];
              const int j = cc % NB, il = (cc / NB) % IB, l = (cc / (NB * IB)) % NB;
              const int k = (cc / (NB * IB * NB)) % NB, t = cc / (NB * IB * NB * NB);
              const double w = s.om[il], op = s.Eii[j], odp = s.Ejj[t][k], otp = s.Ess[t][l];
              const double delta = mismatch(ch, w, op, odp, otp);
              // the filter's exact sigma (recomputed for large cells, same operations)
              double sigma;
              if constexpr (BIG)
                sigma = tile_sigma(m, ch, s.vjj[t][k], s.vss[t][l]);
              else
                sigma = s.sig[t][k][l];
              if (sigma <= (double)kClamp) sigma = 1.0 / sqrt(kPi);
              const double wpv = wp4_given(ch, w, op, odp, otp, s.fb1[j], s.fb3[t][k], s.fb4[t][l], delta, sigma);
              double g = 0.0;
              if (!a.onlyharmonic) {
                const double vr = (double)V[u].re, vi = (double)V[u].im;
                g = gpre * wpv * (vr * vr + vi * vi);
              }
              if (a.wop && g != 0.0)
                wop_add(a, ll0 * NB + ib * IB + il, ch, ii0 * NB + j, s.jj[t] * NB + k, s.ss[t] * NB + l, g);
#pragma unroll
              for (int xi = 0; xi < IB; xi++)
                if (xi == il) {
                  acc_g[xi] += g;
                  acc_w[xi] += wpv;
                  acc_n[xi] += 1;
                }
            }
          }
        };

        // exact filter, fused with sigma: one thread per (t, k, l) computes
        // the exact sigma of phonons 3 and 4 and tests every (i, j) of the
        // tile, each candidate being two FP64 additions in the reference's
        // order, ((w +- w') +- w'') - w''', and one comparison.  Survivors
        // are queued in ascending code order (t, k, l, i, j); the queue is
        // contracted whenever it holds QMAX codes and at the end.
        constexpr int NKL = TQ * NB * NB;
        static_assert(IB * NB <= 64, "candidate mask is 64 bits");
        int qlen = 0, x0 = 0, done = 0, tot2 = 0, pos2 = 0, x = 0;
        unsigned long long mask = 0ull;
        for (;;) {
          if (done == tot2 && x0 < NKL) {
            // next round of (t, k, l) tasks, spread evenly over the waves
            // (task order still follows thread order)
            const int per = min(64, (NKL - x0 + NW - 1) / NW);
            x = (lane < per) ? x0 + wave * per + lane : NKL;
            mask = 0ull;
            if (x < NKL) {
              const int l = x % NB, k = (x / NB) % NB, t = x / (NB * NB);
              double sg = -1.0;
              if (t < tq) {
                const bool okkl = (s.Ejj[t][k] != 0.0) && (s.Ess[t][l] != 0.0) &&
                                  !(ch != 0 && s.ss[t] == s.jj[t] && l < k);
                if (okkl) sg = tile_sigma(m, ch, s.vjj[t][k], s.vss[t][l]);
              }
              if constexpr (!BIG) s.sig[t][k][l] = sg;
              if (sg >= 0.0) {
                const int jj0 = s.jj[t], ss0 = s.ss[t];
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
                const unsigned long long allowed = base_ok & ~excl;
                // delta = ((w +- w') +- w'') - w''' in the reference's order; x - w''
                // is evaluated as x + (-w''), which IEEE defines identically.
                const double odp = (ch == 0) ? s.Ejj[t][k] : -s.Ejj[t][k];
                const double otp = s.Ess[t][l];
                unsigned long long hitm = 0ull;
#pragma unroll 1
                for (int il = 0; il < IB; il++) {
                  double xv[NB];
#pragma unroll
                  for (int j = 0; j < NB; j++) xv[j] = s.x1[il][j];
#pragma unroll
                  for (int j = 0; j < NB; j++) {
                    const double delta = (xv[j] + odp) - otp;
                    hitm |= (unsigned long long)(fabs(delta) <= sg) << (il * NB + j);
                  }
                }
                mask = hitm & allowed;
                if (mask) s.hit[t] = 1;
              }
            }
            pos2 = block_excl_scan<NW>(__popcll(mask), s.wsum, tot2);
            done = 0;
            x0 += NW * min(64, (NKL - x0 + NW - 1) / NW);
          }
          // queue this round's survivors with positions done .. done+take-1
          const int take = min(QMAX - qlen, tot2 - done);
          {
            const int code0 = x * (IB * NB);   // ((t*NB + k)*NB + l) * IB * NB
            unsigned long long mb = mask;
            int p = pos2;
            while (mb) {
              const int b = __builtin_ctzll(mb);
              mb &= mb - 1;
              if (p >= done && p < done + take) s.q[qlen + p - done] = (typename S::Code)(code0 + b);
              p++;
            }
          }
          qlen += take;
          done += take;
          const bool last = (done == tot2) && (x0 >= NKL);
          if (qlen == QMAX || (last && qlen > 0)) {
            __syncthreads();
            contract(qlen);
            __syncthreads();
            qlen = 0;
          }
          if (last) break;
        }
        if (tid < tq && s.hit[tid]) nhit++;
      }
    }
  }
  // fixed-order block reduction of the per-thread partial sums
#pragma unroll
  for (int il = 0; il < IB; il++) {
    const double g = wave_sum(acc_g[il]);
    const double w = wave_sum(acc_w[il]);
    const double n = wave_sum((double)acc_n[il]);
    if (lane == 0) {
      s.red[wave][3 * il] = g;
      s.red[wave][3 * il + 1] = w;
      s.red[wave][3 * il + 2] = n;
    }
  }
  __syncthreads();
  if (tid < IB) {
    double g = 0.0, w = 0.0, n = 0.0;
#pragma unroll
    for (int wv = 0; wv < NW; wv++) {
      g = g + s.red[wv][3 * tid];
      w = w + s.red[wv][3 * tid + 1];
      n = n + s.red[wv][3 * tid + 2];
    }
    a.out_g[item * IB + tid] = g;
    a.out_w[item * IB + tid] = w;
    a.out_n[item * IB + tid] = (int)n;
  }
  const int nh = wave_sum_i(nhit);
  if (tid == 0) a.out_t[item] = nh;
  __syncthreads();
}

template <class T, int NB, int IB, int TQ, int MF, bool BATCH = false>
__global__ void __launch_bounds__(NT_DENSE) k_dense(DenseArgs ka) {
  using S = DenseSmem<T, NB, IB, TQ>;
  extern __shared__ __align__(16) unsigned char dyn[];
  S &s = *reinterpret_cast<S *>(dyn);
  int cur_key = -1;
  for (;;) {
    if (threadIdx.x == 0) s.item = atomicAdd(ka.counter, 1);
    __syncthreads();
    const long gitem = s.item;
    __syncthreads();
    if (gitem >= ka.nitems) break;
    long item = gitem;
    int cr = 0;
    if constexpr (BATCH) {
      const long long o = ka.order[gitem];
      cr = __builtin_amdgcn_readfirstlane((int)(o >> 40));
      item = (long)(o & 0xFFFFFFFFFFLL);
    }
    dense_item<T, NB, IB, TQ, MF, BATCH>(ka, cr, item, cr, cur_key);
  }
}

// Every crystal of a batch in one persistent launch, whatever its Nbands:
// the crystals are grouped in classes by Nbands (slot 0..3 for Nbands 6, 12,
// 18, 24), each class with its own launch arguments (the LDS layout of its
// largest crystal, its crystals' argument table), and one queue holds all
// classes' items, packed as class << 58 | crystal << 40 | item, costliest
// first.  A block runs each item with its class's kernel shape; the dynamic
// LDS is the largest class's.  MFS is the MFMA form of the 2- and 4-atom
// classes (FP32 or FP16X2); the 6- and 8-atom classes run FP32 MFMA with A'
// in device memory.
struct DenseMultiArgs {
  const DenseArgs *cls;     // [kMultiClasses] launch arguments of each class
  const long long *order;   // the launch's work items in queue order
  long nitems;
  int *counter;
};
constexpr int kMultiClasses = 4;
constexpr int kMultiNb[kMultiClasses] = {6, 12, 18, 24};
constexpr int kMultiIb[kMultiClasses] = {6, 1, 1, 1};
constexpr int kMultiShift = 58;

// One class's item as a call of its own: the four inlined bodies together
// would not fit the 256 VGPRs of two waves per SIMD.
template <int NB, int IB, int MF>
__device__ __noinline__ void dense_item_call(const DenseArgs &ka, int cr, long item, int key, int &cur_key) {
  dense_item<float, NB, IB, 8, MF, true>(ka, cr, item, key, cur_key);
}

template <int MFS>
__global__ void __launch_bounds__(NT_DENSE) k_dense_multi(DenseMultiArgs ma) {
  __shared__ long s_next;
  int cur_key = -1;
  for (;;) {
    if (threadIdx.x == 0) s_next = atomicAdd(ma.counter, 1);
    __syncthreads();
    const long gitem = s_next;
    __syncthreads();
    if (gitem >= ma.nitems) break;
    const long long o = ma.order[gitem];
    const int cl = __builtin_amdgcn_readfirstlane((int)(o >> kMultiShift));
    const int cr = __builtin_amdgcn_readfirstlane((int)((o >> 40) & ((1LL << (kMultiShift - 40)) - 1)));
    const long item = (long)(o & 0xFFFFFFFFFFLL);
    const DenseArgs &ka = ma.cls[cl];
    const int key = (cl << (kMultiShift - 40)) | cr;
    if (cl == 0)
      dense_item_call<6, 6, MFS>(ka, cr, item, key, cur_key);
    else if (cl == 1)
      dense_item_call<12, 1, MFS>(ka, cr, item, key, cur_key);
    else if (cl == 2)
      dense_item_call<18, 1, 1>(ka, cr, item, key, cur_key);
    else
      dense_item_call<24, 1, 1>(ka, cr, item, key, cur_key);
  }
}

// Sum over the q' window of each (channel, point, mode-block, mode), in order.
// Bose-Einstein factor of every mode (the bose() the per-process form uses).
__global__ void k_bose(const double *energy, long n, double T, double *out) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) out[t] = bose(energy[t], T);
}

__global__ void k_dense_reduce(const double *og, const double *ow, const int *on, long nseg, int ib_sz, int nii,
                               double *sg, double *sw, long long *sn) {
  const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nseg) return;
  const long blk = s / ib_sz, il = s % ib_sz;
  double g = 0.0, w = 0.0;
  long long n = 0;
  for (int t = 0; t < nii; t++) {
    const long x = (blk * nii + t) * ib_sz + il;
    g = g + og[x];
    w = w + ow[x];
    n += on[x];
  }
  sg[s] = g;
  sw[s] = w;
  sn[s] = n;
}

// ------------------------------------------------------------------- host
DenseState *dense_create(const std::vector<ClusterRec> &cl, const std::vector<double> &phi, int natoms, int nb,
                         int L, int *rc) {
  *rc = 0;
  DenseState *s = new DenseState();
  s->natoms = natoms;
  s->nb = nb;
  s->L = L;
  const int n = (int)cl.size();
  std::vector<int> order(n);
  for (int t = 0; t < n; t++) order[t] = t;
  auto key = [&](int t) {
    const ClusterRec &c = cl[t];
    return std::make_tuple(c.atom[2], c.atom[3], c.pk[0] - c.pl[0], c.pk[1] - c.pl[1], c.pk[2] - c.pl[2], c.atom[0],
                           c.atom[1], t);
  };
  std::sort(order.begin(), order.end(), [&](int x, int y) { return key(x) < key(y); });
  std::vector<DGroup> G;
  std::vector<FGroup> F;
  std::vector<DCluster> Cs(n);
  std::vector<double> P((size_t)n * 81);
  for (int t = 0; t < n; t++) {
    const ClusterRec &c = cl[order[t]];
    DCluster &d = Cs[t];
    for (int a = 0; a < 3; a++) {
      d.pjl[a] = c.pj[a] - c.pl[a];
      d.pl[a] = c.pl[a];
    }
    d.minv = c.minv;
    std::memcpy(&P[(size_t)t * 81], &phi[(size_t)order[t] * 81], sizeof(double) * 81);
    const int dk[3] = {c.pk[0] - c.pl[0], c.pk[1] - c.pl[1], c.pk[2] - c.pl[2]};
    const bool newg = G.empty() || G.back().K != c.atom[2] || G.back().L != c.atom[3] || G.back().d[0] != dk[0] ||
                      G.back().d[1] != dk[1] || G.back().d[2] != dk[2];
    if (newg) {
      DGroup g;
      g.K = c.atom[2];
      g.L = c.atom[3];
      for (int a = 0; a < 3; a++) g.d[a] = dk[a];
      g.f_first = (int)F.size();
      g.f_count = 0;
      G.push_back(g);
    }
    const bool newf = newg || F.back().I != c.atom[0] || F.back().J != c.atom[1];
    if (newf) {
      FGroup f;
      f.I = c.atom[0];
      f.J = c.atom[1];
      f.c_first = t;
      f.c_count = 0;
      F.push_back(f);
      G.back().f_count++;
    }
    F.back().c_count++;
  }
  std::vector<int> klf(natoms * natoms + 1, 0);
  {
    int gi = 0;
    for (int kl = 0; kl < natoms * natoms; kl++) {
      klf[kl] = gi;
      while (gi < (int)G.size() && G[gi].K * natoms + G[gi].L == kl) gi++;
    }
    klf[natoms * natoms] = (int)G.size();
  }
  s->ng = (int)G.size();
  s->nf = (int)F.size();
  {
    // |A'_g[i][j][s][r]| <= sum_{c in g} minv_c sum_{u,t} |Phi_c[u,t,s,r]|
    // (eigenvector components and phases have modulus <= 1)
    double bmax = 0.0;
    for (const DGroup &g : G)
      for (int sr = 0; sr < 9; sr++) {
        double b = 0.0;
        for (int fi = g.f_first; fi < g.f_first + g.f_count; fi++)
          for (int c = F[fi].c_first; c < F[fi].c_first + F[fi].c_count; c++) {
            double x = 0.0;
            for (int ut = 0; ut < 9; ut++) x += std::fabs(P[(size_t)c * 81 + ut + 9 * sr]);
            b += std::fabs(Cs[c].minv) * x;
          }
        bmax = std::max(bmax, b);
      }
    int e = 0;
    if (bmax > 0.0) (void)std::frexp(bmax, &e);   // bmax < 2^e
    s->a_scale = (float)std::ldexp(1.0, 14 - e);
  }
  s->nc = n;
  bool ok = hipMalloc(&s->g, sizeof(DGroup) * G.size()) == hipSuccess &&
            hipMalloc(&s->f, sizeof(FGroup) * F.size()) == hipSuccess &&
            hipMalloc(&s->c, sizeof(DCluster) * n) == hipSuccess &&
            hipMalloc(&s->phi, sizeof(double) * 81 * n) == hipSuccess &&
            hipMalloc(&s->counter, sizeof(int)) == hipSuccess && hipEventCreate(&s->e0) == hipSuccess &&
            hipEventCreate(&s->e1) == hipSuccess;
  ok = ok && hipMemcpy(s->g, G.data(), sizeof(DGroup) * G.size(), hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(s->f, F.data(), sizeof(FGroup) * F.size(), hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(s->c, Cs.data(), sizeof(DCluster) * n, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(s->phi, P.data(), sizeof(double) * 81 * n, hipMemcpyHostToDevice) == hipSuccess;
  if (natoms <= kDenseMaxAtoms)
    for (size_t q = 0; q < klf.size(); q++) s->klf[q] = klf[q];
  if (!ok) {
    dense_destroy(s);
    *rc = derr(FP4_ERR_NOMEM, "dense_create: device allocation or copy failed");
    return nullptr;
  }
  return s;
}

void dense_destroy(DenseState *s) {
  if (!s) return;
  void *ps[] = {s->g, s->f, s->c, s->phi, s->scratch, s->outb, s->fbose, s->counter, s->prof};
  for (void *p : ps)
    if (p) (void)hipFree(p);
  if (s->e0) (void)hipEventDestroy(s->e0);
  if (s->e1) (void)hipEventDestroy(s->e1);
  delete s;
}

int dense_ngroups(const DenseState *s) { return s ? s->ng : 0; }

int dense_selftest_mfma(double *err) {
  double *d = nullptr;
  if (hipMalloc(&d, sizeof(double)) != hipSuccess) return derr(FP4_ERR_NOMEM, "selftest: hipMalloc failed");
  double e32 = 0.0, e16 = 0.0;
  hipLaunchKernelGGL(k_mfma_selftest, dim3(1), dim3(64), 0, 0, d);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpy(&e32, d, sizeof(double), hipMemcpyDeviceToHost);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_mfma_bf16_selftest, dim3(1), dim3(64), 0, 0, d);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(&e16, d, sizeof(double), hipMemcpyDeviceToHost);
  double ef16 = 0.0;
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_mfma_f16_selftest, dim3(1), dim3(64), 0, 0, d);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(&ef16, d, sizeof(double), hipMemcpyDeviceToHost);
  double ex32 = 0.0;
  if (e == hipSuccess && dense2_selftest_mfma(&ex32) != 0) e = hipErrorUnknown;
  *err = std::max(std::max(std::max(e32, e16), ef16), ex32);
  (void)hipFree(d);
  return e == hipSuccess ? 0 : derr(FP4_ERR_DEVICE, "selftest: kernel failed");
}

void dense_shape(const DenseState *s, int *mfma, int *tq, int *ib, int *kg) {
  if (kg) *kg = s ? s->last_kg : 0;
  *mfma = s ? s->last_mfma : 0;
  *tq = s ? s->last_tq : 0;
  *ib = s ? s->last_ib : 0;
}

void dense_info(const DenseState *s, long long *triples, long long *items, int *ngroups, int *nfine) {
  *triples = s ? s->last_triples : 0;
  *items = s ? s->last_items : 0;
  *ngroups = s ? s->ng : 0;
  *nfine = s ? s->nf : 0;
}

namespace {

template <class T, int NB, int IB, int TQ>
size_t static_bytes() {
  return (sizeof(DenseSmem<T, NB, IB, TQ>) + 15) & ~size_t(15);
}

template <class T, int NB, int IB, int TQ, int MF>
size_t lds_bytes(int ng, int ntri, int L) {
  return dyn_layout<T, NB, IB, TQ, MF>(static_bytes<T, NB, IB, TQ>(), ng, ntri, L <= kRootsLds ? L : 0).end;
}

template <class T, int NB, int IB, int TQ, int MF>
size_t lds_bytes_geom(const DenseGeom &g) {
  return lds_bytes<T, NB, IB, TQ, MF>(g.ng, g.ntri, g.L);
}

template <class T, int NB, int IB, int TQ, int MF>
void launch_shape(const DenseArgs &a, int slots, size_t dyn, hipStream_t st) {
  (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_dense<T, NB, IB, TQ, MF>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
  hipLaunchKernelGGL((k_dense<T, NB, IB, TQ, MF>), dim3(slots), dim3(NT_DENSE), dyn, st, a);
}

template <class T, int NB, int IB, int TQ, int MF>
int per_cu_shape(size_t dyn) {
  (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_dense<T, NB, IB, TQ, MF>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, reinterpret_cast<const void *>(&k_dense<T, NB, IB, TQ, MF>),
                                                   NT_DENSE, dyn) != hipSuccess)
    return 1;
  return std::max(n, 1);
}

#define FP4_SHAPE(T, P, NB, IB, TQ, MF)                                                                          \
  {P, NB, IB, TQ, MF, 0, &lds_bytes_geom<T, NB, IB, TQ, MF>, &static_bytes<T, NB, IB, TQ>,                           \
   &launch_shape<T, NB, IB, TQ, MF>, &per_cu_shape<T, NB, IB, TQ, MF>}

// k_dense shapes.  The FP32 shapes with 8 q'' per tile run the Psi GEMM on
// MFMA.
const DenseShape kShapes1[] = {
    FP4_SHAPE(float, FP4_PREC_FP32, 3, 3, 8, 1),             FP4_SHAPE(float, FP4_PREC_FP32, 6, 6, 8, 1),
    FP4_SHAPE(float, FP4_PREC_FP32, 6, 2, 8, 1),             FP4_SHAPE(float, FP4_PREC_FP32, 9, 3, 8, 1),
    FP4_SHAPE(float, FP4_PREC_FP32, 12, 1, 8, 1),            FP4_SHAPE(float, FP4_PREC_FP32, 12, 1, 4, 0),
    FP4_SHAPE(float, FP4_PREC_FP32, 18, 1, 8, 1),            FP4_SHAPE(float, FP4_PREC_FP32, 24, 1, 8, 1),
    FP4_SHAPE(float, FP4_PREC_BF16X3, 3, 3, 8, 2),           FP4_SHAPE(float, FP4_PREC_BF16X3, 6, 6, 8, 2),
    FP4_SHAPE(float, FP4_PREC_BF16X3, 6, 2, 8, 2),           FP4_SHAPE(float, FP4_PREC_BF16X3, 9, 3, 8, 2),
    FP4_SHAPE(float, FP4_PREC_BF16X3, 12, 1, 8, 2),
    FP4_SHAPE(float, FP4_PREC_FP16X3, 3, 3, 8, 3),           FP4_SHAPE(float, FP4_PREC_FP16X3, 6, 6, 8, 3),
    FP4_SHAPE(float, FP4_PREC_FP16X3, 6, 2, 8, 3),           FP4_SHAPE(float, FP4_PREC_FP16X3, 9, 3, 8, 3),
    FP4_SHAPE(float, FP4_PREC_FP16X3, 12, 1, 8, 3),
    FP4_SHAPE(float, FP4_PREC_FP16X2, 3, 3, 8, 4),           FP4_SHAPE(float, FP4_PREC_FP16X2, 6, 6, 8, 4),
    FP4_SHAPE(float, FP4_PREC_FP16X2, 9, 3, 8, 4),           FP4_SHAPE(float, FP4_PREC_FP16X2, 12, 1, 8, 4),
    FP4_SHAPE(double, FP4_PREC_FP64_DENSE, 3, 3, 4, 0),      FP4_SHAPE(double, FP4_PREC_FP64_DENSE, 6, 3, 4, 0),
    FP4_SHAPE(double, FP4_PREC_FP64_DENSE, 6, 1, 4, 0),      FP4_SHAPE(double, FP4_PREC_FP64_DENSE, 9, 1, 2, 0),
    FP4_SHAPE(double, FP4_PREC_FP64_DENSE, 12, 1, 2, 0),
};

// Class shapes of k_dense_multi, in its class-slot order (Nbands 6, 12, 18,
// 24), for the FP32 and the FP16X2 batch; only their LDS layout functions
// are used (the launch is k_dense_multi's).
#define FP4_CLASS_SHAPE(P, NB, IB, MF) \
  {P, NB, IB, 8, MF, 0, &lds_bytes_geom<float, NB, IB, 8, MF>, &static_bytes<float, NB, IB, 8>, nullptr, nullptr}
const DenseShape kMultiShapes[2][kMultiClasses] = {
    {FP4_CLASS_SHAPE(FP4_PREC_FP32, 6, 6, 1), FP4_CLASS_SHAPE(FP4_PREC_FP32, 12, 1, 1),
     FP4_CLASS_SHAPE(FP4_PREC_FP32, 18, 1, 1), FP4_CLASS_SHAPE(FP4_PREC_FP32, 24, 1, 1)},
    {FP4_CLASS_SHAPE(FP4_PREC_FP16X2, 6, 6, 4), FP4_CLASS_SHAPE(FP4_PREC_FP16X2, 12, 1, 4),
     FP4_CLASS_SHAPE(FP4_PREC_FP32, 18, 1, 1), FP4_CLASS_SHAPE(FP4_PREC_FP32, 24, 1, 1)},
};
static_assert(kMultiNb[0] == 6 && kMultiNb[1] == 12 && kMultiNb[2] == 18 && kMultiNb[3] == 24 &&
                  kMultiIb[0] == 6 && kMultiIb[1] == 1 && kMultiIb[2] == 1 && kMultiIb[3] == 1,
              "class shapes of k_dense_multi");
// static LDS of k_dense_multi (its queue slot) ahead of the dynamic region
constexpr size_t kMultiStatic = 16;

template <int MFS>
void launch_multi(const DenseMultiArgs &ma, int slots, size_t dyn, hipStream_t st) {
  (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_dense_multi<MFS>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
  hipLaunchKernelGGL((k_dense_multi<MFS>), dim3(slots), dim3(NT_DENSE), dyn, st, ma);
}

template <int MFS>
int per_cu_multi(size_t dyn) {
  (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_dense_multi<MFS>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, reinterpret_cast<const void *>(&k_dense_multi<MFS>),
                                                   NT_DENSE, dyn) != hipSuccess)
    return 1;
  return std::max(n, 1);
}

__global__ void k_zero_int(int *p, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0;
}

// All shapes in order of preference; the first whose LDS fits is used when
// no tuning is given: k_dense2 with every branch of a point in one block
// (IB = Nb) first, then k_dense, then the remaining k_dense2 shapes (fewer
// modes per block share less work; the autotuner still times them).
const std::vector<DenseShape> &all_shapes() {
  static const std::vector<DenseShape> v = [] {
    std::vector<DenseShape> r;
    int n2 = 0;
    const DenseShape *s2 = dense2_shapes(&n2);
    for (int i = 0; i < n2; i++)
      if (s2[i].IB == s2[i].nb) r.push_back(s2[i]);
    r.insert(r.end(), std::begin(kShapes1), std::end(kShapes1));
    for (int i = 0; i < n2; i++)
      if (s2[i].IB != s2[i].nb) r.push_back(s2[i]);
    return r;
  }();
  return v;
}

DenseGeom geom_of(const DenseState *s) {
  DenseGeom g{};
  g.ng = s->ng;
  g.ntri = s->nc;
  g.L = s->L;
  g.npairs = s->natoms * s->natoms;
  for (int q = 0; q <= kDenseMaxPairs; q++) g.klf[q] = s->klf[q];
  return g;
}

const DenseShape *pick_shape(int prec, int nb, const DenseGeom &g, size_t lds_max) {
  for (const DenseShape &s : all_shapes())
    if (s.prec == prec && s.nb == nb && s.bytes(g) <= lds_max) return &s;
  return nullptr;
}

int ensure(void *&p, size_t &bytes, size_t need) {
  if (need <= bytes) return 0;
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
  if (hipMalloc(&p, need) != hipSuccess) return derr(FP4_ERR_NOMEM, "dense: hipMalloc failed");
  bytes = need;
  return 0;
}

// Device memory per resident block for A' of the large-cell shapes (0 for
// shapes that keep A' in LDS).
size_t ag_bytes(const DenseShape &sh, int ng) {
  if (sh.nb <= 12) return 0;
  const size_t el = (sh.prec == FP4_PREC_FP64_DENSE) ? 2 * sizeof(double) : 2 * sizeof(float);
  return ((size_t)ng * sh.IB * sh.nb * 9 * el + 255) & ~(size_t)255;
}

}  // namespace

// One crystal's launch arguments and the buffers its results go to.
struct DensePrep {
  DenseArgs a{};
  const DenseShape *sh = nullptr;
  DenseGeom geo{};
  std::vector<int> lls;
  int nls = 0, nib = 0, nii = 0;
  long nitems = 0, nseg = 0;
  double *og = nullptr, *ow = nullptr, *sg = nullptr, *sw = nullptr;
  int *on = nullptr, *ot = nullptr;
  long long *sn = nullptr;
};

namespace {

// Chooses the kernel shape of one call (tuned, else the first that fits).
const DenseShape *choose_shape(DenseState *s, const DenseRun &r, size_t ldsmax) {
  const int nb = r.m.nb;
  const bool tuned = r.tuning && r.tuning->shape_prec == r.precision;
  const DenseGeom geo = geom_of(s);
  const std::vector<DenseShape> &shapes = all_shapes();
  if (tuned && r.tuning->shape >= 0 && r.tuning->shape < (int)shapes.size()) {
    const DenseShape &c = shapes[r.tuning->shape];
    if (c.prec == r.precision && c.nb == nb && c.bytes(geo) <= ldsmax) return &c;
  }
  return pick_shape(r.precision, nb, geo, ldsmax);
}

// Everything of a dense call up to the launch: mode selection, output
// buffers, uploads, the Bose-Einstein table and the kernel arguments (all
// on stream r.stream).
int dense_prepare(DenseState *s, const DenseRun &r, const std::vector<int> &modes, int ii_lo0, int nii,
                  const DenseShape &sh, DensePrep &p) {
  const int nb = r.m.nb;
  p.sh = &sh;
  p.geo = geom_of(s);
  const DenseGeom &geo = p.geo;
  s->last_mfma = sh.mfma;
  s->last_tq = sh.TQ;
  s->last_ib = sh.IB;
  s->last_kg = sh.kg;
  const int nmodes_all = nb * r.m.nlist;
  std::vector<unsigned char> sel(nmodes_all, 0);
  p.lls.clear();
  for (int mm : modes) {
    sel[mm - 1] = 1;
    const int ll = (mm - 1) / nb;
    if (p.lls.empty() || p.lls.back() != ll) p.lls.push_back(ll);
  }
  const int nls = (int)p.lls.size(), nib = nb / sh.IB;
  const long nitems = 3L * nls * nib * nii;
  p.nls = nls;
  p.nib = nib;
  p.nii = nii;
  p.nitems = nitems;
  int rc;
  const size_t per_item = sizeof(double) * 2 + sizeof(int);
  const long nseg = 3L * nls * nib * sh.IB;
  p.nseg = nseg;
  // outputs: og, ow, on per item*IB, then sg, sw, sn per segment, then lls and sel
  const size_t need = per_item * nitems * sh.IB + sizeof(int) * nitems + (sizeof(double) * 2 + sizeof(long long)) * nseg +
                      sizeof(int) * (nls + s->ng) + nmodes_all + 1024;
  if ((rc = ensure(s->outb, s->out_bytes, need))) return rc;
  char *q = reinterpret_cast<char *>(s->outb);
  p.og = reinterpret_cast<double *>(q);
  q += sizeof(double) * nitems * sh.IB;
  p.ow = reinterpret_cast<double *>(q);
  q += sizeof(double) * nitems * sh.IB;
  p.sg = reinterpret_cast<double *>(q);
  q += sizeof(double) * nseg;
  p.sw = reinterpret_cast<double *>(q);
  q += sizeof(double) * nseg;
  p.sn = reinterpret_cast<long long *>(q);
  q += sizeof(long long) * nseg;
  p.on = reinterpret_cast<int *>(q);
  q += sizeof(int) * nitems * sh.IB;
  p.ot = reinterpret_cast<int *>(q);
  q += sizeof(int) * nitems;
  int *dlls = reinterpret_cast<int *>(q);
  q += sizeof(int) * nls;
  int *dgslot = reinterpret_cast<int *>(q);
  q += sizeof(int) * s->ng;
  unsigned char *dsel = reinterpret_cast<unsigned char *>(q);
  // k_dense2: phase-table slots of the coarse groups for this k-step shape
  int pair_off[17] = {};
  std::vector<int> gslot(std::max(s->ng, 1), 0);
  if (sh.kg > 0) dense2_slots(geo, sh.kg, pair_off, gslot.data());
  hipStream_t st = r.stream;
  if (hipMemcpyAsync(dlls, p.lls.data(), sizeof(int) * nls, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(dgslot, gslot.data(), sizeof(int) * s->ng, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(dsel, sel.data(), nmodes_all, hipMemcpyHostToDevice, st) != hipSuccess)
    return derr(FP4_ERR_DEVICE, "dense: upload failed");
  const long nmt = (long)nb * r.m.nptk;
  if ((rc = ensure(s->fbose, s->fbose_bytes, sizeof(double) * nmt))) return rc;
  hipLaunchKernelGGL(k_bose, dim3((unsigned)((nmt + 255) / 256)), dim3(256), 0, st, r.m.energy, nmt, r.T,
                     reinterpret_cast<double *>(s->fbose));
  if (hipGetLastError() != hipSuccess) return derr(FP4_ERR_DEVICE, "dense: Bose-Einstein table launch failed");
  DenseArgs &a = p.a;
  a = DenseArgs{};
  a.m = r.m;
  a.roots = r.roots;
  a.L = s->L;
  a.g = s->g;
  a.f = s->f;
  a.c = s->c;
  a.phi = s->phi;
  for (int k = 0; k <= kDenseMaxPairs; k++) a.klf[k] = s->klf[k];
  a.roots_lds = s->L <= kRootsLds;
  a.fbose = reinterpret_cast<const double *>(s->fbose);
  a.ng = s->ng;
  a.lls = dlls;
  a.nls = nls;
  a.nib = nib;
  a.ii_lo0 = ii_lo0;
  a.nii = nii;
  a.sel = dsel;
  a.T = r.T;
  a.onlyharmonic = r.onlyharmonic;
  a.nitems = nitems;
  a.counter = s->counter;
  a.ntri = s->nc;
  a.smem_static = (int)sh.stat();
  {
    const int ng3[3] = {r.m.n1, r.m.n2, r.m.n3};
    for (int nu = 0; nu < 3; nu++)
      for (int c = 0; c < 3; c++) a.bf[3 * nu + c] = (float)(r.m.b[3 * nu + c] / ng3[nu]);
    a.sbf = (float)std::fabs(r.m.scalebroad);
    a.a_scale = s->a_scale;
  }
  for (int k = 0; k < 17; k++) a.pair_off[k] = pair_off[k];
  a.gpt = sh.kg > 0 ? dense2_gpt(geo, sh.kg) : 0;
  a.gslot = dgslot;
  a.lmask = (s->L > 0 && (s->L & (s->L - 1)) == 0) ? s->L - 1 : -1;
#ifdef FP4_PHASE_PROF
  if (!s->prof && hipMalloc(&s->prof, sizeof(unsigned long long) * kDensePhases) != hipSuccess)
    return derr(FP4_ERR_NOMEM, "dense: hipMalloc failed");
  if (hipMemsetAsync(s->prof, 0, sizeof(unsigned long long) * kDensePhases, st) != hipSuccess)
    return derr(FP4_ERR_DEVICE, "dense: memset failed");
#endif
  a.prof = s->prof;
  a.wop = r.wop;
  a.wcols = r.wcols;
  a.wscale = (double)kGammaUnits * 1.e9 / r.m.nptk / r.m.nptk;
  a.out_g = p.og;
  a.out_w = p.ow;
  a.out_n = p.on;
  a.out_t = p.ot;
  return 0;
}

// The per-segment reduction after the launch (queued on st).
void dense_reduce_launch(const DensePrep &p, hipStream_t st) {
  hipLaunchKernelGGL(k_dense_reduce, dim3((unsigned)((p.nseg + 127) / 128)), dim3(128), 0, st, p.og, p.ow, p.on,
                     p.nseg, p.sh->IB, p.nii, p.sg, p.sw, p.sn);
}

// Copies one crystal's segment sums back and maps them to (channel, mode).
int dense_collect(DenseState *s, const DensePrep &p, int nb, const std::vector<int> &modes, hipStream_t st,
                  std::vector<double> &g, std::vector<double> &w, std::vector<long long> &n) {
  const long nseg = p.nseg, nitems = p.nitems;
  const int nls = p.nls, nib = p.nib, IB = p.sh->IB;
  std::vector<double> hg(nseg), hw(nseg);
  std::vector<long long> hn(nseg);
  if (hipMemcpyAsync(hg.data(), p.sg, sizeof(double) * nseg, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(hw.data(), p.sw, sizeof(double) * nseg, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(hn.data(), p.sn, sizeof(long long) * nseg, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return derr(FP4_ERR_DEVICE, "dense: kernel or copy failed");
  {
    std::vector<int> ht(nitems);
    if (hipMemcpy(ht.data(), p.ot, sizeof(int) * nitems, hipMemcpyDeviceToHost) != hipSuccess)
      return derr(FP4_ERR_DEVICE, "dense: copy failed");
    long long tt = 0;
    for (int x : ht) tt += x;
    s->last_triples = tt;
    s->last_items = nitems;
#ifdef FP4_PHASE_PROF
    if (hipMemcpy(s->last_prof, s->prof, sizeof(unsigned long long) * kDensePhases, hipMemcpyDeviceToHost) !=
        hipSuccess)
      return derr(FP4_ERR_DEVICE, "dense: copy failed");
#endif
  }
  // segments are (ch, ls, ib, il); map to (ch, mode-in-subset)
  const int nsub = (int)modes.size();
  g.assign(3 * nsub, 0.0);
  w.assign(3 * nsub, 0.0);
  n.assign(3 * nsub, 0);
  std::map<int, int> mpos;
  for (int t = 0; t < nsub; t++) mpos[modes[t]] = t;
  for (int ch = 0; ch < 3; ch++)
    for (int ls = 0; ls < nls; ls++)
      for (int ib = 0; ib < nib; ib++)
        for (int il = 0; il < IB; il++) {
          const int mm = p.lls[ls] * nb + ib * IB + il + 1;
          auto it = mpos.find(mm);
          if (it == mpos.end()) continue;
          const long sidx = (((long)ch * nls + ls) * nib + ib) * IB + il;
          g[ch * nsub + it->second] = hg[sidx];
          w[ch * nsub + it->second] = hw[sidx];
          n[ch * nsub + it->second] = hn[sidx];
        }
  return 0;
}

}  // namespace

int dense_rates(DenseState *s, const DenseRun &r, const std::vector<int> &modes, int ii_lo0, int nii,
                std::vector<double> &g, std::vector<double> &w, std::vector<long long> &n, double *ms) {
  if (!s) return derr(FP4_ERR_STATE, "dense state missing");
  if (r.precision != FP4_PREC_FP32 && r.precision != FP4_PREC_FP64_DENSE && r.precision != FP4_PREC_BF16X3 &&
      r.precision != FP4_PREC_FP16X3 && r.precision != FP4_PREC_FP16X2)
    return derr(FP4_ERR_UNSUPPORTED, "dense form: precision not built");
  int dev = 0, ldsmax = 0, ncu = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&ldsmax, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  if (s->natoms > kDenseMaxAtoms) return derr(FP4_ERR_UNSUPPORTED, "dense form: more than 8 atoms per cell");
  const DenseShape *shp = choose_shape(s, r, (size_t)ldsmax);
  if (!shp) return derr(FP4_ERR_UNSUPPORTED, "dense form: no kernel shape fits this Nbands / cluster set in LDS");
  const bool tuned = r.tuning && r.tuning->shape_prec == r.precision;
  DensePrep p;
  int rc;
  if ((rc = dense_prepare(s, r, modes, ii_lo0, nii, *shp, p))) return rc;
  int slots = (tuned && r.tuning->slots > 0) ? r.tuning->slots : std::max(1, ncu);
  if (slots > p.nitems) slots = (int)std::max(1L, p.nitems);
  hipStream_t st = r.stream;
  if (const size_t agb = ag_bytes(*shp, s->ng)) {
    if ((rc = ensure(s->scratch, s->scratch_bytes, agb * (size_t)slots))) return rc;
    p.a.ag = s->scratch;
    p.a.ag_stride = (long long)agb;
  }
  if (hipMemsetAsync(s->counter, 0, sizeof(int), st) != hipSuccess) return derr(FP4_ERR_DEVICE, "dense: memset failed");
  if (hipEventRecord(s->e0, st) != hipSuccess) return derr(FP4_ERR_DEVICE, "event record failed");
  shp->launch(p.a, slots, shp->bytes(p.geo), st);
  if (hipGetLastError() != hipSuccess) return derr(FP4_ERR_DEVICE, "dense kernel launch failed");
  if (hipEventRecord(s->e1, st) != hipSuccess) return derr(FP4_ERR_DEVICE, "event record failed");
  dense_reduce_launch(p, st);
  if ((rc = dense_collect(s, p, r.m.nb, modes, st, g, w, n))) return rc;
  float kms = 0.f;
  (void)hipEventElapsedTime(&kms, s->e0, s->e1);
  if (ms) *ms = kms;
  return 0;
}

// ------------------------------------------------------------ batches
struct DenseBatchState {
  // crystals of one Nbands (one class slot of k_dense_multi)
  struct Class {
    int nb = 0;
    const DenseShape *sh = nullptr;
    DenseArgs ka{};                      // the class's launch arguments (LDS layout, argument table)
    std::vector<int> members;
    DenseArgs *d_args = nullptr;         // per-crystal arguments on the device
    size_t args_cap = 0;
    std::vector<DenseArgs> h_args;       // as uploaded (graph replays are valid while unchanged)
  };
  // one set of device tables and one graph per q' part of the crystals
  struct Part {
    Class classes[kMultiClasses];
    DenseArgs *d_cls = nullptr;          // [kMultiClasses] class arguments of k_dense_multi
    std::vector<DenseArgs> h_cls;
    long long *d_order = nullptr;        // every class's items, costliest first
    size_t order_cap = 0;
    std::vector<long long> h_order;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graph_prec = -1;
    int ag_gen = -1;                     // DenseBatchState::ag_gen the graph was captured with
    bool graph_valid = false;
  };
  std::map<int, Part> parts;
  int *counter = nullptr;                // the queue counter of k_dense_multi
  void *ag = nullptr;                    // A' of the large-cell classes (DenseArgs::ag)
  size_t ag_bytes = 0;
  int ag_gen = 0;                        // bumped when ag moves (captured graphs are stale)
  hipEvent_t e0 = nullptr, e1 = nullptr;
};

DenseBatchState *dense_batch_create() {
  DenseBatchState *b = new DenseBatchState();
  if (hipEventCreate(&b->e0) != hipSuccess || hipEventCreate(&b->e1) != hipSuccess ||
      hipMalloc(&b->counter, sizeof(int)) != hipSuccess) {
    dense_batch_destroy(b);
    return nullptr;
  }
  return b;
}

void dense_batch_destroy(DenseBatchState *b) {
  if (!b) return;
  for (auto &kv : b->parts) {
    for (auto &c : kv.second.classes)
      if (c.d_args) (void)hipFree(c.d_args);
    if (kv.second.d_cls) (void)hipFree(kv.second.d_cls);
    if (kv.second.d_order) (void)hipFree(kv.second.d_order);
    if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
    if (kv.second.graph) (void)hipGraphDestroy(kv.second.graph);
  }
  if (b->counter) (void)hipFree(b->counter);
  if (b->ag) (void)hipFree(b->ag);
  if (b->e0) (void)hipEventDestroy(b->e0);
  if (b->e1) (void)hipEventDestroy(b->e1);
  delete b;
}

namespace {
// (Re)allocates a device table of n elements of T when it has fewer; marks
// *changed when the pointer moves.
template <class T>
int ensure_table(T *&p, size_t &cap, size_t n, bool *changed) {
  if (n <= cap) return 0;
  if (p) (void)hipFree(p);
  p = nullptr;
  cap = 0;
  if (hipMalloc(&p, sizeof(T) * n) != hipSuccess) return derr(FP4_ERR_NOMEM, "batch: hipMalloc failed");
  cap = n;
  *changed = true;
  return 0;
}
}  // namespace

int dense_rates_batch(DenseBatchState *b, std::vector<DenseBatchMember> &mem, int precision, hipStream_t st,
                      bool use_graph, int part, double *kernel_ms, int *launches) {
  if (!b) return derr(FP4_ERR_STATE, "batch state missing");
  DenseBatchState::Part &P = b->parts[part];
  if (precision != FP4_PREC_FP32 && precision != FP4_PREC_FP16X2)
    return derr(FP4_ERR_UNSUPPORTED, "batch: FP32 or FP16X2 MFMA forms only");
  const int form = (precision == FP4_PREC_FP16X2) ? 1 : 0;
  int dev = 0, ldsmax = 0, ncu = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&ldsmax, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  // classes by Nbands, in k_dense_multi's slot order; each class's LDS layout
  // is that of the largest group / cluster / root counts of its crystals
  DenseBatchState::Class cls[kMultiClasses];
  for (int i = 0; i < (int)mem.size(); i++) {
    DenseBatchMember &x = mem[i];
    if (!x.s || x.s->natoms > kDenseMaxAtoms)
      return derr(FP4_ERR_UNSUPPORTED, ("batch: crystal " + std::to_string(i) + " has more than 8 atoms").c_str());
    const int nb = x.r.m.nb;
    int k = 0;
    while (k < kMultiClasses && kMultiNb[k] != nb) k++;
    if (k == kMultiClasses)
      return derr(FP4_ERR_UNSUPPORTED, ("batch: no class shape for Nbands=" + std::to_string(nb)).c_str());
    cls[k].nb = nb;
    cls[k].members.push_back(i);
  }
  size_t dyn = 0;
  for (int k = 0; k < kMultiClasses; k++) {
    auto &c = cls[k];
    if (c.members.empty()) continue;
    c.sh = &kMultiShapes[form][k];
    DenseGeom gmax{};
    for (int i : c.members) {
      const DenseGeom g = geom_of(mem[i].s);
      gmax.ng = std::max(gmax.ng, g.ng);
      gmax.ntri = std::max(gmax.ntri, g.ntri);
      gmax.L = std::max(gmax.L, g.L);
    }
    const size_t bytes = c.sh->bytes(gmax);
    if (bytes + kMultiStatic > (size_t)ldsmax)
      return derr(FP4_ERR_UNSUPPORTED,
                  ("batch: the Nbands=" + std::to_string(c.nb) + " class does not fit in LDS").c_str());
    dyn = std::max(dyn, bytes);
    c.ka = DenseArgs{};
    c.ka.ng = gmax.ng;
    c.ka.ntri = gmax.ntri;
    c.ka.L = gmax.L;
    c.ka.roots_lds = gmax.L <= kRootsLds ? 1 : 0;   // the roots go to LDS only when every table fits
    c.ka.smem_static = (int)c.sh->stat();
  }
  const int per_cu = form ? per_cu_multi<4>(dyn) : per_cu_multi<1>(dyn);
  const int slots = std::max(1, ncu) * std::min(per_cu, 2);
  // A' of the large-cell classes in device memory: one region per resident
  // block, as large as the largest class needs (a block runs items of every
  // class)
  bool changed = !P.graph_valid || P.graph_prec != precision;
  {
    size_t agb = 0;
    for (auto &c : cls)
      if (c.sh) agb = std::max(agb, ag_bytes(*c.sh, c.ka.ng));
    const size_t need = agb * (size_t)slots;
    if (need > b->ag_bytes) {
      if (b->ag) (void)hipFree(b->ag);
      b->ag = nullptr;
      b->ag_bytes = 0;
      if (hipMalloc(&b->ag, need) != hipSuccess) return derr(FP4_ERR_NOMEM, "batch: hipMalloc failed");
      b->ag_bytes = need;
      b->ag_gen++;
    }
    for (auto &c : cls)
      if (c.sh && ag_bytes(*c.sh, c.ka.ng) > 0) {
        c.ka.ag = b->ag;
        c.ka.ag_stride = (long long)agb;
      }
    changed = changed || P.ag_gen != b->ag_gen;
  }
  // per crystal: uploads, Bose-Einstein table and arguments (outside the graph)
  std::vector<DensePrep> prep(mem.size());
  int rc;
  for (auto &c : cls)
    for (int i : c.members) {
      DenseBatchMember &x = mem[i];
      DenseRun r = x.r;
      r.stream = st;
      r.precision = c.sh->prec;
      if ((rc = dense_prepare(x.s, r, x.modes, x.ii_lo0, x.nii, *c.sh, prep[i]))) return rc;
      prep[i].a.roots_lds = c.ka.roots_lds;
    }
  // cost of an item (channel, point, mode block, q'): its processes (the
  // prefilter counts of its modes at that q') plus the filter work every
  // candidate q'' costs, about Nbands^2 candidates' worth per q''
  std::vector<std::pair<double, long long>> items;
  for (int k = 0; k < kMultiClasses; k++) {
    auto &c = cls[k];
    auto &old = P.classes[k];
    c.h_args.clear();
    for (int i : c.members) c.h_args.push_back(prep[i].a);
    for (size_t mi = 0; mi < c.members.size(); mi++) {
      const int i = c.members[mi];
      const DenseBatchMember &x = mem[i];
      const DensePrep &p = prep[i];
      const int nsub = (int)x.modes.size(), IB = c.sh->IB;
      const bool have = x.counts.size() == (size_t)3 * nsub * x.nii;
      std::map<int, int> mpos;
      for (int t = 0; t < nsub; t++) mpos[x.modes[t]] = t;
      for (long t = 0; t < p.nitems; t++) {
        const int ii = (int)(t % p.nii);
        long r_ = t / p.nii;
        const int ib = (int)(r_ % p.nib);
        r_ /= p.nib;
        const int ls = (int)(r_ % p.nls), ch = (int)(r_ / p.nls);
        double cost = (double)x.r.m.nptk * (c.nb / 6) * (c.nb / 6);
        if (have)
          for (int il = 0; il < IB; il++) {
            auto it = mpos.find(p.lls[ls] * c.nb + ib * IB + il + 1);
            if (it != mpos.end()) cost += (double)x.counts[((size_t)ch * nsub + it->second) * x.nii + ii];
          }
        items.push_back({cost, ((long long)k << kMultiShift) | ((long long)mi << 40) | (long long)t});
      }
    }
    // keep the device table of the previous call when it fits
    c.d_args = old.d_args;
    c.args_cap = old.args_cap;
    old.d_args = nullptr;
    old.args_cap = 0;
    if ((rc = ensure_table(c.d_args, c.args_cap, c.h_args.size(), &changed))) return rc;
    if (old.h_args.size() != c.h_args.size() ||
        std::memcmp(old.h_args.data(), c.h_args.data(), sizeof(DenseArgs) * c.h_args.size()) != 0 || old.sh != c.sh)
      changed = true;
    if (!c.h_args.empty() && hipMemcpyAsync(c.d_args, c.h_args.data(), sizeof(DenseArgs) * c.h_args.size(),
                                            hipMemcpyHostToDevice, st) != hipSuccess)
      return derr(FP4_ERR_DEVICE, "batch: upload failed");
    c.ka.batch = c.d_args;
    c.ka.nbatch = (int)c.members.size();
    c.ka.counter = b->counter;
  }
  std::stable_sort(items.begin(), items.end(), [](const auto &u, const auto &v) { return u.first > v.first; });
  std::vector<long long> order(items.size());
  for (size_t t = 0; t < items.size(); t++) order[t] = items[t].second;
  std::vector<DenseArgs> hcls(kMultiClasses);
  for (int k = 0; k < kMultiClasses; k++) hcls[k] = cls[k].ka;
  size_t cls_cap = P.d_cls ? kMultiClasses : 0;
  if ((rc = ensure_table(P.d_cls, cls_cap, kMultiClasses, &changed))) return rc;
  if ((rc = ensure_table(P.d_order, P.order_cap, std::max<size_t>(order.size(), 1), &changed))) return rc;
  if (P.h_order != order || P.h_cls.size() != hcls.size() ||
      std::memcmp(P.h_cls.data(), hcls.data(), sizeof(DenseArgs) * hcls.size()) != 0)
    changed = true;
  if (hipMemcpyAsync(P.d_cls, hcls.data(), sizeof(DenseArgs) * kMultiClasses, hipMemcpyHostToDevice, st) !=
          hipSuccess ||
      (!order.empty() && hipMemcpyAsync(P.d_order, order.data(), sizeof(long long) * order.size(),
                                        hipMemcpyHostToDevice, st) != hipSuccess))
    return derr(FP4_ERR_DEVICE, "batch: upload failed");
  for (int k = 0; k < kMultiClasses; k++) {
    if (P.classes[k].d_args) (void)hipFree(P.classes[k].d_args);
    P.classes[k] = std::move(cls[k]);
  }
  P.h_cls = std::move(hcls);
  P.h_order = std::move(order);
  DenseMultiArgs ma{};
  ma.cls = P.d_cls;
  ma.order = P.d_order;
  ma.nitems = (long)P.h_order.size();
  ma.counter = b->counter;
  // the device work of a step: the queue counter, one persistent launch over
  // every crystal, the per-crystal reductions
  auto enqueue = [&](hipStream_t s2) {
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, s2, b->counter, 1);
    if (ma.nitems > 0) {
      const int nblk = (int)std::min<long>(slots, ma.nitems);
      if (form)
        launch_multi<4>(ma, nblk, dyn, s2);
      else
        launch_multi<1>(ma, nblk, dyn, s2);
    }
    for (auto &c : P.classes)
      for (int i : c.members) dense_reduce_launch(prep[i], s2);
  };
  int nl = 1 + (ma.nitems > 0 ? 1 : 0);
  for (auto &c : P.classes) nl += (int)c.members.size();
  if (hipEventRecord(b->e0, st) != hipSuccess) return derr(FP4_ERR_DEVICE, "event record failed");
  if (use_graph) {
    if (changed) {
      if (P.exec) (void)hipGraphExecDestroy(P.exec);
      if (P.graph) (void)hipGraphDestroy(P.graph);
      P.exec = nullptr;
      P.graph = nullptr;
      P.graph_valid = false;
      if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess)
        return derr(FP4_ERR_DEVICE, "batch: stream capture failed");
      enqueue(st);
      if (hipStreamEndCapture(st, &P.graph) != hipSuccess ||
          hipGraphInstantiate(&P.exec, P.graph, nullptr, nullptr, 0) != hipSuccess)
        return derr(FP4_ERR_DEVICE, "batch: graph capture failed");
      P.graph_valid = true;
      P.graph_prec = precision;
      P.ag_gen = b->ag_gen;
    }
    if (hipGraphLaunch(P.exec, st) != hipSuccess) return derr(FP4_ERR_DEVICE, "batch: graph launch failed");
  } else {
    enqueue(st);
    P.graph_valid = false;
  }
  if (hipGetLastError() != hipSuccess) return derr(FP4_ERR_DEVICE, "batch: launch failed");
  if (hipEventRecord(b->e1, st) != hipSuccess) return derr(FP4_ERR_DEVICE, "event record failed");
  for (auto &c : P.classes)
    for (int i : c.members) {
      DenseBatchMember &x = mem[i];
      if ((rc = dense_collect(x.s, prep[i], x.r.m.nb, x.modes, st, x.g, x.w, x.n))) return rc;
    }
  float kms = 0.f;
  (void)hipEventElapsedTime(&kms, b->e0, b->e1);
  if (kernel_ms) *kernel_ms = kms;
  if (launches) *launches = nl;
  return 0;
}

int dense_autotune(DenseState *s, const DenseRun &r, const std::vector<int> &modes, int ii_lo0, int nii,
                   DenseTuning *out, double *best_ms, int *ncand) {
  if (!s || !out) return derr(FP4_ERR_ARG, "autotune: null argument");
  int dev = 0, ldsmax = 0, ncu = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&ldsmax, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  const std::vector<DenseShape> &shapes = all_shapes();
  const DenseGeom geo = geom_of(s);
  const int nshapes = (int)shapes.size();
  s->cands.clear();
  DenseTuning best;
  double tbest = 0.0;
  int n = 0;
  std::vector<double> g, w;
  std::vector<long long> cnt;
  for (int i = 0; i < nshapes; i++) {
    const DenseShape &c = shapes[i];
    if (c.prec != r.precision || c.nb != r.m.nb) continue;
    const size_t dyn = c.bytes(geo);
    if (dyn > (size_t)ldsmax) continue;
    const int per_cu = std::min(c.per_cu(dyn), 2);
    for (int k = 1; k <= per_cu; k++) {
      DenseTuning t;
      t.shape = i;
      t.shape_prec = r.precision;
      t.slots = k * std::max(ncu, 1);
      DenseRun rr = r;
      rr.tuning = &t;
      double ms = 0.0;
      int rc;
      // the first run of a shape pays its one-time set-up; the second is timed
      if ((rc = dense_rates(s, rr, modes, ii_lo0, nii, g, w, cnt, &ms))) return rc;
      if ((rc = dense_rates(s, rr, modes, ii_lo0, nii, g, w, cnt, &ms))) return rc;
      n++;
      s->cands.push_back({i, t.slots, ms});
      if (best.shape < 0 || ms < tbest) {
        best = t;
        tbest = ms;
      }
    }
  }
  if (best.shape < 0) return derr(FP4_ERR_UNSUPPORTED, "autotune: no kernel shape fits this Nbands / cluster set");
  *out = best;
  if (best_ms) *best_ms = tbest;
  if (ncand) *ncand = n;
  s->last_mfma = shapes[best.shape].mfma;
  s->last_tq = shapes[best.shape].TQ;
  s->last_ib = shapes[best.shape].IB;
  s->last_kg = shapes[best.shape].kg;
  return 0;
}

int dense_phase_profile(const DenseState *s, int cap, double *cycles) {
#ifdef FP4_PHASE_PROF
  if (!s) return 0;
  for (int p = 0; p < kDensePhases && p < cap; p++) cycles[p] = (double)s->last_prof[p];
  return kDensePhases;
#else
  (void)s;
  (void)cap;
  (void)cycles;
  return -FP4_ERR_UNSUPPORTED;
#endif
}

int dense_candidates(const DenseState *s, int cap, int *kg, int *ib, int *tq, int *slots, double *ms) {
  if (!s) return 0;
  const std::vector<DenseShape> &shapes = all_shapes();
  const int n = (int)s->cands.size();
  for (int i = 0; i < n && i < cap; i++) {
    const DenseShape &c = shapes[s->cands[i].shape];
    if (kg) kg[i] = c.kg;
    if (ib) ib[i] = c.IB;
    if (tq) tq[i] = c.TQ;
    if (slots) slots[i] = s->cands[i].slots;
    if (ms) ms[i] = s->cands[i].ms;
  }
  return n;
}

int dense_pick(const DenseState *s, int precision, int nb, int kg, int ib, int slots, DenseTuning *out) {
  if (!s || !out) return derr(FP4_ERR_ARG, "dense_pick: null argument");
  int dev = 0, ldsmax = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&ldsmax, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
  const std::vector<DenseShape> &shapes = all_shapes();
  const DenseGeom geo = geom_of(s);
  for (int i = 0; i < (int)shapes.size(); i++) {
    const DenseShape &c = shapes[i];
    if (c.prec == precision && c.nb == nb && c.kg == kg && c.IB == ib && c.bytes(geo) <= (size_t)ldsmax) {
      out->shape = i;
      out->shape_prec = precision;
      out->slots = slots;
      return 0;
    }
  }
  return derr(FP4_ERR_UNSUPPORTED, "dense_pick: no built shape with this precision, Nbands, IB and k-step fits");
}

}  // namespace fp4
