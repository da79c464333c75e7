// Kernel (a): the energy-conservation prefilter and the process-list
// compaction of NP_driver_4ph / Ind_only_driver_4ph
// (Src/processes.f90:2263-2476, 2854-3246).
//
// Work item = (channel, mode, second-phonon q index ii).  One 64-wide
// wavefront owns an item; lane t walks jj = jj0 + t, jj0 + t + 64, ...
// and, for each jj, all Nb^3 branch triples (j,k,l) in the reference's
// order.  Per candidate:
//   1. exact zero-frequency tests and the FP64 mismatch delta (reference
//      expression order);
//   2. a conservative float cull: sigma <= vb(jj,k) + vb(ss,l) where
//      vb = scalebroad * max_nu(|b_nu|/N_nu) * |v| / sqrt(2), padded so that
//      float rounding can only keep a candidate, never drop one;
//   3. a float estimate of sigma (relative error < 1e-6) that settles every
//      candidate whose |delta| is not within 1e-4 of it;
//   4. for the rest, the exact FP64 window test |delta| <= sigma with
//      base_sigma computed in the reference's operation order.
// Steps 2 and 3 only decide where the answer cannot differ from step 4, so
// counts and lists are bit-exact.  Before step 1, a lane skips a jj whose
// (k, l) pairs all put w''' -+ w'' farther than the sigma bounds from the
// item's range of w +- w'_j (triple_may_survive): such a jj has no survivor.
//
// Count pass: per-item survivor count (int32) and phase-space partial sum.
// Emit pass: per-lane counts, a wavefront exclusive scan in lane (= jj)
// order, then each lane writes its survivors at item_offset + lane_offset;
// the concatenation is the reference's (ii, jj, j, k, l) order.
// Hash pass: the emit pass's positions without the writes.  The hash of a
// (channel, mode) list of n entries is H = sum_p key_p M^(n-1-p) mod 2^64,
// key = (i2 << 42) ^ (i3 << 21) ^ i4.  A lane's run of c survivors starting
// at list position p0 contributes M^(n-p0-c) times its Horner sum
// (h = h*M + key over the run), so one candidate scan per lane suffices;
// the contributions are added with 64-bit atomics, whose sum mod 2^64 does
// not depend on their order.
#include "fp4_device.hpp"
#include "fp4_kernels.hpp"

namespace fp4 {

// multiplier of the order-sensitive list hash (the 64-bit golden ratio)
constexpr unsigned long long kListHashM = 0x9E3779B97F4A7C15ULL;

struct ItemDecode {
  int ch, mi, ii0;
};

__device__ __forceinline__ ItemDecode decode_item(long item, int nsub, int nii, int ii_lo0) {
  ItemDecode d;
  d.ii0 = (int)(item % nii) + ii_lo0;
  long r = item / nii;
  d.mi = (int)(r % nsub);
  d.ch = (int)(r / nsub);
  return d;
}

// Visits the survivors of one (item, lane-jj) in reference order.
template <int NB, typename F>
__device__ __forceinline__ void scan_candidates(const ModeView &m, int ch, int i0, int qi0, double w, int ii0,
                                                int jj0, int ss0, const double *Eii, F &&f) {
  double Ejj[NB], Ess[NB];
  float bjj[NB], bss[NB];
#pragma unroll
  for (int b = 0; b < NB; b++) {
    Ejj[b] = E(m, jj0, b);
    Ess[b] = E(m, ss0, b);
    bjj[b] = m.vbound[(size_t)b * m.nptk + jj0];
    bss[b] = m.vbound[(size_t)b * m.nptk + ss0];
  }
  const bool same23 = (ch != 1) && (jj0 == ii0);
  const bool same34 = (ch != 0) && (ss0 == jj0);
  // Two stages per j: the unrolled loop over (k, l) keeps only the cheap
  // tests (loop exclusions, zero frequencies, FP64 mismatch against the float
  // bound) and records the candidates that pass in a (k, l) bit mask; the
  // rest of the tests run for those bits in ascending order, which is the
  // reference's (k, l) order.  Keeping the rare path out of the unrolled body
  // is what keeps the register count (and occupancy) reasonable.
  if constexpr (NB > 8) {
    // Nb > 8: the (k, l) block is too large to unroll; k runs as a loop
    // (frequencies and bounds of q'' re-read from cache), l is unrolled into
    // a one-word mask, and the mask's bits are finished before the next k,
    // which keeps the (j, k, l) order.
    for (int j = 0; j < NB; j++) {
      const double op = Eii[j];
      if (!(op != 0.0)) continue;
#pragma unroll 1
      for (int k = (same23 ? j : 0); k < NB; k++) {
        const double odp = E(m, jj0, k);
        const float bk = m.vbound[(size_t)k * m.nptk + jj0];
        unsigned lm = 0u;
#pragma unroll
        for (int l = 0; l < NB; l++) {
          if (same34 && l < k) continue;
          const double otp = Ess[l];
          if (!((odp != 0.0) && (otp != 0.0))) continue;
          const double delta = mismatch(ch, w, op, odp, otp);
          const float bound = (bk + bss[l]) * 1.0001f + 1e-30f;
          if ((float)fabs(delta) > bound) continue;
          lm |= 1u << l;
        }
        while (lm) {
          const int l = __builtin_ctz(lm);
          lm &= lm - 1u;
          if (ch == 1) {
            if (!((qi0 != jj0 || i0 != k) && (qi0 != ss0 || i0 != l) && (ii0 != jj0 || j != k) &&
                  (ii0 != ss0 || j != l)))
              continue;
          }
          const double delta = mismatch(ch, w, op, odp, E(m, ss0, l));
          double v3[3], v4[3];
#pragma unroll
          for (int a = 0; a < 3; a++) {
            v3[a] = Vel(m, jj0, k, a);
            v4[a] = Vel(m, ss0, l, a);
          }
          const float ad = (float)fabs(delta);
          const float sf = sigma_estimate(m, ch, v3, v4);
          if (ad > sf * 1.0001f + 1e-30f) continue;
          if (ad < sf * 0.9999f) {
            f(j, k, l, delta);
          } else {
            const double sigma = pair_sigma(m, ch, jj0, k, ss0, l);
            if (fabs(delta) <= sigma) f(j, k, l, delta);
          }
        }
      }
    }
    return;
  }
  constexpr int MW = (NB * NB + 63) / 64;
  for (int j = 0; j < NB; j++) {
    const double op = Eii[j];
    if (!(op != 0.0)) continue;
    unsigned long long pass[MW];
#pragma unroll
    for (int w = 0; w < MW; w++) pass[w] = 0ull;
#pragma unroll
    for (int k = 0; k < NB; k++) {
      if (same23 && k < j) continue;
      const double odp = Ejj[k];
#pragma unroll
      for (int l = 0; l < NB; l++) {
        if (same34 && l < k) continue;
        const double otp = Ess[l];
        if (!((odp != 0.0) && (otp != 0.0))) continue;
        const double delta = mismatch(ch, w, op, odp, otp);
        const float bound = (bjj[k] + bss[l]) * 1.0001f + 1e-30f;
        if ((float)fabs(delta) > bound) continue;
        pass[(k * NB + l) >> 6] |= 1ull << ((k * NB + l) & 63);
      }
    }
#pragma unroll
    for (int wd = 0; wd < MW; wd++) {
      unsigned long long bits = pass[wd];
      while (bits) {
        const int bit = 64 * wd + __builtin_ctzll(bits);
        bits &= bits - 1ull;
        const int k = bit / NB, l = bit % NB;
        if (ch == 1) {
          if (!((qi0 != jj0 || i0 != k) && (qi0 != ss0 || i0 != l) && (ii0 != jj0 || j != k) &&
                (ii0 != ss0 || j != l)))
            continue;
        }
        const double delta = mismatch(ch, w, op, E(m, jj0, k), E(m, ss0, l));
        double v3[3], v4[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
          v3[a] = Vel(m, jj0, k, a);
          v4[a] = Vel(m, ss0, l, a);
        }
        const float ad = (float)fabs(delta);
        const float sf = sigma_estimate(m, ch, v3, v4);
        if (ad > sf * 1.0001f + 1e-30f) continue;
        if (ad < sf * 0.9999f) {
          f(j, k, l, delta);
        } else {
          const double sigma = pair_sigma(m, ch, jj0, k, ss0, l);
          if (fabs(delta) <= sigma) f(j, k, l, delta);
        }
      }
    }
  }
}

// For Nb <= 8 the kernel is asked to fit 3 waves per SIMD (168 VGPRs for
// Nb = 6): measured at 16^3, the count pass takes 4.53 s against 6.06 s at
// the compiler's 2 waves; larger Nb keep the compiler's choice.
template <int NB, int MODE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NB * NB <= 64 ? 3 : 1)))
k_prefilter(PrefilterArgs a) {
  const ModeView &m = a.m;
  const int lane = threadIdx.x & 63;
  const long nitems = 3L * a.nsub * a.nii;
  const int wpb = blockDim.x >> 6;
  for (long item = (long)blockIdx.x * wpb + (threadIdx.x >> 6) + a.item_lo; item < a.item_hi;
       item += (long)gridDim.x * wpb) {
    if (item >= nitems) break;
    const ItemDecode d = decode_item(item, a.nsub, a.nii, a.ii_lo0);
    const int ch = d.ch, ii0 = d.ii0;
    const int mm = a.modes[d.mi];
    const int i0 = (mm - 1) % NB;
    const int ll0 = (mm - 1) / NB;
    const int qi0 = m.list[ll0] - 1;
    const double w = E(m, qi0, i0);
    int q[3], qp[3];
#pragma unroll
    for (int t = 0; t < 3; t++) {
      q[t] = m.ijk[3 * qi0 + t];
      qp[t] = m.ijk[3 * ii0 + t];
    }
    double Eii[NB];
#pragma unroll
    for (int b = 0; b < NB; b++) Eii[b] = E(m, ii0, b);
    const bool active = (w != 0.0) && (w <= m.omega_max);
    // range of x1 = w +- w'_j over j, for the q'' screen (triple_may_survive)
    double xlo = (ch == 2) ? w - Eii[0] : w + Eii[0], xhi = xlo;
#pragma unroll
    for (int b = 1; b < NB; b++) {
      const double x = (ch == 2) ? w - Eii[b] : w + Eii[b];
      xlo = fmin(xlo, x);
      xhi = fmax(xhi, x);
    }
    const double xabs = fmax(fabs(xlo), fabs(xhi));
    const int jj_first = (ch == 1) ? 0 : ii0;

    int cnt = 0;
    double psum = 0.0;
    const double dn3 = (double)m.nptk * (double)m.nptk * (double)m.nptk;
    const double dnb4 = (double)(NB * NB * NB * NB);
    if (MODE == 0) {
      if (active) {
        for (int jj0 = jj_first + lane; jj0 < m.nptk; jj0 += 64) {
          int qdp[3];
#pragma unroll
          for (int t = 0; t < 3; t++) qdp[t] = m.ijk[3 * jj0 + t];
          const int ss0 = partner(m, ch, q, qp, qdp);
          if (ch != 0 && ss0 < jj0) continue;
          if (!triple_may_survive_lean<NB>(m, ch, jj0, ss0, xlo, xhi, xabs)) continue;
          scan_candidates<NB>(m, ch, i0, qi0, w, ii0, jj0, ss0, Eii, [&](int, int k, int l, double delta) {
            cnt += 1;
            double sigma = pair_sigma(m, ch, jj0, k, ss0, l);
            if (sigma <= (double)kClamp) sigma = 1.0 / sqrt(kPi);
            const double num = exp(-(delta * delta) / (sigma * sigma));
            const double den = sigma * sqrt(kPi) * dn3 * dnb4;
            psum = psum + num / den;
          });
        }
      }
      cnt = wave_sum_i(cnt);
      psum = wave_sum(psum);
      if (lane == 0) {
        a.cnt[item] = cnt;
        a.psum[item] = psum;
      }
    } else if (MODE == 3) {
      if (!active) continue;
      const long s = item / a.nii;
      const long long seg0 = a.offsets[s * a.nii], nseg = a.offsets[(s + 1) * a.nii] - seg0;
      long long base = a.offsets[item] - seg0;
      unsigned long long acc = 0;
      const int niter = (m.nptk - jj_first + 63) / 64;
      for (int it = 0; it < niter; it++) {
        const int jj0 = jj_first + it * 64 + lane;
        int c = 0;
        unsigned long long h = 0;
        bool ok = jj0 < m.nptk;
        int ss0 = 0;
        if (ok) {
          int qdp[3];
#pragma unroll
          for (int t = 0; t < 3; t++) qdp[t] = m.ijk[3 * jj0 + t];
          ss0 = partner(m, ch, q, qp, qdp);
          if (ch != 0 && ss0 < jj0) ok = false;
          if (ok && !triple_may_survive_lean<NB>(m, ch, jj0, ss0, xlo, xhi, xabs)) ok = false;
        }
        if (ok)
          scan_candidates<NB>(m, ch, i0, qi0, w, ii0, jj0, ss0, Eii, [&](int j, int k, int l, double) {
            const unsigned long long v2 = (unsigned long long)ii0 * NB + j + 1;
            const unsigned long long v3 = (unsigned long long)jj0 * NB + k + 1;
            const unsigned long long v4 = (unsigned long long)ss0 * NB + l + 1;
            h = h * kListHashM + ((v2 << 42) ^ (v3 << 21) ^ v4);
            c++;
          });
        const int off = wave_excl_scan(c, lane);
        const int tot = wave_sum_i(c);
        unsigned long long contrib = 0;
        if (c > 0) {
          // M^(n - p0 - c) by square-and-multiply
          unsigned long long e = (unsigned long long)(nseg - (base + off) - c), x = kListHashM, pw = 1;
          while (e) {
            if (e & 1) pw *= x;
            x *= x;
            e >>= 1;
          }
          contrib = h * pw;
        }
        acc += wave_sum_u64(contrib);
        base += tot;
      }
      if (lane == 0 && acc) atomicAdd(a.seg_hash + s, acc);
    } else {
      if (!active) continue;
      // pass 1: survivors per lane
      const int niter = (m.nptk - jj_first + 63) / 64;
      long base = a.offsets[item] - a.out_base;
      for (int it = 0; it < niter; it++) {
        const int jj0 = jj_first + it * 64 + lane;
        int c = 0;
        int ss0 = 0;
        bool ok = jj0 < m.nptk;
        if (ok) {
          int qdp[3];
#pragma unroll
          for (int t = 0; t < 3; t++) qdp[t] = m.ijk[3 * jj0 + t];
          ss0 = partner(m, ch, q, qp, qdp);
          if (ch != 0 && ss0 < jj0) ok = false;
          if (ok && !triple_may_survive_lean<NB>(m, ch, jj0, ss0, xlo, xhi, xabs)) ok = false;
        }
        if (ok) scan_candidates<NB>(m, ch, i0, qi0, w, ii0, jj0, ss0, Eii, [&](int, int, int, double) { c++; });
        const int off = wave_excl_scan(c, lane);
        const int tot = wave_sum_i(c);
        if (ok && c > 0) {
          long p = base + off;
          scan_candidates<NB>(m, ch, i0, qi0, w, ii0, jj0, ss0, Eii, [&](int j, int k, int l, double) {
            const long v2 = (long)ii0 * NB + j + 1, v3 = (long)jj0 * NB + k + 1, v4 = (long)ss0 * NB + l + 1;
            if (MODE == 1) {
              a.out32[0][p] = (int)v2;
              a.out32[1][p] = (int)v3;
              a.out32[2][p] = (int)v4;
            } else {
              a.out64[0][p] = v2;
              a.out64[1][p] = v3;
              a.out64[2][p] = v4;
            }
            p++;
          });
        }
        base += tot;
      }
    }
  }
}

// Per-(channel, mode) reductions of the count pass: N = sum of counts and
// P = sum of the partial phase-space values, both in ii order.
__global__ void k_reduce_counts(const int *cnt, const double *psum, int nseg, int nii, long long *N, double *P) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nseg) return;
  long long n = 0;
  double p = 0.0;
  for (int t = 0; t < nii; t++) {
    n += cnt[(size_t)s * nii + t];
    p = p + psum[(size_t)s * nii + t];
  }
  N[s] = n;
  P[s] = p;
}

__global__ void k_widen(const int *cnt, long long *out, long n) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) out[t] = cnt[t];
}

template <int NB>
static void launch_nb(const PrefilterArgs &a, int mode, hipStream_t st) {
  const long items = a.item_hi - a.item_lo;
  if (items <= 0) return;
  long blocks = (items + 3) / 4;
  if (blocks > 262144) blocks = 262144;
  if (mode == 0)
    hipLaunchKernelGGL((k_prefilter<NB, 0>), dim3((unsigned)blocks), dim3(256), 0, st, a);
  else if (mode == 1)
    hipLaunchKernelGGL((k_prefilter<NB, 1>), dim3((unsigned)blocks), dim3(256), 0, st, a);
  else if (mode == 2)
    hipLaunchKernelGGL((k_prefilter<NB, 2>), dim3((unsigned)blocks), dim3(256), 0, st, a);
  else
    hipLaunchKernelGGL((k_prefilter<NB, 3>), dim3((unsigned)blocks), dim3(256), 0, st, a);
}

int launch_prefilter(const PrefilterArgs &a, int mode, hipStream_t st) {
  switch (a.m.nb) {
    case 3: launch_nb<3>(a, mode, st); break;
    case 6: launch_nb<6>(a, mode, st); break;
    case 9: launch_nb<9>(a, mode, st); break;
    case 12: launch_nb<12>(a, mode, st); break;
    case 18: launch_nb<18>(a, mode, st); break;
    case 24: launch_nb<24>(a, mode, st); break;
    default: return -1;
  }
  return 0;
}

void launch_reduce_counts(const int *cnt, const double *psum, int nseg, int nii, long long *N, double *P,
                          hipStream_t st) {
  hipLaunchKernelGGL(k_reduce_counts, dim3((nseg + 127) / 128), dim3(128), 0, st, cnt, psum, nseg, nii, N, P);
}

void launch_widen(const int *cnt, long long *out, long n, hipStream_t st) {
  hipLaunchKernelGGL(k_widen, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, cnt, out, n);
}

}  // namespace fp4
