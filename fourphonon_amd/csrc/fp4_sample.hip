// Sampling estimators of the four-phonon phase space and rates: the
// num_sample_process_4ph[_phase_space] branches of NP_driver_4ph and
// RTA_driver_4ph (Src/processes.f90:2480-2746, 4811-5171, dispatched at
// 2807-2826 and 5257-5290).
//
// Per (channel, mode) the estimator draws nsample candidates
// (q', j, q'', k, l) uniformly from the Nb*nptk*Nb*nptk*Nb box, derives q'''
// by momentum conservation, keeps a candidate when it lies in the part of
// the box the full enumeration visits (q'' >= q' with k >= j on the
// diagonal for ++ and --, q''' >= q'' with l >= k for +- and --), passes the
// zero-frequency tests, the +- pair exclusions and the exact FP64 window of
// the full path, and scales the sums by total/nsample.  Every listed
// process of the full path is hit with probability nsample/total per draw,
// so the estimates are unbiased for the full path's N, P, Gamma and WP4.
//
// Random numbers: counter-based SplitMix64 (Steele, Lea & Flood, "Fast
// splittable pseudorandom number generators", OOPSLA 2014).  The stream of
// a (channel, mode) segment is seeded with mix64(seed + gamma*(3*(mm-1)+ch+1))
// and draw t of sample s is mix64(key + gamma*(5*s+t+1)); an index below n
// is the high 64 bits of draw*n.  Results depend only on (seed, nsample)
// and the mode, not on the launch shape or on which other modes are asked.
//
// The reference draws with Fortran random_number after random_seed() (a
// different stream on every run), takes num+1 samples and reads
// rand_matrix(-4..0) for the first; here exactly nsample samples are drawn
// and every draw is in bounds.
//
// Count pass: one 256-thread block per segment; thread t takes samples
// t, t+256, ...; hits and the phase-space terms are summed per thread in
// sample order and over the block in a fixed tree.  Emit pass: the same
// walk in chunks of 256 samples; a block scan of the hit flags writes the
// hits in sample order as a list in the reference encoding, which the FP64
// per-process rate kernel then evaluates.
#include "fp4_device.hpp"
#include "fp4_kernels.hpp"

namespace fp4 {

struct SampleHit {
  int ii0, j, jj0, k, ss0, l;
  double delta;
};

// One draw of segment (ch, mode) at sample s: true when it is a listed
// process of the full enumeration.
template <int NB>
__device__ __forceinline__ bool sample_one(const ModeView &m, int ch, int i0, int qi0, double w, const int *q,
                                           unsigned long long key, long long s, SampleHit &h) {
  const unsigned long long base = key + kGolden * (unsigned long long)(5 * s + 1);
  h.ii0 = below(mix64(base), m.nptk);
  h.j = below(mix64(base + kGolden), NB);
  h.jj0 = below(mix64(base + 2 * kGolden), m.nptk);
  h.k = below(mix64(base + 3 * kGolden), NB);
  h.l = below(mix64(base + 4 * kGolden), NB);
  int qp[3], qdp[3];
#pragma unroll
  for (int t = 0; t < 3; t++) {
    qp[t] = m.ijk[3 * h.ii0 + t];
    qdp[t] = m.ijk[3 * h.jj0 + t];
  }
  h.ss0 = partner(m, ch, q, qp, qdp);
  if (ch != 1 && (h.jj0 < h.ii0 || (h.jj0 == h.ii0 && h.k < h.j))) return false;
  if (ch != 0 && (h.ss0 < h.jj0 || (h.ss0 == h.jj0 && h.l < h.k))) return false;
  const double op = E(m, h.ii0, h.j), odp = E(m, h.jj0, h.k), otp = E(m, h.ss0, h.l);
  if (!((op != 0.0) && (odp != 0.0) && (otp != 0.0))) return false;
  if (ch == 1 && !((qi0 != h.jj0 || i0 != h.k) && (qi0 != h.ss0 || i0 != h.l) && (h.ii0 != h.jj0 || h.j != h.k) &&
                   (h.ii0 != h.ss0 || h.j != h.l)))
    return false;
  h.delta = mismatch(ch, w, op, odp, otp);
  const double sigma = pair_sigma(m, ch, h.jj0, h.k, h.ss0, h.l);
  return fabs(h.delta) <= sigma;
}

template <int NB, int MODE>
__global__ void __launch_bounds__(256) k_sample4(SampleArgs a) {
  const ModeView &m = a.m;
  const int seg = blockIdx.x + a.seg_lo;
  if (seg >= a.seg_hi) return;
  const int ch = seg / a.nsub, mi = seg % a.nsub;
  const int mm = a.modes[mi];
  const int i0 = (mm - 1) % NB;
  const int qi0 = m.list[(mm - 1) / NB] - 1;
  const double w = E(m, qi0, i0);
  int q[3];
#pragma unroll
  for (int t = 0; t < 3; t++) q[t] = m.ijk[3 * qi0 + t];
  const bool active = (w != 0.0) && (w <= m.omega_max);
  const unsigned long long key = mix64(a.seed + kGolden * (unsigned long long)(3 * (mm - 1) + ch + 1));
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  __shared__ double sp[4];
  __shared__ int sc[4];
  if (MODE == 0) {
    int cnt = 0;
    double psum = 0.0;
    const double dn3 = (double)m.nptk * (double)m.nptk * (double)m.nptk;
    const double dnb4 = (double)(NB * NB * NB * NB);
    if (active) {
      for (long long s = tid; s < a.nsample; s += 256) {
        SampleHit h;
        if (!sample_one<NB>(m, ch, i0, qi0, w, q, key, s, h)) continue;
        cnt += 1;
        double sigma = pair_sigma(m, ch, h.jj0, h.k, h.ss0, h.l);
        if (sigma <= (double)kClamp) sigma = 1.0 / sqrt(kPi);
        const double num = exp(-(h.delta * h.delta) / (sigma * sigma));
        const double den = sigma * sqrt(kPi) * dn3 * dnb4;
        psum = psum + num / den;
      }
    }
    cnt = wave_sum_i(cnt);
    psum = wave_sum(psum);
    if (lane == 0) {
      sc[wv] = cnt;
      sp[wv] = psum;
    }
    __syncthreads();
    if (tid == 0) {
      a.cnt[seg] = ((sc[0] + sc[1]) + sc[2]) + sc[3];
      a.psum[seg] = ((sp[0] + sp[1]) + sp[2]) + sp[3];
    }
  } else {
    if (!active) return;
    long long p = a.offsets[seg] - a.out_base;
    const long long end = a.offsets[seg + 1] - a.out_base;   // the count pass's total for this segment
    for (long long s0 = 0; s0 < a.nsample; s0 += 256) {
      const long long s = s0 + tid;
      SampleHit h;
      const bool hit = s < a.nsample && sample_one<NB>(m, ch, i0, qi0, w, q, key, s, h);
      const int f = hit ? 1 : 0;
      const int off = wave_excl_scan(f, lane);
      const int tot = wave_sum_i(f);
      __syncthreads();   // the previous chunk has read sc
      if (lane == 0) sc[wv] = tot;
      __syncthreads();
      int before = 0;
      for (int t = 0; t < wv; t++) before += sc[t];
      const long long at = p + before + off;
      if (hit && at < end) {
        a.out32[0][at] = h.ii0 * NB + h.j + 1;
        a.out32[1][at] = h.jj0 * NB + h.k + 1;
        a.out32[2][at] = h.ss0 * NB + h.l + 1;
      }
      p += ((sc[0] + sc[1]) + sc[2]) + sc[3];
    }
  }
}

template <int NB>
static void launch_nb(const SampleArgs &a, int mode, hipStream_t st) {
  const int n = a.seg_hi - a.seg_lo;
  if (n <= 0) return;
  if (mode == 0)
    hipLaunchKernelGGL((k_sample4<NB, 0>), dim3(n), dim3(256), 0, st, a);
  else
    hipLaunchKernelGGL((k_sample4<NB, 1>), dim3(n), dim3(256), 0, st, a);
}

int launch_sample4(const SampleArgs &a, int mode, hipStream_t st) {
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

}  // namespace fp4
