// Three-phonon path: the energy-conservation scan behind NP_driver /
// Ind_only_driver (Src/processes.f90:712-823, 983-1309) and the FP64
// per-process matrix elements and rates of Vp_plus / Vp_minus, RTA_plus /
// RTA_minus and Ind_plus / Ind_minus (Src/processes.f90:25-139, 361-521,
// 1313-1471).
//
// Channel 0 = "+" (absorption, q'' = q + q'), 1 = "-" (emission,
// q'' = q - q').  A candidate (q', j, k) survives when omega', omega'' are
// non-zero and |delta| <= 2 sigma, sigma = scalebroad * base_sigma(v(q',j) -
// v(q'',k)) with no clamp; delta is (omega + omega') - omega'' or
// (omega - omega') - omega''.
//
// Scan: one 64-wide wavefront per item (channel, mode, j); lane t takes
// q' = lo + t, lo + t + 64, ... and tests the Nb branches k.  The count pass
// writes per-item survivor counts and phase-space partial sums; the emit
// pass scans lane counts in lane (= q') order and writes int32 list entries,
// so the concatenation over items is the reference's j-outer, q'-inner,
// k-innermost list order (Ind_only_plus, processes.f90:1105-1127).
//
// Rates: one thread per listed process, 256 consecutive processes of one
// (channel, mode) per block, exact root-of-unity phases as in the
// four-phonon FP64 form, block sums in a fixed tree order.
#include "fp4_device.hpp"
#include "fp4_kernels.hpp"

namespace fp4 {

__device__ __forceinline__ int partner3(const ModeView &m, int ch, const int *q, const int *qp) {
  int t0, t1, t2;
  if (ch == 0) {
    t0 = q[0] + qp[0]; t1 = q[1] + qp[1]; t2 = q[2] + qp[2];
  } else {
    t0 = q[0] - qp[0]; t1 = q[1] - qp[1]; t2 = q[2] - qp[2];
  }
  t0 = fmod_pos(t0, m.n1); t1 = fmod_pos(t1, m.n2); t2 = fmod_pos(t2, m.n3);
  return (t2 * m.n2 + t1) * m.n1 + t0;
}

// sigma of the (q',j) / (q'',k) pair, processes.f90:750-752
__device__ __forceinline__ double sigma3(const ModeView &m, int ii0, int j0, int ss0, int k0) {
  const double d0 = Vel(m, ii0, j0, 0) - Vel(m, ss0, k0, 0);
  const double d1 = Vel(m, ii0, j0, 1) - Vel(m, ss0, k0, 1);
  const double d2 = Vel(m, ii0, j0, 2) - Vel(m, ss0, k0, 2);
  return m.scalebroad * base_sigma(m, d0, d1, d2);
}

__device__ __forceinline__ double delta3(int ch, double w, double wp, double wdp) {
  return (ch == 0) ? w + wp - wdp : w - wp - wdp;
}

template <int NB, int MODE>
__global__ void __launch_bounds__(256) k3_scan(Scan3Args a) {
  const ModeView &m = a.m;
  const int lane = threadIdx.x & 63;
  const long nitems = 2L * a.nsub * NB;
  const int wpb = blockDim.x >> 6;
  // sigma*sqrt(Pi)*nptk**2*nbands**3 with integer nptk**2 and nbands**3 (processes.f90:757)
  const double den0 = (double)((long long)m.nptk * m.nptk);
  const double nb3 = (double)(NB * NB * NB);
  for (long item = (long)blockIdx.x * wpb + (threadIdx.x >> 6) + a.item_lo; item < a.item_hi;
       item += (long)gridDim.x * wpb) {
    if (item >= nitems) break;
    const int j0 = (int)(item % NB);
    const long r = item / NB;
    const int mi = (int)(r % a.nsub);
    const int ch = (int)(r / a.nsub);
    const int mm = a.modes[mi];
    const int i0 = (mm - 1) % NB;
    const int qi0 = m.list[(mm - 1) / NB] - 1;
    const double w = E(m, qi0, i0);
    const bool active = (w != 0.0) && (w <= m.omega_max);
    int q[3];
#pragma unroll
    for (int t = 0; t < 3; t++) q[t] = m.ijk[3 * qi0 + t];
    const int niter = (a.nii + 63) / 64;
    int cnt = 0;
    double psum = 0.0;
    long base = (MODE == 1) ? (long)(a.offsets[item] - a.out_base) : 0;
    if (MODE == 1 && !active) continue;
    for (int it = 0; it < niter; it++) {
      const int ii0 = a.ii_lo0 + it * 64 + lane;
      const bool ok = active && (it * 64 + lane < a.nii);
      int c = 0;
      int ss0 = 0;
      double op = 0.0;
      unsigned hits = 0u;   // bit k: branch k of q'' survives
      if (ok) {
        int qp[3];
#pragma unroll
        for (int t = 0; t < 3; t++) qp[t] = m.ijk[3 * ii0 + t];
        ss0 = partner3(m, ch, q, qp);
        op = E(m, ii0, j0);
#pragma unroll
        for (int k = 0; k < NB; k++) {
          const double odp = E(m, ss0, k);
          if (!((op != 0.0) && (odp != 0.0))) continue;
          const double sigma = sigma3(m, ii0, j0, ss0, k);
          const double delta = delta3(ch, w, op, odp);
          if (fabs(delta) <= 2.0 * sigma) {
            hits |= 1u << k;
            c++;
            if (MODE == 0) {
              const double num = exp(-(delta * delta) / (sigma * sigma));
              const double den = sigma * sqrt(kPi) * den0 * nb3;
              psum = psum + num / den;
            }
          }
        }
      }
      if (MODE == 0) {
        cnt += c;
      } else {
        const int off = wave_excl_scan(c, lane);
        const int tot = wave_sum_i(c);
        if (c > 0) {
          long p = base + off;
          const int v2 = ii0 * NB + j0 + 1;
#pragma unroll
          for (int k = 0; k < NB; k++)
            if (hits & (1u << k)) {
              a.out[0][p] = v2;
              a.out[1][p] = ss0 * NB + k + 1;
              p++;
            }
        }
        base += tot;
      }
    }
    if (MODE == 0) {
      cnt = wave_sum_i(cnt);
      psum = wave_sum(psum);
      if (lane == 0) {
        a.cnt[item] = cnt;
        a.psum[item] = psum;
      }
    }
  }
}

template <int NB>
__device__ double2 vp3_64(const ModeView &m, const Ifc3View &f, int ch, int q0, int i0, int qp0, int j0, int qdp0,
                          int k0) {
  int n2[3], n3[3];
#pragma unroll
  for (int t = 0; t < 3; t++) {
    n2[t] = m.ijk[3 * qp0 + t];
    n3[t] = m.ijk[3 * qdp0 + t];
  }
  const double2 *e1 = m.eig + ((size_t)q0 * NB + i0) * NB;
  const double2 *e2 = m.eig + ((size_t)qp0 * NB + j0) * NB;
  const double2 *e3 = m.eig + ((size_t)qdp0 * NB + k0) * NB;
  const int s2 = (ch == 0) ? 1 : -1;
  const double c2 = (ch == 0) ? 1.0 : -1.0;   // conj on phonon 2 for "-"
  double Vre = 0.0, Vim = 0.0;
  for (int t = 0; t < f.ntri; t++) {
    const Cluster3Rec c = f.cl[t];
    const int kj = n2[0] * c.pj[0] + n2[1] * c.pj[1] + n2[2] * c.pj[2];
    const int kk = n3[0] * c.pk[0] + n3[1] * c.pk[1] + n3[2] * c.pk[2];
    const double2 ph = f.roots[fmod_pos(s2 * kj - kk, f.L)];
    double a1r[3], a1i[3], a2r[3], a2i[3], a3r[3], a3i[3];
#pragma unroll
    for (int u = 0; u < 3; u++) {
      const double2 x1 = e1[3 * c.atom[0] + u], x2 = e2[3 * c.atom[1] + u], x3 = e3[3 * c.atom[2] + u];
      a1r[u] = x1.x; a1i[u] = x1.y;
      a2r[u] = x2.x; a2i[u] = c2 * x2.y;
      a3r[u] = x3.x; a3i[u] = -x3.y;
    }
    const double *phi = f.phi + (size_t)t * 27;
    double v0r = 0.0, v0i = 0.0;
#pragma unroll
    for (int rr = 0; rr < 3; rr++)
#pragma unroll
      for (int ss = 0; ss < 3; ss++)
#pragma unroll
        for (int tt = 0; tt < 3; tt++) {
          const double p = phi[tt + 3 * ss + 9 * rr];
          double xr = p * a1r[tt], xi = p * a1i[tt];
          double yr = xr * a2r[ss] - xi * a2i[ss], yi = xr * a2i[ss] + xi * a2r[ss];
          xr = yr * a3r[rr] - yi * a3i[rr];
          xi = yr * a3i[rr] + yi * a3r[rr];
          v0r = v0r + xr;
          v0i = v0i + xi;
        }
    const double pr = c.minv * ph.x, pi = c.minv * ph.y;
    Vre = Vre + (pr * v0r - pi * v0i);
    Vim = Vim + (pr * v0i + pi * v0r);
  }
  return make_double2(Vre, Vim);
}

template <int NB>
__global__ void __launch_bounds__(256) k3_rate64(Rate3Args a) {
  const ModeView &m = a.m;
  const BlockDesc bd = a.blocks[blockIdx.x];
  const int tid = threadIdx.x;
  double g = 0.0, wp = 0.0;
  if (tid < bd.count) {
    const long long p = bd.start + tid;
    const int v2 = a.lst[0][p], v3 = a.lst[1][p];
    const int ch = bd.ch;
    const int i0 = (bd.mm - 1) % NB;
    const int qi0 = m.list[(bd.mm - 1) / NB] - 1;
    const int ii0 = (v2 - 1) / NB, j0 = (v2 - 1) % NB;
    const int ss0 = (v3 - 1) / NB, k0 = (v3 - 1) % NB;
    const double w = E(m, qi0, i0), op = E(m, ii0, j0), odp = E(m, ss0, k0);
    const double delta = delta3(ch, w, op, odp);
    const double sigma = sigma3(m, ii0, j0, ss0, k0);
    const double fp = bose(op, a.T), fdp = bose(odp, a.T);
    const double occ = (ch == 0) ? fp - fdp : fp + fdp + 1;
    wp = occ * exp(-(delta * delta) / (sigma * sigma)) / sigma / sqrt(kPi) / (w * op * odp);
    if (a.store_gamma || !a.onlyharmonic) {
      const double2 V = vp3_64<NB>(m, a.f, ch, qi0, i0, ii0, j0, ss0, k0);
      const double hbarp = kHbar * (double)k1e22;
      const double av = hypot(V.x, V.y);
      g = hbarp * kPi / 4.0 * wp * (av * av);
      if (a.store_gamma) {
        const double gs = g * (double)kGamma3Units * 1.e8 / m.nptk;
        a.proc_gamma[p] = gs;
        g = gs;
      }
    }
  }
  __shared__ double sg[4], sw[4];
  g = wave_sum(g);
  wp = wave_sum(wp);
  if ((tid & 63) == 0) {
    sg[tid >> 6] = g;
    sw[tid >> 6] = wp;
  }
  __syncthreads();
  if (tid == 0) {
    a.blk_gamma[blockIdx.x] = ((sg[0] + sg[1]) + sg[2]) + sg[3];
    a.blk_wp[blockIdx.x] = ((sw[0] + sw[1]) + sw[2]) + sw[3];
  }
}

template <int NB>
__global__ void __launch_bounds__(256) k3_vp64(Vp3Args a) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n) return;
  const int *r = a.idx + 6 * t;
  a.out[t] = vp3_64<NB>(a.m, a.f, a.ch, r[3] - 1, r[0] - 1, r[4] - 1, r[1] - 1, r[5] - 1, r[2] - 1);
}

template <int NB>
static void scan3_nb(const Scan3Args &a, int mode, hipStream_t st) {
  const long items = a.item_hi - a.item_lo;
  if (items <= 0) return;
  long blocks = (items + 3) / 4;
  if (blocks > 262144) blocks = 262144;
  if (mode == 0)
    hipLaunchKernelGGL((k3_scan<NB, 0>), dim3((unsigned)blocks), dim3(256), 0, st, a);
  else
    hipLaunchKernelGGL((k3_scan<NB, 1>), dim3((unsigned)blocks), dim3(256), 0, st, a);
}

int launch_scan3(const Scan3Args &a, int mode, hipStream_t st) {
  switch (a.m.nb) {
    case 3: scan3_nb<3>(a, mode, st); break;
    case 6: scan3_nb<6>(a, mode, st); break;
    case 9: scan3_nb<9>(a, mode, st); break;
    case 12: scan3_nb<12>(a, mode, st); break;
    case 18: scan3_nb<18>(a, mode, st); break;
    case 24: scan3_nb<24>(a, mode, st); break;
    default: return -1;
  }
  return 0;
}

// Sampling estimators of NP_driver / RTA_driver (Src/processes.f90:828-979,
// 1474-1673): per (channel, mode) segment nsample draws of (q', j, k), each
// uniform over the Nb*nptk*Nb candidates of the full loops (index
// ((ii0*Nb + j)*Nb + k), decoded as the reference does), kept when they pass
// the full path's tests.  Draw s: mix64(key + kGolden*(s+1)) with
// key = mix64(seed + kGolden*(2*(mm-1)+ch+1)).  Count pass: hits and
// phase-space sums per thread in sample order, fixed-tree block sums; emit
// pass: hits in sample order as int32 list entries for k3_rate64.
template <int NB>
__device__ __forceinline__ bool sample3_one(const ModeView &m, int ch, double w, const int *q,
                                            unsigned long long key, long long s, int &ii0, int &j0, int &ss0,
                                            int &k0, double &delta, double &sigma) {
  const int r = below(mix64(key + kGolden * (unsigned long long)(s + 1)), NB * m.nptk * NB);
  k0 = r % NB;
  j0 = (r / NB) % NB;
  ii0 = r / (NB * NB);
  int qp[3];
#pragma unroll
  for (int t = 0; t < 3; t++) qp[t] = m.ijk[3 * ii0 + t];
  ss0 = partner3(m, ch, q, qp);
  const double op = E(m, ii0, j0), odp = E(m, ss0, k0);
  if (!((op != 0.0) && (odp != 0.0))) return false;
  sigma = sigma3(m, ii0, j0, ss0, k0);
  delta = delta3(ch, w, op, odp);
  return fabs(delta) <= 2.0 * sigma;
}

template <int NB, int MODE>
__global__ void __launch_bounds__(256) k3_sample(SampleArgs a) {
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
  const unsigned long long key = mix64(a.seed + kGolden * (unsigned long long)(2 * (mm - 1) + ch + 1));
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  __shared__ double sp[4];
  __shared__ int sc[4];
  if (MODE == 0) {
    int cnt = 0;
    double psum = 0.0;
    const double den0 = (double)((long long)m.nptk * m.nptk);
    const double nb3 = (double)(NB * NB * NB);
    if (active) {
      for (long long s = tid; s < a.nsample; s += 256) {
        int ii0, j0, ss0, k0;
        double delta, sigma;
        if (!sample3_one<NB>(m, ch, w, q, key, s, ii0, j0, ss0, k0, delta, sigma)) continue;
        cnt += 1;
        const double num = exp(-(delta * delta) / (sigma * sigma));
        const double den = sigma * sqrt(kPi) * den0 * nb3;
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
      int ii0 = 0, j0 = 0, ss0 = 0, k0 = 0;
      double delta, sigma;
      const bool hit = s < a.nsample && sample3_one<NB>(m, ch, w, q, key, s, ii0, j0, ss0, k0, delta, sigma);
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
        a.out32[0][at] = ii0 * NB + j0 + 1;
        a.out32[1][at] = ss0 * NB + k0 + 1;
      }
      p += ((sc[0] + sc[1]) + sc[2]) + sc[3];
    }
  }
}

template <int NB>
static void sample3_nb(const SampleArgs &a, int mode, hipStream_t st) {
  const int n = a.seg_hi - a.seg_lo;
  if (n <= 0) return;
  if (mode == 0)
    hipLaunchKernelGGL((k3_sample<NB, 0>), dim3(n), dim3(256), 0, st, a);
  else
    hipLaunchKernelGGL((k3_sample<NB, 1>), dim3(n), dim3(256), 0, st, a);
}

int launch_sample3(const SampleArgs &a, int mode, hipStream_t st) {
  switch (a.m.nb) {
    case 3: sample3_nb<3>(a, mode, st); break;
    case 6: sample3_nb<6>(a, mode, st); break;
    case 9: sample3_nb<9>(a, mode, st); break;
    case 12: sample3_nb<12>(a, mode, st); break;
    case 18: sample3_nb<18>(a, mode, st); break;
    case 24: sample3_nb<24>(a, mode, st); break;
    default: return -1;
  }
  return 0;
}

int launch_rate3(const Rate3Args &a, hipStream_t st) {
  if (a.nblocks <= 0) return 0;
  const dim3 g((unsigned)a.nblocks), b(256);
  switch (a.m.nb) {
    case 3: hipLaunchKernelGGL((k3_rate64<3>), g, b, 0, st, a); break;
    case 6: hipLaunchKernelGGL((k3_rate64<6>), g, b, 0, st, a); break;
    case 9: hipLaunchKernelGGL((k3_rate64<9>), g, b, 0, st, a); break;
    case 12: hipLaunchKernelGGL((k3_rate64<12>), g, b, 0, st, a); break;
    case 18: hipLaunchKernelGGL((k3_rate64<18>), g, b, 0, st, a); break;
    case 24: hipLaunchKernelGGL((k3_rate64<24>), g, b, 0, st, a); break;
    default: return -1;
  }
  return 0;
}

int launch_vp3(const Vp3Args &a, hipStream_t st) {
  if (a.n <= 0) return 0;
  const dim3 g((unsigned)((a.n + 255) / 256)), b(256);
  switch (a.m.nb) {
    case 3: hipLaunchKernelGGL((k3_vp64<3>), g, b, 0, st, a); break;
    case 6: hipLaunchKernelGGL((k3_vp64<6>), g, b, 0, st, a); break;
    case 9: hipLaunchKernelGGL((k3_vp64<9>), g, b, 0, st, a); break;
    case 12: hipLaunchKernelGGL((k3_vp64<12>), g, b, 0, st, a); break;
    case 18: hipLaunchKernelGGL((k3_vp64<18>), g, b, 0, st, a); break;
    case 24: hipLaunchKernelGGL((k3_vp64<24>), g, b, 0, st, a); break;
    default: return -1;
  }
  return 0;
}

}  // namespace fp4
