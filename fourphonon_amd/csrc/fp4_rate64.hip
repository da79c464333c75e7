// FP64 per-process form: the matrix elements Vp_pp/pm/mm
// (Src/processes.f90:142-330) and the Gaussian / Bose-Einstein weighted rate
// accumulation of RTA_* and Ind_* (Src/processes.f90:3828-4129, 3254-3545),
// evaluated for every process of a compacted list.
//
// One thread per process, 256 consecutive processes of one (channel, mode)
// per block.  The cluster loop is uniform across the block, so the 81 IFC
// values and the first phonon's eigenvector components are broadcast loads;
// each thread gathers the 3 Cartesian components of its phonons 2-4 for the
// cluster's atoms from the [q][branch][component] eigenvector layout
// (48 contiguous bytes per phonon).  Phases are exact roots of unity:
// q.R = 2 pi (n.m)/N for grid points n and lattice vectors m, so the three
// phase factors of a cluster collapse to one table entry.  The 81-term sum
// keeps the reference's order ((((Phi e1) e2) e3) e4).
//
// Partial sums of Gamma and WP4 are reduced per block in a fixed tree order
// and per (channel, mode) in block order, so results are bitwise
// reproducible for a given list.
#include "fp4_device.hpp"
#include "fp4_kernels.hpp"

namespace fp4 {

struct cd {
  double re, im;
};
__device__ __forceinline__ cd cmul(cd a, cd b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }

template <int NB>
__device__ double2 vp64(const ModeView &m, const IfcView &f, int ch, int q0, int i0, int qp0, int j0, int qdp0,
                        int k0, int qt0, int l0) {
  int n2[3], n3[3], n4[3];
#pragma unroll
  for (int t = 0; t < 3; t++) {
    n2[t] = m.ijk[3 * qp0 + t];
    n3[t] = m.ijk[3 * qdp0 + t];
    n4[t] = m.ijk[3 * qt0 + t];
  }
  const double2 *e1 = m.eig + ((size_t)q0 * NB + i0) * NB;
  const double2 *e2 = m.eig + ((size_t)qp0 * NB + j0) * NB;
  const double2 *e3 = m.eig + ((size_t)qdp0 * NB + k0) * NB;
  const double2 *e4 = m.eig + ((size_t)qt0 * NB + l0) * NB;
  const int s2 = (ch == 2) ? -1 : 1;
  const int s3 = (ch == 0) ? 1 : -1;
  const double c2 = (ch == 2) ? -1.0 : 1.0;   // conj on phonon 2 (--)
  const double c3 = (ch == 0) ? 1.0 : -1.0;   // conj on phonon 3 (+-, --)
  double Vre = 0.0, Vim = 0.0;
  for (int t = 0; t < f.ntri; t++) {
    const ClusterRec c = f.cl[t];
    const int kj = n2[0] * c.pj[0] + n2[1] * c.pj[1] + n2[2] * c.pj[2];
    const int kk = n3[0] * c.pk[0] + n3[1] * c.pk[1] + n3[2] * c.pk[2];
    const int kl = n4[0] * c.pl[0] + n4[1] * c.pl[1] + n4[2] * c.pl[2];
    const double2 ph = f.roots[fmod_pos(s2 * kj + s3 * kk - kl, f.L)];
    cd a1[3], a2[3], a3[3], a4[3];
#pragma unroll
    for (int u = 0; u < 3; u++) {
      const double2 x1 = e1[3 * c.atom[0] + u], x2 = e2[3 * c.atom[1] + u];
      const double2 x3 = e3[3 * c.atom[2] + u], x4 = e4[3 * c.atom[3] + u];
      a1[u] = {x1.x, x1.y};
      a2[u] = {x2.x, c2 * x2.y};
      a3[u] = {x3.x, c3 * x3.y};
      a4[u] = {x4.x, -x4.y};
    }
    const double *phi = f.phi + (size_t)t * 81;
    cd v0 = {0.0, 0.0};
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
      for (int s = 0; s < 3; s++)
#pragma unroll
        for (int tt = 0; tt < 3; tt++)
#pragma unroll
          for (int u = 0; u < 3; u++) {
            const double p = phi[u + 3 * tt + 9 * s + 27 * r];
            cd x = {p * a1[u].re, p * a1[u].im};
            x = cmul(x, a2[tt]);
            x = cmul(x, a3[s]);
            x = cmul(x, a4[r]);
            v0.re = v0.re + x.re;
            v0.im = v0.im + x.im;
          }
    const cd pre = {c.minv * ph.x, c.minv * ph.y};
    const cd y = cmul(pre, v0);
    Vre = Vre + y.re;
    Vim = Vim + y.im;
  }
  return make_double2(Vre, Vim);
}

template <int NB>
__global__ void __launch_bounds__(256) k_rate64(RateArgs a) {
  const ModeView &m = a.m;
  const BlockDesc bd = a.blocks[blockIdx.x];
  const int tid = threadIdx.x;
  double g = 0.0, wp = 0.0;
  if (tid < bd.count) {
    const long long p = bd.start + tid;
    long long v2, v3, v4;
    if (a.lst[0]) {
      v2 = a.lst[0][p]; v3 = a.lst[1][p]; v4 = a.lst[2][p];
    } else {
      v2 = a.lst64[0][p]; v3 = a.lst64[1][p]; v4 = a.lst64[2][p];
    }
    const int ch = bd.ch;
    const int i0 = (bd.mm - 1) % NB;
    const int qi0 = m.list[(bd.mm - 1) / NB] - 1;
    const int ii0 = (int)((v2 - 1) / NB), j0 = (int)((v2 - 1) % NB);
    const int jj0 = (int)((v3 - 1) / NB), k0 = (int)((v3 - 1) % NB);
    const int ss0 = (int)((v4 - 1) / NB), l0 = (int)((v4 - 1) % NB);
    const double w = E(m, qi0, i0), op = E(m, ii0, j0), odp = E(m, jj0, k0), otp = E(m, ss0, l0);
    const double delta = mismatch(ch, w, op, odp, otp);
    double sigma = pair_sigma(m, ch, jj0, k0, ss0, l0);
    if (sigma <= (double)kClamp) sigma = 1.0 / sqrt(kPi);
    wp = wp4_of(ch, w, op, odp, otp, delta, sigma, a.T);
    if (a.store_gamma || !a.onlyharmonic) {
      const double2 V = vp64<NB>(m, a.f, ch, qi0, i0, ii0, j0, jj0, k0, ss0, l0);
      const double hbarp = kHbar * (double)k1e22;
      const double av = hypot(V.x, V.y);
      g = hbarp * hbarp * kPi / 8.0 * wp * (av * av);
      if (a.store_gamma) {
        const double gs = g * (double)kGammaUnits * 1.e9 / m.nptk / m.nptk;
        a.proc_gamma[p] = gs;
        g = gs;
      }
    }
  }
  // fixed-order block reduction: wave butterflies, then the 4 wave sums in order
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
    a.blk_wp4[blockIdx.x] = ((sw[0] + sw[1]) + sw[2]) + sw[3];
  }
}

__global__ void k_seg_reduce(const double *bg, const double *bw, const int *seg_first, int nseg, double *sg,
                             double *sw) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nseg) return;
  double g = 0.0, w = 0.0;
  for (int b = seg_first[s]; b < seg_first[s + 1]; b++) {
    g = g + bg[b];
    w = w + bw[b];
  }
  sg[s] = g;
  sw[s] = w;
}

template <int NB>
__global__ void __launch_bounds__(256) k_vp64(VpArgs a) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n) return;
  const int *r = a.idx + 8 * t;
  a.out[t] = vp64<NB>(a.m, a.f, a.ch, r[4] - 1, r[0] - 1, r[5] - 1, r[1] - 1, r[6] - 1, r[2] - 1, r[7] - 1, r[3] - 1);
}

template <int NB>
static void rate_nb(const RateArgs &a, hipStream_t st) {
  hipLaunchKernelGGL((k_rate64<NB>), dim3(a.nblocks), dim3(256), 0, st, a);
}

int launch_rate64(const RateArgs &a, hipStream_t st) {
  if (a.nblocks <= 0) return 0;
  switch (a.m.nb) {
    case 3: rate_nb<3>(a, st); break;
    case 6: rate_nb<6>(a, st); break;
    case 9: rate_nb<9>(a, st); break;
    case 12: rate_nb<12>(a, st); break;
    case 18: rate_nb<18>(a, st); break;
    case 24: rate_nb<24>(a, st); break;
    default: return -1;
  }
  return 0;
}

void launch_seg_reduce(const double *bg, const double *bw, const int *seg_first, int nseg, double *sg, double *sw,
                       hipStream_t st) {
  if (nseg <= 0) return;
  hipLaunchKernelGGL(k_seg_reduce, dim3((nseg + 127) / 128), dim3(128), 0, st, bg, bw, seg_first, nseg, sg, sw);
}

int launch_vp64(const VpArgs &a, hipStream_t st) {
  if (a.n <= 0) return 0;
  const unsigned blocks = (unsigned)((a.n + 255) / 256);
  switch (a.m.nb) {
    case 3: hipLaunchKernelGGL((k_vp64<3>), dim3(blocks), dim3(256), 0, st, a); break;
    case 6: hipLaunchKernelGGL((k_vp64<6>), dim3(blocks), dim3(256), 0, st, a); break;
    case 9: hipLaunchKernelGGL((k_vp64<9>), dim3(blocks), dim3(256), 0, st, a); break;
    case 12: hipLaunchKernelGGL((k_vp64<12>), dim3(blocks), dim3(256), 0, st, a); break;
    case 18: hipLaunchKernelGGL((k_vp64<18>), dim3(blocks), dim3(256), 0, st, a); break;
    case 24: hipLaunchKernelGGL((k_vp64<24>), dim3(blocks), dim3(256), 0, st, a); break;
    default: return -1;
  }
  return 0;
}

// Aggregated operator: iteration_4ph's sum over a mode's processes,
// sum_p Gamma_p * (+-F(partner 2) +-F(partner 3) [+F(partner 4)]), regrouped
// by partner: W[mode][partner] collects the signed Gamma of every process
// that has that partner.  Signs: 3ph + -F(j)+F(k), 3ph - (F(j)+F(k))/2
// (Src/iterations.f90:152-166); 4ph ++ -F(j)-F(k)+F(l), +- -F(j)+F(k)+F(l),
// -- F(j)+F(k)+F(l) (:177-203).  The additions are atomic, so W agrees
// between runs to rounding, not bit for bit.
__global__ void __launch_bounds__(256) k_w_scatter(WScatterArgs a) {
  const BlockDesc d = a.blocks[blockIdx.x];
  const int t = threadIdx.x;
  if (t >= d.count) return;
  const long long p = d.start + t;
  const double g = a.gamma[p];
  if (g == 0.0) return;
  double *r = a.W + (size_t)(d.mm - 1) * (size_t)a.wcols;
  const int c2 = a.lst[0][p] - 1, c3 = a.lst[1][p] - 1;
  if (a.order == 3) {
    if (d.ch == 0) {
      unsafeAtomicAdd(r + c2, -g);
      unsafeAtomicAdd(r + c3, g);
    } else {
      const double h = g * 5.e-1;
      unsafeAtomicAdd(r + c2, h);
      unsafeAtomicAdd(r + c3, h);
    }
  } else {
    const int c4 = a.lst[2][p] - 1;
    unsafeAtomicAdd(r + c2, d.ch == 2 ? g : -g);
    unsafeAtomicAdd(r + c3, d.ch == 0 ? -g : g);
    unsafeAtomicAdd(r + c4, g);
  }
}

int launch_w_scatter(const WScatterArgs &a, hipStream_t st) {
  if (a.nblocks <= 0) return 0;
  hipLaunchKernelGGL(k_w_scatter, dim3((unsigned)a.nblocks), dim3(256), 0, st, a);
  return 0;
}

}  // namespace fp4
