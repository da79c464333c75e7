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
// Scheduling: one persistent block per slot pulls work items from an atomic
// counter; each block owns a global scratch slot for A'.  Results are
// written per item and reduced over q' in a fixed order, and each block's
// partial sums are reduced in a fixed tree order, so the output does not
// depend on which block ran which item.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <tuple>

#include "../../include/fp4.h"
#include "fp4_dense.hpp"

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
  int *kl_first = nullptr;
  void *scratch = nullptr;
  size_t scratch_bytes = 0;
  void *outb = nullptr;
  size_t out_bytes = 0;
  int *counter = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  long long last_triples = 0, last_items = 0;
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
  const int *kl_first;
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
  void *scratch;
  long slot_elems;       // complex elements of A' per slot
  double *out_g;
  double *out_w;
  int *out_n;
  int *out_t;            // [nitems] visited (q, q', q'') triples
};

template <class T>
__device__ __forceinline__ C2<T> cm(C2<T> a, C2<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ double2 zm(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// block-wide exclusive scan of 0/1 flags in thread order; returns the total
__device__ __forceinline__ int block_flag_scan(int flag, int *s_wsum, int &pos) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const unsigned long long bal = __ballot(flag);
  const int before = __popcll(bal & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))));
  if (lane == 0) s_wsum[wv] = __popcll(bal);
  __syncthreads();
  int base = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < 4; w++) {
    if (w < wv) base += s_wsum[w];
    tot += s_wsum[w];
  }
  pos = base + before;
  __syncthreads();
  return tot;
}

// block-wide exclusive scan of per-thread counts in thread order
__device__ __forceinline__ int block_excl_scan(int v, int *s_wsum, int &tot) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int incl_w = wave_excl_scan(v, lane) + v;
  if (lane == 63) s_wsum[wv] = incl_w;
  __syncthreads();
  int base = 0;
  tot = 0;
#pragma unroll
  for (int w = 0; w < 4; w++) {
    if (w < wv) base += s_wsum[w];
    tot += s_wsum[w];
  }
  __syncthreads();
  return base + incl_w - v;
}

template <class T, int NB, int IB, int TQ>
__global__ void __launch_bounds__(256) k_dense(DenseArgs a) {
  constexpr int NA = NB / 3;
  constexpr int NCOL = IB * NB * NB * NB;
  constexpr int QMAX = 1280;
  constexpr int NOUT2 = IB * NB * 9;            // A' entries per group
  constexpr int OPT = (NOUT2 + 255) / 256;      // A' entries owned per thread
  extern __shared__ unsigned char dyn[];
  C2<T> *s_ph = reinterpret_cast<C2<T> *>(dyn);  // [TQ][ng]
  __shared__ double2 s_e1[NB][IB];
  __shared__ double2 s_e2[NB][NB];
  __shared__ double2 s_af[81];
  __shared__ double2 s_b1[IB][27];
  __shared__ C2<T> s_psi[TQ][NCOL];
  __shared__ C2<T> s_e3[TQ][NB][NB];
  __shared__ C2<T> s_e4[TQ][NB][NB];
  __shared__ int s_jj[TQ], s_ss[TQ], s_n3[TQ][3];
  __shared__ double s_Ejj[TQ][NB], s_Ess[TQ][NB];
  __shared__ int s_vjj[256];
  __shared__ int s_q[QMAX];
  __shared__ double s_qw[QMAX];
  __shared__ int s_wsum[4];
  __shared__ double s_om[IB];
  __shared__ int s_act[IB];
  __shared__ double s_Eii[NB];
  __shared__ long s_item;
  __shared__ double s_red[4][3 * IB];

  const ModeView &m = a.m;
  const int tid = threadIdx.x;
  C2<T> *Ap = reinterpret_cast<C2<T> *>(a.scratch) + (size_t)blockIdx.x * a.slot_elems;
  const double hbarp = kHbar * (double)k1e22;
  const double gpre = hbarp * hbarp * kPi / 8.0;

  for (;;) {
    if (tid == 0) s_item = atomicAdd(a.counter, 1);
    __syncthreads();
    const long item = s_item;
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
      const int mm0 = ll0 * NB + i0;
      s_om[tid] = w;
      s_act[tid] = (a.sel[mm0] != 0) && (w != 0.0) && (w <= m.omega_max);
    }
    if (tid < NB) s_Eii[tid] = E(m, ii0, tid);
    __syncthreads();
    int any = 0;
#pragma unroll
    for (int il = 0; il < IB; il++) any |= s_act[il];

    double acc_g[IB], acc_w[IB];
    int acc_n[IB];
    int ntrip = 0;
#pragma unroll
    for (int il = 0; il < IB; il++) {
      acc_g[il] = 0.0;
      acc_w[il] = 0.0;
      acc_n[il] = 0;
    }

    if (any) {
      // ---------------- setup: eigenvectors of phonons 1 and 2
      for (int x = tid; x < NB * IB; x += 256) {
        const int comp = x / IB, il = x % IB;
        s_e1[comp][il] = m.eig[((size_t)qi0 * NB + ib * IB + il) * NB + comp];
      }
      for (int x = tid; x < NB * NB; x += 256) {
        const int comp = x / NB, j = x % NB;
        double2 e = m.eig[((size_t)ii0 * NB + j) * NB + comp];
        if (ch == 2) e.y = -e.y;
        s_e2[comp][j] = e;
      }
      __syncthreads();
      // ---------------- setup: A'[gamma] for every coarse group
      for (int gi = 0; gi < a.ng; gi++) {
        const DGroup G = a.g[gi];
        double2 acc[OPT];
#pragma unroll
        for (int o = 0; o < OPT; o++) acc[o] = make_double2(0.0, 0.0);
        for (int fi = G.f_first; fi < G.f_first + G.f_count; fi++) {
          const FGroup F = a.f[fi];
          if (tid < 81) {
            double2 s = make_double2(0.0, 0.0);
            for (int ci = F.c_first; ci < F.c_first + F.c_count; ci++) {
              const DCluster c = a.c[ci];
              const int k = s2 * (n2[0] * c.pjl[0] + n2[1] * c.pjl[1] + n2[2] * c.pjl[2]) -
                            (n1[0] * c.pl[0] + n1[1] * c.pl[1] + n1[2] * c.pl[2]);
              const double2 rt = a.roots[fmod_pos(k, a.L)];
              const double p = a.phi[(size_t)ci * 81 + tid] * c.minv;
              s.x += p * rt.x;
              s.y += p * rt.y;
            }
            s_af[tid] = s;
          }
          __syncthreads();
          for (int x = tid; x < IB * 27; x += 256) {
            const int il = x / 27, tsr = x % 27;
            double2 s = make_double2(0.0, 0.0);
#pragma unroll
            for (int u = 0; u < 3; u++) {
              const double2 v = zm(s_af[u + 3 * tsr], s_e1[3 * F.I + u][il]);
              s.x += v.x;
              s.y += v.y;
            }
            s_b1[il][tsr] = s;
          }
          __syncthreads();
#pragma unroll
          for (int o = 0; o < OPT; o++) {
            const int x = tid + 256 * o;
            if (x < NOUT2) {
              const int il = x / (NB * 9), rem = x % (NB * 9), j = rem / 9, sr = rem % 9;
              double2 s = acc[o];
#pragma unroll
              for (int t = 0; t < 3; t++) {
                const double2 v = zm(s_b1[il][t + 3 * sr], s_e2[3 * F.J + t][j]);
                s.x += v.x;
                s.y += v.y;
              }
              acc[o] = s;
            }
          }
          __syncthreads();
        }
#pragma unroll
        for (int o = 0; o < OPT; o++) {
          const int x = tid + 256 * o;
          if (x < NOUT2) Ap[(size_t)gi * NOUT2 + x] = C2<T>{(T)acc[o].x, (T)acc[o].y};
        }
      }
      __syncthreads();

      // ---------------- q'' loop
      const int jj_first = (ch == 1) ? 0 : ii0;
      for (int base = jj_first; base < m.nptk; base += 256) {
        // valid q'' of this chunk, in ascending order
        const int jj = base + tid;
        int ok = 0;
        if (jj < m.nptk) {
          int n3[3];
#pragma unroll
          for (int t = 0; t < 3; t++) n3[t] = m.ijk[3 * jj + t];
          const int ss = partner(m, ch, n1, n2, n3);
          ok = (ch == 0) || (ss >= jj);
        }
        int pos;
        const int nv = block_flag_scan(ok, s_wsum, pos);
        ntrip += nv;
        if (ok) s_vjj[pos] = jj;
        __syncthreads();
        for (int t0 = 0; t0 < nv; t0 += TQ) {
          const int tq = min(TQ, nv - t0);
          // tile data
          if (tid < TQ) {
            if (tid < tq) {
              const int jv = s_vjj[t0 + tid];
              int n3[3];
#pragma unroll
              for (int t = 0; t < 3; t++) {
                n3[t] = m.ijk[3 * jv + t];
                s_n3[tid][t] = n3[t];
              }
              s_jj[tid] = jv;
              s_ss[tid] = partner(m, ch, n1, n2, n3);
            }
          }
          __syncthreads();
          for (int x = tid; x < TQ * NB * NB; x += 256) {
            const int t = x / (NB * NB), comp = (x / NB) % NB, b = x % NB;
            if (t < tq) {
              double2 e3 = m.eig[((size_t)s_jj[t] * NB + b) * NB + comp];
              double2 e4 = m.eig[((size_t)s_ss[t] * NB + b) * NB + comp];
              if (ch != 0) e3.y = -e3.y;
              e4.y = -e4.y;
              s_e3[t][comp][b] = C2<T>{(T)e3.x, (T)e3.y};
              s_e4[t][comp][b] = C2<T>{(T)e4.x, (T)e4.y};
            }
          }
          for (int x = tid; x < TQ * NB; x += 256) {
            const int t = x / NB, b = x % NB;
            if (t < tq) {
              s_Ejj[t][b] = E(m, s_jj[t], b);
              s_Ess[t][b] = E(m, s_ss[t], b);
            }
          }
          for (int x = tid; x < TQ * a.ng; x += 256) {
            const int t = x / a.ng, gi = x % a.ng;
            if (t < tq) {
              const DGroup G = a.g[gi];
              const int k = s3 * (s_n3[t][0] * G.d[0] + s_n3[t][1] * G.d[1] + s_n3[t][2] * G.d[2]);
              const double2 rt = a.roots[fmod_pos(k, a.L)];
              s_ph[t * a.ng + gi] = C2<T>{(T)rt.x, (T)rt.y};
            }
          }
          __syncthreads();
          // skinny GEMM: Psi[t][col] = sum_{gamma in (K,L)} ph[t][gamma] A'[gamma][i j s r]
          for (int col = tid; col < NCOL; col += 256) {
            const int D = col % NB, Cc = (col / NB) % NB, ij = col / (NB * NB);
            const int K = Cc / 3, s = Cc % 3, L = D / 3, r = D % 3;
            const int g0 = a.kl_first[K * NA + L], g1 = a.kl_first[K * NA + L + 1];
            C2<T> acc[TQ];
#pragma unroll
            for (int t = 0; t < TQ; t++) acc[t] = C2<T>{(T)0, (T)0};
            for (int gi = g0; gi < g1; gi++) {
              const C2<T> av = Ap[(size_t)gi * NOUT2 + ij * 9 + s + 3 * r];
#pragma unroll
              for (int t = 0; t < TQ; t++) {
                const C2<T> p = s_ph[t * a.ng + gi];
                acc[t].re += p.re * av.re - p.im * av.im;
                acc[t].im += p.re * av.im + p.im * av.re;
              }
            }
#pragma unroll
            for (int t = 0; t < TQ; t++) s_psi[t][col] = acc[t];
          }
          __syncthreads();
          // exact filter over (t, i, j, k, l): each thread tests R consecutive
          // candidates, a block scan of the per-thread counts keeps the
          // survivors in candidate order, and the queue is contracted when
          // it is about to overflow and at the end of the tile.
          constexpr int NCAND = TQ * IB * NB * NB * NB;
          constexpr int R = 4, CH = 256 * R;
          int qlen = 0;
          for (int cb = 0; cb < NCAND; cb += CH) {
            double wps[R];
            int mask = 0;
#pragma unroll
            for (int rr = 0; rr < R; rr++) {
              const int cidx = cb + tid * R + rr;
              wps[rr] = 0.0;
              if (cidx < NCAND) {
                const int l = cidx % NB, k = (cidx / NB) % NB, j = (cidx / (NB * NB)) % NB;
                const int il = (cidx / (NB * NB * NB)) % IB, t = cidx / (NB * NB * NB * IB);
                if (t < tq && s_act[il]) {
                  const int jj0 = s_jj[t], ss0 = s_ss[t];
                  const int i0 = ib * IB + il;
                  const double w = s_om[il], op = s_Eii[j], odp = s_Ejj[t][k], otp = s_Ess[t][l];
                  bool c1 = (op != 0.0) && (odp != 0.0) && (otp != 0.0);
                  if (ch != 1 && jj0 == ii0 && k < j) c1 = false;
                  if (ch != 0 && ss0 == jj0 && l < k) c1 = false;
                  if (ch == 1 && c1)
                    c1 = (qi0 != jj0 || i0 != k) && (qi0 != ss0 || i0 != l) && (ii0 != jj0 || j != k) &&
                         (ii0 != ss0 || j != l);
                  if (c1) {
                    const double delta = mismatch(ch, w, op, odp, otp);
                    const float bound =
                        (m.vbound[(size_t)k * m.nptk + jj0] + m.vbound[(size_t)l * m.nptk + ss0]) * 1.0001f + 1e-30f;
                    if (!((float)fabs(delta) > bound)) {
                      double sigma = pair_sigma(m, ch, jj0, k, ss0, l);
                      if (fabs(delta) <= sigma) {
                        if (sigma <= (double)kClamp) sigma = 1.0 / sqrt(kPi);
                        wps[rr] = wp4_of(ch, w, op, odp, otp, delta, sigma, a.T);
                        mask |= 1 << rr;
                      }
                    }
                  }
                }
              }
            }
            int tot;
            int o = qlen + block_excl_scan(__popc(mask), s_wsum, tot);
#pragma unroll
            for (int rr = 0; rr < R; rr++)
              if ((mask >> rr) & 1) {
                s_q[o] = cb + tid * R + rr;
                s_qw[o] = wps[rr];
                o++;
              }
            qlen += tot;
            const bool last = (cb + CH >= NCAND);
            if (qlen > QMAX - CH || (last && qlen > 0)) {
              __syncthreads();
              for (int qi = tid; qi < qlen; qi += 256) {
                const int cc = s_q[qi];
                const int l = cc % NB, k = (cc / NB) % NB, j = (cc / (NB * NB)) % NB;
                const int il = (cc / (NB * NB * NB)) % IB, t = cc / (NB * NB * NB * IB);
                const double wpv = s_qw[qi];
                double g = 0.0;
                if (!a.onlyharmonic) {
                  const C2<T> *psi = &s_psi[t][(il * NB + j) * NB * NB];
                  C2<T> V = {(T)0, (T)0};
#pragma unroll
                  for (int Cc = 0; Cc < NB; Cc++) {
                    C2<T> inner = {(T)0, (T)0};
#pragma unroll
                    for (int D = 0; D < NB; D++) {
                      const C2<T> x = cm(psi[Cc * NB + D], s_e4[t][D][l]);
                      inner.re += x.re;
                      inner.im += x.im;
                    }
                    const C2<T> y = cm(inner, s_e3[t][Cc][k]);
                    V.re += y.re;
                    V.im += y.im;
                  }
                  const double vr = (double)V.re, vi = (double)V.im;
                  g = gpre * wpv * (vr * vr + vi * vi);
                }
#pragma unroll
                for (int x = 0; x < IB; x++)
                  if (x == il) {
                    acc_g[x] += g;
                    acc_w[x] += wpv;
                    acc_n[x] += 1;
                  }
              }
              qlen = 0;
              __syncthreads();
            }
          }
          __syncthreads();
        }
      }
    }
    // fixed-order block reduction of the per-thread partial sums
#pragma unroll
    for (int il = 0; il < IB; il++) {
      double g = wave_sum(acc_g[il]);
      double w = wave_sum(acc_w[il]);
      double n = wave_sum((double)acc_n[il]);
      if ((tid & 63) == 0) {
        s_red[tid >> 6][3 * il] = g;
        s_red[tid >> 6][3 * il + 1] = w;
        s_red[tid >> 6][3 * il + 2] = n;
      }
    }
    __syncthreads();
    if (tid < IB) {
      double g = ((s_red[0][3 * tid] + s_red[1][3 * tid]) + s_red[2][3 * tid]) + s_red[3][3 * tid];
      double w = ((s_red[0][3 * tid + 1] + s_red[1][3 * tid + 1]) + s_red[2][3 * tid + 1]) + s_red[3][3 * tid + 1];
      double n = ((s_red[0][3 * tid + 2] + s_red[1][3 * tid + 2]) + s_red[2][3 * tid + 2]) + s_red[3][3 * tid + 2];
      a.out_g[item * IB + tid] = g;
      a.out_w[item * IB + tid] = w;
      a.out_n[item * IB + tid] = (int)n;
    }
    if (tid == 0) a.out_t[item] = ntrip;
    __syncthreads();
  }
}

// Sum over the q' window of each (channel, point, mode-block, mode), in order.
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
  s->nc = n;
  bool ok = hipMalloc(&s->g, sizeof(DGroup) * G.size()) == hipSuccess &&
            hipMalloc(&s->f, sizeof(FGroup) * F.size()) == hipSuccess &&
            hipMalloc(&s->c, sizeof(DCluster) * n) == hipSuccess &&
            hipMalloc(&s->phi, sizeof(double) * 81 * n) == hipSuccess &&
            hipMalloc(&s->kl_first, sizeof(int) * klf.size()) == hipSuccess &&
            hipMalloc(&s->counter, sizeof(int)) == hipSuccess && hipEventCreate(&s->e0) == hipSuccess &&
            hipEventCreate(&s->e1) == hipSuccess;
  ok = ok && hipMemcpy(s->g, G.data(), sizeof(DGroup) * G.size(), hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(s->f, F.data(), sizeof(FGroup) * F.size(), hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(s->c, Cs.data(), sizeof(DCluster) * n, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(s->phi, P.data(), sizeof(double) * 81 * n, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(s->kl_first, klf.data(), sizeof(int) * klf.size(), hipMemcpyHostToDevice) == hipSuccess;
  if (!ok) {
    dense_destroy(s);
    *rc = derr(FP4_ERR_NOMEM, "dense_create: device allocation or copy failed");
    return nullptr;
  }
  return s;
}

void dense_destroy(DenseState *s) {
  if (!s) return;
  void *ps[] = {s->g, s->f, s->c, s->phi, s->kl_first, s->scratch, s->outb, s->counter};
  for (void *p : ps)
    if (p) (void)hipFree(p);
  if (s->e0) (void)hipEventDestroy(s->e0);
  if (s->e1) (void)hipEventDestroy(s->e1);
  delete s;
}

int dense_ngroups(const DenseState *s) { return s ? s->ng : 0; }

void dense_info(const DenseState *s, long long *triples, long long *items, int *ngroups, int *nfine) {
  *triples = s ? s->last_triples : 0;
  *items = s ? s->last_items : 0;
  *ngroups = s ? s->ng : 0;
  *nfine = s ? s->nf : 0;
}

namespace {

struct Shape {
  int IB, TQ;
};

template <class T, int NB, int IB, int TQ>
int launch_dense(DenseArgs &a, int slots, hipStream_t st) {
  const size_t dyn = sizeof(C2<T>) * TQ * a.ng;
  hipLaunchKernelGGL((k_dense<T, NB, IB, TQ>), dim3(slots), dim3(256), dyn, st, a);
  return hipGetLastError() == hipSuccess ? 0 : derr(FP4_ERR_DEVICE, "dense kernel launch failed");
}

// (precision, Nb) -> (modes per block, q'' per tile) under 64 KiB of LDS
bool pick_shape(int prec, int nb, Shape &s) {
  const bool f32 = prec == FP4_PREC_FP32;
  switch (nb) {
    case 3: s = {3, f32 ? 16 : 8}; return true;
    case 6: s = {6, f32 ? 4 : 2}; return true;
    case 9: s = {3, f32 ? 2 : 1}; return true;
    case 12: s = {1, f32 ? 2 : 1}; return true;
    default: return false;
  }
}

int dispatch(int prec, int nb, DenseArgs &a, int slots, hipStream_t st) {
  if (prec == FP4_PREC_FP32) {
    switch (nb) {
      case 3: return launch_dense<float, 3, 3, 16>(a, slots, st);
      case 6: return launch_dense<float, 6, 6, 4>(a, slots, st);
      case 9: return launch_dense<float, 9, 3, 2>(a, slots, st);
      case 12: return launch_dense<float, 12, 1, 2>(a, slots, st);
    }
  } else {
    switch (nb) {
      case 3: return launch_dense<double, 3, 3, 8>(a, slots, st);
      case 6: return launch_dense<double, 6, 6, 2>(a, slots, st);
      case 9: return launch_dense<double, 9, 3, 1>(a, slots, st);
      case 12: return launch_dense<double, 12, 1, 1>(a, slots, st);
    }
  }
  return derr(FP4_ERR_UNSUPPORTED, "dense form: Nbands not supported");
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

}  // namespace

int dense_rates(DenseState *s, const DenseRun &r, const std::vector<int> &modes, int ii_lo0, int nii,
                std::vector<double> &g, std::vector<double> &w, std::vector<long long> &n, double *ms) {
  if (!s) return derr(FP4_ERR_STATE, "dense state missing");
  if (r.precision != FP4_PREC_FP32 && r.precision != FP4_PREC_FP64_DENSE)
    return derr(FP4_ERR_UNSUPPORTED, "dense form: precision not built");
  const int nb = r.m.nb;
  Shape sh;
  if (!pick_shape(r.precision, nb, sh)) return derr(FP4_ERR_UNSUPPORTED, "dense form: Nbands not supported");
  const int nmodes_all = nb * r.m.nlist;
  std::vector<unsigned char> sel(nmodes_all, 0);
  std::vector<int> lls;
  for (int mm : modes) {
    sel[mm - 1] = 1;
    const int ll = (mm - 1) / nb;
    if (lls.empty() || lls.back() != ll) lls.push_back(ll);
  }
  const int nls = (int)lls.size(), nib = nb / sh.IB;
  const long nitems = 3L * nls * nib * nii;
  int dev = 0, ncu = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  int slots = (r.tuning && r.tuning->slots > 0) ? r.tuning->slots : std::max(1, ncu) * 2;
  if (slots > nitems) slots = (int)std::max(1L, nitems);
  const size_t tsz = (r.precision == FP4_PREC_FP32) ? sizeof(float) : sizeof(double);
  const long slot_elems = (long)s->ng * sh.IB * nb * 9;
  int rc;
  if ((rc = ensure(s->scratch, s->scratch_bytes, (size_t)slots * slot_elems * 2 * tsz))) return rc;
  const size_t per_item = sizeof(double) * 2 + sizeof(int);
  const long nseg = 3L * nls * nib * sh.IB;
  // outputs: og, ow, on per item*IB, then sg, sw, sn per segment, then lls and sel
  const size_t need = per_item * nitems * sh.IB + sizeof(int) * nitems + (sizeof(double) * 2 + sizeof(long long)) * nseg +
                      sizeof(int) * nls + nmodes_all + 1024;
  if ((rc = ensure(s->outb, s->out_bytes, need))) return rc;
  char *p = reinterpret_cast<char *>(s->outb);
  double *og = reinterpret_cast<double *>(p);
  p += sizeof(double) * nitems * sh.IB;
  double *ow = reinterpret_cast<double *>(p);
  p += sizeof(double) * nitems * sh.IB;
  double *sg = reinterpret_cast<double *>(p);
  p += sizeof(double) * nseg;
  double *sw = reinterpret_cast<double *>(p);
  p += sizeof(double) * nseg;
  long long *sn = reinterpret_cast<long long *>(p);
  p += sizeof(long long) * nseg;
  int *on = reinterpret_cast<int *>(p);
  p += sizeof(int) * nitems * sh.IB;
  int *ot = reinterpret_cast<int *>(p);
  p += sizeof(int) * nitems;
  int *dlls = reinterpret_cast<int *>(p);
  p += sizeof(int) * nls;
  unsigned char *dsel = reinterpret_cast<unsigned char *>(p);
  hipStream_t st = r.stream;
  if (hipMemcpyAsync(dlls, lls.data(), sizeof(int) * nls, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(dsel, sel.data(), nmodes_all, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemsetAsync(s->counter, 0, sizeof(int), st) != hipSuccess)
    return derr(FP4_ERR_DEVICE, "dense: upload failed");
  DenseArgs a{};
  a.m = r.m;
  a.roots = r.roots;
  a.L = s->L;
  a.g = s->g;
  a.f = s->f;
  a.c = s->c;
  a.phi = s->phi;
  a.kl_first = s->kl_first;
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
  a.scratch = s->scratch;
  a.slot_elems = slot_elems;
  a.out_g = og;
  a.out_w = ow;
  a.out_n = on;
  a.out_t = ot;
  if (hipEventRecord(s->e0, st) != hipSuccess) return derr(FP4_ERR_DEVICE, "event record failed");
  if ((rc = dispatch(r.precision, nb, a, slots, st))) return rc;
  if (hipEventRecord(s->e1, st) != hipSuccess) return derr(FP4_ERR_DEVICE, "event record failed");
  hipLaunchKernelGGL(k_dense_reduce, dim3((unsigned)((nseg + 127) / 128)), dim3(128), 0, st, og, ow, on, nseg, sh.IB,
                     nii, sg, sw, sn);
  std::vector<double> hg(nseg), hw(nseg);
  std::vector<long long> hn(nseg);
  if (hipMemcpyAsync(hg.data(), sg, sizeof(double) * nseg, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(hw.data(), sw, sizeof(double) * nseg, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(hn.data(), sn, sizeof(long long) * nseg, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return derr(FP4_ERR_DEVICE, "dense: kernel or copy failed");
  {
    std::vector<int> ht(nitems);
    if (hipMemcpy(ht.data(), ot, sizeof(int) * nitems, hipMemcpyDeviceToHost) != hipSuccess)
      return derr(FP4_ERR_DEVICE, "dense: copy failed");
    long long tt = 0;
    for (int x : ht) tt += x;
    s->last_triples = tt;
    s->last_items = nitems;
  }
  float kms = 0.f;
  (void)hipEventElapsedTime(&kms, s->e0, s->e1);
  if (ms) *ms = kms;
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
        for (int il = 0; il < sh.IB; il++) {
          const int mm = lls[ls] * nb + ib * sh.IB + il + 1;
          auto it = mpos.find(mm);
          if (it == mpos.end()) continue;
          const long sidx = (((long)ch * nls + ls) * nib + ib) * sh.IB + il;
          g[ch * nsub + it->second] = hg[sidx];
          w[ch * nsub + it->second] = hw[sidx];
          n[ch * nsub + it->second] = hn[sidx];
        }
  return 0;
}

}  // namespace fp4
