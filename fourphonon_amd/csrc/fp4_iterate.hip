// Iterative BTE on the device:
//   iteration0     Src/iterations.f90:13-30
//   iteration_4ph  Src/iterations.f90:95-215
//   TConduct       Src/conductivity.f90:11-37
//   convergence    Src/ShengBTE.f90:1250-1307 (with symmetrize_tensor,
//                  Src/config.f90:503-518, and twonorm3x3, Src/misc.f90:52-65)
//
// iteration_4ph updates F_n in place while later steps read it, so its
// result depends on the visiting order (irreducible point ll, branch i, star
// member kk).  That order is kept: one launch per (ll, i, kk) step.  Inside a
// step the sum over the mode's processes (3ph +, 3ph -, 4ph ++, +-, --) is a
// parallel reduction with a fixed assignment of processes to threads and a
// fixed tree; the last block to finish (atomic ticket) adds the block
// partials in block order, writes F_n(i, q_kk, :) and applies the
// symmetriser of q_kk to all branches, as the reference does inside its
// loop.  Everything is FP64; results are bitwise reproducible run to run.
//
// Device layout: F_n is [q][b][alpha] (one mode's 3-vector in 24 contiguous
// bytes, so a gather touches one cache line); the lists are int32 (the
// encoded (q-1)*Nb+branch fits for nptk*Nb < 2^31) and Gamma double, 20 B
// per 4ph process.
//
// Aggregated operator.  The sum over a mode's processes is linear in F_n and
// each term reads F_n at one partner mode, so it regroups by partner:
// DeltaF(i, q_kk) = sum_c W[(ll,i)][c] F_n(b, ID_equi(S_kk, q), :) over the
// full-mesh modes c = q*Nb + b, with W the signed Gamma summed over the
// processes that have partner c (3ph and 4ph alike).  W has Nb*Nlist rows of
// nptk*Nb doubles whatever the number of processes (590 MB at 20^3, Nb = 6,
// against about 10 TB of lists), so it is assembled once per temperature
// on the device and the in-place sweep keeps the reference's order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/fp4.h"
#include "fp4_device.hpp"
#include "fp4_iterate.hpp"

namespace fp4 {

namespace {

thread_local std::string b_err;

int bfail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  b_err = buf;
  return code;
}

#define BCHECK(x)                                                                           \
  do {                                                                                      \
    hipError_t e_ = (x);                                                                    \
    if (e_ != hipSuccess) return bfail(FP4_ERR_DEVICE, "%s: %s", #x, hipGetErrorString(e_)); \
  } while (0)

struct DBuf {
  void *p = nullptr;
  size_t bytes = 0;
  int ensure(size_t need) {
    if (need <= bytes && p) return 0;
    release();
    if (hipMalloc(&p, std::max<size_t>(need, 64)) != hipSuccess) {
      p = nullptr;
      return bfail(FP4_ERR_NOMEM, "hipMalloc(%zu) failed", need);
    }
    bytes = std::max<size_t>(need, 64);
    return 0;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <typename T>
  T *as() const {
    return reinterpret_cast<T *>(p);
  }
};

constexpr int kStepThreads = 256;
constexpr int kMaxStepBlocks = 2048;
constexpr long long kPerThread = 8;        // processes per thread before another block is added
constexpr long long kStageEntries = 1LL << 24;

}  // namespace

struct BteState {
  int nptk = 0, nb = 0, nlist = 0, nsymm = 0;
  bool have_symm = false, have_orbits = false, have_types = false, have_op = false;
  std::vector<int> h_ideq, h_nequi, h_alleq, h_type;
  DBuf ideq, symm, crot, nequi, alleq;
  DBuf e2[5], e3[5], e4[3], g[5];
  std::vector<long long> h_n[5], h_off[5];  // per mode
  DBuf F, stage, part, ticket, kappa, mode, tau, hist, rel, flag;
  DBuf W;                      // aggregated operator [Nb*Nlist][nptk*Nb]
  bool use_w = false;
  bool persistent = true;      // aggregated operator: one cooperative launch per sweep
  DBuf steps, taus, bar;       // persistent sweep: step table, tau per step, barrier counter and error flag
  long long wcols = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0.0;
  long long last_launches = 0;
};

#define FQ(F, nb, q, b, a) (F)[((size_t)(q) * (nb) + (b)) * 3 + (a)]

// -------------------------------------------------------------- kernels
__device__ __forceinline__ int map_q(const IterStep &s, long long enc, int *branch) {
  const int e0 = (int)(enc - 1);
  *branch = e0 % s.nb;
  const int q0 = e0 / s.nb;
  return s.ideq[(size_t)s.sym + (size_t)s.nsymm * q0] - 1;
}

__device__ __forceinline__ void add_term(const IterStep &s, int type, long long p, double acc[3]) {
  int j, k, l = 0;
  const int qm = map_q(s, s.e2[type][p], &j);
  const int qn = map_q(s, s.e3[type][p], &k);
  const double gam = s.g[type][p];
  const double *Fj = &FQ(s.F, s.nb, qm, j, 0);
  const double *Fk = &FQ(s.F, s.nb, qn, k, 0);
  if (type < 2) {
#pragma unroll
    for (int a = 0; a < 3; a++) {
      // iterations.f90:152-155 (+) and 163-166 (-)
      const double t = (type == 0) ? gam * (Fk[a] - Fj[a]) : gam * (Fk[a] + Fj[a]) * 5.e-1;
      acc[a] = acc[a] + t;
    }
  } else {
    const int qs = map_q(s, s.e4[type - 2][p], &l);
    const double *Fl = &FQ(s.F, s.nb, qs, l, 0);
#pragma unroll
    for (int a = 0; a < 3; a++) {
      // iterations.f90:177-179 (++), 189-191 (+-), 201-203 (--)
      double t;
      if (type == 2)
        t = gam * (Fl[a] - Fj[a] - Fk[a]);
      else if (type == 3)
        t = gam * (Fl[a] - Fj[a] + Fk[a]);
      else
        t = gam * (Fl[a] + Fj[a] + Fk[a]);
      acc[a] = acc[a] + t;
    }
  }
}

__global__ void __launch_bounds__(kStepThreads) k_iter_step(IterStep s) {
  __shared__ double red[kStepThreads / 64][3];
  __shared__ int last;
  const int tid = threadIdx.x;
  double acc[3] = {0.0, 0.0, 0.0};
  const long long stride = (long long)gridDim.x * kStepThreads;
  const long long me = (long long)blockIdx.x * kStepThreads + tid;
  if (s.W) {
    // aggregated operator: sum over partner modes c = q*Nb + b of
    // W[mode][c] * F_n(b, ID_equi(sym, q), :)
    const double *row = s.W + (size_t)s.wrow * (size_t)s.wcols;
    for (long long c = me; c < s.wcols; c += stride) {
      const double w = row[c];
      if (w == 0.0) continue;
      const int q = (int)(c / s.nb), b = (int)(c - (long long)q * s.nb);
      const int qr = s.ideq[(size_t)s.sym + (size_t)s.nsymm * q] - 1;
      const double *Fp = &FQ(s.F, s.nb, qr, b, 0);
#pragma unroll
      for (int a = 0; a < 3; a++) acc[a] = acc[a] + w * Fp[a];
    }
  } else {
    for (int type = 0; type < 5; type++) {
      const long long n = s.n[type], off = s.off[type];
      for (long long p = me; p < n; p += stride) add_term(s, type, off + p, acc);
    }
  }
#pragma unroll
  for (int a = 0; a < 3; a++) acc[a] = wave_sum(acc[a]);
  if ((tid & 63) == 0)
#pragma unroll
    for (int a = 0; a < 3; a++) red[tid >> 6][a] = acc[a];
  __syncthreads();
  if (tid == 0) {
    for (int a = 0; a < 3; a++) {
      double x = red[0][a];
      for (int w = 1; w < kStepThreads / 64; w++) x = x + red[w][a];
      s.part[3 * (size_t)blockIdx.x + a] = x;
    }
    __threadfence();
    const unsigned int t = atomicAdd(s.ticket, 1u);
    last = (t == gridDim.x - 1) ? 1 : 0;
  }
  __syncthreads();
  if (!last || tid != 0) return;
  __threadfence();
  double dF[3];
  for (int a = 0; a < 3; a++) {
    double x = 0.0;
    for (unsigned int b = 0; b < gridDim.x; b++) x = x + s.part[3 * (size_t)b + a];
    dF[a] = x;
  }
  *s.ticket = 0u;
  // F_n(i, q_kk, :) = tau_zero*v*omega + tau_zero*DeltaF   (iterations.f90:206-207)
  const double w = s.omega[(size_t)s.i * s.nptk + s.qk];
  for (int a = 0; a < 3; a++) {
    const double v = s.vel[((size_t)a * s.nb + s.i) * s.nptk + s.qk];
    FQ(s.F, s.nb, s.qk, s.i, a) = s.tau * v * w + s.tau * dF[a];
  }
  // F_n(:, q_kk, :) <- symmetrizers(:,:,q_kk) applied to every branch   (iterations.f90:208)
  const double *S = s.symm + 9 * (size_t)s.qk;
  for (int b = 0; b < s.nb; b++) {
    double x[3], y[3];
    for (int a = 0; a < 3; a++) x[a] = FQ(s.F, s.nb, s.qk, b, a);
    for (int a = 0; a < 3; a++) {
      double t = 0.0;
      for (int c = 0; c < 3; c++) t = t + S[a + 3 * c] * x[c];
      y[a] = t;
    }
    for (int a = 0; a < 3; a++) FQ(s.F, s.nb, s.qk, b, a) = y[a];
  }
}

// A whole iteration_4ph sweep with the aggregated operator in one
// cooperative launch of G blocks (the G of k_iter_step for this operator):
// the (ll, i, kk) steps run in order, each block computing the same
// partial sums as k_iter_step's block of that index.  A grid barrier per
// step (a monotonic arrival counter, device-scope fences around it) makes
// the partials visible; every block then adds them in block order, as
// k_iter_step's last block does, and writes the same F_n(:, q_kk, :)
// update itself, so the next step's gathers read the new values without a
// second barrier (every block writes bit-identical values).  The old row
// F_n(:, q_kk, :) the symmetriser reads is loaded before the block arrives
// at the step's barrier, when no block can have written the step's update
// yet; after the barrier a block computes only from that copy and the
// partials (reading the row again could see another block's finished update
// and apply the symmetriser to it twice).  Partials are double-buffered by
// step parity.  The barrier waits are bounded: after
// spin_limit polls a block sets *err and every block leaves at its next
// barrier.  Results are bitwise those of one k_iter_step launch per step.
struct SweepW {
  const int4 *steps;      // per step: W row (ll*Nb + i), i, q_kk, symmetry operation (0-based)
  const double *tau;      // tau_zero of the step's mode
  int nsteps;
  const double *W;
  long long wcols;
  const int *ideq;
  const double *symm, *omega, *vel;
  double *F;
  double *part;           // [2][G][3]
  unsigned int *bar;      // arrivals, 0 at launch
  int *err;               // set when a barrier wait exceeds spin_limit
  int nb, nptk, nsymm;
  long long spin_limit;
};

__device__ __forceinline__ bool sweep_barrier(const SweepW &a, unsigned target) {
  __threadfence();
  atomicAdd(a.bar, 1u);
  long long n = 0;
  while (__hip_atomic_load(a.bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
    if (__hip_atomic_load(a.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return false;
    if (++n > a.spin_limit) {
      atomicExch(a.err, 1);
      return false;
    }
    __builtin_amdgcn_s_sleep(1);
  }
  __threadfence();
  return true;
}

constexpr int kSweepMaxNb = 96;   // branches of one F_n row a block keeps in LDS

__global__ void __launch_bounds__(kStepThreads) k_sweep_w(SweepW a) {
  __shared__ double red[kStepThreads / 64][3];
  __shared__ double row0[kSweepMaxNb * 3];   // F_n(:, q_kk, :) before the step
  __shared__ int stop;
  const int tid = threadIdx.x;
  const int G = (int)gridDim.x;
  const long long stride = (long long)G * kStepThreads;
  const long long me = (long long)blockIdx.x * kStepThreads + tid;
  for (int st = 0; st < a.nsteps; st++) {
    const int4 sp = a.steps[st];
    double acc[3] = {0.0, 0.0, 0.0};
    const double *row = a.W + (size_t)sp.x * (size_t)a.wcols;
    for (long long c = me; c < a.wcols; c += stride) {
      const double w = row[c];
      if (w == 0.0) continue;
      const int q = (int)(c / a.nb), b = (int)(c - (long long)q * a.nb);
      const int qr = a.ideq[(size_t)sp.w + (size_t)a.nsymm * q] - 1;
      const double *Fp = &FQ(a.F, a.nb, qr, b, 0);
#pragma unroll
      for (int x = 0; x < 3; x++) acc[x] = acc[x] + w * Fp[x];
    }
#pragma unroll
    for (int x = 0; x < 3; x++) acc[x] = wave_sum(acc[x]);
    if ((tid & 63) == 0)
#pragma unroll
      for (int x = 0; x < 3; x++) red[tid >> 6][x] = acc[x];
    __syncthreads();
    if (tid == 0) {
      double *part = a.part + (size_t)(st & 1) * 3 * G;
      for (int x = 0; x < 3; x++) {
        double v = red[0][x];
        for (int w = 1; w < kStepThreads / 64; w++) v = v + red[w][x];
        part[3 * (size_t)blockIdx.x + x] = v;
      }
      for (int b = 0; b < a.nb; b++)
        for (int x = 0; x < 3; x++) row0[3 * b + x] = FQ(a.F, a.nb, sp.z, b, x);
      stop = sweep_barrier(a, (unsigned)(st + 1) * (unsigned)G) ? 0 : 1;
      if (!stop) {
        double dF[3];
        for (int x = 0; x < 3; x++) {
          double v = 0.0;
          for (int b = 0; b < G; b++) v = v + part[3 * (size_t)b + x];
          dF[x] = v;
        }
        // F_n(i, q_kk, :) = tau_zero*v*omega + tau_zero*DeltaF   (iterations.f90:206-207)
        const int i = sp.y, qk = sp.z;
        const double tau = a.tau[st];
        const double w = a.omega[(size_t)i * a.nptk + qk];
        for (int x = 0; x < 3; x++) {
          const double v = a.vel[((size_t)x * a.nb + i) * a.nptk + qk];
          row0[3 * i + x] = tau * v * w + tau * dF[x];
        }
        // F_n(:, q_kk, :) <- symmetrizers(:,:,q_kk) applied to every branch   (iterations.f90:208)
        const double *S = a.symm + 9 * (size_t)qk;
        for (int b = 0; b < a.nb; b++) {
          double xv[3], y[3];
          for (int x = 0; x < 3; x++) xv[x] = row0[3 * b + x];
          for (int x = 0; x < 3; x++) {
            double t = 0.0;
            for (int c = 0; c < 3; c++) t = t + S[x + 3 * c] * xv[c];
            y[x] = t;
          }
          for (int x = 0; x < 3; x++) FQ(a.F, a.nb, qk, b, x) = y[x];
        }
      }
    }
    __syncthreads();
    if (stop) return;
  }
}

__global__ void k_zero_u32(unsigned int *p, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0u;
}

// iteration0: block per irreducible point, threads over (branch, star member).
__global__ void k_iteration0(int nb, int nptk, int nsymm, const int *Nequi, const int *alleq, const double *omega,
                             const double *vel, const double *tau, double *F) {
  const int ll = blockIdx.x;
  const int ne = Nequi[ll];
  for (int x = threadIdx.x; x < nb * ne; x += blockDim.x) {
    const int ii = x % nb, kk = x / nb;
    const int q = alleq[(size_t)kk + (size_t)nsymm * ll] - 1;
    const double t = tau[ii + (size_t)nb * ll];
    const double w = omega[(size_t)ii * nptk + q];
    for (int a = 0; a < 3; a++) {
      const double v = vel[((size_t)a * nb + ii) * nptk + q];
      FQ(F, nb, q, ii, a) = t * v * w;
    }
  }
}

// Symmetriser on every q (Src/ShengBTE.f90:1288-1290): thread per (q, b).
__global__ void k_symmetrize_all(int nb, int nptk, const double *symm, double *F) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nb * nptk) return;
  const int q = (int)(t / nb), b = (int)(t % nb);
  const double *S = symm + 9 * (size_t)q;
  double x[3], y[3];
  for (int a = 0; a < 3; a++) x[a] = FQ(F, nb, q, b, a);
  for (int a = 0; a < 3; a++) {
    double acc = 0.0;
    for (int c = 0; c < 3; c++) acc = acc + S[a + 3 * c] * x[c];
    y[a] = acc;
  }
  for (int a = 0; a < 3; a++) FQ(F, nb, q, b, a) = y[a];
}

// Layout changes between the reference's F_n(Nb,nptk,3) and [q][b][alpha].
__global__ void k_F_from_fortran(int nb, int nptk, const double *src, double *F) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n = (long long)nb * nptk;
  if (t >= n) return;
  for (int a = 0; a < 3; a++) F[3 * t + a] = src[t + n * a];
}
__global__ void k_F_to_fortran(int nb, int nptk, const double *F, double *dst) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n = (long long)nb * nptk;
  if (t >= n) return;
  for (int a = 0; a < 3; a++) dst[t + n * a] = F[3 * t + a];
}

// TConduct: block per branch, q = 2..nptk, fixed-order reduction; the
// prefactor 1e21(float) hbar^2 / (kB T^2 V nptk) is applied in the
// reference's expression order.
__global__ void __launch_bounds__(256) k_tconduct(int nb, int nptk, double T, double V, const double *omega,
                                                   const double *vel, const double *F, double *kappa, double *mode) {
  const int jj = blockIdx.x;
  const int tid = threadIdx.x;
  const double pre = (double)1e21f * (kHbar * kHbar);
  const double den = kKb * T * T * V * (double)nptk;
  double acc[9];
#pragma unroll
  for (int t = 0; t < 9; t++) acc[t] = 0.0;
  if (mode && tid == 0)
    for (int t = 0; t < 9; t++) mode[(size_t)nptk * ((size_t)jj + (size_t)nb * t)] = 0.0;
  for (int ii = 1 + tid; ii < nptk; ii += blockDim.x) {
    const double w = omega[(size_t)jj * nptk + ii];
    const double f = 1.0 / (exp(kHbar * w / kKb / T) - 1.0);
#pragma unroll
    for (int d2 = 0; d2 < 3; d2++) {
      const double Fd = FQ(F, nb, ii, jj, d2);
#pragma unroll
      for (int d1 = 0; d1 < 3; d1++) {
        const double tmp = vel[((size_t)d1 * nb + jj) * nptk + ii] * Fd;
        const double md = f * (f + 1) * w * tmp;
        acc[d1 + 3 * d2] = acc[d1 + 3 * d2] + md;
        if (mode) mode[(size_t)ii + (size_t)nptk * ((size_t)jj + (size_t)nb * (d1 + 3 * d2))] = pre * md / den;
      }
    }
  }
  __shared__ double red[4][9];
#pragma unroll
  for (int t = 0; t < 9; t++) {
    const double v = wave_sum(acc[t]);
    if ((tid & 63) == 0) red[tid >> 6][t] = v;
  }
  __syncthreads();
  if (tid < 9) {
    const double sum = ((red[0][tid] + red[1][tid]) + red[2][tid]) + red[3][tid];
    kappa[jj + (size_t)nb * tid] = pre * sum / den;
  }
}

// Largest singular value of a 3x3 matrix (column-major): square root of the
// largest eigenvalue of A^T A by cyclic Jacobi rotations.
__device__ double twonorm3x3(const double *A) {
  double B[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double s = 0.0;
      for (int k = 0; k < 3; k++) s = s + A[k + 3 * i] * A[k + 3 * j];
      B[i][j] = s;
    }
  for (int sweep = 0; sweep < 12; sweep++) {
    const double off = fabs(B[0][1]) + fabs(B[0][2]) + fabs(B[1][2]);
    if (off == 0.0) break;
    for (int p = 0; p < 2; p++)
      for (int q = p + 1; q < 3; q++) {
        const double apq = B[p][q];
        if (apq == 0.0) continue;
        const double theta = (B[q][q] - B[p][p]) / (2.0 * apq);
        const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < 3; k++) {
          const double bkp = B[k][p], bkq = B[k][q];
          B[k][p] = c * bkp - s * bkq;
          B[k][q] = s * bkp + c * bkq;
        }
        for (int k = 0; k < 3; k++) {
          const double bpk = B[p][k], bqk = B[q][k];
          B[p][k] = c * bpk - s * bqk;
          B[q][k] = s * bpk + c * bqk;
        }
      }
  }
  const double lmax = fmax(fmax(B[0][0], B[1][1]), B[2][2]);
  return sqrt(fmax(lmax, 0.0));
}

// symmetrize_tensor on every branch, then sum over branches into hist[it]
// and, for it > 0, the relative change against hist[it-1].
__global__ void k_kappa_finish(int nb, int nsymm, const double *crot, double *kappa, double *hist, int it,
                               double *rel) {
  const int b = threadIdx.x;
  if (b < nb) {
    double Tn[9], tmp[9];
    for (int t = 0; t < 9; t++) {
      Tn[t] = kappa[b + (size_t)nb * t];
      tmp[t] = 0.0;
    }
    for (int is = 0; is < nsymm; is++) {
      const double *C = crot + 9 * (size_t)is;
      double M1[9];
      for (int a = 0; a < 3; a++)
        for (int c = 0; c < 3; c++) {  // (T C^T)(a,c)
          double s = 0.0;
          for (int k = 0; k < 3; k++) s = s + Tn[a + 3 * k] * C[c + 3 * k];
          M1[a + 3 * c] = s;
        }
      for (int a = 0; a < 3; a++)
        for (int c = 0; c < 3; c++) {  // (C M1)(a,c)
          double s = 0.0;
          for (int k = 0; k < 3; k++) s = s + C[a + 3 * k] * M1[k + 3 * c];
          tmp[a + 3 * c] = tmp[a + 3 * c] + s;
        }
    }
    for (int t = 0; t < 9; t++) kappa[b + (size_t)nb * t] = tmp[t] / (double)nsymm;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double ks[9];
  for (int t = 0; t < 9; t++) {
    double s = 0.0;
    for (int bb = 0; bb < nb; bb++) s = s + kappa[bb + (size_t)nb * t];
    ks[t] = s;
    hist[9 * (size_t)it + t] = s;
  }
  if (it > 0) {
    double d[9];
    const double *old = hist + 9 * (size_t)(it - 1);
    for (int t = 0; t < 9; t++) d[t] = ks[t] - old[t];
    rel[it] = twonorm3x3(d) / twonorm3x3(old);
  } else {
    rel[0] = 0.0;
  }
}

// int32/int64 list entries -> int32 with a range check 1..maxv.
template <typename T>
__global__ void k_narrow(const T *src, int *dst, long long n, long long maxv, int *bad) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const long long v = (long long)src[t];
  if (v < 1 || v > maxv) {
    *bad = 1;
    dst[t] = 1;
  } else {
    dst[t] = (int)v;
  }
}

// -------------------------------------------------------------- host side
const char *bte_error() { return b_err.c_str(); }

BteState *bte_create() {
  BteState *s = new BteState();
  if (hipEventCreate(&s->ev0) != hipSuccess || hipEventCreate(&s->ev1) != hipSuccess) {
    bte_destroy(s);
    return nullptr;
  }
  return s;
}

void bte_destroy(BteState *s) {
  if (!s) return;
  DBuf *bufs[] = {&s->ideq, &s->symm, &s->crot, &s->nequi, &s->alleq, &s->F,    &s->stage, &s->part,
                  &s->ticket, &s->kappa, &s->mode, &s->tau, &s->hist, &s->rel, &s->flag, &s->W};
  for (DBuf *b : bufs) b->release();
  for (int t = 0; t < 5; t++) {
    s->e2[t].release();
    s->e3[t].release();
    s->g[t].release();
  }
  for (int t = 0; t < 3; t++) s->e4[t].release();
  if (s->ev0) (void)hipEventDestroy(s->ev0);
  if (s->ev1) (void)hipEventDestroy(s->ev1);
  delete s;
}

double bte_last_ms(const BteState *s) { return s ? s->last_ms : 0.0; }
long long bte_last_launches(const BteState *s) { return s ? s->last_launches : 0; }

static int check_dims(BteState *s, const BteModes &m) {
  if (s->nptk && (s->nptk != m.nptk || s->nb != m.nb || s->nlist != m.nlist)) {
    // mode data changed shape: everything derived from it is stale
    s->have_symm = s->have_orbits = s->have_types = s->have_op = false;
    s->use_w = false;
  }
  s->nptk = m.nptk;
  s->nb = m.nb;
  s->nlist = m.nlist;
  return 0;
}

int bte_set_symmetry(BteState *s, const BteModes &m, int nsymm, const int32_t *ID_equi, const double *symmetrizers,
                     const double *crotations) {
  if (!ID_equi || !symmetrizers || !crotations) return bfail(FP4_ERR_ARG, "null symmetry array");
  if (nsymm < 1 || nsymm > 1024) return bfail(FP4_ERR_ARG, "Nsymm_rot=%d out of range", nsymm);
  check_dims(s, m);
  const size_t np = m.nptk;
  for (size_t t = 0; t < (size_t)nsymm * np; t++)
    if (ID_equi[t] != -1 && (ID_equi[t] < 1 || ID_equi[t] > m.nptk))
      return bfail(FP4_ERR_ARG, "ID_equi entry %zu = %d out of range", t + 1, ID_equi[t]);
  int rc;
  if ((rc = s->ideq.ensure(sizeof(int) * nsymm * np)) || (rc = s->symm.ensure(sizeof(double) * 9 * np)) ||
      (rc = s->crot.ensure(sizeof(double) * 9 * nsymm)))
    return rc;
  BCHECK(hipMemcpy(s->ideq.p, ID_equi, sizeof(int) * nsymm * np, hipMemcpyHostToDevice));
  BCHECK(hipMemcpy(s->symm.p, symmetrizers, sizeof(double) * 9 * np, hipMemcpyHostToDevice));
  BCHECK(hipMemcpy(s->crot.p, crotations, sizeof(double) * 9 * nsymm, hipMemcpyHostToDevice));
  s->h_ideq.assign(ID_equi, ID_equi + (size_t)nsymm * np);
  // orbits were validated against the previous ID_equi
  s->have_orbits = s->have_types = s->have_op = false;
  s->nsymm = nsymm;
  s->have_symm = true;
  return 0;
}

int bte_set_orbits(BteState *s, const BteModes &m, const int32_t *Nequi, const int32_t *ALLEquiList,
                   const int32_t *TypeofSymmetry) {
  check_dims(s, m);
  if (!s->have_symm) return bfail(FP4_ERR_STATE, "symmetry data not set");
  if (!Nequi || !ALLEquiList) return bfail(FP4_ERR_ARG, "null orbit array");
  const int ns = s->nsymm;
  std::vector<char> seen(m.nptk, 0);
  for (int ll = 0; ll < m.nlist; ll++) {
    if (Nequi[ll] < 1 || Nequi[ll] > ns) return bfail(FP4_ERR_ARG, "Nequi(%d)=%d out of range", ll + 1, Nequi[ll]);
    for (int kk = 0; kk < Nequi[ll]; kk++) {
      const size_t at = (size_t)kk + (size_t)ns * ll;
      const int q = ALLEquiList[at];
      if (q < 1 || q > m.nptk) return bfail(FP4_ERR_ARG, "ALLEquiList(%d,%d)=%d out of range", kk + 1, ll + 1, q);
      if (seen[q - 1]) return bfail(FP4_ERR_ARG, "q-point %d appears in two stars", q);
      seen[q - 1] = 1;
      if (TypeofSymmetry) {
        const int sy = TypeofSymmetry[at];
        if (sy < 1 || sy > ns) return bfail(FP4_ERR_ARG, "TypeofSymmetry(%d,%d)=%d out of range", kk + 1, ll + 1, sy);
        const int *row = s->h_ideq.data();
        for (int q0 = 0; q0 < m.nptk; q0++)
          if (row[(size_t)(sy - 1) + (size_t)ns * q0] < 1)
            return bfail(FP4_ERR_ARG, "symmetry operation %d used by the star of point %d leaves the grid", sy,
                         ll + 1);
      }
    }
  }
  const size_t na = (size_t)ns * m.nptk;
  s->h_nequi.assign(Nequi, Nequi + m.nlist);
  s->h_alleq.assign(ALLEquiList, ALLEquiList + na);
  if (TypeofSymmetry) s->h_type.assign(TypeofSymmetry, TypeofSymmetry + na);
  int rc;
  if ((rc = s->nequi.ensure(sizeof(int) * m.nlist)) || (rc = s->alleq.ensure(sizeof(int) * na))) return rc;
  BCHECK(hipMemcpy(s->nequi.p, Nequi, sizeof(int) * m.nlist, hipMemcpyHostToDevice));
  BCHECK(hipMemcpy(s->alleq.p, ALLEquiList, sizeof(int) * na, hipMemcpyHostToDevice));
  s->have_orbits = true;
  s->have_types = TypeofSymmetry != nullptr;
  return 0;
}

template <typename T>
static int upload_list(BteState *s, const BteModes &m, const T *src, long long n, DBuf &dst) {
  int rc;
  if ((rc = dst.ensure(sizeof(int) * std::max(n, 1LL)))) return rc;
  if (n == 0) return 0;
  if (!src) return bfail(FP4_ERR_ARG, "null list");
  if ((rc = s->stage.ensure(sizeof(T) * std::min(n, kStageEntries)))) return rc;
  if ((rc = s->flag.ensure(sizeof(int)))) return rc;
  BCHECK(hipMemsetAsync(s->flag.p, 0, sizeof(int), m.stream));
  const long long maxv = (long long)m.nptk * m.nb;
  for (long long p0 = 0; p0 < n; p0 += kStageEntries) {
    const long long c = std::min(kStageEntries, n - p0);
    BCHECK(hipMemcpyAsync(s->stage.p, src + p0, sizeof(T) * c, hipMemcpyHostToDevice, m.stream));
    hipLaunchKernelGGL(k_narrow<T>, dim3((unsigned)((c + 255) / 256)), dim3(256), 0, m.stream, s->stage.as<T>(),
                       dst.as<int>() + p0, c, maxv, s->flag.as<int>());
    BCHECK(hipGetLastError());
  }
  int bad = 0;
  BCHECK(hipMemcpyAsync(&bad, s->flag.p, sizeof(int), hipMemcpyDeviceToHost, m.stream));
  BCHECK(hipStreamSynchronize(m.stream));
  if (bad) return bfail(FP4_ERR_ARG, "list entry outside 1..nptk*Nbands");
  return 0;
}

int bte_set_operator(BteState *s, const BteModes &m, const int32_t *const N3[2], const int32_t *const L3[2][2],
                     const double *const G3[2], const int64_t *const N4[3], const int64_t *const L4[3][3],
                     const double *const G4[3]) {
  check_dims(s, m);
  if ((long long)m.nptk * m.nb >= (1LL << 31)) return bfail(FP4_ERR_UNSUPPORTED, "nptk*Nbands must be < 2^31");
  const int nm = m.nb * m.nlist;
  s->have_op = false;
  s->use_w = false;
  for (int t = 0; t < 5; t++) {
    s->h_n[t].assign(nm, 0);
    s->h_off[t].assign(nm + 1, 0);
    for (int mm = 0; mm < nm; mm++) {
      long long v = 0;
      if (t < 2 && N3[t]) v = N3[t][mm];
      if (t >= 2 && N4[t - 2]) v = N4[t - 2][mm];
      if (v < 0) return bfail(FP4_ERR_ARG, "negative count for mode %d", mm + 1);
      s->h_n[t][mm] = v;
      s->h_off[t][mm + 1] = s->h_off[t][mm] + v;
    }
  }
  int rc;
  for (int t = 0; t < 5; t++) {
    const long long n = s->h_off[t][nm];
    if (t < 2) {
      if ((rc = upload_list(s, m, L3[t][0], n, s->e2[t])) || (rc = upload_list(s, m, L3[t][1], n, s->e3[t])))
        return rc;
    } else {
      if ((rc = upload_list(s, m, L4[t - 2][0], n, s->e2[t])) || (rc = upload_list(s, m, L4[t - 2][1], n, s->e3[t])) ||
          (rc = upload_list(s, m, L4[t - 2][2], n, s->e4[t - 2])))
        return rc;
    }
    const double *gsrc = t < 2 ? G3[t] : G4[t - 2];
    if ((rc = s->g[t].ensure(sizeof(double) * std::max(n, 1LL)))) return rc;
    if (n > 0) {
      if (!gsrc) return bfail(FP4_ERR_ARG, "null Gamma array");
      BCHECK(hipMemcpyAsync(s->g[t].p, gsrc, sizeof(double) * n, hipMemcpyHostToDevice, m.stream));
    }
  }
  BCHECK(hipStreamSynchronize(m.stream));
  s->have_op = true;
  return 0;
}

static int ensure_F(BteState *s, const BteModes &m) {
  int rc;
  const size_t n = (size_t)m.nptk * m.nb * 3;
  if ((rc = s->F.ensure(sizeof(double) * n)) || (rc = s->stage.ensure(sizeof(double) * n))) return rc;
  return 0;
}

static int upload_tau(BteState *s, const BteModes &m, const double *tau_zero) {
  if (!tau_zero) return bfail(FP4_ERR_ARG, "null tau_zero");
  int rc;
  if ((rc = s->tau.ensure(sizeof(double) * m.nb * m.nlist))) return rc;
  BCHECK(hipMemcpyAsync(s->tau.p, tau_zero, sizeof(double) * m.nb * m.nlist, hipMemcpyHostToDevice, m.stream));
  return 0;
}

int bte_put_F(BteState *s, const BteModes &m, const double *F_n) {
  if (!F_n) return bfail(FP4_ERR_ARG, "null F_n");
  int rc;
  if ((rc = ensure_F(s, m))) return rc;
  const long long n = (long long)m.nptk * m.nb;
  BCHECK(hipMemcpyAsync(s->stage.p, F_n, sizeof(double) * 3 * n, hipMemcpyHostToDevice, m.stream));
  hipLaunchKernelGGL(k_F_from_fortran, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, m.stream, m.nb, m.nptk,
                     s->stage.as<double>(), s->F.as<double>());
  BCHECK(hipGetLastError());
  return 0;
}

int bte_get_F(BteState *s, const BteModes &m, double *F_n) {
  if (!F_n) return bfail(FP4_ERR_ARG, "null F_n");
  if (!s->F.p) return bfail(FP4_ERR_STATE, "no F_n on the device");
  int rc;
  if ((rc = ensure_F(s, m))) return rc;
  const long long n = (long long)m.nptk * m.nb;
  hipLaunchKernelGGL(k_F_to_fortran, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, m.stream, m.nb, m.nptk,
                     s->F.as<double>(), s->stage.as<double>());
  BCHECK(hipGetLastError());
  BCHECK(hipMemcpyAsync(F_n, s->stage.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, m.stream));
  BCHECK(hipStreamSynchronize(m.stream));
  return 0;
}

int bte_w_begin(BteState *s, const BteModes &m, double **W, long long *cols) {
  check_dims(s, m);
  if (!s->have_orbits) return bfail(FP4_ERR_STATE, "orbit data not set");
  s->use_w = false;
  const long long rows = (long long)m.nb * m.nlist, nc = (long long)m.nptk * m.nb;
  int rc;
  if ((rc = s->W.ensure(sizeof(double) * (size_t)rows * (size_t)nc))) return rc;
  BCHECK(hipMemsetAsync(s->W.p, 0, sizeof(double) * (size_t)rows * (size_t)nc, m.stream));
  s->wcols = nc;
  *W = s->W.as<double>();
  *cols = nc;
  return 0;
}

int bte_w_finish(BteState *s, const BteModes &m) {
  check_dims(s, m);
  if (!s->W.p || s->wcols != (long long)m.nptk * m.nb) return bfail(FP4_ERR_STATE, "operator not assembled");
  BCHECK(hipStreamSynchronize(m.stream));
  s->use_w = true;
  return 0;
}

int bte_iteration0(BteState *s, const BteModes &m, const double *tau_zero) {
  check_dims(s, m);
  if (!s->have_orbits) return bfail(FP4_ERR_STATE, "orbit data not set");
  int rc;
  if ((rc = ensure_F(s, m)) || (rc = upload_tau(s, m, tau_zero))) return rc;
  BCHECK(hipMemsetAsync(s->F.p, 0, sizeof(double) * 3 * m.nptk * m.nb, m.stream));
  hipLaunchKernelGGL(k_iteration0, dim3(m.nlist), dim3(256), 0, m.stream, m.nb, m.nptk, s->nsymm, s->nequi.as<int>(),
                     s->alleq.as<int>(), m.omega, m.vel, s->tau.as<double>(), s->F.as<double>());
  BCHECK(hipGetLastError());
  return 0;
}

int bte_sweep(BteState *s, const BteModes &m, const double *tau_zero) {
  check_dims(s, m);
  if (!(s->have_op || s->use_w) || !s->have_orbits || !s->have_types || !s->have_symm)
    return bfail(FP4_ERR_STATE, "operator, orbits or symmetry data not set");
  if (!tau_zero) return bfail(FP4_ERR_ARG, "null tau_zero");
  if (!s->F.p) return bfail(FP4_ERR_STATE, "no F_n on the device");
  int rc;
  if ((rc = s->part.ensure(sizeof(double) * 3 * kMaxStepBlocks))) return rc;
  if (!s->ticket.p) {
    if ((rc = s->ticket.ensure(sizeof(unsigned int)))) return rc;
    BCHECK(hipMemsetAsync(s->ticket.p, 0, sizeof(unsigned int), m.stream));
  }
  IterStep st{};
  for (int t = 0; t < 5; t++) {
    st.e2[t] = s->e2[t].as<int>();
    st.e3[t] = s->e3[t].as<int>();
    st.g[t] = s->g[t].as<double>();
  }
  for (int t = 0; t < 3; t++) st.e4[t] = s->e4[t].as<int>();
  st.ideq = s->ideq.as<int>();
  st.symm = s->symm.as<double>();
  st.omega = m.omega;
  st.vel = m.vel;
  st.F = s->F.as<double>();
  st.part = s->part.as<double>();
  st.ticket = s->ticket.as<unsigned int>();
  st.nb = m.nb;
  st.nptk = m.nptk;
  st.nsymm = s->nsymm;
  st.W = s->use_w ? s->W.as<double>() : nullptr;
  st.wcols = s->wcols;
  long long launches = 0;
  if (s->use_w && s->persistent && m.nb <= kSweepMaxNb) {
    // the whole sweep in one cooperative launch (k_sweep_w)
    std::vector<int4> steps;
    std::vector<double> taus;
    steps.reserve((size_t)m.nptk * m.nb);
    for (int ll = 0; ll < m.nlist; ll++)
      for (int i = 0; i < m.nb; i++)
        for (int kk = 0; kk < s->h_nequi[ll]; kk++) {
          const size_t at = (size_t)kk + (size_t)s->nsymm * ll;
          steps.push_back(make_int4(ll * m.nb + i, i, s->h_alleq[at] - 1, s->h_type[at] - 1));
          taus.push_back(tau_zero[i + (size_t)m.nb * ll]);
        }
    const long long want = (s->wcols + kStepThreads * kPerThread - 1) / (kStepThreads * kPerThread);
    const int G = (int)std::min<long long>(std::max<long long>(want, 1), kMaxStepBlocks);
    int dev = 0, ncu = 0, per_cu = 0;
    BCHECK(hipGetDevice(&dev));
    BCHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    BCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(&k_sweep_w),
                                                        kStepThreads, 0));
    if ((long long)per_cu * ncu < G) return bfail(FP4_ERR_DEVICE, "sweep: %d blocks cannot be co-resident", G);
    if ((rc = s->steps.ensure(sizeof(int4) * steps.size())) || (rc = s->taus.ensure(sizeof(double) * taus.size())) ||
        (rc = s->bar.ensure(2 * sizeof(unsigned int))) || (rc = s->part.ensure(sizeof(double) * 6 * (size_t)G)))
      return rc;
    BCHECK(hipMemcpyAsync(s->steps.p, steps.data(), sizeof(int4) * steps.size(), hipMemcpyHostToDevice, m.stream));
    BCHECK(hipMemcpyAsync(s->taus.p, taus.data(), sizeof(double) * taus.size(), hipMemcpyHostToDevice, m.stream));
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, m.stream, s->bar.as<unsigned int>(), 2);
    SweepW sw{};
    sw.steps = s->steps.as<int4>();
    sw.tau = s->taus.as<double>();
    sw.nsteps = (int)steps.size();
    sw.W = s->W.as<double>();
    sw.wcols = s->wcols;
    sw.ideq = s->ideq.as<int>();
    sw.symm = s->symm.as<double>();
    sw.omega = m.omega;
    sw.vel = m.vel;
    sw.F = s->F.as<double>();
    sw.part = s->part.as<double>();
    sw.bar = s->bar.as<unsigned int>();
    sw.err = reinterpret_cast<int *>(s->bar.as<unsigned int>() + 1);
    sw.nb = m.nb;
    sw.nptk = m.nptk;
    sw.nsymm = s->nsymm;
    sw.spin_limit = 1LL << 24;
    BCHECK(hipEventRecord(s->ev0, m.stream));
    void *args[] = {&sw};
    BCHECK(hipLaunchCooperativeKernel(reinterpret_cast<const void *>(&k_sweep_w), dim3(G), dim3(kStepThreads), args,
                                      0, m.stream));
    BCHECK(hipEventRecord(s->ev1, m.stream));
    int herr = 0;
    BCHECK(hipMemcpyAsync(&herr, sw.err, sizeof(int), hipMemcpyDeviceToHost, m.stream));
    BCHECK(hipStreamSynchronize(m.stream));
    if (herr) return bfail(FP4_ERR_DEVICE, "sweep: a grid barrier wait timed out");
    float ms = 0.f;
    BCHECK(hipEventElapsedTime(&ms, s->ev0, s->ev1));
    s->last_ms = ms;
    s->last_launches = 1;
    return 0;
  }
  BCHECK(hipEventRecord(s->ev0, m.stream));
  for (int ll = 0; ll < m.nlist; ll++)
    for (int i = 0; i < m.nb; i++) {
      const int mm = ll * m.nb + i;
      long long tot = 0;
      if (s->use_w) {
        st.wrow = mm;
        tot = s->wcols;
      } else {
        for (int t = 0; t < 5; t++) {
          st.off[t] = s->h_off[t][mm];
          st.n[t] = s->h_n[t][mm];
          tot += st.n[t];
        }
      }
      st.i = i;
      st.tau = tau_zero[i + (size_t)m.nb * ll];
      const long long want = (tot + kStepThreads * kPerThread - 1) / (kStepThreads * kPerThread);
      const int nblocks = (int)std::min<long long>(std::max<long long>(want, 1), kMaxStepBlocks);
      for (int kk = 0; kk < s->h_nequi[ll]; kk++) {
        const size_t at = (size_t)kk + (size_t)s->nsymm * ll;
        st.qk = s->h_alleq[at] - 1;
        st.sym = s->h_type[at] - 1;
        hipLaunchKernelGGL(k_iter_step, dim3(nblocks), dim3(kStepThreads), 0, m.stream, st);
        launches++;
      }
    }
  BCHECK(hipGetLastError());
  BCHECK(hipEventRecord(s->ev1, m.stream));
  BCHECK(hipEventSynchronize(s->ev1));
  float ms = 0.f;
  BCHECK(hipEventElapsedTime(&ms, s->ev0, s->ev1));
  s->last_ms = ms;
  s->last_launches = launches;
  return 0;
}

int bte_set_persistent(BteState *s, bool on) {
  if (!s) return bfail(FP4_ERR_STATE, "no iteration state");
  s->persistent = on;
  return 0;
}

int bte_symmetrize_all(BteState *s, const BteModes &m) {
  if (!s->have_symm) return bfail(FP4_ERR_STATE, "symmetry data not set");
  const long long n = (long long)m.nptk * m.nb;
  hipLaunchKernelGGL(k_symmetrize_all, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, m.stream, m.nb, m.nptk,
                     s->symm.as<double>(), s->F.as<double>());
  BCHECK(hipGetLastError());
  return 0;
}

static int launch_tconduct(BteState *s, const BteModes &m, double T, double V, bool with_mode) {
  int rc;
  if ((rc = s->kappa.ensure(sizeof(double) * 9 * m.nb))) return rc;
  if (with_mode && (rc = s->mode.ensure(sizeof(double) * 9 * (size_t)m.nb * m.nptk))) return rc;
  hipLaunchKernelGGL(k_tconduct, dim3(m.nb), dim3(256), 0, m.stream, m.nb, m.nptk, T, V, m.omega, m.vel,
                     s->F.as<double>(), s->kappa.as<double>(), with_mode ? s->mode.as<double>() : nullptr);
  BCHECK(hipGetLastError());
  return 0;
}

int bte_tconduct(BteState *s, const BteModes &m, double T, double V, double *kappa, double *mode) {
  if (!kappa) return bfail(FP4_ERR_ARG, "null kappa");
  if (!(T > 0) || !(V > 0)) return bfail(FP4_ERR_ARG, "T and V must be positive");
  if (!s->F.p) return bfail(FP4_ERR_STATE, "no F_n on the device");
  int rc;
  if ((rc = launch_tconduct(s, m, T, V, mode != nullptr))) return rc;
  BCHECK(hipMemcpyAsync(kappa, s->kappa.p, sizeof(double) * 9 * m.nb, hipMemcpyDeviceToHost, m.stream));
  if (mode)
    BCHECK(hipMemcpyAsync(mode, s->mode.p, sizeof(double) * 9 * (size_t)m.nb * m.nptk, hipMemcpyDeviceToHost,
                          m.stream));
  BCHECK(hipStreamSynchronize(m.stream));
  return 0;
}

int bte_solve(BteState *s, const BteModes &m, double T, double V, const double *tau_zero, int maxiter, double eps,
              double *kappa, double *kappa_iter, double *relchange, int *niter) {
  if (maxiter < 1) return bfail(FP4_ERR_ARG, "maxiter must be > 0");
  if (!kappa || !niter) return bfail(FP4_ERR_ARG, "null output");
  if (!(T > 0) || !(V > 0)) return bfail(FP4_ERR_ARG, "T and V must be positive");
  int rc;
  if ((rc = s->hist.ensure(sizeof(double) * 9 * (size_t)(maxiter + 1))) ||
      (rc = s->rel.ensure(sizeof(double) * (size_t)(maxiter + 1))))
    return rc;
  if ((rc = bte_iteration0(s, m, tau_zero))) return rc;
  if ((rc = launch_tconduct(s, m, T, V, false))) return rc;
  hipLaunchKernelGGL(k_kappa_finish, dim3(1), dim3(64), 0, m.stream, m.nb, s->nsymm, s->crot.as<double>(),
                     s->kappa.as<double>(), s->hist.as<double>(), 0, s->rel.as<double>());
  BCHECK(hipGetLastError());
  double total_ms = 0.0;
  long long total_launches = 0;
  int it = 0;
  for (it = 1; it <= maxiter; it++) {
    if ((rc = bte_sweep(s, m, tau_zero))) return rc;
    total_ms += s->last_ms;
    total_launches += s->last_launches;
    if ((rc = bte_symmetrize_all(s, m)) || (rc = launch_tconduct(s, m, T, V, false))) return rc;
    hipLaunchKernelGGL(k_kappa_finish, dim3(1), dim3(64), 0, m.stream, m.nb, s->nsymm, s->crot.as<double>(),
                       s->kappa.as<double>(), s->hist.as<double>(), it, s->rel.as<double>());
    BCHECK(hipGetLastError());
    double r = 0.0;
    BCHECK(hipMemcpyAsync(&r, s->rel.as<double>() + it, sizeof(double), hipMemcpyDeviceToHost, m.stream));
    BCHECK(hipStreamSynchronize(m.stream));
    if (r < eps) break;
  }
  const int done = std::min(it, maxiter);
  *niter = done;
  BCHECK(hipMemcpyAsync(kappa, s->kappa.p, sizeof(double) * 9 * m.nb, hipMemcpyDeviceToHost, m.stream));
  if (kappa_iter)
    BCHECK(hipMemcpyAsync(kappa_iter, s->hist.p, sizeof(double) * 9 * (size_t)(done + 1), hipMemcpyDeviceToHost,
                          m.stream));
  if (relchange)
    BCHECK(hipMemcpyAsync(relchange, s->rel.p, sizeof(double) * (size_t)(done + 1), hipMemcpyDeviceToHost, m.stream));
  BCHECK(hipStreamSynchronize(m.stream));
  s->last_ms = total_ms;
  s->last_launches = total_launches;
  return 0;
}

}  // namespace fp4
