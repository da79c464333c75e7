// Device-side data layout and exact FP64 helpers shared by the fp4 kernels.
//
// HBM layout (all device arrays are owned by fp4_ctx):
//   energy   double[Nb][nptk]          (the reference's energy(nptk,Nb))
//   vel      double[3][Nb][nptk]       (velocity(nptk,Nb,3))
//   vbound   float [Nb][nptk]          conservative upper bound of the
//                                      broadening a mode can contribute
//   eig      double2[nptk][Nb][Nb]     e(q, branch, component), component
//                                      fastest (the transpose of the
//                                      reference's eigenvect(nptk,Nb,Nb), so
//                                      one mode's vector is 16*Nb contiguous B)
//   ijk      int[nptk][3]
//   clusters ClusterRec[Ntri] + double phi[Ntri][81]
//
// The window test |delta| <= sigma decides membership in the process list,
// so it is evaluated with the reference's operation order, no contraction,
// IEEE division and a correctly rounded square root (config.f90:481-499,
// processes.f90:2312-2314).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace fp4 {

constexpr double kPi = 3.141592653589793238;
constexpr double kKb = 1.380648813e-23;
constexpr double kHbar = 1.05457172647e-22;
// unsuffixed (default REAL, float32) literals of the reference
constexpr float kClamp = 0.001f;           // processes.f90:2315
constexpr float kGammaUnits = 3.37617087f; // processes.f90:3922
constexpr float k1e22 = 1e22f;             // processes.f90:20
constexpr float kGamma3Units = 5.60626442f; // processes.f90:433, 512, 1389, 1469

struct ClusterRec {
  int atom[4];     // 0-based atoms of phonons 1..4
  int pj[3];       // phase multipliers: q.R_j = 2 pi (sum_nu n_nu pj[nu]) / L
  int pk[3];
  int pl[3];
  int pad;
  double minv;     // 1/sqrt(m_i m_j m_k m_l), reference expression order
};

// Third-order cluster: atoms of phonons 1..3, phase multipliers of R_j, R_k
// (q.R = 2 pi (sum_nu n_nu p[nu]) / L) and 1/sqrt(m_i m_j m_k).
struct Cluster3Rec {
  int atom[3];
  int pj[3];
  int pk[3];
  int pad;
  double minv;
};

struct ModeView {
  const double *energy;
  const double *vel;
  const float *vbound;
  const double2 *eig;
  const int *ijk;
  const int *list;
  int nptk, nb, nlist;
  int n1, n2, n3;
  double b[9];     // rlattvec(3,3) column-major
  double scalebroad;
  double omega_max;
  float bf[9];     // rlattvec(:,nu)/ngrid(nu) in float, for the sigma estimate
  float sbf;       // |scalebroad| in float
  double inv_n[3]; // 1/ngrid(nu) when ngrid(nu) is a power of two, else 0
};

struct IfcView {
  const ClusterRec *cl;
  const double *phi;     // [Ntri][81], (u,t,s,r) with u fastest
  const double2 *roots;  // [L] e^{2 pi i k / L}
  int ntri;
  int L;
};

struct Ifc3View {
  const Cluster3Rec *cl;
  const double *phi;     // [Ntri][27], (t,s,r) with t fastest
  const double2 *roots;  // [L] e^{2 pi i k / L}
  int ntri;
  int L;
};

__device__ __forceinline__ double E(const ModeView &m, int q0, int b0) { return m.energy[(size_t)b0 * m.nptk + q0]; }
__device__ __forceinline__ double Vel(const ModeView &m, int q0, int b0, int a) {
  return m.vel[((size_t)a * m.nb + b0) * m.nptk + q0];
}

// Correctly rounded sqrt: hardware estimate, then the Tuckerman test
// y*pred(y) < x <= y*succ(y) decided exactly with fma.
__device__ __forceinline__ double sqrt_rn(double x) {
  double y = __builtin_sqrt(x);
  if (!(x > 0.0) || !__builtin_isfinite(x)) return y;
  const double lo = __longlong_as_double(__double_as_longlong(y) - 1);
  const double hi = __longlong_as_double(__double_as_longlong(y) + 1);
  if (__builtin_fma(y, lo, -x) >= 0.0) return lo;
  if (__builtin_fma(y, hi, -x) < 0.0) return hi;
  return y;
}

// base_sigma(v), config.f90:481-499
__device__ __forceinline__ double base_sigma(const ModeView &m, double v0, double v1, double v2) {
  double s[3];
  const int ng[3] = {m.n1, m.n2, m.n3};
#pragma unroll
  for (int nu = 0; nu < 3; nu++) {
    double d = 0.0;
    d = d + m.b[3 * nu + 0] * v0;
    d = d + m.b[3 * nu + 1] * v1;
    d = d + m.b[3 * nu + 2] * v2;
    // d / N; for N = 2^k the product with 2^-k rounds the same exact value
    const double t = (m.inv_n[nu] != 0.0) ? d * m.inv_n[nu] : d / (double)ng[nu];
    s[nu] = t * t;
  }
  double mx = s[0];
  if (s[1] > mx) mx = s[1];
  if (s[2] > mx) mx = s[2];
  return sqrt_rn(mx / 2.0);
}

// sigma of the (jj,k) / (ss,l) pair: ++ takes -v3+v4, +- and -- take v3-v4
__device__ __forceinline__ double pair_sigma(const ModeView &m, int ch, int jj0, int k0, int ss0, int l0) {
  double d[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const double v3 = Vel(m, jj0, k0, a), v4 = Vel(m, ss0, l0, a);
    d[a] = (ch == 0) ? (-v3 + v4) : (v3 - v4);
  }
  return m.scalebroad * base_sigma(m, d[0], d[1], d[2]);
}

// Float estimate of sigma used to cull candidates: velocity differences
// formed in FP64 and rounded once, then float dot products.  Its relative
// error is below 1e-6, so any candidate whose |delta| lies outside
// [0.9999, 1.0001] * estimate is decided by the estimate; the others take the
// exact FP64 test.
__device__ __forceinline__ float sigma_estimate(const ModeView &m, int ch, const double *v3, const double *v4) {
  float d[3];
#pragma unroll
  for (int a = 0; a < 3; a++) d[a] = (float)((ch == 0) ? (-v3[a] + v4[a]) : (v3[a] - v4[a]));
  float mx = 0.f;
#pragma unroll
  for (int nu = 0; nu < 3; nu++) {
    const float x = m.bf[3 * nu] * d[0] + m.bf[3 * nu + 1] * d[1] + m.bf[3 * nu + 2] * d[2];
    mx = fmaxf(mx, x * x);
  }
  return m.sbf * sqrtf(0.5f * mx);
}

__device__ __forceinline__ double mismatch(int ch, double w, double wp, double wdp, double wtp) {
  if (ch == 0) return w + wp + wdp - wtp;
  if (ch == 1) return w + wp - wdp - wtp;
  return w - wp - wdp - wtp;
}

__device__ __forceinline__ int fmod_pos(int a, int n) {
  int r = a % n;
  return r < 0 ? r + n : r;
}

// Third-partner grid index (0-based) for channel ch, processes.f90:2297-2298
__device__ __forceinline__ int partner(const ModeView &m, int ch, const int *q, const int *qp, const int *qdp) {
  int t0, t1, t2;
  if (ch == 0) {
    t0 = q[0] + qp[0] + qdp[0]; t1 = q[1] + qp[1] + qdp[1]; t2 = q[2] + qp[2] + qdp[2];
  } else if (ch == 1) {
    t0 = q[0] + qp[0] - qdp[0]; t1 = q[1] + qp[1] - qdp[1]; t2 = q[2] + qp[2] - qdp[2];
  } else {
    t0 = q[0] - qp[0] - qdp[0]; t1 = q[1] - qp[1] - qdp[1]; t2 = q[2] - qp[2] - qdp[2];
  }
  t0 = fmod_pos(t0, m.n1); t1 = fmod_pos(t1, m.n2); t2 = fmod_pos(t2, m.n3);
  return (t2 * m.n2 + t1) * m.n1 + t0;
}

// Whether the triple (q, q', q'') of a work item can have a surviving
// process, conservatively: some (k, l) with c = w''' -+ w'' within
// vbound''_k + vbound'''_l (each vbound is a rounded-up bound of that mode's
// share of sigma, sigma_kl <= vbound''_k + vbound'''_l) of the item's range
// [xlo, xhi] of x1 = w_i +- w'_j.  2^-40 (|x1|max + |w''| + |w'''|) covers
// the FP64 rounding of the exact test and of this one, so a triple this
// rejects has no survivor and the visited processes are unchanged.
template <int NB>
__device__ __forceinline__ bool triple_may_survive(const ModeView &m, int ch, int jj, int ss, double xlo, double xhi,
                                                   double xabs) {
  double od[NB], vk[NB];
#pragma unroll
  for (int k = 0; k < NB; k++) {
    const double e = E(m, jj, k);
    od[k] = (ch == 0) ? e : -e;
    vk[k] = (double)m.vbound[(size_t)k * m.nptk + jj];
  }
  bool any = false;
#pragma unroll 1
  for (int l = 0; l < NB; l++) {
    const double ot = E(m, ss, l);
    const double vl = (double)m.vbound[(size_t)l * m.nptk + ss];
#pragma unroll
    for (int k = 0; k < NB; k++) {
      const double c = ot - od[k];
      const double W = (vk[k] + vl) + 0x1p-40 * (xabs + fabs(od[k]) + fabs(ot));
      any = any || ((c + W >= xlo) && (c - W <= xhi));
    }
  }
  return any;
}

// The same test with rolled loops and an early exit, for kernels that run
// close to their register budget (the prefilter at three waves per SIMD).
template <int NB>
__device__ __forceinline__ bool triple_may_survive_lean(const ModeView &m, int ch, int jj, int ss, double xlo,
                                                        double xhi, double xabs) {
#pragma unroll 1
  for (int k = 0; k < NB; k++) {
    const double e = E(m, jj, k);
    const double od = (ch == 0) ? e : -e;
    const double vk = (double)m.vbound[(size_t)k * m.nptk + jj];
#pragma unroll 1
    for (int l = 0; l < NB; l++) {
      const double ot = E(m, ss, l);
      const double c = ot - od;
      const double W = (vk + (double)m.vbound[(size_t)l * m.nptk + ss]) + 0x1p-40 * (xabs + fabs(od) + fabs(ot));
      if ((c + W >= xlo) && (c - W <= xhi)) return true;
    }
  }
  return false;
}

__device__ __forceinline__ double bose(double w, double T) { return 1.0 / (exp(kHbar * w / kKb / T) - 1.0); }

// WP4 of one process (processes.f90:3895-3901, 3988-3994, 4095-4101), with
// the Bose-Einstein factors fp, fdp, ftp of phonons 2..4 given (bose() of
// their frequencies).
__device__ __forceinline__ double wp4_given(int ch, double w, double wp, double wdp, double wtp, double fp,
                                            double fdp, double ftp, double delta, double sigma) {
  double occ;
  if (ch == 0)
    occ = fp * fdp * (1 + ftp) - (1 + fp) * (1 + fdp) * ftp;
  else if (ch == 1)
    occ = fp * (1 + fdp) * (1 + ftp) - (1 + fp) * fdp * ftp;
  else
    occ = (1 + fp) * (1 + fdp) * (1 + ftp) - fp * fdp * ftp;
  double r = occ * exp(-(delta * delta) / (sigma * sigma)) / sigma / sqrt(kPi) / (w * wp * wdp * wtp);
  if (wp <= 1.25 || wdp <= 1.25 || wtp <= 1.25) r = 0.0;
  return r;
}

__device__ __forceinline__ double wp4_of(int ch, double w, double wp, double wdp, double wtp, double delta,
                                         double sigma, double T) {
  return wp4_given(ch, w, wp, wdp, wtp, bose(wp, T), bose(wdp, T), bose(wtp, T), delta, sigma);
}

// Counter-based random numbers of the sampling estimators: the SplitMix64
// output function (Steele, Lea & Flood, OOPSLA 2014) applied to
// key + kGolden*position, and an index below n as the high 64 bits of u*n.
constexpr unsigned long long kGolden = 0x9E3779B97F4A7C15ull;

__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ int below(unsigned long long u, int n) {
  return (int)__umul64hi(u, (unsigned long long)n);
}

// Deterministic wave-wide sums (fixed butterfly order).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ int wave_sum_i(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
// exclusive prefix sum over the 64 lanes
__device__ __forceinline__ int wave_excl_scan(int v, int lane) {
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  return x - v;
}

}  // namespace fp4
