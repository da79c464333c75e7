// Harmonic front end on the device: eigenDM for Quantum Espresso and phonopy
// 2nd-order constants (Src/phonon_routines.f90:12-307, 311-677).
//
// One 64-wide wavefront per grid point q:
//   1. D(q) and dD/dq from the weighted supercell images (the image list
//      with its weights is built on the host: Wigner-Seitz weights for
//      espresso, 1/neq over the nearest images for phonopy):
//      D_ij += Phi_ij(image) w exp(-i k.t), dD_ij += -i t Phi_ij w exp(-i k.t),
//      both divided by sqrt(m_i m_j), held in LDS; with phonopy's dielectric
//      correction, its dyn_nac / ncell enters every image's constants alike,
//      so D_ij gets S dyn_nac_ij / ncell and dD_ij gets
//      (-i Zr dyn_nac_ij + S ddyn_nac_ij) / ncell, with S = sum w exp(-i k.t)
//      and Zr = sum w t exp(-i k.t) over the pair's images;
//   2. the Hermitian eigenproblem by cyclic complex Jacobi rotations in FP64
//      (one rotation at a time, the row and column updates spread over the
//      lanes), eigenvalues sorted ascending;
//   3. omega = sign(sqrt|w2|, w2), v_b = Re(e_b^H dD e_b) / (2 omega_b), the
//      reference's unit factors, the q-point symmetriser applied to every
//      branch's velocity, and exactly zero acoustic modes at Gamma;
//   4. outputs written in the reference's layouts omega(nptk,Nb),
//      eigenvect(nptk,Nb,Nb) (component last), velocity(nptk,Nb,3).
// Eigenvectors of degenerate modes are one orthonormal basis of their
// subspace (any Hermitian solver picks its own).
#include "fp4_device.hpp"
#include "fp4_kernels.hpp"

namespace fp4 {

__device__ __forceinline__ double2 cm(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }

template <int NB>
__global__ void __launch_bounds__(64) k_eigen_dm(EigenDMArgs a) {
  constexpr int NN = NB * NB;
  __shared__ double2 A[NN];
  __shared__ double2 V[NN];
  __shared__ double2 dA[3 * NN];
  __shared__ double w2[NB];
  __shared__ int perm[NB];
  __shared__ double red[64];
  __shared__ double vraw[3 * NB];
  const int lane = threadIdx.x;
  const int q0 = blockIdx.x;
  const int i1 = q0 % a.n1, i2 = (q0 / a.n1) % a.n2, i3 = q0 / (a.n1 * a.n2);
  // kspace of eigenDM, rlattvec(:,1)*(ii-1.0)/ngrid(1) + ...: evaluated left
  // to right, so each term is (b * (i-1)) / N in double (Src/phonon_routines.f90:29-31)
  double k[3];
#pragma unroll
  for (int c = 0; c < 3; c++)
    k[c] = (((a.b[c] * (double)i1) / (double)a.n1 + (a.b[3 + c] * (double)i2) / (double)a.n2) +
            (a.b[6 + c] * (double)i3) / (double)a.n3) * a.bohr2nm;

  // 1. D and dD
  for (int e = lane; e < NN; e += 64) {
    const int i = e / NB, j = e % NB;
    const int iat = i / 3, ia = i % 3, jat = j / 3, ja = j % 3;
    const int pair = iat * a.nat + jat;
    double2 d = make_double2(0.0, 0.0), g0 = d, g1 = d, g2 = d;
    double2 S = d, Z0 = d, Z1 = d, Z2 = d;
    for (int m = a.img_start[pair]; m < a.img_start[pair + 1]; m++) {
      const double *t = a.img_t + 3 * (size_t)m;
      const double kt = k[0] * t[0] + k[1] * t[1] + k[2] * t[2];
      double sn, cs;
      sincos(kt, &sn, &cs);
      const double f = a.img_fc[9 * (size_t)m + 3 * ia + ja];
      // f * exp(-i kt) and -i t f exp(-i kt)
      d.x += f * cs;
      d.y -= f * sn;
      g0.x -= t[0] * f * sn; g0.y -= t[0] * f * cs;
      g1.x -= t[1] * f * sn; g1.y -= t[1] * f * cs;
      g2.x -= t[2] * f * sn; g2.y -= t[2] * f * cs;
      if (a.nac_q) {
        // w exp(-i kt) and w t exp(-i kt)
        const double w = a.img_w[m];
        S.x += w * cs; S.y -= w * sn;
        Z0.x += w * t[0] * cs; Z0.y -= w * t[0] * sn;
        Z1.x += w * t[1] * cs; Z1.y -= w * t[1] * sn;
        Z2.x += w * t[2] * cs; Z2.y -= w * t[2] * sn;
      }
    }
    const double mm = a.mm[pair];
    d = make_double2(d.x / mm, d.y / mm);
    g0 = make_double2(g0.x / mm, g0.y / mm);
    g1 = make_double2(g1.x / mm, g1.y / mm);
    g2 = make_double2(g2.x / mm, g2.y / mm);
    if (a.nac_q) {
      const double *qs = a.nac_q + 3 * (size_t)q0;
      if (qs[0] != 0.0 || qs[1] != 0.0 || qs[2] != 0.0) {
        // dyn_nac and ddyn_nac of phonon_phonopy (Src/phonon_routines.f90:168-197)
        double eq[3], t1 = 0.0, t2 = 0.0;
#pragma unroll
        for (int x = 0; x < 3; x++) {
          eq[x] = a.eps[x] * qs[0] + a.eps[x + 3] * qs[1] + a.eps[x + 6] * qs[2];
          t1 += qs[x] * a.born[x + 3 * (ia + 3 * iat)];
          t2 += qs[x] * a.born[x + 3 * (ja + 3 * jat)];
        }
        const double tmp3 = qs[0] * eq[0] + qs[1] * eq[1] + qs[2] * eq[2];
        const double nm = a.nac_mm[pair];
        const double dn = t1 * t2 / nm * a.nac_scale / tmp3;
        double dd[3];
#pragma unroll
        for (int x = 0; x < 3; x++)
          dd[x] = (t1 * a.born[x + 3 * (ja + 3 * jat)] + t2 * a.born[x + 3 * (ia + 3 * iat)] -
                   2.0 * t1 * t2 * eq[x] / tmp3) / nm * a.nac_scale / tmp3;
        d.x += S.x * dn;
        d.y += S.y * dn;
        // -i Zr dn + S dd
        g0.x += Z0.y * dn + S.x * dd[0]; g0.y += -Z0.x * dn + S.y * dd[0];
        g1.x += Z1.y * dn + S.x * dd[1]; g1.y += -Z1.x * dn + S.y * dd[1];
        g2.x += Z2.y * dn + S.x * dd[2]; g2.y += -Z2.x * dn + S.y * dd[2];
      }
    }
    A[e] = d;
    dA[e] = g0;
    dA[NN + e] = g1;
    dA[2 * NN + e] = g2;
    V[e] = make_double2(i == j ? 1.0 : 0.0, 0.0);
  }
  __syncthreads();

  // 2. cyclic Jacobi
  for (int sweep = 0; sweep < 40; sweep++) {
    double off = 0.0, dia = 0.0;
    for (int e = lane; e < NN; e += 64) {
      const int i = e / NB, j = e % NB;
      const double v = A[e].x * A[e].x + A[e].y * A[e].y;
      if (i < j) off += v;
      if (i == j) dia += v;
    }
    off = wave_sum(off);
    dia = wave_sum(dia);
    if (!(off > 1e-34 * dia)) break;
    for (int p = 0; p < NB - 1; p++)
      for (int q = p + 1; q < NB; q++) {
        const double2 apq = A[p * NB + q];
        const double r = hypot(apq.x, apq.y);
        if (!(r > 0.0)) continue;
        const double app = A[p * NB + p].x, aqq = A[q * NB + q].x;
        const double theta = (aqq - app) / (2.0 * r);
        const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        const double2 ph = make_double2(apq.x / r, apq.y / r);   // e^{i phi}
        const double2 phc = cconj(ph);
        __syncthreads();
        if (lane < NB) {
          // columns: A <- A J, V <- V J
          const double2 x = A[lane * NB + p], y = cm(phc, A[lane * NB + q]);
          A[lane * NB + p] = make_double2(c * x.x - s * y.x, c * x.y - s * y.y);
          A[lane * NB + q] = make_double2(s * x.x + c * y.x, s * x.y + c * y.y);
          const double2 vx = V[lane * NB + p], vy = cm(phc, V[lane * NB + q]);
          V[lane * NB + p] = make_double2(c * vx.x - s * vy.x, c * vx.y - s * vy.y);
          V[lane * NB + q] = make_double2(s * vx.x + c * vy.x, s * vx.y + c * vy.y);
        }
        __syncthreads();
        if (lane < NB) {
          // rows: A <- J^H A
          const double2 x = A[p * NB + lane], y = cm(ph, A[q * NB + lane]);
          A[p * NB + lane] = make_double2(c * x.x - s * y.x, c * x.y - s * y.y);
          A[q * NB + lane] = make_double2(s * x.x + c * y.x, s * x.y + c * y.y);
        }
        __syncthreads();
        if (lane == 0) {
          A[p * NB + q] = make_double2(0.0, 0.0);
          A[q * NB + p] = make_double2(0.0, 0.0);
          A[p * NB + p] = make_double2(app - t * r, 0.0);
          A[q * NB + q] = make_double2(aqq + t * r, 0.0);
        }
        __syncthreads();
      }
  }

  // 3. sort ascending, frequencies, velocities
  if (lane == 0) {
    for (int i = 0; i < NB; i++) {
      w2[i] = A[i * NB + i].x;
      perm[i] = i;
    }
    for (int i = 0; i < NB; i++) {
      int m = i;
      for (int j = i + 1; j < NB; j++)
        if (w2[perm[j]] < w2[perm[m]]) m = j;
      const int tmp = perm[i];
      perm[i] = perm[m];
      perm[m] = tmp;
    }
  }
  __syncthreads();
  const size_t np = (size_t)a.nptk;
  for (int b = lane; b < NB; b += 64) {
    const double x = w2[perm[b]];
    const double om = (x >= 0.0 ? 1.0 : -1.0) * sqrt(fabs(x));
    red[b] = om;
    double out = om * a.tothz;
    if (q0 == 0 && b < 3) out = 0.0;
    a.omega[(size_t)b * np + q0] = out;
    for (int c = 0; c < NB; c++) a.eig[((size_t)c * NB + b) * np + q0] = V[c * NB + perm[b]];
  }
  __syncthreads();
  for (int e = lane; e < 3 * NB; e += 64) {
    const int b = e / 3, x = e % 3;
    const int col = perm[b];
    double acc = 0.0;
    for (int i = 0; i < NB; i++) {
      double2 s = make_double2(0.0, 0.0);
      for (int j = 0; j < NB; j++) {
        const double2 t = cm(dA[x * NN + i * NB + j], V[j * NB + col]);
        s.x += t.x;
        s.y += t.y;
      }
      const double2 vi = V[i * NB + col];
      acc += vi.x * s.x + vi.y * s.y;   // Re(conj(v_i) s)
    }
    vraw[e] = acc / (2.0 * red[b]);
  }
  __syncthreads();
  for (int e = lane; e < 3 * NB; e += 64) {
    const int b = e / 3, x = e % 3;
    double v;
    if (a.symm) {
      const double *S = a.symm + 9 * (size_t)q0;
      v = S[x] * vraw[3 * b] + S[x + 3] * vraw[3 * b + 1] + S[x + 6] * vraw[3 * b + 2];
    } else {
      v = vraw[e];
    }
    v = v * a.tothz * a.bohr2nm;
    if (q0 == 0 && b < 3) v = 0.0;
    a.vel[((size_t)x * NB + b) * np + q0] = v;
  }
}

int launch_eigen_dm(const EigenDMArgs &a, hipStream_t st) {
  const dim3 g((unsigned)a.nptk), b(64);
  switch (a.nb) {
    case 3: hipLaunchKernelGGL((k_eigen_dm<3>), g, b, 0, st, a); break;
    case 6: hipLaunchKernelGGL((k_eigen_dm<6>), g, b, 0, st, a); break;
    case 9: hipLaunchKernelGGL((k_eigen_dm<9>), g, b, 0, st, a); break;
    case 12: hipLaunchKernelGGL((k_eigen_dm<12>), g, b, 0, st, a); break;
    case 18: hipLaunchKernelGGL((k_eigen_dm<18>), g, b, 0, st, a); break;
    case 24: hipLaunchKernelGGL((k_eigen_dm<24>), g, b, 0, st, a); break;
    default: return -1;
  }
  return 0;
}

}  // namespace fp4
