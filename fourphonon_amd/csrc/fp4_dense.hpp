// Dense (fold + contraction) form of the four-phonon rates; see fp4_dense.hip.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "fp4_device.hpp"

namespace fp4 {

struct DenseTuning {
  int qblock = 0;   // q'' per tile; 0 = choose automatically
};

struct DenseRun {
  ModeView m;
  double T;
  int onlyharmonic;
  int natoms;
  int precision;
  hipStream_t stream;
  void **scratch;
  size_t *scratch_bytes;
  DenseTuning *tuning;
};

int dense_prepare(const std::vector<ClusterRec> &cl, const std::vector<double> &phi, int natoms, int nb);
int dense_rates(const DenseRun &r, const std::vector<int> &modes, int ii_lo0, int nii, std::vector<double> &g,
                std::vector<double> &w, double *ms);
const char *dense_error();

}  // namespace fp4
