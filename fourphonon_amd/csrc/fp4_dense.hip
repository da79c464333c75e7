#include "fp4_dense.hpp"
#include "../../include/fp4.h"

namespace fp4 {
int dense_prepare(const std::vector<ClusterRec> &, const std::vector<double> &, int, int) { return 0; }
int dense_rates(const DenseRun &, const std::vector<int> &, int, int, std::vector<double> &, std::vector<double> &,
                double *) {
  return FP4_ERR_UNSUPPORTED;
}
const char *dense_error() { return "dense form not built yet"; }
}  // namespace fp4
