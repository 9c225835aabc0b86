// Placeholder until the split-precision kernels land.
#include "psmma_common.h"
namespace psmma {
int launch_right_split(int, int, int, int, int, int, double, const double*, int, const GenPtrs&,
                       int, int, double, double*, int, hipStream_t) {
  return -15;
}
}  // namespace psmma
