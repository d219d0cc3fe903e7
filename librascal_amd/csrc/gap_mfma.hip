#include "gap_mfma.hh"

namespace rgpu {

bool mfma_available() { return false; }

void prepare_mixed_sparse(MixedSparse& ms, const std::vector<double>&, int64_t M, int64_t D, hipStream_t) {
  ms.M = M;
  ms.D = D;
}

void run_gap_mixed(const GapBlock<float>&, const MixedSparse&, double*, float*, float*, hipStream_t) {
  throw RuntimeError("mixed-precision matrix-core GAP kernel not built");
}

}  // namespace rgpu
