// kernels_br_u64.hip -- the blind-rotation launcher for uint64_t device words.
#include "kernels_impl.hpp"

namespace fheacc {
template hipError_t launch_blind_rotate<uint64_t>(const BRArgs<uint64_t> &, int, int, int64_t, size_t, hipStream_t,
                                                  int, int, int64_t *);
}  // namespace fheacc
