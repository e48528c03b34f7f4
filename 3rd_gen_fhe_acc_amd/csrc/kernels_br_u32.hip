// kernels_br_u32.hip -- the blind-rotation launcher for uint32_t device words.
#include "kernels_impl.hpp"

namespace fheacc {
template hipError_t launch_blind_rotate<uint32_t>(const BRArgs<uint32_t> &, int, int, int64_t, size_t, hipStream_t,
                                                  int, int, int64_t *);
}  // namespace fheacc
