// kernels_cmux_u32.hip -- the CMUX-chain launcher for uint32_t device words.
#include "kernels_impl.hpp"

namespace fheacc {
template hipError_t launch_cmux_chain<uint32_t>(const CmuxArgs<uint32_t> &, int, int, int64_t, size_t, hipStream_t);
}  // namespace fheacc
