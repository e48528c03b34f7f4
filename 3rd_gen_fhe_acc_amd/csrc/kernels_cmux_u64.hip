// kernels_cmux_u64.hip -- the CMUX-chain launcher for uint64_t device words.
#include "kernels_impl.hpp"

namespace fheacc {
template hipError_t launch_cmux_chain<uint64_t>(const CmuxArgs<uint64_t> &, int, int, int64_t, size_t, hipStream_t);
}  // namespace fheacc
