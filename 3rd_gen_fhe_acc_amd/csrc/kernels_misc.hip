// kernels_misc.hip -- every other launcher (key switch, NTT, external
// product, substitution, expansion, key conversion and generation) for both
// device word widths, and the key digest.
#include "kernels_impl.hpp"

namespace fheacc {

// FNV-1a style digest over 64-bit words (order dependent), one partial per block
__global__ void k_digest(const uint64_t *w, int64_t count, uint64_t *partials) {
    __shared__ uint64_t sh[256];
    uint64_t h = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        h += mix64(w[i] ^ (0x9E3779B97F4A7C15ULL * (uint64_t)(i + 1)));
    sh[threadIdx.x] = h;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = sh[0];
}


hipError_t launch_digest(const uint64_t *w, int64_t count, uint64_t *partials, int blocks, hipStream_t st) {
    hipLaunchKernelGGL(k_digest, dim3(blocks), dim3(256), 0, st, w, count, partials);
    return hipGetLastError();
}

template size_t ks_staged_lds<uint32_t>(uint32_t, int);
template size_t ks_staged_lds<uint64_t>(uint32_t, int);

#define FHEACC_INSTANTIATE(W)                                                                                     \
    template hipError_t launch_lwe_keyswitch<W>(const KSArgs<W> &, bool, hipStream_t);                            \
    template hipError_t launch_ntt<W>(const NTTArgs<W> &, int, hipStream_t);                                      \
    template hipError_t launch_external_product<W>(const EPArgs<W> &, int, int64_t, hipStream_t);                 \
    template hipError_t launch_subs_keyswitch<W>(const SubsArgs<W> &, int, int64_t, hipStream_t);                 \
    template hipError_t launch_expand_level<W>(const SubsArgs<W> &, int, int64_t, hipStream_t);                   \
    template hipError_t launch_rgsw_to_device<W>(const uint64_t *, int64_t, W *, int64_t, int, int, ShoupGen,     \
                                                 const int64_t *, int64_t, hipStream_t, bool);                    \
    template hipError_t launch_rgsw_to_ref<W>(const W *, int64_t, uint64_t *, int64_t, int, int, const int64_t *, \
                                              int64_t, hipStream_t, bool);                                        \
    template hipError_t launch_ksk_convert<W>(const uint64_t *, W *, uint64_t *, int, int, int, int, int,         \
                                              hipStream_t);                                                       \
    template hipError_t launch_keygen<W>(const KeygenArgs<W> &, const KskGenArgs<W> &, int, const uint64_t *,     \
                                         hipStream_t, bool);                                                      \
    template hipError_t launch_subs_convert<W>(const void *, void *, int, int64_t, bool, ShoupGen, hipStream_t); \
    template hipError_t launch_secret_ntt<W>(const KeygenArgs<W> &, int, const uint64_t *, hipStream_t);          \
    template hipError_t launch_encrypt<W>(const EncArgs<W> &, int, int, int64_t, hipStream_t);

FHEACC_INSTANTIATE(uint32_t)
FHEACC_INSTANTIATE(uint64_t)

}  // namespace fheacc
