// device_checks.hpp -- device-side bounds checks of the checked library.
//
// Compiled with -DFHEACC_DEVICE_CHECKS=1 (build_lib.py's checked build,
// lib/libfheacc_checked.so), every LDS access of the transforms and of the
// external products is checked against the workgroup's LDS allocation (the
// group segment size of the dispatch packet), and each gate workgroup's
// ciphertext against its launch's range.  A failed check never traps: it
// ORs a bit into this translation unit's violation word, and the access it
// checked runs as it would without the check.  fheacc_device_check_read()
// synchronizes the device, then returns and clears the OR of every
// translation unit's word.  Without the flag the checks compile to nothing.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#ifndef FHEACC_DEVICE_CHECKS
#define FHEACC_DEVICE_CHECKS 0
#endif

namespace fheacc {

// violation bits, as fheacc_device_check_read() reports them
constexpr unsigned kCheckLds = 1u;         // an LDS access outside the allocation
constexpr unsigned kCheckCiphertext = 2u;  // a gate workgroup outside its launch's ciphertexts

// Collects the checked translation units' readers (defined in api.cpp, which
// every build links); a reader returns and clears its unit's word.
void register_check_reader(unsigned (*reader)());

#if FHEACC_DEVICE_CHECKS
// this translation unit's violation word
static __device__ unsigned int g_check_bits;

__device__ __forceinline__ void device_check(bool ok, unsigned bit) {
    if (!ok) atomicOr(&g_check_bits, bit);
}
// bytes of LDS this workgroup holds: group_segment_size of the HSA kernel
// dispatch packet (32-bit word 7), static and dynamic LDS together
__device__ __forceinline__ uint32_t lds_bytes() {
    typedef const __attribute__((address_space(4))) uint32_t packet_word;
    return ((packet_word *)__builtin_amdgcn_dispatch_ptr())[7];
}
// p: a generic pointer into LDS; its offset is the LDS address
template <typename T>
__device__ __forceinline__ void check_lds(const T *p) {
    const uint32_t off = (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) T *)p;
    device_check(off + sizeof(T) <= lds_bytes(), kCheckLds);
}

namespace {
unsigned read_and_clear_checks() {
    unsigned v = 0;
    const unsigned zero = 0;
    if (hipMemcpyFromSymbol(&v, HIP_SYMBOL(g_check_bits), sizeof v) != hipSuccess) return ~0u;
    if (hipMemcpyToSymbol(HIP_SYMBOL(g_check_bits), &zero, sizeof zero) != hipSuccess) return ~0u;
    return v;
}
struct CheckRegistration {
    CheckRegistration() { register_check_reader(&read_and_clear_checks); }
};
const CheckRegistration check_registration;
}  // namespace

#define FHEACC_CHECK_LDS(...) ::fheacc::check_lds(__VA_ARGS__)
#define FHEACC_CHECK(ok, bit) ::fheacc::device_check((ok), (bit))
#else
#define FHEACC_CHECK_LDS(...) ((void)0)
#define FHEACC_CHECK(ok, bit) ((void)0)
#endif

}  // namespace fheacc
