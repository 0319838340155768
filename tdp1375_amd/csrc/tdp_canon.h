// Configuration canonicalisation kernel (Configuration.java:53-198).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace tdp {

// Largest configuration the canonicalisation kernel takes (one wave, one
// lane per symbol).
constexpr int kMaxCanonN = 64;

// count configurations, each as a one-line spi (spis[c * 64 + x] = image of x)
// and pi (pis[c * 64 + i] = symbol at position i) of ns[c] symbols; writes the
// canonical signature (64 floats per configuration), its hash, the index of
// the equivalent signature it came from (2 * rotation + mirror) and the
// canonical spi (one-line, pi the identity).  Device pointers; max_n (host)
// bounds every ns[c] and sizes the kernel's LDS.
void launch_canon(const uint8_t *spis, const uint8_t *pis, const int32_t *ns, int count, int max_n, float *out_sig,
                  int32_t *out_hash, int32_t *out_index, uint8_t *out_spi, hipStream_t stream);

}  // namespace tdp
