// radix.h -- stable LSD radix sort of (uint64 key, uint32 value) pairs.
#pragma once
#include "common.h"

namespace gdsj {

size_t radix_tmp_bytes(uint32_t n);

// Sorts the low `bits` bits of keys (ping-pong between the two buffer
// pairs).  *result_in_alt = 1 when the sorted data ended in keys_alt/vals_alt.
int radix_sort_pairs(uint64_t *keys, uint32_t *vals, uint64_t *keys_alt, uint32_t *vals_alt, uint32_t n,
                     int bits, void *tmp, hipStream_t s, int *result_in_alt);

}  // namespace gdsj
