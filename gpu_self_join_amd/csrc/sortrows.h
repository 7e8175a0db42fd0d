// sortrows.h -- per-point ascending order of CSR neighbor rows.
#pragma once
#include "common.h"

namespace gdsj {

// Sorts every row [offs[i], offs[i+1]) of ids ascending (n rows).
int sort_rows(const uint64_t *offs, uint32_t n, uint32_t *ids, hipStream_t s, DevBuf &work);

}  // namespace gdsj
