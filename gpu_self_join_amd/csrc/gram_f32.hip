// gram_f32.hip -- Gram-filter kernels for FP32 data (bf16 operands, split form).
#include "gram_kernels.h"

namespace gdsj {

int gram_dispatch_f32(const GramData &gd, const JoinArgs &a, int mode, int pass, unsigned long long *stats,
                      bool err, hipStream_t s) {
    return gram_dispatch<float, 0>(gd, a, mode, pass, stats, err, s);
}

}  // namespace gdsj
