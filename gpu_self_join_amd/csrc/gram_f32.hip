// gram_f32.hip -- Gram-filter kernels for FP32 data (fp16 operands, or bf16
// operands in the split form).
#include "gram_kernels.h"

namespace gdsj {

int gram_dispatch_f32(const GramData &gd, const JoinArgs &a, int mode, int pass, unsigned long long *stats,
                      bool err, hipStream_t s) {
    if (gd.lowp) return gram_dispatch<float, 1>(gd, a, mode, pass, stats, err, s);
    return gram_dispatch<float, 0>(gd, a, mode, pass, stats, err, s);
}

}  // namespace gdsj
