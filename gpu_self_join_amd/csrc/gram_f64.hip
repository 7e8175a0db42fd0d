// gram_f64.hip -- Gram-filter kernels for FP64 data (fp16 operands).
#include "gram_kernels.h"

namespace gdsj {

int gram_dispatch_f64(const GramData &gd, const JoinArgs &a, int mode, int pass, unsigned long long *stats,
                      bool err, hipStream_t s) {
    return gram_dispatch<double, 1>(gd, a, mode, pass, stats, err, s);
}

}  // namespace gdsj
