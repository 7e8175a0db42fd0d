// batch_f32.hip -- FP32 instantiations of the persistent batch kernels.
#define BT float
#define BNAME batch_pass_f32
#include "batch_inst.h"
