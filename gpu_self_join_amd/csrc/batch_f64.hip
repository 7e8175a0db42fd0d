// batch_f64.hip -- FP64 instantiations of the persistent batch kernels.
#define BT double
#define BNAME batch_pass_f64
#include "batch_inst.h"
