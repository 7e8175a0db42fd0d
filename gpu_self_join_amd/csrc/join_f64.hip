// join_f64.hip -- FP64 instantiations of the join kernels.
#define JT double
#define JNAME join_pass_f64
#define JNAME_SPECIALISED join_specialised_f64
#define JOIN_DIMS GDSJ_DIMCASE(2) GDSJ_DIMCASE(4) GDSJ_DIMCASE(8) GDSJ_DIMCASE(16) GDSJ_DIMCASE(32) \
    GDSJ_DIMCASE(64)
#define JOIN_DIMQ GDSJ_DIMQ(2) GDSJ_DIMQ(4) GDSJ_DIMQ(8) GDSJ_DIMQ(16) GDSJ_DIMQ(32) GDSJ_DIMQ(64)
#define JOIN_GV_DIM(D) false
#include "join_inst.h"
