// join_f32.hip -- FP32 instantiations of the join kernels.
#define JT float
#define JNAME join_pass_f32
#define JNAME_SPECIALISED join_specialised_f32
#define JOIN_DIMS GDSJ_DIMCASE(2) GDSJ_DIMCASE(3) GDSJ_DIMCASE(4) GDSJ_DIMCASE(5) GDSJ_DIMCASE(6) \
    GDSJ_DIMCASE(8) GDSJ_DIMCASE(16) GDSJ_DIMCASE(32) GDSJ_DIMCASE(64)
#define JOIN_DIMQ GDSJ_DIMQ(2) GDSJ_DIMQ(3) GDSJ_DIMQ(4) GDSJ_DIMQ(5) GDSJ_DIMQ(6) GDSJ_DIMQ(8) \
    GDSJ_DIMQ(16) GDSJ_DIMQ(32) GDSJ_DIMQ(64)
#define JOIN_GV_DIM(D) ((D) >= 3 && (D) <= 8)
#include "join_inst.h"
