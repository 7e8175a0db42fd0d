// join.hip -- dispatch of the join pass by DTYPE.
#include "join.h"

namespace gdsj {

int join_pass_f32(const Grid &G, const JoinArgs &a, int mode, int pass, hipStream_t s);
int join_pass_f64(const Grid &G, const JoinArgs &a, int mode, int pass, hipStream_t s);
bool join_specialised_f32(uint32_t d, int mode, uint32_t ilp);
bool join_specialised_f64(uint32_t d, int mode, uint32_t ilp);

int join_pass(const Grid &G, const JoinArgs &a, int mode, int pass, hipStream_t s) {
    if (G.dtype == GDSJ_DTYPE_F32) return join_pass_f32(G, a, mode, pass, s);
    return join_pass_f64(G, a, mode, pass, s);
}

const char *join_variant(int dtype, uint32_t d, int mode, uint32_t ilp) {
    bool sp = dtype == GDSJ_DTYPE_F32 ? join_specialised_f32(d, mode, ilp) : join_specialised_f64(d, mode, ilp);
    return sp ? "k_join<register-resident>" : "k_join<generic>";
}

}  // namespace gdsj
