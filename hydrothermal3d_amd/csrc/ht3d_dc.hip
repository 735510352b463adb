// ht3d_dc.hip -- mixed-precision solve: FP32 DYREL inner sweeps inside an FP64
// defect-correction loop.
#include "ht3d_internal.hpp"

namespace ht3d {
int dc_solve_impl(const ht3d_dyrel_field *f, const ht3d_dyrel_params *p, ht3d_dyrel_report *rep, double *nu_out,
                  int *outer, int *inner) {
    (void)f; (void)p; (void)rep; (void)nu_out; (void)outer; (void)inner;
    return arg_fail("FP32 defect-correction solve not built yet");
}
}  // namespace ht3d

extern "C" int ht3d_dyrel_solve_dc(const ht3d_dyrel_field *f, const ht3d_dyrel_params *p, ht3d_dyrel_report *rep,
                                   int *outer_iters, int *inner_iters) {
    int o = 0, i = 0;
    int rc = ht3d::dc_solve_impl(f, p, rep, nullptr, &o, &i);
    if (outer_iters) *outer_iters = o;
    if (inner_iters) *inner_iters = i;
    return rc;
}
