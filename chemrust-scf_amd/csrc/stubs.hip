// Entry points whose implementation lands in later translation units.
#include "impl.hpp"
#include "ops.hpp"

namespace chs {

int nonlocal_apply(chs_context* ctx, int, const void*, int, void*, void*) {
  return set_error(ctx, CHS_ERR_NOT_IMPLEMENTED, "projectors not implemented");
}
int sinv_apply(chs_context* ctx, int, void*, int) {
  return set_error(ctx, CHS_ERR_NOT_IMPLEMENTED, "S^-1 not implemented");
}

}  // namespace chs

using namespace chs;

extern "C" {
int chs_bind_projectors(chs_context* ctx, int, const double*, const double*, const double*) {
  return set_error(ctx, CHS_ERR_NOT_IMPLEMENTED, "not implemented");
}
int chs_set_dion(chs_context* ctx, const double*) {
  return set_error(ctx, CHS_ERR_NOT_IMPLEMENTED, "not implemented");
}
int chs_density(chs_context* ctx, const double*, int, const double*, double, double*) {
  return set_error(ctx, CHS_ERR_NOT_IMPLEMENTED, "not implemented");
}
int chs_scf_run(chs_context* ctx, const chs_scf_params*, const double*, chs_scf_result*, double*,
                double*, double*) {
  return set_error(ctx, CHS_ERR_NOT_IMPLEMENTED, "not implemented");
}
}
