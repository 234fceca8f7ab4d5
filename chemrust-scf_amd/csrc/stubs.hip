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
}
