// sht_split.hip -- split-precision (INT8-sliced) Legendre transforms.
#include "sht_internal.h"

namespace ksht {

int split_init(sht_plan *p) {
  (void)p;
  set_error("split precision: not available in this build");
  return SHT_ERR_UNSUPPORTED;
}
void split_free(sht_plan *p) { (void)p; }
int launch_leg_bwd_split(sht_plan *p, int, int, const double *, SpecAddr, double *, GridAddr) {
  (void)p;
  set_error("split precision: not available in this build");
  return SHT_ERR_STATE;
}
int launch_leg_fwd_split(sht_plan *p, int, int, const double *, GridAddr, double *, SpecAddr) {
  (void)p;
  set_error("split precision: not available in this build");
  return SHT_ERR_STATE;
}

}  // namespace ksht
