// Mohr-Coulomb P4 triangle (A4) -- kernels added below.
#include "fne_common.h"

int fne_launch_p4(fne_plan* plan, FneOp op, const FneCall& c) {
  (void)plan; (void)op; (void)c;
  fne_set_error("Mohr-Coulomb kernels not built yet");
  return FNE_EUNSUPPORTED;
}
