// Neo-Hookean P1 tetrahedron (A3) -- kernels added below.
#include "fne_common.h"

int fne_launch_tet(fne_plan* plan, FneOp op, const FneCall& c) {
  (void)plan; (void)op; (void)c;
  fne_set_error("Neo-Hookean kernels not built yet");
  return FNE_EUNSUPPORTED;
}
