// Mixed-precision SP2 products (gfx950): placeholder until the split and
// sliced kernels land; the selector rejects the modes loudly.
#include "prg_sp2.h"

namespace prg {

int sp2_step_mixed(const DMat& x0, const DMat& x1, Sp2State* d_state, double* d_part,
                   int* ntiles, int prec) {
  (void)x0; (void)x1; (void)d_state; (void)d_part; (void)ntiles; (void)prec;
  set_error(ERR_UNSUPPORTED, "mixed-precision SP2 products are not built yet");
  return ERR_UNSUPPORTED;
}

}  // namespace prg
