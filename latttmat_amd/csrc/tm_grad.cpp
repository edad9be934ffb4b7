// tm_grad.cpp -- Hellmann-Feynman derivative kernels (filled in below).
#include "tm_internal.hpp"

namespace ltm {

void grad_numerators(tm_op *, const double *, const double *, double *)
{
  throw Error(TM_EUNSUPPORTED, "Hellmann-Feynman gradient is not available for this operator");
}

}  // namespace ltm
