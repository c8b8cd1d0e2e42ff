// ref_basis.cpp -- test-only driver that exposes the reference's own shape
// functions (src/basis_functions.h, compiled unmodified from the reference
// tree) through a C entry point, to pin the oracle's restatement.
#include <cmath>

#include "basis_functions.h"

extern "C" double ref_basis(int shape, int derivative, double r, int ntype, double inv_h) {
  if (shape == 0)
    return derivative ? BasisFunction::derivative_linear(r, ntype, inv_h)
                      : BasisFunction::linear(r, ntype);
  return derivative ? BasisFunction::derivative_cubic_spline(r, ntype, inv_h)
                    : BasisFunction::cubic_spline(r, ntype);
}
