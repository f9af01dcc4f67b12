// Min / max of bounding-box coordinates, shared by the single and batched paths.
#pragma once

#include "qpip_internal.hpp"

namespace qpip {

// geometry/vec_2d.hpp:185-205 (device ::min/::max ignore a NaN operand); -0.0
// orders before +0.0 so the result does not depend on the reduction order.
__device__ __forceinline__ double box_min(double a, double b)
{
  if (isnan(a)) return b;
  if (isnan(b)) return a;
  if (a < b) return a;
  if (b < a) return b;
  return signbit(a) ? a : b;
}
__device__ __forceinline__ double box_max(double a, double b)
{
  if (isnan(a)) return b;
  if (isnan(b)) return a;
  if (a > b) return a;
  if (b > a) return b;
  return signbit(a) ? b : a;
}

}  // namespace qpip
