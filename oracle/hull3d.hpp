// TEST INFRASTRUCTURE ONLY — the oracle's hull narrowphase (wgsl-hull.ts), filled in with the
// hull shape support.
#pragma once
#include "gpu3d.hpp"

namespace gpuo {
template <class T>
typename Model<T>::Found Model<T>::collideHull(uint32_t, uint32_t, bool, bool, Sat &) const {
  Found f; f.count = 0;
  return f;
}
}  // namespace gpuo
