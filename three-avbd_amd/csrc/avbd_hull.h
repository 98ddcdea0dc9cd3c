// Convex hull narrowphase (wgsl-hull.ts): hull-hull, hull-box and sphere-hull contacts.
#pragma once
#include "avbd_device.h"

namespace avbd {
/** Pairs with a hull on either side; sat.n points from A towards B. */
__device__ __noinline__ void collideHullPair(const uint4 *hulls, const Box &A, Q qA, float shA, const Box &B, Q qB, float shB,
                                             bool sphereA, bool sphereB, Sat &sat, Found &found) {
  found.count = 0;
}
}  // namespace avbd
