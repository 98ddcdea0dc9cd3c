// Host-side scene planning done once before a solver exists, as GpuSolver3D's constructor does it:
// the spatial body order (mortonOrder), the CPU estimate of the starting pairs and touching pairs
// (estimatePairs), and the capacities and starting colour cap sized from them. Host code only
// (plain C++ over body records); the arithmetic is f64, as the reference's JavaScript numbers are.
#pragma once
#include <cmath>
#include <cstdint>
#include <unordered_map>
#include <vector>

namespace avbd_plan {

/** A body record's fields the planning reads (the 40-float body record). */
struct PlanBody {
  double pos[3], rot[4], size[3], mass, radius;
  bool sphere;
};
inline PlanBody planBody(const float *r) {
  PlanBody b;
  for (int k = 0; k < 3; k++) { b.pos[k] = r[k]; b.size[k] = r[16 + k]; }
  for (int k = 0; k < 4; k++) b.rot[k] = r[4 + k];
  b.mass = r[19];
  b.radius = r[23];
  b.sphere = r[39] == 1.0f;
  return b;
}
/** rotate (ref/math.ts:115-126): v by the unit quaternion q = (x, y, z, w). */
inline void rotate(const double q[4], const double v[3], double out[3]) {
  double ux = q[0], uy = q[1], uz = q[2], w = q[3];
  double tx = (uy * v[2] - uz * v[1]) * 2, ty = (uz * v[0] - ux * v[2]) * 2, tz = (ux * v[1] - uy * v[0]) * 2;
  out[0] = v[0] + tx * w + (uy * tz - uz * ty);
  out[1] = v[1] + ty * w + (uz * tx - ux * tz);
  out[2] = v[2] + tz * w + (ux * ty - uy * tx);
}

/** mortonOrder (solver.ts:463-485): body indices sorted along a Z-order curve of their positions,
 *  10 bits per axis over the scene's bounding box, ties by index. order[g] is the input index of
 *  the body placed g-th. */
inline void mortonOrder(const float *rec, uint32_t n, uint32_t *order) {
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t i = 0; i < n; i++)
    for (int k = 0; k < 3; k++) { lo[k] = std::min(lo[k], (double)rec[40ull * i + k]); hi[k] = std::max(hi[k], (double)rec[40ull * i + k]); }
  auto spread = [](uint32_t v) {
    uint32_t x = v & 0x3ffu;
    x = (x | (x << 16)) & 0x30000ffu;
    x = (x | (x << 8)) & 0x300f00fu;
    x = (x | (x << 4)) & 0x30c30c3u;
    x = (x | (x << 2)) & 0x9249249u;
    return x;
  };
  std::vector<uint32_t> code(n);
  for (uint32_t i = 0; i < n; i++) {
    uint32_t q[3];
    for (int k = 0; k < 3; k++) {
      double t = std::floor(((double)rec[40ull * i + k] - lo[k]) / std::max(hi[k] - lo[k], 1e-9) * 1024);
      q[k] = (uint32_t)std::min(1023.0, t);
    }
    code[i] = spread(q[0]) | (spread(q[1]) << 1) | (spread(q[2]) << 2);
  }
  for (uint32_t i = 0; i < n; i++) order[i] = i;
  std::stable_sort(order, order + n, [&](uint32_t a, uint32_t b) { return code[a] < code[b]; });
}

/** estimatePairs (solver.ts:350-460): the pairs the broadphase will report for these bodies
 *  (bounding spheres and world AABBs overlap, not both static), counted on a CPU hash grid whose
 *  cells are sized by the largest small radius (bodies above it are tested against everything),
 *  and how many of them touch (the separating-axis test with a 1e-3 tolerance); each touching
 *  pair counts once into both bodies' `degree`. Two cells sharing a hash bucket count their pairs
 *  twice, as in the reference (an estimate). */
struct Estimate { uint64_t pairs = 0, touching = 0; };
inline Estimate estimatePairs(const float *rec, uint32_t n, int32_t *degree, double (*largestSmallOf)(std::vector<double>)) {
  const double TOUCH = 1e-3;
  std::vector<PlanBody> B(n);
  std::vector<double> axes(9ull * n), half(3ull * n), radii(n);
  for (uint32_t i = 0; i < n; i++) {
    B[i] = planBody(rec + 40ull * i);
    radii[i] = B[i].radius;
    for (int k = 0; k < 3; k++) {
      double e[3] = {0, 0, 0};
      e[k] = 1;
      rotate(B[i].rot, e, &axes[9ull * i + 3 * k]);
    }
    for (int c = 0; c < 3; c++) {
      if (B[i].sphere) { half[3ull * i + c] = B[i].radius; continue; }
      double h = 0;
      for (int k = 0; k < 3; k++) h += B[i].size[k] * 0.5 * std::fabs(axes[9ull * i + 3 * k + c]);
      half[3ull * i + c] = h;
    }
  }
  auto ax = [&](uint32_t i, int k) { return &axes[9ull * i + 3 * k]; };
  auto dot = [](const double *u, const double *v) { return u[0] * v[0] + u[1] * v[1] + u[2] * v[2]; };
  double maxSmall = largestSmallOf(radii);
  double cell = std::max(2 * maxSmall, 1e-3);
  auto overlap = [&](uint32_t i, uint32_t j) {
    const PlanBody &a = B[i], &b = B[j];
    if (a.mass <= 0 && b.mass <= 0) return false;
    double d[3] = {a.pos[0] - b.pos[0], a.pos[1] - b.pos[1], a.pos[2] - b.pos[2]};
    double r = a.radius + b.radius;
    if (d[0] * d[0] + d[1] * d[1] + d[2] * d[2] > r * r) return false;
    for (int k = 0; k < 3; k++)
      if (!(std::fabs(d[k]) <= half[3ull * i + k] + half[3ull * j + k])) return false;
    return true;
  };
  auto touches = [&](uint32_t i, uint32_t j) {
    const PlanBody &a = B[i], &b = B[j];
    double d[3] = {b.pos[0] - a.pos[0], b.pos[1] - a.pos[1], b.pos[2] - a.pos[2]};
    if (a.sphere && b.sphere) return std::hypot(d[0], d[1], d[2]) <= a.radius + b.radius + TOUCH;
    if (a.sphere || b.sphere) {  // the sphere centre's distance to the box
      uint32_t box = a.sphere ? j : i;
      double r = a.sphere ? a.radius : b.radius, sign = a.sphere ? -1 : 1, distSq = 0;
      for (int k = 0; k < 3; k++) {
        double out = std::fabs(sign * dot(d, ax(box, k))) - B[box].size[k] * 0.5;
        if (out > 0) distSq += out * out;
      }
      return distSq <= (r + TOUCH) * (r + TOUCH);
    }
    auto radius = [&](uint32_t body, const double *axis) {
      return B[body].size[0] * 0.5 * std::fabs(dot(ax(body, 0), axis)) + B[body].size[1] * 0.5 * std::fabs(dot(ax(body, 1), axis)) +
             B[body].size[2] * 0.5 * std::fabs(dot(ax(body, 2), axis));
    };
    auto separated = [&](const double *axis) {
      double len = std::hypot(axis[0], axis[1], axis[2]);
      if (len < 1e-6) return false;
      return std::fabs(dot(d, axis)) - radius(i, axis) - radius(j, axis) > TOUCH * len;
    };
    for (int k = 0; k < 3; k++)
      if (separated(ax(i, k)) || separated(ax(j, k))) return false;
    for (int p = 0; p < 3; p++)
      for (int q = 0; q < 3; q++) {
        const double *u = ax(i, p), *v = ax(j, q);
        double c[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
        if (separated(c)) return false;
      }
    return true;
  };
  // the hash of the reference's grid, with JavaScript's ToInt32 on each product
  auto key = [](int64_t x, int64_t y, int64_t z) {
    uint32_t hx = (uint32_t)(uint64_t)(x * 73856093), hy = (uint32_t)(uint64_t)(y * 19349663), hz = (uint32_t)(uint64_t)(z * 83492791);
    return hx ^ hy ^ hz;
  };
  auto cellOf = [&](uint32_t i, int64_t c[3]) {
    for (int k = 0; k < 3; k++) c[k] = (int64_t)std::floor(B[i].pos[k] / cell);
  };
  std::vector<uint8_t> large(n, 0);
  std::vector<uint32_t> largeList;
  std::unordered_map<uint32_t, std::vector<uint32_t>> grid;
  for (uint32_t i = 0; i < n; i++) {
    if (B[i].radius > maxSmall) { large[i] = 1; largeList.push_back(i); continue; }
    int64_t c[3];
    cellOf(i, c);
    grid[key(c[0], c[1], c[2])].push_back(i);
  }
  Estimate e;
  auto count = [&](uint32_t i, uint32_t j) {
    e.pairs++;
    if (touches(i, j)) {
      e.touching++;
      if (degree) { degree[i]++; degree[j]++; }
    }
  };
  for (uint32_t i = 0; i < n; i++) {
    if (large[i]) continue;
    int64_t c[3];
    cellOf(i, c);
    for (int dz = -1; dz <= 1; dz++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          auto it = grid.find(key(c[0] + dx, c[1] + dy, c[2] + dz));
          if (it == grid.end()) continue;
          for (uint32_t j : it->second)
            if (j < i && overlap(i, j)) count(i, j);
        }
  }
  for (uint32_t l : largeList)
    for (uint32_t j = 0; j < n; j++)
      if (j != l && (!large[j] || j < l) && overlap(l, j)) count(l, j);
  return e;
}

}  // namespace avbd_plan
