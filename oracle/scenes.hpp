// TEST INFRASTRUCTURE ONLY — the oracle's copy of the demo scenes.
// The 14 avbd-demo3d scenes in the demo's order (src/avbd3d/ref/scenes.ts:243-258), z up, built
// into the FP64 reference solver. The index is the golden fixture's scene index.
#pragma once
#include "ref3d.hpp"

namespace refo {

static inline Rigid *ground(Solver &s, R z = 0) { return s.addBody(v3(100, 100, 1), 0, 0.5, v3(0, 0, z)); }  // scenes.ts:8-10

static void buildScene(Solver &s, int id) {
  s.clear();
  switch (id) {
    case 0: break;  // Empty
    case 1:  // Ground, scenes.ts:14-18
      ground(s);
      s.addBody(v3(1, 1, 1), 1, 0.5, v3(0, 0, 4));
      break;
    case 2:  // Dynamic Friction, scenes.ts:20-26
      ground(s);
      for (int x = 0; x <= 10; x++) s.addBody(v3(1, 1, 0.5), 1, 5 - (x / 10.0) * 5, v3(0, -30 + x * 2, 0.75), v3(10, 0, 0));
      break;
    case 3: {  // Static Friction, scenes.ts:28-47
      ground(s);
      R angle = 30 * 0.017453292519943295;
      Rigid *ramp = s.addBody(v3(40, 24, 1), 0, 1, v3(0, 0, 3));
      ramp->positionAng = q4(0, std::sin(angle * 0.5), 0, std::cos(angle * 0.5));
      V3 tangent = normalize3(rotate(ramp->positionAng, v3(1, 0, 0)));
      V3 normal = normalize3(rotate(ramp->positionAng, v3(0, 0, 1)));
      for (int i = 0; i <= 10; i++) {
        R friction = (i / 10.0) * 0.25 + 0.25;
        R y = -10 + i * 2;
        V3 pos = addScaled(ramp->positionLin, tangent, -12);
        pos.y += y;
        pos = addScaled(pos, normal, 1.05);
        s.addBody(v3(1, 1, 1), 1, friction, pos);
      }
      break;
    }
    case 4: {  // Pyramid (16 rows), scenes.ts:50-58
      int size = 16;
      ground(s, -0.5);
      for (int y = 0; y < size; y++)
        for (int x = 0; x < size - y; x++) s.addBody(v3(1, 0.5, 0.5), 1, 0.5, v3(x * 1.01 + y * 0.5 - size / 2.0, 0, y * 0.85 + 0.5));
      break;
    }
    case 5: {  // Rope, scenes.ts:60-69
      ground(s, -20);
      Rigid *prev = nullptr;
      for (int i = 0; i < 20; i++) {
        Rigid *cur = s.addBody(v3(1, 0.5, 0.5), i == 0 ? 0 : 1, 0.5, v3(i, 0, 10));
        if (prev) s.addJoint(prev, cur, v3(0.5, 0, 0), v3(-0.5, 0, 0));
        prev = cur;
      }
      break;
    }
    case 6: {  // Heavy Rope, scenes.ts:71-83
      const int N = 20; const R SIZE = 5;
      ground(s, -20);
      Rigid *prev = nullptr;
      for (int i = 0; i < N; i++) {
        bool last = i == N - 1;
        Rigid *cur = s.addBody(last ? v3(SIZE, SIZE, SIZE) : v3(1, 0.5, 0.5), i == 0 ? 0 : 1, 0.5, v3(i + (last ? SIZE / 2 : 0), 0, 10));
        if (prev) s.addJoint(prev, cur, v3(0.5, 0, 0), last ? v3(-SIZE / 2, 0, 0) : v3(-0.5, 0, 0));
        prev = cur;
      }
      break;
    }
    case 7: {  // Spring, scenes.ts:85-91
      ground(s);
      Rigid *anchor = s.addBody(v3(1, 1, 1), 0, 0.5, v3(0, 0, 14));
      Rigid *block = s.addBody(v3(2, 2, 2), 1, 0.5, v3(0, 0, 8));
      s.addSpring(anchor, block, V3(), V3(), 100, 4);
      break;
    }
    case 8: {  // Spring Ratio, scenes.ts:93-104
      const int N = 8;
      ground(s, -10);
      Rigid *prev = nullptr;
      for (int i = 0; i < N; i++) {
        R x = (i - (N - 1) * 0.5) * 3;
        Rigid *cur = s.addBody(v3(1, 0.75, 0.75), (i == 0 || i == N - 1) ? 0 : 1, 0.5, v3(x, 0, 12));
        if (prev) s.addSpring(prev, cur, v3(0.5, 0, 0), v3(-0.5, 0, 0), i % 2 == 0 ? 10 : 10000, 3);
        prev = cur;
      }
      break;
    }
    case 9:  // Stack, scenes.ts:106-110
      ground(s);
      for (int i = 0; i < 10; i++) s.addBody(v3(1, 1, 1), 1, 0.5, v3(0, 0, i * 1.5 + 1));
      break;
    case 10: {  // Stack Ratio, scenes.ts:112-125
      R t = 1;
      s.addBody(v3(100, 100, t), 0, 0.5, v3(0, 0, 0));
      R topZ = t * 0.5, sz = 1;
      for (int i = 0; i < 4; i++) {
        R half = sz * 0.5, centerZ = topZ + half;
        s.addBody(v3(sz, sz, sz), 1, 0.5, v3(0, 0, centerZ));
        topZ = centerZ + half;
        sz *= 2;
      }
      break;
    }
    case 11: {  // Soft Body, scenes.ts:127-188
      ground(s);
      const R Klin = 1000, Kang = 250, size = 0.8, half = size * 0.5, baseZ = 8, stackGap = 2;
      const int W = 4, D = 4, H = 4, N = 3;
      for (int i = 0; i < N; i++) {
        R stackZ = i * (H * size + stackGap);
        Rigid *g[4][4][4];
        for (int x = 0; x < W; x++)
          for (int y = 0; y < D; y++)
            for (int z = 0; z < H; z++)
              g[x][y][z] = s.addBody(v3(size, size, size), 1, 0.5, v3((x - (W - 1) * 0.5) * size, (y - (D - 1) * 0.5) * size, baseZ + stackZ + z * size));
        for (int x = 1; x < W; x++) for (int y = 0; y < D; y++) for (int z = 0; z < H; z++) s.addJoint(g[x - 1][y][z], g[x][y][z], v3(half, 0, 0), v3(-half, 0, 0), Klin, Kang);
        for (int x = 0; x < W; x++) for (int y = 1; y < D; y++) for (int z = 0; z < H; z++) s.addJoint(g[x][y - 1][z], g[x][y][z], v3(0, half, 0), v3(0, -half, 0), Klin, Kang);
        for (int x = 0; x < W; x++) for (int y = 0; y < D; y++) for (int z = 1; z < H; z++) s.addJoint(g[x][y][z - 1], g[x][y][z], v3(0, 0, half), v3(0, 0, -half), Klin, Kang);
        for (int x = 1; x < W; x++) for (int y = 0; y < D; y++) for (int z = 1; z < H; z++) { s.addIgnore(g[x - 1][y][z - 1], g[x][y][z]); s.addIgnore(g[x][y][z - 1], g[x - 1][y][z]); }
        for (int x = 0; x < W; x++) for (int y = 1; y < D; y++) for (int z = 1; z < H; z++) { s.addIgnore(g[x][y - 1][z - 1], g[x][y][z]); s.addIgnore(g[x][y][z - 1], g[x][y - 1][z]); }
        for (int x = 1; x < W; x++) for (int y = 1; y < D; y++) for (int z = 0; z < H; z++) { s.addIgnore(g[x - 1][y - 1][z], g[x][y][z]); s.addIgnore(g[x][y - 1][z], g[x - 1][y][z]); }
      }
      break;
    }
    case 12: {  // Bridge, scenes.ts:190-214
      const int N = 40;
      const R L = 1, Wd = 4, Ht = 0.5, hl = L * 0.5, hw = Wd * 0.5;
      ground(s);
      Rigid *prev = nullptr;
      for (int i = 0; i < N; i++) {
        Rigid *cur = s.addBody(v3(L, Wd, Ht), (i == 0 || i == N - 1) ? 0 : 1, 0.5, v3(i - N / 2.0, 0, 10));
        if (prev) {
          s.addJoint(prev, cur, v3(hl, hw, 0), v3(-hl, hw, 0), INF, 0);
          s.addJoint(prev, cur, v3(hl, -hw, 0), v3(-hl, -hw, 0), INF, 0);
        }
        prev = cur;
      }
      for (int x = 0; x < N / 4; x++)
        for (int y = 0; y < N / 8; y++) s.addBody(v3(1, 1, 1), 1, 0.5, v3(x - N / 8.0, 0, y + 12));
      break;
    }
    case 13: {  // Breakable, scenes.ts:216-235
      const int N = 10, M = 5;
      const R breakForce = 90;
      ground(s);
      Rigid *prev = nullptr;
      for (int i = 0; i <= N; i++) {
        Rigid *cur = s.addBody(v3(1, 1, 0.5), 1, 0.5, v3(i - N / 2.0, 0, 6));
        if (prev) s.addJoint(prev, cur, v3(0.5, 0, 0), v3(-0.5, 0, 0), INF, INF, breakForce);
        prev = cur;
      }
      s.addBody(v3(1, 1, 5), 0, 0.5, v3(-N / 2.0, 0, 2.5));
      s.addBody(v3(1, 1, 5), 0, 0.5, v3(N / 2.0, 0, 2.5));
      for (int i = 0; i < M; i++) s.addBody(v3(2, 1, 1), 1, 0.5, v3(0, 0, i * 2 + 8));
      break;
    }
    default: break;
  }
}

}  // namespace refo
