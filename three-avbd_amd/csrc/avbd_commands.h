// World's command passes on body records between steps: impulses (Pusher, src/lib/push.ts),
// kinematic moves (BodyMover, src/avbd3d/gpu/body-mover.ts) and resets (BodyReset,
// src/avbd3d/gpu/body-reset.ts). One thread per commanded body; the host checks that no body is
// commanded twice in one call, so no two threads touch a body.
#pragma once
#include "avbd_kernels.h"

namespace avbd {

/** v rotated by quaternion q as the passes' turn: t = 2 q.xyz × v, v + w t + q.xyz × t. */
__device__ __forceinline__ V turnQ(float4 q, V v) {
  V u = vv(q.x, q.y, q.z);
  V t = mul(cross(u, v), 2.0f);
  return add(add(v, mul(t, q.w)), cross(u, t));
}

/** Pusher's pass (push.ts:28-45): per record (lin xyz + body bits, offLin xyz, moment xyz) the
 *  velocity += lin / m and the angular velocity += I⁻¹ L, L = moment - x × offLin taken about the
 *  centre where the GPU has it, I⁻¹ applied in the principal frame. */
__global__ void k_push(Dev d, const float4 *push, uint32_t count) {
  GRID_STRIDE(k, count) {
    float4 lin = push[3 * k], off = push[3 * k + 1], mom = push[3 * k + 2];
    uint32_t i = __float_as_uint(lin.w);
    if (i >= d.n || !bodyEnabled(d, i)) continue;
    float mass = d.b.size[i].w;
    if (mass <= 0.0f) continue;
    float4 v = d.b.vel[i];
    d.b.vel[i] = make_float4(v.x + lin.x / mass, v.y + lin.y / mass, v.z + lin.z / mass, v.w);
    V L = sub(xyz(mom), cross(xyz(d.b.pos[i]), xyz(off)));
    float4 q = d.b.rot[i];
    float4 mo = d.b.moment[i];
    V local = turnQ(make_float4(-q.x, -q.y, -q.z, q.w), L);
    local = vv(local.x / mx(mo.x, 1e-12f), local.y / mx(mo.y, 1e-12f), local.z / mx(mo.z, 1e-12f));
    V dw = turnQ(q, local);
    float4 w = d.b.angvel[i];
    d.b.angvel[i] = make_float4(w.x + dw.x, w.y + dw.y, w.z + dw.z, w.w);
  }
}

__device__ __forceinline__ float4 qmulF(float4 a, float4 b) {
  return make_float4(a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y, a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
                     a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z);
}
/** BodyMover's pass (body-mover.ts:10-31): per move (target xyz + body bits, target rotation,
 *  (steps left, dt, has rotation)) the velocity that reaches the target in `left` steps and the
 *  angular velocity whose qadd step covers 1/left of the shortest arc (2 tan(half angle / left) / dt
 *  about its axis). */
__global__ void k_move(Dev d, const float4 *move, uint32_t count) {
  GRID_STRIDE(k, count) {
    float4 pos = move[3 * k], rot = move[3 * k + 1], tm = move[3 * k + 2];
    uint32_t i = __float_as_uint(pos.w);
    if (i >= d.n || !bodyEnabled(d, i)) continue;
    float4 p = d.b.pos[i], q = d.b.rot[i], v = d.b.vel[i], w = d.b.angvel[i];
    float time = tm.x * tm.y;
    d.b.vel[i] = make_float4((pos.x - p.x) / time, (pos.y - p.y) / time, (pos.z - p.z) / time, v.w);
    V omega = vv(0.0f, 0.0f, 0.0f);
    if (tm.z != 0.0f) {
      float4 dq = qmulF(rot, make_float4(-q.x, -q.y, -q.z, q.w));
      float len = sqrtf(dq.x * dq.x + dq.y * dq.y + dq.z * dq.z + dq.w * dq.w);
      dq = make_float4(dq.x / len, dq.y / len, dq.z / len, dq.w / len);
      if (dq.w < 0.0f) dq = make_float4(-dq.x, -dq.y, -dq.z, -dq.w);
      float sine = sqrtf(dq.x * dq.x + dq.y * dq.y + dq.z * dq.z);
      if (sine > 1e-7f) {
        float half = atan2f(sine, mx(dq.w, 0.0f)) / tm.x;
        float s = 2.0f * tanf(mn(half, 1.5706963f)) / tm.y;
        omega = vv(dq.x / sine * s, dq.y / sine * s, dq.z / sine * s);
      }
    }
    d.b.angvel[i] = f4(omega, w.w);
  }
}

/** BodyReset's pass (body-reset.ts:9-17): x- and the inertial pose to the current pose (their w's
 *  kept), velocity and angular velocity to zero (the shape word kept). */
__global__ void k_reset(Dev d, const uint32_t *index, uint32_t count) {
  GRID_STRIDE(k, count) {
    uint32_t i = index[k];
    if (i >= d.n) continue;
    float4 p = d.b.pos[i], q = d.b.rot[i], ip = d.b.ipos[i], inp = d.b.inpos[i];
    d.b.ipos[i] = make_float4(p.x, p.y, p.z, ip.w);
    d.b.inpos[i] = make_float4(p.x, p.y, p.z, inp.w);
    d.b.irot[i] = q;
    d.b.inrot[i] = q;
    d.b.vel[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    d.b.angvel[i] = make_float4(0.0f, 0.0f, 0.0f, d.b.angvel[i].w);
  }
}

}  // namespace avbd
