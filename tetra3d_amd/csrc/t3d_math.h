// t3d_math.h — float32 Matrix4/Vector3 arithmetic of Tetra3D, callable from
// host and device code.
//
// Every function keeps the reference's operation order so host uniforms and
// device per-vertex math round exactly as the Go code does.  The library is
// compiled with -ffp-contract=off, so a*b+c is never fused.  Division and
// square root are IEEE correctly rounded on both sides: on the device `/` and
// llvm.sqrt.f32 are lowered to correctly rounded sequences (HIP's default
// -fhip-fp32-correctly-rounded-divide-sqrt); HIP's __fsqrt_rn is the ~1 ulp
// hardware square root and is not used.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/t3d.h"

#define T3D_HD __host__ __device__ __forceinline__

namespace t3d {

struct V3 {
  float x, y, z;
};
struct V4 {
  float x, y, z, w;
};
typedef t3d_mat4 M4;

T3D_HD float fdiv(float a, float b) { return a / b; }
// math32.Sqrt (math32/math32.go:121-123): float64 sqrt rounded to float32 is
// the correctly rounded float32 sqrt.
T3D_HD float fsqrt(float a) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_sqrtf(a);
#else
  return (float)__builtin_sqrt((double)a);
#endif
}

T3D_HD V3 v3(float x, float y, float z) { V3 r = {x, y, z}; return r; }
T3D_HD V3 sub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
T3D_HD V3 add(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
T3D_HD V3 hmul(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }    // Vector3.Mult, vector.go:250-255
T3D_HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }    // vector.go:526-528
T3D_HD float magnitude(V3 a) { return fsqrt(a.x * a.x + a.y * a.y + a.z * a.z); }  // vector.go:188-190
// DistanceSquaredTo: vector.go:242-247 (subtract in place, then square-sum).
T3D_HD float dist_sq(V3 a, V3 b) {
  float x = a.x - b.x, y = a.y - b.y, z = a.z - b.z;
  return x * x + y * y + z * z;
}
// Vector3.Unit: vector.go:258-266 — unchanged when the length is < 1e-8 or exactly 1.
T3D_HD V3 unit(V3 v) {
  float l = magnitude(v);
  if (l < 1e-8f || l == 1.0f) return v;
  return v3(fdiv(v.x, l), fdiv(v.y, l), fdiv(v.z, l));
}
// Vector4.Unit: vector.go:1869-1877 — xyz normalised, w kept.
T3D_HD V4 unit4(V4 v) {
  float l = fsqrt(v.x * v.x + v.y * v.y + v.z * v.z);
  if (l < 1e-8f || l == 1.0f) return v;
  V4 r = {fdiv(v.x, l), fdiv(v.y, l), fdiv(v.z, l), v.w};
  return r;
}
T3D_HD float magnitude4(V4 v) { return fsqrt(v.x * v.x + v.y * v.y + v.z * v.z); }

T3D_HD M4 identity() {
  M4 r;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) r.m[i][j] = i == j ? 1.0f : 0.0f;
  return r;
}
T3D_HD M4 zero4() {
  M4 r;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) r.m[i][j] = 0.0f;
  return r;
}
// Matrix4.Mult: matrix_none.go:6-30 and the SSE form matrix_amd64.s:4-81 both
// accumulate ((a_i0·b_0j + a_i1·b_1j) + a_i2·b_2j) + a_i3·b_3j.
T3D_HD M4 mult(const M4& a, const M4& b) {
  M4 r;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++)
      r.m[i][j] = a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j] + a.m[i][2] * b.m[2][j] + a.m[i][3] * b.m[3][j];
  return r;
}
T3D_HD M4 transposed(const M4& a) {                                         // matrix.go:218-230
  M4 r;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) r.m[i][j] = a.m[j][i];
  return r;
}
T3D_HD M4 translate(float x, float y, float z) {                           // matrix.go:82-88
  M4 r = identity();
  r.m[3][0] = x;
  r.m[3][1] = y;
  r.m[3][2] = z;
  return r;
}
T3D_HD M4 scale4(float x, float y, float z) {                              // matrix.go:90-96
  M4 r = identity();
  r.m[0][0] = x;
  r.m[1][1] = y;
  r.m[2][2] = z;
  return r;
}
// Matrix4.Inverted: matrix.go:255-307, the 2x2-minor cofactor expansion with
// its exact grouping.
T3D_HD M4 inverted(const M4& M) {
  const float(*m)[4] = M.m;
  float A2323 = m[2][2] * m[3][3] - m[2][3] * m[3][2];
  float A1323 = m[2][1] * m[3][3] - m[2][3] * m[3][1];
  float A1223 = m[2][1] * m[3][2] - m[2][2] * m[3][1];
  float A0323 = m[2][0] * m[3][3] - m[2][3] * m[3][0];
  float A0223 = m[2][0] * m[3][2] - m[2][2] * m[3][0];
  float A0123 = m[2][0] * m[3][1] - m[2][1] * m[3][0];
  float A2313 = m[1][2] * m[3][3] - m[1][3] * m[3][2];
  float A1313 = m[1][1] * m[3][3] - m[1][3] * m[3][1];
  float A1213 = m[1][1] * m[3][2] - m[1][2] * m[3][1];
  float A2312 = m[1][2] * m[2][3] - m[1][3] * m[2][2];
  float A1312 = m[1][1] * m[2][3] - m[1][3] * m[2][1];
  float A1212 = m[1][1] * m[2][2] - m[1][2] * m[2][1];
  float A0313 = m[1][0] * m[3][3] - m[1][3] * m[3][0];
  float A0213 = m[1][0] * m[3][2] - m[1][2] * m[3][0];
  float A0312 = m[1][0] * m[2][3] - m[1][3] * m[2][0];
  float A0212 = m[1][0] * m[2][2] - m[1][2] * m[2][0];
  float A0113 = m[1][0] * m[3][1] - m[1][1] * m[3][0];
  float A0112 = m[1][0] * m[2][1] - m[1][1] * m[2][0];
  float det = m[0][0] * (m[1][1] * A2323 - m[1][2] * A1323 + m[1][3] * A1223) -
              m[0][1] * (m[1][0] * A2323 - m[1][2] * A0323 + m[1][3] * A0223) +
              m[0][2] * (m[1][0] * A1323 - m[1][1] * A0323 + m[1][3] * A0123) -
              m[0][3] * (m[1][0] * A1223 - m[1][1] * A0223 + m[1][2] * A0123);
  det = fdiv(1.0f, det);
  M4 r = identity();
  r.m[0][0] = det * (m[1][1] * A2323 - m[1][2] * A1323 + m[1][3] * A1223);
  r.m[0][1] = det * -(m[0][1] * A2323 - m[0][2] * A1323 + m[0][3] * A1223);
  r.m[0][2] = det * (m[0][1] * A2313 - m[0][2] * A1313 + m[0][3] * A1213);
  r.m[0][3] = det * -(m[0][1] * A2312 - m[0][2] * A1312 + m[0][3] * A1212);
  r.m[1][0] = det * -(m[1][0] * A2323 - m[1][2] * A0323 + m[1][3] * A0223);
  r.m[1][1] = det * (m[0][0] * A2323 - m[0][2] * A0323 + m[0][3] * A0223);
  r.m[1][2] = det * -(m[0][0] * A2313 - m[0][2] * A0313 + m[0][3] * A0213);
  r.m[1][3] = det * (m[0][0] * A2312 - m[0][2] * A0312 + m[0][3] * A0212);
  r.m[2][0] = det * (m[1][0] * A1323 - m[1][1] * A0323 + m[1][3] * A0123);
  r.m[2][1] = det * -(m[0][0] * A1323 - m[0][1] * A0323 + m[0][3] * A0123);
  r.m[2][2] = det * (m[0][0] * A1313 - m[0][1] * A0313 + m[0][3] * A0113);
  r.m[2][3] = det * -(m[0][0] * A1312 - m[0][1] * A0312 + m[0][3] * A0112);
  r.m[3][0] = det * -(m[1][0] * A1223 - m[1][1] * A0223 + m[1][2] * A0123);
  r.m[3][1] = det * (m[0][0] * A1223 - m[0][1] * A0223 + m[0][2] * A0123);
  r.m[3][2] = det * -(m[0][0] * A1213 - m[0][1] * A0213 + m[0][2] * A0113);
  r.m[3][3] = det * (m[0][0] * A1212 - m[0][1] * A0212 + m[0][2] * A0112);
  return r;
}
T3D_HD V4 row(const M4& a, int i) { V4 r = {a.m[i][0], a.m[i][1], a.m[i][2], a.m[i][3]}; return r; }
T3D_HD void set_row(M4& a, int i, V4 v) {
  a.m[i][0] = v.x;
  a.m[i][1] = v.y;
  a.m[i][2] = v.z;
  a.m[i][3] = v.w;
}
T3D_HD V3 row3(const M4& a) { return v3(a.m[3][0], a.m[3][1], a.m[3][2]); }   // WorldPosition, node.go:737-739
// Matrix4.Decompose: matrix.go:181-196.
T3D_HD void decompose(const M4& a, V3& pos, V3& scale, M4& rot) {
  pos = row3(a);
  rot = identity();
  set_row(rot, 0, unit4(row(a, 0)));
  set_row(rot, 1, unit4(row(a, 1)));
  set_row(rot, 2, unit4(row(a, 2)));
  M4 rt = transposed(rot);
  M4 in = mult(a, rt);
  scale = v3(magnitude4(row(in, 0)), magnitude4(row(in, 1)), magnitude4(row(in, 2)));
}
// Matrix4.MultVec: matrix.go:619-627 (translation row included).
T3D_HD V3 mult_vec(const M4& a, V3 v) {
  return v3(a.m[0][0] * v.x + a.m[1][0] * v.y + a.m[2][0] * v.z + a.m[3][0],
            a.m[0][1] * v.x + a.m[1][1] * v.y + a.m[2][1] * v.z + a.m[3][1],
            a.m[0][2] * v.x + a.m[1][2] * v.y + a.m[2][2] * v.z + a.m[3][2]);
}
// Matrix4.MultVecW: matrix.go:630-637, inlined in model.go:772-775, 795-798.
T3D_HD V4 mult_vec_w(const M4& a, V3 v) {
  V4 r;
  r.x = a.m[0][0] * v.x + a.m[1][0] * v.y + a.m[2][0] * v.z + a.m[3][0];
  r.y = a.m[0][1] * v.x + a.m[1][1] * v.y + a.m[2][1] * v.z + a.m[3][1];
  r.z = a.m[0][2] * v.x + a.m[1][2] * v.y + a.m[2][2] * v.z + a.m[3][2];
  r.w = a.m[0][3] * v.x + a.m[1][3] * v.y + a.m[2][3] * v.z + a.m[3][3];
  return r;
}
// Forward/Right/Up: matrix.go:142-176.
T3D_HD V3 forward(const M4& a) { return unit(v3(a.m[2][0], a.m[2][1], a.m[2][2])); }
T3D_HD V3 right(const M4& a) { return unit(v3(a.m[0][0], a.m[0][1], a.m[0][2])); }
T3D_HD V3 up(const M4& a) { return unit(v3(a.m[1][0], a.m[1][1], a.m[1][2])); }

// Vector3.Cross (vector.go:164-176), components in the reference's order.
T3D_HD V3 cross(V3 a, V3 o) {
  V3 r;
  r.z = a.x * o.y - o.x * a.y;
  r.y = a.z * o.x - o.z * a.x;
  r.x = a.y * o.z - o.y * a.z;
  return r;
}
// Vector3.Equals (vector.go:415-425): within 1e-4 per component.
T3D_HD bool equals3(V3 a, V3 b) {
  const float eps = 1e-4f;
  return !(fabsf(a.x - b.x) > eps || fabsf(a.y - b.y) > eps || fabsf(a.z - b.z) > eps);
}
// NewMatrix4LookAt (matrix.go:788-815).
T3D_HD M4 look_at(V3 from, V3 to, V3 up) {
  if (equals3(from, to)) return identity();
  const V3 z = unit(sub(to, from));
  up = unit(up);
  if (equals3(z, up) || equals3(z, v3(-up.x, -up.y, -up.z)))
    up = !equals3(up, v3(1, 0, 0)) ? v3(1, 0, 0) : v3(0, 0, 1);   // WorldRight, else WorldBackward (vector.go:15, 27)
  const V3 x = unit(cross(up, z));
  const V3 y = cross(z, x);
  M4 r = identity();
  r.m[0][0] = x.x; r.m[0][1] = x.y; r.m[0][2] = x.z;
  r.m[1][0] = y.x; r.m[1][1] = y.y; r.m[1][2] = y.z;
  r.m[2][0] = z.x; r.m[2][1] = z.y; r.m[2][2] = z.z;
  return r;
}
// The billboarded base transform of ProcessVertices (model.go:634-693): the
// look-at from the model's world position to the camera (up: WorldUp, or the
// camera's up for BillboardYUpCameraY), the locks in X, Y, Z order (each sets
// its row to the axis and zeroes that component of the other two rows before
// renormalising them, W kept), then Decompose's rotation times Scale(scale)
// times the look-at, with the position put back in row 3.
T3D_HD M4 billboard_base(const M4& model, V3 cam_pos, V3 up, bool lock_x, bool lock_y, bool lock_z) {
  M4 la = look_at(row3(model), cam_pos, up);
  const bool lock[3] = {lock_x, lock_y, lock_z};
  for (int axis = 0; axis < 3; axis++) {
    if (!lock[axis]) continue;
    V4 e = {axis == 0 ? 1.0f : 0.0f, axis == 1 ? 1.0f : 0.0f, axis == 2 ? 1.0f : 0.0f, 0.0f};
    set_row(la, axis, e);
    for (int r = 0; r < 3; r++) {
      if (r == axis) continue;
      V4 w = row(la, r);
      if (axis == 0) w.x = 0;
      else if (axis == 1) w.y = 0;
      else w.z = 0;
      set_row(la, r, unit4(w));
    }
  }
  V3 p, s;
  M4 rot;
  decompose(model, p, s, rot);
  M4 base = mult(mult(rot, scale4(s.x, s.y, s.z)), la);
  V4 r3 = {p.x, p.y, p.z, 1.0f};
  set_row(base, 3, r3);
  return base;
}

// invertedCamPos (model.go:629-630) and PointLight.workingPosition
// (light.go:175-189): r·x + p ⊙ (1/s) with (p, s, r) = Decompose(Inverted(M)).
T3D_HD V3 inverted_position(const M4& model, V3 world) {
  M4 inv = inverted(model);
  V3 p, s;
  M4 r;
  decompose(inv, p, s, r);
  V3 rs = mult_vec(r, world);
  V3 is = v3(fdiv(1.0f, s.x), fdiv(1.0f, s.y), fdiv(1.0f, s.z));
  return add(rs, hmul(p, is));
}

}  // namespace t3d
