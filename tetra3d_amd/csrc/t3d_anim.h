// t3d_anim.h — keyframe sampling of AnimationPlayer.updateValues
// (animation.go:63-160, 674-720) and the math it needs, callable from host
// and device code with the reference's rounding.
//
// Quaternion.Slerp calls math32.Sin and math32.Atan2, which are float32
// wrappers of Go's float64 math.Sin and math.Atan2 (math32/math32.go:137-139,
// 204-206).  Those follow the Cephes library (S. Moshier): sin.c's octant
// reduction with pi/4 split in three and its degree-6 sine and cosine
// polynomials, atan.c's rational approximation on [0, 0.66] with the
// tan(3pi/8) and pi/4 range splits.  They are restated below in float64,
// operation for operation (no fused multiply-adds: the library is built with
// -ffp-contract=off), so host and device round every step identically.
#pragma once

#include "t3d_math.h"

namespace t3d {

// Go math.Sin for |x| < 2^29.  Slerp's arguments lie in [0, pi/2]; Go's
// Payne-Hanek branch for |x| >= 2^29 is never reached and returns NaN here.
T3D_HD double go_sin(double x) {
  const double PI4A = 7.85398125648498535156e-1, PI4B = 3.77489470793079817668e-8,
               PI4C = 2.69515142907905952645e-15;
  if (x == 0 || x != x) return x;
  if (__builtin_isinf(x)) return __builtin_nan("");
  bool sign = false;
  if (x < 0) {
    x = -x;
    sign = true;
  }
  if (x >= 536870912.0) return __builtin_nan("");
  uint64_t j = (uint64_t)(x * 1.2732395447351628);   // 4/pi rounded to float64
  double y = (double)j;
  if (j & 1) {
    j++;
    y++;
  }
  j &= 7;
  const double z = ((x - y * PI4A) - y * PI4B) - y * PI4C;
  if (j > 3) {
    sign = !sign;
    j -= 4;
  }
  const double zz = z * z;
  if (j == 1 || j == 2)
    y = 1.0 - 0.5 * zz +
        zz * zz *
            ((((((-1.13585365213876817300e-11 * zz) + 2.08757008419747316778e-9) * zz - 2.75573141792967388112e-7) *
                   zz + 2.48015872888517045348e-5) * zz - 1.38888888888730564116e-3) * zz + 4.16666666666665929218e-2);
  else
    y = z + z * zz *
                ((((((1.58962301576546568060e-10 * zz) - 2.50507477628578072866e-8) * zz + 2.75573136213857245213e-6) *
                       zz - 1.98412698295895385996e-4) * zz + 8.33333333332211858878e-3) * zz - 1.66666666666666307295e-1);
  return sign ? -y : y;
}

T3D_HD double go_xatan(double x) {
  double z = x * x;
  z = z *
      ((((-8.750608600031904122785e-01 * z - 1.615753718733365076637e+01) * z - 7.500855792314704667340e+01) * z -
        1.228866684490136173410e+02) * z - 6.485021904942025371773e+01) /
      (((((z + 2.485846490142306297962e+01) * z + 1.650270098316988542046e+02) * z + 4.328810604912902668951e+02) * z +
        4.853903996359136964868e+02) * z + 1.945506571482613964425e+02);
  return x * z + x;
}
T3D_HD double go_satan(double x) {
  const double Morebits = 6.123233995736765886130e-17;   // pi/2 = 1.5707963267948966 + Morebits
  if (x <= 0.66) return go_xatan(x);
  if (x > 2.41421356237309504880) return 1.5707963267948966 - go_xatan(1 / x) + Morebits;
  return 0.7853981633974483 + go_xatan((x - 1) / (x + 1)) + 0.5 * Morebits;
}
T3D_HD double go_atan2(double y, double x) {
  const double PI = 3.141592653589793, PIO2 = 1.5707963267948966, PIO4 = 0.7853981633974483,
               PI3O4 = 2.356194490192345;
  if (y != y || x != x) return __builtin_nan("");
  if (y == 0) return (x >= 0 && !__builtin_signbit(x)) ? __builtin_copysign(0.0, y) : __builtin_copysign(PI, y);
  if (x == 0) return __builtin_copysign(PIO2, y);
  if (__builtin_isinf(x)) {
    if (x > 0) return __builtin_isinf(y) ? __builtin_copysign(PIO4, y) : __builtin_copysign(0.0, y);
    return __builtin_isinf(y) ? __builtin_copysign(PI3O4, y) : __builtin_copysign(PI, y);
  }
  if (__builtin_isinf(y)) return __builtin_copysign(PIO2, y);
  const double t = y / x;
  double q = t == 0 ? t : (t > 0 ? go_satan(t) : -go_satan(-t));   // math.Atan
  if (x < 0) return q <= 0 ? q + PI : q - PI;
  return q;
}

struct Quat {
  float x, y, z, w;
};

// Quaternion.Slerp: quaternion.go:26-59.
T3D_HD Quat slerp(Quat q, Quat e, float percent) {
  float c = q.x * e.x + q.y * e.y + q.z * e.z + q.w * e.w;   // Dot, quaternion.go:82-84
  if (c < 0) {
    e.w *= -1;
    e.x *= -1;
    e.y *= -1;
    e.z *= -1;
    c *= -1;
  }
  float k0, k1;
  if (c > 0.9999f) {
    k0 = 1 - percent;
    k1 = percent;
  } else {
    const float s = fsqrt(1 - (c * c));
    const float omega = (float)go_atan2((double)s, (double)c);
    const float inv = fdiv(1.0f, s);
    k0 = (float)go_sin((double)((1 - percent) * omega)) * inv;
    k1 = (float)go_sin((double)(percent * omega)) * inv;
  }
  q.w = q.w * k0 + e.w * k1;
  q.x = q.x * k0 + e.x * k1;
  q.y = q.y * k0 + e.y * k1;
  q.z = q.z * k0 + e.z * k1;
  return q;
}

// Quaternion.ToMatrix4: quaternion.go:135-183, the product of two matrices.
T3D_HD M4 quat_matrix(Quat q) {
  M4 a = identity(), b = identity();
  a.m[0][0] = q.w;  a.m[0][1] = q.z;  a.m[0][2] = -q.y; a.m[0][3] = q.x;
  a.m[1][0] = -q.z; a.m[1][1] = q.w;  a.m[1][2] = q.x;  a.m[1][3] = q.y;
  a.m[2][0] = q.y;  a.m[2][1] = -q.x; a.m[2][2] = q.w;  a.m[2][3] = q.z;
  a.m[3][0] = -q.x; a.m[3][1] = -q.y; a.m[3][2] = -q.z; a.m[3][3] = q.w;
  b.m[0][0] = q.w;  b.m[0][1] = q.z;  b.m[0][2] = -q.y; b.m[0][3] = -q.x;
  b.m[1][0] = -q.z; b.m[1][1] = q.w;  b.m[1][2] = q.x;  b.m[1][3] = -q.y;
  b.m[2][0] = q.y;  b.m[2][1] = -q.x; b.m[2][2] = q.w;  b.m[2][3] = -q.z;
  b.m[3][0] = q.x;  b.m[3][1] = q.y;  b.m[3][2] = q.z;  b.m[3][3] = q.w;
  return mult(a, b);
}

// Keyframe of a track at `time` (animation.go:63-93, 116-147): 0 no
// keyframes; 1 keyframe k as it is; 2 interpolate from k to k + 1 at t.
// The reference scans for the first keyframe at or after `time`; with times
// in non-decreasing order (checked at upload) a binary search finds the same.
T3D_HD int track_find(const t3d_anim_track& tr, float time, uint32_t& k, float& t) {
  const uint32_t n = tr.keyframe_count;
  if (n == 0) return 0;
  const float* T = tr.times;
  if (time <= T[0]) {
    k = 0;
    return 1;
  }
  if (time >= T[n - 1]) {
    k = n - 1;
    return 1;
  }
  uint32_t lo = 1, hi = n - 1;   // T[0] < time < T[n-1]
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (T[mid] < time) lo = mid + 1;
    else hi = mid;
  }
  const uint32_t first = lo - 1;
  if (time == T[lo]) {
    k = lo;
    return 1;
  }
  k = first;
  if (tr.interpolation == T3D_INTERPOLATION_CONSTANT) return 1;
  t = fdiv(time - T[first], T[lo] - T[first]);
  return 2;
}

// AnimationPlayer.forceUpdate without a blend, absolute motion
// (animation.go:902-983): each present track sets the bone's local position
// (Vector3.Lerp, vector.go:559-565), scale, or rotation (Slerp, then
// ToMatrix4; SetLocalRotation ignores an all-zero matrix, node.go:906-912).
T3D_HD void animate_bone(const t3d_anim_channel& ch, float time, float pos[3], float scl[3], M4& rot) {
  uint32_t k = 0;
  float t = 0;
  const t3d_anim_track* vt[2] = {&ch.position, &ch.scale};
  float* dst[2] = {pos, scl};
  for (int w = 0; w < 2; w++) {
    const int r = track_find(*vt[w], time, k, t);
    if (!r) continue;
    const float* a = vt[w]->values + 3 * (size_t)k;
    if (r == 1) {
      dst[w][0] = a[0];
      dst[w][1] = a[1];
      dst[w][2] = a[2];
    } else {
      dst[w][0] = a[0] + ((a[3] - a[0]) * t);
      dst[w][1] = a[1] + ((a[4] - a[1]) * t);
      dst[w][2] = a[2] + ((a[5] - a[2]) * t);
    }
  }
  const int r = track_find(ch.rotation, time, k, t);
  if (r) {
    const float* a = ch.rotation.values + 4 * (size_t)k;
    Quat q = {a[0], a[1], a[2], a[3]};
    if (r == 2) q = slerp(q, Quat{a[4], a[5], a[6], a[7]}, t);
    const M4 m = quat_matrix(q);
    bool zero = true;
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++) zero = zero && m.m[i][j] == 0.0f;
    if (!zero) rot = m;
  }
}

}  // namespace t3d
