/*
 * t3d_oracle.c — CPU restatement of Tetra3D's per-frame geometry path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Only the tests,
 * the smoke check and the benchmark's CPU-baseline leg load this code.
 *
 * Build: gcc -O2 -ffp-contract=off -fno-fast-math (see oracle/Makefile).  Go on
 * amd64 (GOAMD64=v1, the default) evaluates float32 expressions left to right,
 * rounds every operation to float32 and never fuses a*b+c; the same holds for
 * C on x86-64 with contraction off.  math32 functions are float64 round trips
 * (math32/math32.go:114-149, 476-478): Sqrt and Round are exact in both
 * languages; Tan/Cos/Atan only feed camera-level constants and use the C
 * library's double routines here (Go's differ by at most one double ulp,
 * which survives rounding to float32 only in rare boundary cases).
 *
 * Every function cites the reference file:line it restates.  Parity status of
 * each part is recorded next to its tests.  Parity unpinned: go_sin and
 * go_atan2 (Go's math.Sin/Atan2 restated; checked against the C library and
 * Go's special cases, not against Go's own outputs), the product order inside
 * q_to_m4 (its two factors commute), and go_stable_sort under the two
 * reference comparators that are not strict orders.
 */
#include "t3d_oracle.h"

#include <float.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MAXF32 3.40282346638528859811704183484516925440e+38f /* math.MaxFloat32 */
#define NBINS 512                                          /* sortingTriangles.go:166-168 */

typedef struct { float x, y, z; } v3;
typedef struct { float x, y, z, w; } v4;
typedef t3d_mat4 m4;

/* ------------------------------------------------------------------------ */
/* math32 (math32/math32.go)                                                 */
/* ------------------------------------------------------------------------ */

static float sqrt32(float x) { return (float)sqrt((double)x); }   /* math32.go:121-123 */
static float round32(float x) { return (float)round((double)x); } /* math32.go:109-111, half away from zero */
static float tan32(float x) { return (float)tan((double)x); }     /* math32.go:476-478 */
static float cos32(float x) { return (float)cos((double)x); }     /* math32.go:143-145 */
static float atan32(float x) { return (float)atan((double)x); }
static float abs32(float x) { return (float)fabs((double)x); }    /* math32.go:227-229 */
static float clamp32(float v, float lo, float hi) {               /* math32.go:38-45 */
  if (v < lo) return lo;
  else if (v > hi) return hi;
  return v;
}
static int isnan32(float x) { return x != x; }
static int isinf32(float x) { return x == INFINITY || x == -INFINITY; }
static const float PI32 = (float)3.14159265358979323846264338327950288419716939937510582097494459; /* math32.go Pi */

/* Go's int(float32) on amd64 (CVTTSS2SQ): truncation, and the integer
 * indefinite value (minimum int64) for NaN or out-of-range inputs. */
static int64_t go_int(float f) {
  if (!(f >= -9.2233720368547758e18f && f < 9.2233720368547758e18f)) return INT64_MIN;
  return (int64_t)f;
}

/* ------------------------------------------------------------------------ */
/* vectors (vector.go)                                                       */
/* ------------------------------------------------------------------------ */

static v3 v3_sub(v3 a, v3 b) { v3 r = {a.x - b.x, a.y - b.y, a.z - b.z}; return r; }
static v3 v3_add(v3 a, v3 b) { v3 r = {a.x + b.x, a.y + b.y, a.z + b.z}; return r; }
static v3 v3_mul(v3 a, v3 b) { v3 r = {a.x * b.x, a.y * b.y, a.z * b.z}; return r; }     /* vector.go:250-255 */
static float v3_dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }           /* vector.go:526-528 */
static float v3_mag(v3 a) { return sqrt32(a.x * a.x + a.y * a.y + a.z * a.z); }          /* vector.go:188-190 */
static float v3_dist_sq(v3 a, v3 b) {                                                    /* vector.go:242-247 */
  a.x -= b.x;
  a.y -= b.y;
  a.z -= b.z;
  return a.x * a.x + a.y * a.y + a.z * a.z;
}
static v3 v3_unit(v3 v) {                                                                /* vector.go:258-266 */
  float l = v3_mag(v);
  if (l < 1e-8f || l == 1) return v;
  v.x = v.x / l;
  v.y = v.y / l;
  v.z = v.z / l;
  return v;
}
static v4 v4_unit(v4 v) {                                                                /* vector.go:1869-1877, W kept */
  float l = sqrt32(v.x * v.x + v.y * v.y + v.z * v.z);
  if (l < 1e-8f || l == 1) return v;
  v.x = v.x / l;
  v.y = v.y / l;
  v.z = v.z / l;
  return v;
}
static float v4_mag(v4 v) { return sqrt32(v.x * v.x + v.y * v.y + v.z * v.z); }         /* vector.go:1880-1882 */
static v3 v3_from(const float* f) { v3 r = {f[0], f[1], f[2]}; return r; }
/* Vector3.Cross: vector.go:164-176, components in the reference's order. */
static v3 v3_cross(v3 a, v3 o) {
  v3 r;
  r.z = a.x * o.y - o.x * a.y;
  r.y = a.z * o.x - o.z * a.x;
  r.x = a.y * o.z - o.y * a.z;
  return r;
}
/* Vector3.Equals: vector.go:415-425, 1e-4 per component. */
static int v3_equals(v3 a, v3 b) {
  const float eps = 1e-4f;
  return !(fabsf(a.x - b.x) > eps || fabsf(a.y - b.y) > eps || fabsf(a.z - b.z) > eps);
}

/* ------------------------------------------------------------------------ */
/* matrices (matrix.go, matrix_none.go, matrix_amd64.s)                      */
/* ------------------------------------------------------------------------ */

static m4 m4_identity(void) {                                                            /* matrix.go:13-24 */
  m4 r;
  memset(&r, 0, sizeof r);
  r.m[0][0] = r.m[1][1] = r.m[2][2] = r.m[3][3] = 1;
  return r;
}
static m4 m4_zero(void) { m4 r; memset(&r, 0, sizeof r); return r; }

/* Mult: matrix_none.go:6-30; matrix_amd64.s:4-81 accumulates
 * ((a0·r0 + a1·r1) + a2·r2) + a3·r3 with MULPS/ADDPS, the same order. */
static m4 m4_mult(const m4* a, const m4* b) {
  m4 r;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++)
      r.m[i][j] = a->m[i][0] * b->m[0][j] + a->m[i][1] * b->m[1][j] + a->m[i][2] * b->m[2][j] +
                  a->m[i][3] * b->m[3][j];
  return r;
}
static m4 m4_transposed(const m4* a) {                                                   /* matrix.go:218-230 */
  m4 r = m4_identity();
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) r.m[i][j] = a->m[j][i];
  return r;
}
static m4 m4_translate(float x, float y, float z) {                                     /* matrix.go:82-88 */
  m4 r = m4_identity();
  r.m[3][0] = x;
  r.m[3][1] = y;
  r.m[3][2] = z;
  return r;
}
static m4 m4_scale(float x, float y, float z) {                                         /* matrix.go:89-96 */
  m4 r = m4_identity();
  r.m[0][0] = x;
  r.m[1][1] = y;
  r.m[2][2] = z;
  return r;
}
/* Inverted: matrix.go:255-307 (cofactor form), operation order as written. */
static m4 m4_inverted(const m4* mat) {
  const float(*m)[4] = mat->m;
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
  det = 1 / det;

  m4 r = m4_identity();
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
static v4 m4_row(const m4* a, int i) { v4 r = {a->m[i][0], a->m[i][1], a->m[i][2], a->m[i][3]}; return r; }
static void m4_set_row(m4* a, int i, v4 v) {
  a->m[i][0] = v.x;
  a->m[i][1] = v.y;
  a->m[i][2] = v.z;
  a->m[i][3] = v.w;
}
/* Decompose: matrix.go:181-196. */
static void m4_decompose(const m4* a, v3* pos, v3* scale, m4* rot) {
  pos->x = a->m[3][0];
  pos->y = a->m[3][1];
  pos->z = a->m[3][2];
  *rot = m4_identity();
  m4_set_row(rot, 0, v4_unit(m4_row(a, 0)));
  m4_set_row(rot, 1, v4_unit(m4_row(a, 1)));
  m4_set_row(rot, 2, v4_unit(m4_row(a, 2)));
  m4 rt = m4_transposed(rot);
  m4 in = m4_mult(a, &rt);
  scale->x = v4_mag(m4_row(&in, 0));
  scale->y = v4_mag(m4_row(&in, 1));
  scale->z = v4_mag(m4_row(&in, 2));
}
/* MultVec: matrix.go:619-627. */
static v3 m4_mult_vec(const m4* a, v3 v) {
  v3 r;
  r.x = a->m[0][0] * v.x + a->m[1][0] * v.y + a->m[2][0] * v.z + a->m[3][0];
  r.y = a->m[0][1] * v.x + a->m[1][1] * v.y + a->m[2][1] * v.z + a->m[3][1];
  r.z = a->m[0][2] * v.x + a->m[1][2] * v.y + a->m[2][2] * v.z + a->m[3][2];
  return r;
}
/* MultVecW: matrix.go:630-637; also the inlined form in model.go:772-775, 795-798. */
static v4 m4_mult_vec_w(const m4* a, v3 v) {
  v4 r;
  r.x = a->m[0][0] * v.x + a->m[1][0] * v.y + a->m[2][0] * v.z + a->m[3][0];
  r.y = a->m[0][1] * v.x + a->m[1][1] * v.y + a->m[2][1] * v.z + a->m[3][1];
  r.z = a->m[0][2] * v.x + a->m[1][2] * v.y + a->m[2][2] * v.z + a->m[3][2];
  r.w = a->m[0][3] * v.x + a->m[1][3] * v.y + a->m[2][3] * v.z + a->m[3][3];
  return r;
}
/* Forward/Right/Up: matrix.go:142-176. */
static v3 m4_forward(const m4* a) { v3 r = {a->m[2][0], a->m[2][1], a->m[2][2]}; return v3_unit(r); }
static v3 m4_right(const m4* a) { v3 r = {a->m[0][0], a->m[0][1], a->m[0][2]}; return v3_unit(r); }
static v3 m4_up(const m4* a) { v3 r = {a->m[1][0], a->m[1][1], a->m[1][2]}; return v3_unit(r); }
static v3 m4_row3(const m4* a) { v3 r = {a->m[3][0], a->m[3][1], a->m[3][2]}; return r; }  /* WorldPosition, node.go:737-739 */

/* NewProjectionMatrix4Perspective: matrix.go:588-604. */
static m4 projection_perspective(float fovy, float near, float far, float vw, float vh) {
  float aspect = vw / vh;
  float t = tan32(fovy * PI32 / 360);
  float b = -t;
  float r = t * aspect;
  float l = -r;
  m4 p = m4_zero();
  p.m[0][0] = (2 * near) / (r - l);
  p.m[0][2] = (r + l) / (r - l);
  p.m[1][1] = (2 * near) / (t - b);
  p.m[1][2] = (t + b) / (t - b);
  p.m[2][2] = -((far + near) / (far - near));
  p.m[2][3] = -((2 * far * near) / (far - near));
  p.m[3][2] = -1;
  return p;
}
/* NewProjectionMatrix4Orthographic: matrix.go:609-616. */
static m4 projection_orthographic(float near, float far, float right, float left, float top, float bottom) {
  m4 p = m4_zero();
  p.m[0][0] = 2 / (right - left);
  p.m[1][1] = 2 / (top - bottom);
  p.m[2][2] = -2 / (far - near);
  p.m[3][3] = 1;
  return p;
}

/* NewMatrix4LookAt: matrix.go:788-815. */
static m4 m4_lookat(v3 from, v3 to, v3 up) {
  if (v3_equals(from, to)) return m4_identity();
  v3 z = v3_unit(v3_sub(to, from));
  up = v3_unit(up);
  v3 nup = {-up.x, -up.y, -up.z};
  if (v3_equals(z, up) || v3_equals(z, nup)) {
    const v3 right = {1, 0, 0}, backward = {0, 0, 1};   /* WorldRight, WorldBackward (vector.go:15, 27) */
    up = !v3_equals(up, right) ? right : backward;
  }
  v3 x = v3_unit(v3_cross(up, z));
  v3 y = v3_cross(z, x);
  m4 r = m4_identity();
  r.m[0][0] = x.x; r.m[0][1] = x.y; r.m[0][2] = x.z;
  r.m[1][0] = y.x; r.m[1][1] = y.y; r.m[1][2] = y.z;
  r.m[2][0] = z.x; r.m[2][1] = z.y; r.m[2][2] = z.z;
  return r;
}

/* The billboarded base transform of ProcessVertices: model.go:634-693.  The
 * look-at from the model's world position to the camera, the locks in X, Y, Z
 * order (each sets its row to the axis and zeroes that component of the other
 * two rows before renormalising them, W kept), then Decompose's rotation times
 * Scale(scale) times the look-at, with the position put back in row 3. */
static m4 billboard_base(const m4* M, const t3d_material_desc* mat, v3 cam_pos, v3 cam_up) {
  v3 up = {0, 1, 0};                                                /* WorldUp, vector.go:21 */
  if (mat->billboard_up == T3D_BILLBOARD_UP_CAMERA_Y) up = cam_up;
  m4 la = m4_lookat(m4_row3(M), cam_pos, up);
  for (int axis = 0; axis < 3; axis++) {
    if (!mat->billboard_lock[axis]) continue;
    v4 e = {axis == 0 ? 1.0f : 0.0f, axis == 1 ? 1.0f : 0.0f, axis == 2 ? 1.0f : 0.0f, 0.0f};
    m4_set_row(&la, axis, e);
    for (int r = 0; r < 3; r++) {
      if (r == axis) continue;
      v4 row = m4_row(&la, r);
      if (axis == 0) row.x = 0;
      else if (axis == 1) row.y = 0;
      else row.z = 0;
      m4_set_row(&la, r, v4_unit(row));
    }
  }
  v3 p, s;
  m4 rot;
  m4_decompose(M, &p, &s, &rot);
  m4 sc = m4_scale(s.x, s.y, s.z);
  m4 base = m4_mult(&rot, &sc);
  base = m4_mult(&base, &la);
  v4 r3 = {p.x, p.y, p.z, 1};
  m4_set_row(&base, 3, r3);
  return base;
}

/* ------------------------------------------------------------------------ */
/* Go sort.SliceStable (Go 1.24 sort package: stable_func, insertionSort_func, */
/* symMerge_func, rotate_func, swapRange_func).  Needed only where the       */
/* reference passes a comparator that is not a strict weak order            */
/* (camera.go:1458-1467, 1483-1495); elsewhere any stable sort agrees.       */
/* ------------------------------------------------------------------------ */

typedef int (*less_fn)(void* env, int32_t a, int32_t b);
typedef struct { int32_t* d; less_fn less; void* env; } go_sorter;

static int gs_less(go_sorter* s, int i, int j) { return s->less(s->env, s->d[i], s->d[j]); }
static void gs_swap(go_sorter* s, int i, int j) { int32_t t = s->d[i]; s->d[i] = s->d[j]; s->d[j] = t; }
static void gs_insertion(go_sorter* s, int a, int b) {
  for (int i = a + 1; i < b; i++)
    for (int j = i; j > a && gs_less(s, j, j - 1); j--) gs_swap(s, j, j - 1);
}
static void gs_swap_range(go_sorter* s, int a, int b, int n) {
  for (int i = 0; i < n; i++) gs_swap(s, a + i, b + i);
}
static void gs_rotate(go_sorter* s, int a, int m, int b) {
  int i = m - a, j = b - m;
  while (i != j) {
    if (i > j) {
      gs_swap_range(s, m - i, m, j);
      i -= j;
    } else {
      gs_swap_range(s, m - i, m + j - i, i);
      j -= i;
    }
  }
  gs_swap_range(s, m - i, m, i);
}
static void gs_sym_merge(go_sorter* s, int a, int m, int b) {
  if (m - a == 1) {
    int i = m, j = b;
    while (i < j) {
      int h = (int)((unsigned)(i + j) >> 1);
      if (gs_less(s, h, a)) i = h + 1;
      else j = h;
    }
    for (int k = a; k < i - 1; k++) gs_swap(s, k, k + 1);
    return;
  }
  if (b - m == 1) {
    int i = a, j = m;
    while (i < j) {
      int h = (int)((unsigned)(i + j) >> 1);
      if (!gs_less(s, m, h)) i = h + 1;
      else j = h;
    }
    for (int k = m; k > i; k--) gs_swap(s, k, k - 1);
    return;
  }
  int mid = (int)((unsigned)(a + b) >> 1);
  int n = mid + m;
  int start, r;
  if (m > mid) {
    start = n - b;
    r = mid;
  } else {
    start = a;
    r = m;
  }
  int p = n - 1;
  while (start < r) {
    int c = (int)((unsigned)(start + r) >> 1);
    if (!gs_less(s, p - c, c)) start = c + 1;
    else r = c;
  }
  int end = n - start;
  if (start < m && m < end) gs_rotate(s, start, m, end);
  if (a < start && start < mid) gs_sym_merge(s, a, start, mid);
  if (mid < end && end < b) gs_sym_merge(s, mid, end, b);
}
static void go_stable_sort(int32_t* d, int n, less_fn less, void* env) {
  go_sorter s = {d, less, env};
  int block = 20, a = 0, b = block;
  while (b <= n) {
    gs_insertion(&s, a, b);
    a = b;
    b += block;
  }
  gs_insertion(&s, a, n);
  while (block < n) {
    a = 0;
    b = 2 * block;
    while (b <= n) {
      gs_sym_merge(&s, a, a + block, b);
      a = b;
      b += 2 * block;
    }
    int m = a + block;
    if (m < n) gs_sym_merge(&s, a, m, n);
    block *= 2;
  }
}

/* ------------------------------------------------------------------------ */
/* oracle context: registered meshes and light volumes                       */
/* ------------------------------------------------------------------------ */

typedef struct {
  uint64_t handle;
  t3d_mesh_desc d;   /* pointers below own copies */
  float* pos;
  float* nrm;
  float* col;
  float* uv;
  uint16_t* bones;
  float* weights;
  float* sk_pos;
  float* sk_nrm;
  uint32_t* idx;
  t3d_part_range* parts;
  v3* tri_center;    /* Triangle.Center (mesh.go:1980-1982) */
} omesh;

typedef struct {
  uint64_t handle;
  float* cells;
  uint32_t n;
  uint32_t* first;        /* cell light lists (LightVolumeCell.useLights): [n + 1] offsets, NULL for none */
  uint32_t* index;
  uint32_t need;          /* 1 + the largest light index the lists name */
} ovolume;

typedef struct {
  uint64_t handle;
  t3d_anim_channel* ch;   /* tracks point into the copies below */
  float* data;
  uint32_t n;
} oanim;

typedef struct {
  uint64_t handle;
  uint32_t w, h;
  uint8_t* px;            /* RGBA8, premultiplied, row-major */
} otexture;

struct t3do_ctx {
  omesh* meshes;
  int nmeshes, cap_meshes;
  ovolume* vols;
  int nvols, cap_vols;
  oanim* anims;
  int nanims, cap_anims;
  otexture* texs;
  int ntexs, cap_texs;
  char err[512];
};

static int set_err(t3do_ctx* c, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(c->err, sizeof c->err, fmt, ap);
  va_end(ap);
  return T3D_ERR_INVALID_ARGUMENT;
}

t3do_ctx* t3do_create(void) { return (t3do_ctx*)calloc(1, sizeof(t3do_ctx)); }

static void omesh_free(omesh* m) {
  free(m->pos); free(m->nrm); free(m->col); free(m->uv); free(m->bones); free(m->weights);
  free(m->sk_pos); free(m->sk_nrm); free(m->idx); free(m->parts); free(m->tri_center);
}

void t3do_destroy(t3do_ctx* c) {
  if (!c) return;
  for (int i = 0; i < c->nmeshes; i++) omesh_free(&c->meshes[i]);
  for (int i = 0; i < c->nvols; i++) {
    free(c->vols[i].cells);
    free(c->vols[i].first);
    free(c->vols[i].index);
  }
  for (int i = 0; i < c->nanims; i++) {
    free(c->anims[i].ch);
    free(c->anims[i].data);
  }
  for (int i = 0; i < c->ntexs; i++) free(c->texs[i].px);
  free(c->meshes);
  free(c->vols);
  free(c->anims);
  free(c->texs);
  free(c);
}

const char* t3do_last_error(const t3do_ctx* c) { return c ? c->err : "null context"; }

static void* dup_mem(const void* p, size_t n) {
  if (!p || n == 0) return NULL;
  void* r = malloc(n);
  memcpy(r, p, n);
  return r;
}

int t3do_mesh_register(t3do_ctx* c, uint64_t handle, const t3d_mesh_desc* d) {
  if (!d || !d->positions || !d->normals || (!d->indices && d->triangle_count)) return set_err(c, "mesh: missing arrays");
  for (uint32_t t = 0; t < 3u * d->triangle_count; t++)
    if (d->indices[t] >= d->vertex_count) return set_err(c, "mesh: index out of range");
  if (c->nmeshes == c->cap_meshes) {
    c->cap_meshes = c->cap_meshes ? 2 * c->cap_meshes : 16;
    c->meshes = (omesh*)realloc(c->meshes, sizeof(omesh) * c->cap_meshes);
  }
  omesh* m = &c->meshes[c->nmeshes++];
  memset(m, 0, sizeof *m);
  m->handle = handle;
  m->d = *d;
  size_t V = d->vertex_count, T = d->triangle_count, K = d->shape_key_count;
  m->pos = (float*)dup_mem(d->positions, V * 12);
  m->nrm = (float*)dup_mem(d->normals, V * 12);
  m->col = (float*)dup_mem(d->colors, V * 16);
  m->uv = (float*)dup_mem(d->uvs, V * 8);
  m->bones = (uint16_t*)dup_mem(d->bone_indices, V * 8);
  m->weights = (float*)dup_mem(d->bone_weights, V * 16);
  m->sk_pos = (float*)dup_mem(d->shape_key_positions, K * V * 12);
  m->sk_nrm = (float*)dup_mem(d->shape_key_normals, K * V * 12);
  m->idx = (uint32_t*)dup_mem(d->indices, T * 12);
  m->parts = (t3d_part_range*)dup_mem(d->parts, sizeof(t3d_part_range) * d->part_count);
  m->tri_center = (v3*)malloc(sizeof(v3) * (T ? T : 1));
  /* RecalculateCenter: mesh.go:1980-1982 */
  for (size_t t = 0; t < T; t++) {
    const float* a = m->pos + 3 * (size_t)m->idx[3 * t + 0];
    const float* b = m->pos + 3 * (size_t)m->idx[3 * t + 1];
    const float* cc = m->pos + 3 * (size_t)m->idx[3 * t + 2];
    m->tri_center[t].x = (a[0] + b[0] + cc[0]) / 3;
    m->tri_center[t].y = (a[1] + b[1] + cc[1]) / 3;
    m->tri_center[t].z = (a[2] + b[2] + cc[2]) / 3;
  }
  return T3D_OK;
}

int t3do_volume_register(t3do_ctx* c, uint64_t handle, const float* cells, uint32_t n) {
  ovolume* v = NULL;
  for (int i = 0; i < c->nvols && !v; i++)   /* registering a handle again replaces its cells and lists */
    if (c->vols[i].handle == handle) {
      v = &c->vols[i];
      free(v->cells);
      free(v->first);
      free(v->index);
    }
  if (!v) {
    if (c->nvols == c->cap_vols) {
      c->cap_vols = c->cap_vols ? 2 * c->cap_vols : 8;
      c->vols = (ovolume*)realloc(c->vols, sizeof(ovolume) * c->cap_vols);
    }
    v = &c->vols[c->nvols++];
  }
  v->handle = handle;
  v->n = n;
  v->cells = (float*)dup_mem(cells, (size_t)n * 16);
  v->first = NULL;
  v->index = NULL;
  v->need = 0;
  return T3D_OK;
}

static ovolume* find_vol(t3do_ctx* c, uint64_t h);

/* LightVolumeCell.SetUseLights / ClearUseLights of every cell (light.go:1508-1518). */
int t3do_volume_set_cell_lights(t3do_ctx* c, uint64_t handle, const uint32_t* first, const uint32_t* index) {
  ovolume* v = find_vol(c, handle);
  if (!v) return set_err(c, "volume_set_cell_lights: unknown handle"), T3D_ERR_NOT_FOUND;
  free(v->first);
  free(v->index);
  v->first = NULL;
  v->index = NULL;
  v->need = 0;
  if (!first) return T3D_OK;
  uint32_t total = first[v->n];
  v->first = (uint32_t*)dup_mem(first, sizeof(uint32_t) * ((size_t)v->n + 1));
  v->index = (uint32_t*)dup_mem(index, sizeof(uint32_t) * (total ? total : 1));
  for (uint32_t j = 0; j < total; j++)
    if (index[j] + 1 > v->need) v->need = index[j] + 1;
  return T3D_OK;
}

int t3do_animation_register(t3do_ctx* c, uint64_t handle, const t3d_anim_channel* ch, uint32_t n) {
  size_t floats = 0;
  for (uint32_t i = 0; i < n; i++) {
    const t3d_anim_track* t[3] = {&ch[i].position, &ch[i].scale, &ch[i].rotation};
    for (int k = 0; k < 3; k++) floats += (size_t)t[k]->keyframe_count * (k == 2 ? 5 : 4);
  }
  if (c->nanims == c->cap_anims) {
    c->cap_anims = c->cap_anims ? 2 * c->cap_anims : 8;
    c->anims = (oanim*)realloc(c->anims, sizeof(oanim) * c->cap_anims);
  }
  oanim* a = &c->anims[c->nanims++];
  a->handle = handle;
  a->n = n;
  a->ch = (t3d_anim_channel*)calloc(n ? n : 1, sizeof(t3d_anim_channel));
  a->data = (float*)malloc(sizeof(float) * (floats ? floats : 1));
  float* d = a->data;
  for (uint32_t i = 0; i < n; i++) {
    a->ch[i] = ch[i];
    t3d_anim_track* t[3] = {&a->ch[i].position, &a->ch[i].scale, &a->ch[i].rotation};
    for (int k = 0; k < 3; k++) {
      const uint32_t m = t[k]->keyframe_count, w = k == 2 ? 4 : 3;
      if (!m) continue;
      memcpy(d, t[k]->times, sizeof(float) * m);
      t[k]->times = d;
      d += m;
      memcpy(d, t[k]->values, sizeof(float) * m * w);
      t[k]->values = d;
      d += (size_t)m * w;
    }
  }
  return T3D_OK;
}
static oanim* find_anim(t3do_ctx* c, uint64_t h) {
  for (int i = 0; i < c->nanims; i++)
    if (c->anims[i].handle == h) return &c->anims[i];
  return NULL;
}

static omesh* find_mesh(t3do_ctx* c, uint64_t h) {
  for (int i = 0; i < c->nmeshes; i++)
    if (c->meshes[i].handle == h) return &c->meshes[i];
  return NULL;
}
static ovolume* find_vol(t3do_ctx* c, uint64_t h) {
  for (int i = 0; i < c->nvols; i++)
    if (c->vols[i].handle == h) return &c->vols[i];
  return NULL;
}

/* ------------------------------------------------------------------------ */
/* camera                                                                    */
/* ------------------------------------------------------------------------ */

typedef struct {
  const t3d_camera_desc* d;
  v3 pos;                 /* WorldPosition */
  m4 view, proj, vp;
  v3 forward, right, up;  /* cached at Clear (camera.go:1085-1088) */
  float sf_x, sf_y, sf_tang;
  float aspect;
  int only_visible;       /* ProcessVertices' processOnlyVisible for this call's pairs (model.go:600) */
} ocam;

/* ViewMatrix: camera.go:638-648. */
static m4 view_matrix(const t3d_camera_desc* cam) {
  v3 cp = m4_row3(&cam->transform);
  m4 t = m4_translate(-cp.x, -cp.y, -cp.z);
  m4 rt = m4_transposed(&cam->world_rotation);
  return m4_mult(&t, &rt);
}

/* Projection + sphere factors: camera.go:651-679 (3.1415 quirk at 660). */
static void cam_setup(ocam* c, const t3d_camera_desc* d) {
  c->d = d;
  c->pos = m4_row3(&d->transform);
  c->aspect = (float)d->width / (float)d->height;                   /* camera.go:986-992 */
  float angle = d->fov * 3.1415f / 360;
  c->sf_tang = tan32(angle);
  c->sf_y = 1.0f / cos32(angle);
  c->sf_x = 1.0f / cos32(atan32(c->sf_tang * c->aspect));
  if (d->perspective) {
    c->proj = projection_perspective(d->fov, d->near, d->far, (float)d->width, (float)d->height);
  } else {
    float asr = (float)d->height / (float)d->width;
    c->proj = projection_orthographic(d->near, d->far, 1 * d->ortho_scale, -1 * d->ortho_scale,
                                      asr * d->ortho_scale, -asr * d->ortho_scale);
  }
  c->view = view_matrix(d);
  c->vp = m4_mult(&c->view, &c->proj);                              /* camera.go:1329 */
  c->forward = m4_forward(&d->world_rotation);
  c->right = m4_right(&d->world_rotation);
  c->up = m4_up(&d->world_rotation);
}

/* SphereInFrustum: camera.go:933-983. */
static int sphere_in_frustum(const ocam* c, v3 position, float radius) {
  v3 diff = v3_sub(position, c->pos);
  v3 nf = {-c->forward.x, -c->forward.y, -c->forward.z};
  float pcZ = v3_dot(diff, nf);
  const t3d_camera_desc* d = c->d;
  if (pcZ > d->far + radius || pcZ < d->near - radius) return 0;
  if (d->perspective) {
    float dd = c->sf_y * radius;
    pcZ *= c->sf_tang;
    float pcY = v3_dot(diff, c->up);
    if (pcY > pcZ + dd || pcY < -pcZ - dd) return 0;
    pcZ *= c->aspect;
    dd = c->sf_x * radius;
    float pcX = v3_dot(diff, c->right);
    if (pcX > pcZ + dd || pcX < -pcZ - dd) return 0;
  } else {
    float width = d->ortho_scale;
    float height = width / c->aspect;
    float pcY = v3_dot(diff, c->up);
    if (-height / 2 - radius > pcY || pcY > height / 2 + radius) return 0;
    float pcX = v3_dot(diff, c->right);
    if (-width / 2 - radius > pcX || pcX > width / 2 + radius) return 0;
  }
  return 1;
}

/* Model.onTransformUpdate frustum sphere: model.go:142-182. */
static void model_sphere(const m4* transform, int skinned, const float* dmin, const float* dmax, v3* pos,
                         float* radius) {
  v3 position, scale;
  m4 rotation;
  m4_decompose(transform, &position, &scale, &rotation);
  v3 center = {0, 0, 0};
  if (!skinned) {                                                   /* Dimensions.Center, mesh.go:46-52 */
    center.x = (dmax[0] + dmin[0]) / 2;
    center.y = (dmax[1] + dmin[1]) / 2;
    center.z = (dmax[2] + dmin[2]) / 2;
  }
  center = m4_mult_vec(&rotation, center);
  position.x += center.x * scale.x;
  position.y += center.y * scale.y;
  position.z += center.z * scale.z;
  v3 mn = {dmin[0] * scale.x, dmin[1] * scale.y, dmin[2] * scale.z};
  v3 mx = {dmax[0] * scale.x, dmax[1] * scale.y, dmax[2] * scale.z};
  *pos = position;
  *radius = v3_mag(v3_sub(mx, mn)) / 2;                             /* Dimensions.MaxSpan, mesh.go:41-43 */
}

/* invertedCamPos: model.go:629-630 (also PointLight.beginModel, light.go:175-189). */
static v3 inverted_pos(const m4* model_transform, v3 world) {
  m4 inv = m4_inverted(model_transform);
  v3 p, s;
  m4 r;
  m4_decompose(&inv, &p, &s, &r);
  v3 rs = m4_mult_vec(&r, world);
  v3 is = {1 / s.x, 1 / s.y, 1 / s.z};
  return v3_add(rs, v3_mul(p, is));
}

/* ------------------------------------------------------------------------ */
/* lights (light.go)                                                         */
/* ------------------------------------------------------------------------ */

typedef struct olight_s {
  int type;
  int visible;
  float color[3];
  float energy;
  float range, range_sq;    /* SetRange: light.go:273-276 */
  v3 world_pos;
  float amb[3];             /* AmbientLight.result (light.go:110-112) */
  v3 forward;               /* DirectionalLight.workingForward (light.go:322-324) */
  /* per model (beginModel) */
  v3 working_pos;           /* PointLight.workingPosition */
  m4 working_rot;           /* DirectionalLight.workingModelRotation */
  /* volume */
  const float* cells;
  uint32_t ncells;
  v3 vmin, vmax, cell, offset, bias;
  int cx, cy, cz;
  const uint32_t* cfirst;          /* cell light lists of the volume (NULL: none) */
  const uint32_t* cindex;
  struct olight_s* use;            /* useLightsRenderSet, light.go:392 (owned) */
  const t3d_light_desc* use_descs;
  int nuse;
} olight;

static void light_from_desc(t3do_ctx* c, olight* l, const t3d_light_desc* d) {
  memset(l, 0, sizeof *l);
  l->type = d->type;
  l->visible = d->visible;
  l->color[0] = d->color[0];
  l->color[1] = d->color[1];
  l->color[2] = d->color[2];
  l->energy = d->energy;
  l->range = d->range;
  l->range_sq = d->range * d->range;
  l->world_pos = m4_row3(&d->transform);
  if (d->type == T3D_LIGHT_VOLUME) {
    ovolume* v = find_vol(c, d->volume);
    l->cells = v ? v->cells : NULL;
    l->ncells = v ? v->n : 0;
    l->vmin = v3_from(d->volume_min);
    l->vmax = v3_from(d->volume_max);
    l->cell = v3_from(d->cell_size);
    l->offset = v3_from(d->offset);
    l->bias = v3_from(d->light_bias);
    l->cx = d->cell_count[0];
    l->cy = d->cell_count[1];
    l->cz = d->cell_count[2];
    if (v && v->first) {
      l->cfirst = v->first;
      l->cindex = v->index;
      l->nuse = (int)d->use_light_count;
      l->use_descs = d->use_lights;
      l->use = (olight*)calloc(l->nuse + 1, sizeof(olight));
      for (int i = 0; i < l->nuse; i++) light_from_desc(c, &l->use[i], &d->use_lights[i]);
    }
  }
}

static void light_release(olight* l) {
  for (int i = 0; l->use && i < l->nuse; i++) light_release(&l->use[i]);
  free(l->use);
  l->use = NULL;
}

/* beginRender: light.go:110-112 (ambient), 322-324 (directional). */
static void light_begin_render(olight* l, const t3d_light_desc* d) {
  if (l->type == T3D_LIGHT_AMBIENT) {
    l->amb[0] = l->color[0] * l->energy;
    l->amb[1] = l->color[1] * l->energy;
    l->amb[2] = l->color[2] * l->energy;
  } else if (l->type == T3D_LIGHT_DIRECTIONAL) {
    l->forward = m4_forward(&d->world_rotation);
  } else if (l->type == T3D_LIGHT_VOLUME) {                          /* light.go:566-571 */
    for (int i = 0; i < l->nuse; i++) light_begin_render(&l->use[i], &l->use_descs[i]);
  }
}

/* beginModel: light.go:175-189 (point), 326-330 (directional). */
static void light_begin_model(olight* l, const t3d_model_desc* model) {
  if (l->type == T3D_LIGHT_POINT) {
    if (model->skinned) l->working_pos = l->world_pos;
    else l->working_pos = inverted_pos(&model->transform, l->world_pos);
  } else if (l->type == T3D_LIGHT_DIRECTIONAL) {
    if (!model->skinned) {
      m4 inv = m4_inverted(&model->world_rotation);
      l->working_rot = m4_transposed(&inv);
    }
  } else if (l->type == T3D_LIGHT_VOLUME) {                          /* light.go:574-580 */
    for (int i = 0; i < l->nuse; i++) light_begin_model(&l->use[i], model);
  }
}

typedef struct {
  v3* alt_pos;   /* globalMeshAlteredVertexPositions */
  v3* alt_nrm;   /* globalMeshAlteredVertexNormals */
  v4* clip;      /* globalVertexTransforms */
  float* lights; /* mesh.vertexLights.colors (RGBA) */
  v3* tnrm;      /* globalVertexTransformedNormals (Camera.RenderNormals) */
  float* udepth; /* globalVertexDepthUnbillboarded */
} oscratch;

/* getLightCell with convertToXYZIndices: light.go:717-724, 1267-1272. */
static int64_t volume_cell_index(const olight* l, v3 p) {
  int64_t x = go_int(round32((p.x - l->offset.x - l->vmin.x - (l->cell.x / 2)) / l->cell.x));
  int64_t y = go_int(round32((p.y - l->offset.y - l->vmin.y - (l->cell.y / 2)) / l->cell.y));
  int64_t z = go_int(round32((p.z - l->offset.z - l->vmin.z - (l->cell.z / 2)) / l->cell.z));
  if (x < 0 || y < 0 || z < 0 || x >= l->cx || y >= l->cy || z >= l->cz) return -1;
  int64_t i = x + (y * l->cx) + (z * l->cx * l->cy);
  if (!l->cells || i >= (int64_t)l->ncells) return -1;
  return i;
}

static const float* volume_cell(const olight* l, v3 p) {
  int64_t i = volume_cell_index(l, p);
  return i < 0 ? NULL : l->cells + 4 * i;
}

static int dims_contains(const olight* l, v3 p) {                  /* mesh.go:80-82 */
  return p.x >= l->vmin.x && p.y >= l->vmin.y && p.z >= l->vmin.z && p.x <= l->vmax.x && p.y <= l->vmax.y &&
         p.z <= l->vmax.z;
}

/* VertexLight for one vertex and one light: light.go:116-120 (ambient),
 * 191-258 (point), 332-360 (directional), 583-663 (volume). */
static void vertex_light(const olight* l, uint32_t v, const t3d_model_desc* model, const t3d_material_desc* mat,
                         const oscratch* s, float* out) {
  int lighting_mode = (mat && mat->present) ? mat->lighting_mode : T3D_LIGHTING_DEFAULT;
  if (l->type == T3D_LIGHT_AMBIENT) {
    out[0] += l->amb[0];
    out[1] += l->amb[1];
    out[2] += l->amb[2];
  } else if (l->type == T3D_LIGHT_POINT) {
    v3 vp = s->alt_pos[v];
    v3 vn = s->alt_nrm[v];
    float distance = v3_dist_sq(l->working_pos, vp);
    if (l->range > 0) {
      if (distance > l->range_sq) return;
    } else {
      if (1 / distance * l->energy < 0.001f) return;
    }
    v3 lv = v3_unit(v3_sub(l->working_pos, vp));
    if (lighting_mode == T3D_LIGHTING_FIXED_NORMALS) vn = lv;
    float diffuse = v3_dot(vn, lv);
    if (lighting_mode == T3D_LIGHTING_DOUBLE_SIDED) diffuse = abs32(diffuse);
    if (diffuse > 0) {
      float df = diffuse * (1.0f / (1.0f + (0.1f * distance))) * 2;
      if (l->range_sq > 0) {
        float dc = clamp32((l->range_sq - distance) / distance, 0, 1);
        df *= dc;
      }
      out[0] += l->color[0] * df * l->energy;
      out[1] += l->color[1] * df * l->energy;
      out[2] += l->color[2] * df * l->energy;
    }
  } else if (l->type == T3D_LIGHT_DIRECTIONAL) {
    v3 n;
    if (lighting_mode == T3D_LIGHTING_FIXED_NORMALS) n = l->forward;
    else if (model->skinned) n = s->alt_nrm[v];
    else n = m4_mult_vec(&l->working_rot, s->alt_nrm[v]);
    float df = v3_dot(n, l->forward);
    if (lighting_mode == T3D_LIGHTING_DOUBLE_SIDED) df = abs32(df);
    if (df <= 0) return;
    out[0] += l->color[0] * df * l->energy;
    out[1] += l->color[1] * df * l->energy;
    out[2] += l->color[2] * df * l->energy;
  } else if (l->type == T3D_LIGHT_VOLUME) {
    int mode = (mat && mat->present) ? mat->light_volume_shading_mode : 0;
    if (mode == T3D_LV_OFF) return;                                 /* LightVolume.Light, light.go:671-673 */
    float cc[4] = {0, 0, 0, 1};
    v3 halfc = {l->cell.x / 2, l->cell.y / 2, l->cell.z / 2};
    if (mode == T3D_LV_PER_OBJECT) {
      const float* cell = volume_cell(l, m4_row3(&model->transform));
      if (cell) {
        cc[0] = cell[0];
        cc[1] = cell[1];
        cc[2] = cell[2];
      }
    } else {
      v3 vpos = {0, 0, 0}, vpn = {0, 0, 0};
      if (model->skinned) {
        if (mode == T3D_LV_PER_VERTEX_WITH_NORMAL) vpn = v3_add(v3_add(vpos, v3_mul(s->alt_nrm[v], halfc)), l->bias);
        else vpos = v3_add(s->alt_pos[v], l->bias);
      } else {
        if (mode == T3D_LV_PER_VERTEX_WITH_NORMAL)
          vpn = v3_add(m4_mult_vec(&model->transform, v3_add(s->alt_pos[v], v3_mul(s->alt_nrm[v], halfc))), l->bias);
        else vpos = v3_add(m4_mult_vec(&model->transform, s->alt_pos[v]), l->bias);
      }
      if (!dims_contains(l, vpos)) return;
      cc[0] = cc[1] = cc[2] = 0;
      int64_t ci = volume_cell_index(l, mode == T3D_LV_PER_VERTEX_WITH_NORMAL ? vpn : vpos);
      if (ci >= 0) {
        const float* cell = l->cells + 4 * ci;
        cc[0] = cell[0];
        cc[1] = cell[1];
        cc[2] = cell[2];
        cc[3] = cell[3];
        for (uint32_t j = l->cfirst ? l->cfirst[ci] : 0; l->cfirst && j < l->cfirst[ci + 1]; j++)  /* light.go:650-653 */
          vertex_light(&l->use[l->cindex[j]], v, model, mat, s, out);
      }
    }
    out[0] += l->color[0] * cc[0] * l->energy;
    out[1] += l->color[1] * cc[1] * l->energy;
    out[2] += l->color[2] * cc[2] * l->energy;
  }
}

/* ------------------------------------------------------------------------ */
/* bucket sort (sortingTriangles.go)                                         */
/* ------------------------------------------------------------------------ */

/* Sort's bin formula: sortingTriangles.go:38-64. */
static uint32_t bucket_bin(float depth, float minRange, float maxRange, int sort_mode) {
  float rangeDiff = maxRange - minRange;
  if (rangeDiff == 0) {
    maxRange += 0.001f;
    rangeDiff += 0.001f;
  }
  int64_t target = 0;
  if (sort_mode != T3D_SORT_NONE) {
    float d = (depth - minRange) / rangeDiff * (float)NBINS;
    float t = clamp32(d, 0, (float)(NBINS - 1));
    target = go_int(t);
  }
  return (uint32_t)target;
}

/* Sort + ForEach (sortingTriangles.go:38-64, 93-125): returns the visit order. */
static void bucket_order(const float* depth, uint32_t n, float mn, float mx, int sort_mode, uint32_t* order,
                         uint32_t* bin_scratch, uint32_t* counts) {
  memset(counts, 0, sizeof(uint32_t) * (NBINS + 1));
  for (uint32_t i = 0; i < n; i++) {
    bin_scratch[i] = bucket_bin(depth[i], mn, mx, sort_mode);
    counts[bin_scratch[i]]++;
  }
  /* bins are appended in insertion order; ForEach walks them 511→0 for
   * BackToFront, 0→511 otherwise */
  uint32_t start[NBINS];
  uint32_t acc = 0;
  if (sort_mode == T3D_SORT_BACK_TO_FRONT) {
    for (int b = NBINS - 1; b >= 0; b--) {
      start[b] = acc;
      acc += counts[b];
    }
  } else {
    for (int b = 0; b < NBINS; b++) {
      start[b] = acc;
      acc += counts[b];
    }
  }
  for (uint32_t i = 0; i < n; i++) order[start[bin_scratch[i]]++] = i;
}

/* ------------------------------------------------------------------------ */
/* render                                                                    */
/* ------------------------------------------------------------------------ */

typedef struct {
  uint32_t model, part;
  float depth;
  int32_t order;
  int batched;     /* a dynamic batcher's part (Model.DynamicBatcher(), model.go:272-274) */
} opair;

typedef struct {
  t3d_vertex* color;
  t3d_vertex* depth;
  uint32_t* idx;
  uint32_t* tri_ids;
  float* keys;
  t3d_vertex* normals;      /* normalVertexList, when want_normals */
  int want_normals;
  uint64_t n, cap;          /* vertices */
} obuf;

static void obuf_reserve(obuf* b, uint64_t extra) {
  if (b->n + extra <= b->cap) return;
  uint64_t nc = b->cap ? b->cap : 1024;
  while (nc < b->n + extra) nc *= 2;
  b->color = (t3d_vertex*)realloc(b->color, nc * sizeof(t3d_vertex));
  b->depth = (t3d_vertex*)realloc(b->depth, nc * sizeof(t3d_vertex));
  b->idx = (uint32_t*)realloc(b->idx, nc * sizeof(uint32_t));
  b->tri_ids = (uint32_t*)realloc(b->tri_ids, (nc / 3 + 1) * sizeof(uint32_t));
  b->keys = (float*)realloc(b->keys, (nc / 3 + 1) * sizeof(float));
  if (b->want_normals) b->normals = (t3d_vertex*)realloc(b->normals, nc * sizeof(t3d_vertex));
  b->cap = nc;
}

/* Model.isTransparent: model.go:1274-1282. */
static int is_transparent(const t3d_model_desc* model, const t3d_material_desc* mat) {
  if (mat && mat->present) {
    int matT = mat->transparency_mode == T3D_TRANSPARENCY_TRANSPARENT || !mat->blend_source_over ||
               (mat->transparency_mode == T3D_TRANSPARENCY_AUTO && mat->color[3] < 0.999f);
    int modelT = mat->transparency_mode != T3D_TRANSPARENCY_OPAQUE && model->color[3] < 0.999f;
    return matT || modelT;
  }
  return model->color[3] < 0.999f;
}

static const t3d_material_desc* part_mat(const t3d_model_desc* model, uint32_t p) {
  if (!model->materials) return NULL;
  const t3d_material_desc* m = &model->materials[p];
  return m->present ? m : NULL;
}

typedef struct { const opair* pairs; } pair_env;
/* transparent pass order: camera.go:2050-2055 */
static int less_transparent(void* env, int32_t a, int32_t b) {
  const opair* p = ((pair_env*)env)->pairs;
  if (p[a].order == p[b].order) return p[a].depth > p[b].depth;
  return p[a].order < p[b].order;
}
/* solid order when RenderDepth is off: camera.go:1458-1467 */
static int less_solid_nodepth(void* env, int32_t a, int32_t b) {
  const opair* p = ((pair_env*)env)->pairs;
  if (p[a].order < p[b].order) return 1;
  return p[a].depth > p[b].depth;
}

typedef struct {
  const olight* lights;
  v3 pos;
} light_env;
/* MaxLightCount comparator: camera.go:1483-1495 */
static int less_light(void* env, int32_t a, int32_t b) {
  light_env* e = (light_env*)env;
  const olight* la = &e->lights[a];
  const olight* lb = &e->lights[b];
  if (la->type == T3D_LIGHT_AMBIENT) return 1;
  else if (lb->type == T3D_LIGHT_AMBIENT) return 0;
  else if (la->type == T3D_LIGHT_DIRECTIONAL) return 1;
  else if (lb->type == T3D_LIGHT_DIRECTIONAL) return 0;
  return v3_dist_sq(e->pos, la->world_pos) < v3_dist_sq(e->pos, lb->world_pos);
}

static v3 shape_pos(const omesh* m, const t3d_model_desc* model, uint32_t v) {   /* mesh.go:796-808 */
  v3 out = v3_from(m->pos + 3 * (size_t)v);
  for (uint32_t k = 0; k < m->d.shape_key_count; k++) {
    float w = model->shape_key_weights ? model->shape_key_weights[k] : 0.0f;
    const float* kp = m->sk_pos + ((size_t)k * m->d.vertex_count + v) * 3;
    out.x += kp[0] * w;
    out.y += kp[1] * w;
    out.z += kp[2] * w;
  }
  return out;
}
static v3 shape_nrm(const omesh* m, const t3d_model_desc* model, uint32_t v) {   /* mesh.go:812-822 */
  v3 out = v3_from(m->nrm + 3 * (size_t)v);
  for (uint32_t k = 0; k < m->d.shape_key_count; k++) {
    float w = model->shape_key_weights ? model->shape_key_weights[k] : 0.0f;
    const float* kn = m->sk_nrm + ((size_t)k * m->d.vertex_count + v) * 3;
    out.x += kn[0] * w;
    out.y += kn[1] * w;
    out.z += kn[2] * w;
  }
  return out;
}

/* ------------------------------------------------------------------------
 * Animation sampling.  math32's Sin and Atan2 are float32 wrappers of Go's
 * float64 math.Sin and math.Atan2 (math32/math32.go:137-139, 204-206); Go's
 * own math package is not in the reference, so its published algorithm is
 * restated here: Go 1.2x src/math/sin.go and atan.go/atan2.go, which follow
 * Cephes sin.c and atan.c (S. Moshier).  Constants are the Cephes ones with
 * the IEEE bit patterns Go's source lists beside them.  amd64 Go does not fuse
 * multiply-adds (GOAMD64=v1), and neither does this file.
 * ------------------------------------------------------------------------ */
static double go_sin(double x) {
  const double PI4A = 7.85398125648498535156e-1, PI4B = 3.77489470793079817668e-8,
               PI4C = 2.69515142907905952645e-15;
  static const double S[6] = {1.58962301576546568060e-10, -2.50507477628578072866e-8, 2.75573136213857245213e-6,
                              -1.98412698295895385996e-4, 8.33333333332211858878e-3, -1.66666666666666307295e-1};
  static const double C[6] = {-1.13585365213876817300e-11, 2.08757008419747316778e-9, -2.75573141792967388112e-7,
                              2.48015872888517045348e-5, -1.38888888888730564116e-3, 4.16666666666665929218e-2};
  if (x == 0 || x != x) return x;
  if (isinf(x)) return NAN;
  int sign = 0;
  if (x < 0) { x = -x; sign = 1; }
  /* Go reduces x >= 2^29 with Payne-Hanek (trig_reduce.go); slerp's
   * arguments lie in [0, pi/2], so that branch is never taken here. */
  if (x >= 536870912.0) return NAN;
  uint64_t j = (uint64_t)(x * 1.2732395447351628);   /* 4/Pi rounded to float64: 0x3ff45f306dc9c883 */
  double y = (double)j;
  if (j & 1) { j++; y++; }
  j &= 7;
  double z = ((x - y * PI4A) - y * PI4B) - y * PI4C;
  if (j > 3) { sign = !sign; j -= 4; }
  double zz = z * z;
  if (j == 1 || j == 2)
    y = 1.0 - 0.5 * zz + zz * zz * ((((((C[0] * zz) + C[1]) * zz + C[2]) * zz + C[3]) * zz + C[4]) * zz + C[5]);
  else
    y = z + z * zz * ((((((S[0] * zz) + S[1]) * zz + S[2]) * zz + S[3]) * zz + S[4]) * zz + S[5]);
  return sign ? -y : y;
}
static double go_xatan(double x) {
  const double P0 = -8.750608600031904122785e-01, P1 = -1.615753718733365076637e+01,
               P2 = -7.500855792314704667340e+01, P3 = -1.228866684490136173410e+02,
               P4 = -6.485021904942025371773e+01, Q0 = +2.485846490142306297962e+01,
               Q1 = +1.650270098316988542046e+02, Q2 = +4.328810604912902668951e+02,
               Q3 = +4.853903996359136964868e+02, Q4 = +1.945506571482613964425e+02;
  double z = x * x;
  z = z * ((((P0 * z + P1) * z + P2) * z + P3) * z + P4) / (((((z + Q0) * z + Q1) * z + Q2) * z + Q3) * z + Q4);
  return x * z + x;
}
static double go_satan(double x) {
  const double PIO2 = 1.5707963267948966, PIO4 = 0.7853981633974483;   /* Pi/2, Pi/4 rounded */
  const double Morebits = 6.123233995736765886130e-17, Tan3pio8 = 2.41421356237309504880;
  if (x <= 0.66) return go_xatan(x);
  if (x > Tan3pio8) return PIO2 - go_xatan(1 / x) + Morebits;
  return PIO4 + go_xatan((x - 1) / (x + 1)) + 0.5 * Morebits;
}
static double go_atan(double x) {
  if (x == 0) return x;
  return x > 0 ? go_satan(x) : -go_satan(-x);
}
static double go_atan2(double y, double x) {
  const double PI = 3.141592653589793, PIO2 = 1.5707963267948966, PIO4 = 0.7853981633974483,
               PI3O4 = 2.356194490192345;                                /* 3*Pi/4 rounded */
  if (y != y || x != x) return NAN;
  if (y == 0) return (x >= 0 && !signbit(x)) ? copysign(0.0, y) : copysign(PI, y);
  if (x == 0) return copysign(PIO2, y);
  if (isinf(x)) {
    if (x > 0) return isinf(y) ? copysign(PIO4, y) : copysign(0.0, y);
    return isinf(y) ? copysign(PI3O4, y) : copysign(PI, y);
  }
  if (isinf(y)) return copysign(PIO2, y);
  double q = go_atan(y / x);
  if (x < 0) return q <= 0 ? q + PI : q - PI;
  return q;
}
static float m32_sin(float x) { return (float)go_sin((double)x); }                     /* math32.go:137-139 */
static float m32_atan2(float y, float x) { return (float)go_atan2((double)y, (double)x); } /* math32.go:204-206 */

typedef struct { float x, y, z, w; } oquat;
/* Quaternion.Slerp: quaternion.go:26-59. */
static oquat q_slerp(oquat q, oquat e, float percent) {
  float cosOmega = q.x * e.x + q.y * e.y + q.z * e.z + q.w * e.w;   /* Dot, quaternion.go:82-84 */
  if (cosOmega < 0) {
    e.w *= -1; e.x *= -1; e.y *= -1; e.z *= -1;
    cosOmega *= -1;
  }
  float k0, k1;
  if (cosOmega > 0.9999f) {
    k0 = 1 - percent;
    k1 = percent;
  } else {
    float sinOmega = sqrt32(1 - (cosOmega * cosOmega));
    float omega = m32_atan2(sinOmega, cosOmega);
    float oneOverSinOmega = 1 / sinOmega;
    k0 = m32_sin((1 - percent) * omega) * oneOverSinOmega;
    k1 = m32_sin(percent * omega) * oneOverSinOmega;
  }
  q.w = q.w * k0 + e.w * k1;
  q.x = q.x * k0 + e.x * k1;
  q.y = q.y * k0 + e.y * k1;
  q.z = q.z * k0 + e.z * k1;
  return q;
}
/* Quaternion.ToMatrix4: quaternion.go:135-183, the product of two matrices. */
static m4 q_to_m4(oquat q) {
  m4 a = m4_identity(), b = m4_identity();
  a.m[0][0] = q.w;  a.m[0][1] = q.z;  a.m[0][2] = -q.y; a.m[0][3] = q.x;
  a.m[1][0] = -q.z; a.m[1][1] = q.w;  a.m[1][2] = q.x;  a.m[1][3] = q.y;
  a.m[2][0] = q.y;  a.m[2][1] = -q.x; a.m[2][2] = q.w;  a.m[2][3] = q.z;
  a.m[3][0] = -q.x; a.m[3][1] = -q.y; a.m[3][2] = -q.z; a.m[3][3] = q.w;
  b.m[0][0] = q.w;  b.m[0][1] = q.z;  b.m[0][2] = -q.y; b.m[0][3] = -q.x;
  b.m[1][0] = -q.z; b.m[1][1] = q.w;  b.m[1][2] = q.x;  b.m[1][3] = -q.y;
  b.m[2][0] = q.y;  b.m[2][1] = -q.x; b.m[2][2] = q.w;  b.m[2][3] = -q.z;
  b.m[3][0] = q.x;  b.m[3][1] = q.y;  b.m[3][2] = q.z;  b.m[3][3] = q.w;
  return m4_mult(&a, &b);
}
/* Keyframe search shared by both samplers (animation.go:69-93, 122-147):
 * returns 0 (no keyframes), 1 (use keyframe *k0 as is) or 2 (interpolate
 * between *k0 and *k0 + 1 at *t). */
static int track_find(const t3d_anim_track* tr, float time, uint32_t* k0, float* t) {
  const uint32_t n = tr->keyframe_count;
  if (n == 0) return 0;
  if (time <= tr->times[0]) { *k0 = 0; return 1; }
  if (time >= tr->times[n - 1]) { *k0 = n - 1; return 1; }
  uint32_t last = 0;
  while (tr->times[last] < time) last++;               /* first keyframe at or after time */
  const uint32_t first = last - 1;
  if (time == tr->times[first]) { *k0 = first; return 1; }
  if (time == tr->times[last]) { *k0 = last; return 1; }
  if (tr->interpolation == T3D_INTERPOLATION_CONSTANT) { *k0 = first; return 1; }
  *k0 = first;
  *t = (time - tr->times[first]) / (tr->times[last] - tr->times[first]);
  return 2;
}
static int track_vec(const t3d_anim_track* tr, float time, v3* out) {        /* ValueAsVector, animation.go:63-110 */
  uint32_t k;
  float t;
  const int r = track_find(tr, time, &k, &t);
  if (!r) return 0;
  const float* a = tr->values + 3 * (size_t)k;
  v3 fd = {a[0], a[1], a[2]};
  if (r == 1) { *out = fd; return 1; }
  v3 ld = {a[3], a[4], a[5]};
  v3 o = {fd.x + ((ld.x - fd.x) * t), fd.y + ((ld.y - fd.y) * t), fd.z + ((ld.z - fd.z) * t)};   /* vector.go:559-565 */
  *out = o;
  return 1;
}
static int track_quat(const t3d_anim_track* tr, float time, oquat* out) {    /* ValueAsQuaternion, animation.go:116-160 */
  uint32_t k;
  float t;
  const int r = track_find(tr, time, &k, &t);
  if (!r) return 0;
  const float* a = tr->values + 4 * (size_t)k;
  oquat fd = {a[0], a[1], a[2], a[3]};
  if (r == 1) { *out = fd; return 1; }
  oquat ld = {a[4], a[5], a[6], a[7]};
  *out = q_slerp(fd, ld, t);
  return 1;
}
/* AnimationPlayer.forceUpdate without a blend and with absolute motion
 * (animation.go:902-983): each present track sets the bone's local
 * position, scale or rotation (SetLocalRotation ignores an all-zero matrix,
 * node.go:906-912). */
static void animate_poses(const t3d_anim_channel* ch, uint32_t nch, float time, t3d_bone_pose* poses, uint32_t n) {
  for (uint32_t b = 0; b < n && b < nch; b++) {
    v3 v;
    oquat q;
    if (track_vec(&ch[b].position, time, &v)) {
      poses[b].position[0] = v.x; poses[b].position[1] = v.y; poses[b].position[2] = v.z;
    }
    if (track_vec(&ch[b].scale, time, &v)) {
      poses[b].scale[0] = v.x; poses[b].scale[1] = v.y; poses[b].scale[2] = v.z;
    }
    if (track_quat(&ch[b].rotation, time, &q)) {
      const m4 r = q_to_m4(q);
      int zero = 1;
      for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) zero &= r.m[i][j] == 0;
      if (!zero) poses[b].rotation = r;
    }
  }
}

/* Node.Transform (node.go:572-608) down an armature, bones listed parents
 * first: Scale(s).Mult(rotation).Mult(Translate(p)), then .Mult(parent's
 * Transform()) (the armature parent's for a root bone, when it has one), and
 * boneInfluence = inverseBindMatrix.Mult(transform) (node.go:592-594).
 * world[] receives each bone's Transform().  Returns 0, or -1 for a parent
 * index that is not an earlier bone. */
static int bone_palettes(const t3d_bone_pose* poses, uint32_t n, const m4* root, m4* world, m4* out) {
  for (uint32_t b = 0; b < n; b++) {
    const t3d_bone_pose* B = &poses[b];
    if (B->parent < -1 || B->parent >= (int32_t)b) return -1;
    const m4 sc = m4_scale(B->scale[0], B->scale[1], B->scale[2]);
    const m4 tr = m4_translate(B->position[0], B->position[1], B->position[2]);
    m4 t = m4_mult(&sc, &B->rotation);
    t = m4_mult(&t, &tr);
    if (B->parent >= 0) t = m4_mult(&t, &world[B->parent]);
    else if (root) t = m4_mult(&t, root);
    world[b] = t;
    out[b] = m4_mult(&B->inverse_bind, &t);
  }
  return 0;
}

/* Model.skinVertex: model.go:559-596. */
static void skin_vertex(const omesh* m, const t3d_model_desc* model, uint32_t v, v3* op, v3* on) {
  m4 S = m4_zero();
  for (int k = 0; k < 4; k++) {
    float w = m->weights[4 * (size_t)v + k];
    if (w == 0) continue;
    const m4* B = &model->bone_influences[m->bones[4 * (size_t)v + k]];
    if (w == 1) {
      S = *B;
      break;
    }
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++) S.m[i][j] += B->m[i][j] * w;   /* Add(Scale(w)), matrix.go:639-661 */
  }
  *op = m4_mult_vec(&S, shape_pos(m, model, v));
  S.m[3][0] = 0;
  S.m[3][1] = 0;
  S.m[3][2] = 0;
  S.m[3][3] = 1;
  *on = m4_mult_vec(&S, shape_nrm(m, model, v));
}

static int validate_model(t3do_ctx* c, const t3d_model_desc* model, const omesh* m, const t3d_camera_desc* cam) {
  if (model->has_vertex_transform_function || model->has_vertex_clip_function)
    return set_err(c, "model uses a Go vertex callback") , T3D_ERR_UNSUPPORTED;
  /* A skinned model never writes the unbillboarded depth the depth list would read (model.go:760-775 has no
   * such branch), so the reference draws whatever an earlier model left there: refused. */
  if (model->skinned && model->materials)
    for (uint32_t p = 0; p < m->d.part_count; p++) {
      const t3d_material_desc* mat = &model->materials[p];
      if (mat->present && mat->visible && mat->depth_mode == T3D_DEPTH_MODE_UNBILLBOARDED)
        return set_err(c, "skinned model with an unbillboarded depth mode"), T3D_ERR_UNSUPPORTED;
    }
  if (model->skinned) {
    if (!m->bones || !m->weights || (!model->bone_influences && !model->bone_poses))
      return set_err(c, "skinned model without skin data");
    if (model->animation) {
      if (!model->bone_poses) return set_err(c, "animation without bone poses");
      if (!find_anim(c, model->animation)) return set_err(c, "unknown animation");
      if (model->animation_time != model->animation_time) return set_err(c, "animation time is NaN");
    }
    if (model->bone_poses)
      for (uint32_t b = 0; b < model->bone_count; b++)
        if (model->bone_poses[b].parent < -1 || model->bone_poses[b].parent >= (int32_t)b)
          return set_err(c, "bone pose parent is not an earlier bone");
    for (size_t i = 0; i < 4 * (size_t)m->d.vertex_count; i++)
      if (m->weights[i] != 0 && m->bones[i] >= model->bone_count) return set_err(c, "bone index out of range");
  }
  (void)cam;
  return T3D_OK;
}

/* render closure for one pair: camera.go:1507-1827 with ProcessVertices
 * (model.go:600-995). Appends the flush's arrays to *ob. */
static int render_pair(t3do_ctx* c, const ocam* cam, const t3d_world_desc* world, const t3d_model_desc* model,
                       uint32_t model_index, uint32_t part_index, const omesh* m, olight* scene_lights, int nscene,
                       olight* group_lights, const t3d_light_desc* group_descs, int ngroup, oscratch* s, obuf* ob,
                       t3do_result* res, uint32_t* pair_cap, int64_t flush) {
  const t3d_camera_desc* cd = cam->d;
  const t3d_material_desc* mat = part_mat(model, part_index);
  const t3d_part_range* pr = &m->parts[part_index];

  int lighting = 0;                                                 /* camera.go:1521-1528 */
  if (world && world->present) {
    if (mat) lighting = world->lighting_on && !mat->shadeless && !model->shadeless;
    else lighting = world->lighting_on && !model->shadeless;
  }
  int sort_mode = mat ? mat->triangle_sort_mode : T3D_SORT_BACK_TO_FRONT;

  /* pair record (emitted even when empty, survivor_count 0); a dynamic batch's parts all add to the flush record
   * of their batcher (flush >= 0), their vertices following on in its vertex list (camera.go:1552, 2084-2099) */
  t3d_pair_result* out_pair;
  if (flush >= 0) {
    out_pair = &res->pairs[flush];
    out_pair->triangle_count += pr->triangle_count;
  } else {
    if (res->pair_count == *pair_cap) {
      *pair_cap = *pair_cap ? 2 * *pair_cap : 64;
      res->pairs = (t3d_pair_result*)realloc(res->pairs, sizeof(t3d_pair_result) * *pair_cap);
    }
    out_pair = &res->pairs[res->pair_count++];
    out_pair->model_index = model_index;
    out_pair->part_index = part_index;
    out_pair->triangle_count = pr->triangle_count;
    out_pair->survivor_count = 0;
    out_pair->vertex_offset = ob->n;
  }
  const uint32_t vli0 = 3 * out_pair->survivor_count;              /* vertexListIndex when this render starts */

  /* ProcessVertices early out: model.go:604-606 */
  const int only_visible = cam->only_visible;
  if (only_visible && ((model->color[3] == 0 && is_transparent(model, mat)) || !model->visible)) return T3D_OK;
  if (mat && mat->custom_depth_function)
    return set_err(c, "custom depth function"), T3D_ERR_UNSUPPORTED;

  const m4* M = &model->transform;
  v3 camPos = cam->pos;
  v3 icp = inverted_pos(M, camPos);                                 /* model.go:629-630 */
  m4 bbase = *M;
  if (mat && mat->billboard) bbase = billboard_base(M, mat, camPos, cam->up);
  m4 mvp = m4_mult(&bbase, &cam->vp);                               /* model.go:695 */
  /* DepthModeUnbillboarded (model.go:697-703): the depth list takes Z of the unbillboarded MVP, unsnapped */
  const int unbb = mat && mat->depth_mode == T3D_DEPTH_MODE_UNBILLBOARDED;
  m4 umvp = m4_mult(M, &cam->vp);
  float minDepth = MAXF32, maxDepth = -MAXF32;
  float camNear = cd->near, camFar = cd->far;
  int skinned = model->skinned;
  int snapping = cd->vertex_snapping > 0;
  int backface = mat && mat->backface_culling;
  int render_normals = cd->render_normals;
  m4 mvn = m4_identity();                                           /* mvJustRForNormals, model.go:705-708 */
  if (render_normals) {
    m4 mv = m4_mult(M, &cam->view);
    v3 dp, ds;
    m4_decompose(&mv, &dp, &ds, &mvn);
  }

  uint32_t T = pr->triangle_count;
  uint32_t* bucket_tri = (uint32_t*)malloc(sizeof(uint32_t) * (T ? T : 1));
  float* bucket_depth = (float*)malloc(sizeof(float) * (T ? T : 1));
  uint32_t nb = 0;

  for (uint32_t t = pr->triangle_start; t < pr->triangle_start + T; t++) {   /* forEachTri, mesh.go:2238-2256 */
    const uint32_t* tri = m->idx + 3 * (size_t)t;
    for (int vi = 0; vi < 3; vi++) {                                         /* model.go:751-822 */
      uint32_t v = tri[vi];
      v3 vert, normal;
      if (skinned) {
        skin_vertex(m, model, v, &vert, &normal);
        s->alt_pos[v] = vert;
        s->alt_nrm[v] = normal;
        s->clip[v] = m4_mult_vec_w(&cam->vp, vert);
      } else {
        vert = shape_pos(m, model, v);
        normal = shape_nrm(m, model, v);
        s->alt_pos[v] = vert;
        s->alt_nrm[v] = normal;
        s->clip[v] = m4_mult_vec_w(&mvp, vert);
        if (unbb) s->udepth[v] = m4_mult_vec_w(&umvp, vert).z;           /* model.go:800-803 */
      }
      if (snapping) {                                                        /* model.go:807-812 */
        float vs = cd->vertex_snapping;
        s->clip[v].x = round32(s->clip[v].x * vs) / vs;
        s->clip[v].y = round32(s->clip[v].y * vs) / vs;
        s->clip[v].z = round32(s->clip[v].z * vs) / vs;
      }
      /* model.go:814-818: the rotation of Model.Transform()·ViewMatrix() applied to the (shape-keyed or
       * skinned) vertex position, as the reference writes it (not to the normal) */
      if (render_normals) s->tnrm[v] = m4_mult_vec(&mvn, vert);
    }
    v3 stc = {0, 0, 0};
    int outOfBounds = 1;
    v3 tvp[3];
    for (int i = 0; i < 3; i++) {                                            /* model.go:855-886 */
      uint32_t v = tri[i];
      if (skinned) {
        stc.x += s->alt_pos[v].x;
        stc.y += s->alt_pos[v].y;
        stc.z += s->alt_pos[v].z;
      }
      float w = s->clip[v].w;
      if (w < 0) w = 0.000001f;
      tvp[i].x = s->clip[v].x / w;
      tvp[i].y = s->clip[v].y / w;
      if (s->clip[v].z + 1 >= camNear && s->clip[v].z < camFar) outOfBounds = 0;
    }
    if (only_visible && outOfBounds) continue;                              /* model.go:883-886 */
    if (only_visible && ((tvp[0].x < -0.5f && tvp[1].x < -0.5f && tvp[2].x < -0.5f) ||   /* model.go:890-898 */
                         (tvp[0].x > 0.5f && tvp[1].x > 0.5f && tvp[2].x > 0.5f) ||
                         (tvp[0].y < -0.5f && tvp[1].y < -0.5f && tvp[2].y < -0.5f) ||
                         (tvp[0].y > 0.5f && tvp[1].y > 0.5f && tvp[2].y > 0.5f)))
      continue;
    if (only_visible && backface) {                                          /* model.go:902-925 */
      float n0x = tvp[0].x - tvp[1].x;
      float n0y = tvp[0].y - tvp[1].y;
      float n1x = tvp[1].x - tvp[2].x;
      float n1y = tvp[1].y - tvp[2].y;
      float nor = (n0x * n1y) - (n1x * n0y);
      if (nor < 0) continue;
    }
    float depth;
    if (skinned) {                                                           /* model.go:941-955 */
      float dx = camPos.x - (stc.x / 3);
      float dy = camPos.y - (stc.y / 3);
      float dz = camPos.z - (stc.z / 3);
      depth = dx * dx + dy * dy + dz * dz;
    } else {
      v3 ctr = m->tri_center[t];
      float dx = icp.x - ctr.x;
      float dy = icp.y - ctr.y;
      float dz = icp.z - ctr.z;
      depth = dx * dx + dy * dy + dz * dz;
    }
    if (only_visible && (isnan32(depth) || isinf32(depth))) continue;        /* model.go:957-960 */
    if (depth < minDepth) minDepth = depth;
    if (depth > maxDepth) maxDepth = depth;
    bucket_tri[nb] = t;                                                      /* AddTriangle, sortingTriangles.go:32-36 */
    bucket_depth[nb] = depth;
    nb++;
  }

  if (nb == 0) {                                                             /* camera.go:1556-1558 */
    free(bucket_tri);
    free(bucket_depth);
    return T3D_OK;
  }
  uint32_t* order = (uint32_t*)malloc(sizeof(uint32_t) * nb);
  uint32_t* bins = (uint32_t*)malloc(sizeof(uint32_t) * nb);
  uint32_t counts[NBINS + 1];
  bucket_order(bucket_depth, nb, minDepth, maxDepth, sort_mode, order, bins, counts);

  float srcW = 0, srcH = 0;                                                  /* camera.go:1560-1566 */
  if (mat && mat->use_texture) {
    srcW = mat->texture_size[0];
    srcH = mat->texture_size[1];
  }

  olight* lights = scene_lights;
  int nlights = nscene;
  if (lighting) {                                                            /* camera.go:1573-1588 */
    if (model->light_group_active) {
      lights = group_lights;
      nlights = ngroup;
      for (int i = 0; i < nlights; i++) light_begin_render(&lights[i], &group_descs[i]);
    }
    for (int i = 0; i < nlights; i++) light_begin_model(&lights[i], model);
  }

  v3 mdmin = v3_from(m->d.dimensions_min), mdmax = v3_from(m->d.dimensions_max);
  float maxSpan = v3_mag(v3_sub(mdmax, mdmin));                              /* camera.go:1592 */
  v3 modelPos = m4_row3(M);
  float mp[4] = {model->color[0], model->color[1], model->color[2], model->color[3]};   /* camera.go:1597-1601 */
  if (mat) {
    mp[0] *= mat->color[0];
    mp[1] *= mat->color[1];
    mp[2] *= mat->color[2];
    mp[3] *= mat->color[3];
  }
  float camWidth = (float)cd->width, camHeight = (float)cd->height;
  float halfW = camWidth / 2, halfH = camHeight / 2;
  float dmp = (cd->far - cd->near) * cd->depth_margin;                       /* camera.go:1339-1341 */
  float camSpread = cd->far - cd->near + (dmp * 2);

  if (lighting) {                                                            /* camera.go:1610-1641 */
    for (uint32_t k = 0; k < nb; k++) {
      const uint32_t* tri = m->idx + 3 * (size_t)bucket_tri[order[k]];
      for (int vi = 0; vi < 3; vi++) {
        float* L = s->lights + 4 * (size_t)tri[vi];
        L[0] = 0;
        L[1] = 0;
        L[2] = 0;
        L[3] = 1;
      }
      for (int li = 0; li < nlights; li++) {
        const olight* l = &lights[li];
        if (l->type == T3D_LIGHT_POINT && l->range > 0) {
          float dist = maxSpan + l->range;
          if (v3_dist_sq(modelPos, l->world_pos) > dist * dist) continue;
        }
        for (int vi = 0; vi < 3; vi++) vertex_light(l, tri[vi], model, mat, s, s->lights + 4 * (size_t)tri[vi]);
      }
    }
  }

  /* vertex assembly: camera.go:1646-1820 */
  obuf_reserve(ob, 3 * (uint64_t)nb);
  uint64_t base = ob->n;
  const float* vcol = m->col;
  for (uint32_t k = 0; k < nb; k++) {
    uint32_t t = bucket_tri[order[k]];
    const uint32_t* tri = m->idx + 3 * (size_t)t;
    ob->tri_ids[(base / 3) + k] = t;
    ob->keys[(base / 3) + k] = bucket_depth[order[k]];
    for (int vi = 0; vi < 3; vi++) {
      uint32_t v = tri[vi];
      uint64_t o = base + 3 * (uint64_t)k + vi;
      t3d_vertex* cv = &ob->color[o];
      t3d_vertex* dv = &ob->depth[o];
      memset(cv, 0, sizeof *cv);
      memset(dv, 0, sizeof *dv);
      v4 target = s->clip[v];
      float w = target.w;
      if (!cd->perspective) w = 1.0f;
      if (w < 0) w = 0.001f;
      float dx = (target.x / w) * camWidth + halfW;
      float dy = (target.y / -w) * camHeight + halfH;
      cv->dst_x = dx;
      cv->dst_y = dy;
      dv->dst_x = dx;
      dv->dst_y = dy;
      float U = m->uv ? m->uv[2 * (size_t)v] : 0.0f;
      float Vv = m->uv ? m->uv[2 * (size_t)v + 1] : 0.0f;
      float uvU, uvV;
      if (cd->perspective_corrected_texture_mapping) {
        uvU = (U / w) * srcW;
        uvV = ((1 - Vv) / w) * srcH;
      } else {
        uvU = U * srcW;
        uvV = (1 - Vv) * srcH;
      }
      cv->src_x = uvU;
      cv->src_y = uvV;
      if (cd->perspective_corrected_texture_mapping) {
        float d = 1.0f / w;
        cv->custom0 = d;
        dv->custom0 = d;
      }
      dv->src_x = uvU;
      dv->src_y = uvV;
      if (render_normals) {                                                  /* camera.go:1710-1722 */
        t3d_vertex* nv = &ob->normals[o];
        memset(nv, 0, sizeof *nv);
        nv->dst_x = dx;
        nv->dst_y = dy;
        nv->src_x = uvU;
        nv->src_y = uvV;
        nv->color_r = s->tnrm[v].x * 0.5f + 0.5f;
        nv->color_g = s->tnrm[v].y * 0.5f + 0.5f;
        nv->color_b = s->tnrm[v].z * 0.5f + 0.5f;
      }
      if (vcol) {
        cv->color_r = vcol[4 * (size_t)v + 0] * mp[0];
        cv->color_g = vcol[4 * (size_t)v + 1] * mp[1];
        cv->color_b = vcol[4 * (size_t)v + 2] * mp[2];
        cv->color_a = vcol[4 * (size_t)v + 3] * mp[3];
      } else {
        cv->color_r = mp[0];
        cv->color_g = mp[1];
        cv->color_b = mp[2];
        cv->color_a = mp[3];
      }
      if (lighting) {
        const float* L = s->lights + 4 * (size_t)v;
        cv->color_r *= L[0];
        cv->color_g *= L[1];
        cv->color_b *= L[2];
      }
      if (cd->render_depth) {
        float vd = unbb ? s->udepth[v] : target.z;                           /* camera.go:1765-1770 */
        float depth = (vd + dmp) / camSpread;
        if (depth < 0) depth = 0;
        else if (depth > 1) depth = 1;
        dv->color_r = depth;
        dv->color_g = depth;
        dv->color_b = depth;
        dv->color_a = 1;
      } else if (world && world->present && world->fog_on) {
        float vd = target.z;
        float depth = (vd + dmp) / camSpread;
        if (depth < 0) depth = 0;
        else if (depth > 1) depth = 1;
        depth = world->fog_range[0] + ((world->fog_range[1] - world->fog_range[0]) * 1 - depth);
        if (world->fog_mode == T3D_FOG_ADD) {
          cv->color_r += world->fog_color[0] * depth;
          cv->color_g += world->fog_color[1] * depth;
          cv->color_b += world->fog_color[2] * depth;
        } else if (world->fog_mode == T3D_FOG_SUB) {
          cv->color_r *= world->fog_color[0] * depth;
          cv->color_g *= world->fog_color[1] * depth;
          cv->color_b *= world->fog_color[2] * depth;
        }
      }
      ob->idx[o] = (uint32_t)(vli0 + 3 * k + vi);                            /* indexList, camera.go:1816 */
    }
  }
  ob->n += 3 * (uint64_t)nb;
  out_pair->survivor_count += nb;
  free(order);
  free(bins);
  free(bucket_tri);
  free(bucket_depth);
  return T3D_OK;
}

/* Dynamic batching (model.go:217-274).  The batched models of (owner, part): the models naming that owner and part,
 * in their DynamicBatchModels slice order (batch_order). */
typedef struct { const t3d_render_desc* desc; const float* dist; } batch_env;
static int less_batch_order(void* env, int32_t a, int32_t b) {
  const t3d_model_desc* m = ((batch_env*)env)->desc->models;
  return m[a].batch_order < m[b].batch_order;
}
static int less_farther(void* env, int32_t a, int32_t b) {
  const float* d = ((batch_env*)env)->dist;
  return d[a] > d[b];
}
static int batch_models(const t3d_render_desc* desc, uint32_t owner, uint32_t part, int32_t* out) {
  int n = 0;
  for (uint32_t k = 0; k < desc->model_count; k++) {
    const t3d_model_desc* b = &desc->models[k];
    if (b->dynamic_batched && b->batch_owner == owner && b->batch_part == part) out[n++] = (int32_t)k;
  }
  batch_env env = {desc, NULL};
  go_stable_sort(out, n, less_batch_order, &env);
  return n;
}

/* ProcessVertices + the render closure of one (model, part), with the model's light group and posed palette
 * (camera.go:1507-1827); flush >= 0 adds it to that flush record (a dynamic batch). */
static int render_model_part(t3do_ctx* c, const t3d_render_desc* desc, const ocam* cam, const t3d_world_desc* world,
                             uint32_t mi, uint32_t part, olight* lights, int nl, oscratch* s, obuf* ob, t3do_result* res,
                             uint32_t* pair_cap, int64_t flush) {
  const t3d_model_desc* model = &desc->models[mi];
  omesh* m = find_mesh(c, model->mesh);
  olight* glights = NULL;
  int ng = 0;
  if (model->light_group_active) {
    ng = (int)model->light_group_count;
    glights = (olight*)calloc(ng + 1, sizeof(olight));
    for (int g = 0; g < ng; g++) light_from_desc(c, &glights[g], &model->light_group[g]);
  }
  /* bones given as poses: the palette from Node.Transform down the armature */
  t3d_model_desc posed;
  m4* pal = NULL;
  if (model->skinned && model->bone_poses) {
    const uint32_t nb = model->bone_count;
    pal = (m4*)malloc(sizeof(m4) * 2 * (nb ? nb : 1));
    const t3d_bone_pose* poses = model->bone_poses;
    t3d_bone_pose* animated = NULL;
    if (model->animation) {
      const oanim* an = find_anim(c, model->animation);
      animated = (t3d_bone_pose*)malloc(sizeof(t3d_bone_pose) * (nb ? nb : 1));
      memcpy(animated, poses, sizeof(t3d_bone_pose) * nb);
      animate_poses(an->ch, an->n, model->animation_time, animated, nb);
      poses = animated;
    }
    bone_palettes(poses, nb, model->armature_parent, pal + nb, pal);
    free(animated);
    posed = *model;
    posed.bone_influences = pal;
    model = &posed;
  }
  int rc = render_pair(c, cam, world, model, mi, part, m, lights, nl, glights, model->light_group, ng, s, ob, res,
                       pair_cap, flush);
  free(pal);
  for (int g = 0; g < ng; g++) light_release(&glights[g]);
  free(glights);
  return rc;
}

/* A dynamic batcher's pair in the pass loop (camera.go:2075-2099): its batched models sorted farther first with
 * sort.Slice (insertion sort for up to 12, as Go does; longer slices use a stable sort here, which orders equal
 * distances as pdqsort need not), each visible one inside the frustum rendering every part of its mesh into the
 * one flush of the batcher's pair. */
static int render_batch(t3do_ctx* c, const t3d_render_desc* desc, const ocam* cam, const t3d_world_desc* world,
                        const opair* rp, olight* lights, int nl, oscratch* s, obuf* ob, t3do_result* res,
                        uint32_t* pair_cap) {
  if (res->pair_count == *pair_cap) {
    *pair_cap = *pair_cap ? 2 * *pair_cap : 64;
    res->pairs = (t3d_pair_result*)realloc(res->pairs, sizeof(t3d_pair_result) * *pair_cap);
  }
  const int64_t flush = res->pair_count++;
  t3d_pair_result* fp = &res->pairs[flush];
  memset(fp, 0, sizeof *fp);
  fp->model_index = rp->model;
  fp->part_index = rp->part;
  fp->vertex_offset = ob->n;
  int32_t* kids = (int32_t*)malloc(sizeof(int32_t) * (desc->model_count + 1));
  float* dist = (float*)malloc(sizeof(float) * (desc->model_count + 1));
  int n = batch_models(desc, rp->model, rp->part, kids);
  for (int i = 0; i < n; i++) dist[kids[i]] = v3_dist_sq(cam->pos, m4_row3(&desc->models[kids[i]].transform));
  if (n <= 12) {
    for (int i = 1; i < n; i++)
      for (int j = i; j > 0 && dist[kids[j]] > dist[kids[j - 1]]; j--) {
        int32_t t = kids[j];
        kids[j] = kids[j - 1];
        kids[j - 1] = t;
      }
  } else {
    batch_env env = {desc, dist};
    go_stable_sort(kids, n, less_farther, &env);
  }
  int rc = T3D_OK;
  for (int i = 0; i < n && rc == T3D_OK; i++) {
    const t3d_model_desc* km = &desc->models[kids[i]];
    if (!km->visible) continue;
    omesh* m = find_mesh(c, km->mesh);
    if (!m) { rc = set_err(c, "render: unknown mesh handle"); break; }
    if (km->frustum_culling) {
      v3 sp;
      float sr;
      model_sphere(&km->transform, km->skinned, m->d.dimensions_min, m->d.dimensions_max, &sp, &sr);
      if (!sphere_in_frustum(cam, sp, sr)) continue;
    }
    rc = validate_model(c, km, m, cam->d);
    for (uint32_t q = 0; q < m->d.part_count && rc == T3D_OK; q++)
      rc = render_model_part(c, desc, cam, world, (uint32_t)kids[i], q, lights, nl, s, ob, res, pair_cap, flush);
  }
  free(kids);
  free(dist);
  return rc;
}

/* A volume light's cells and cell lights, a few levels down (a volume may name volumes). */
static int check_volume_light(t3do_ctx* c, const t3d_light_desc* l, int depth) {
  if (l->type != T3D_LIGHT_VOLUME) return T3D_OK;
  ovolume* v = find_vol(c, l->volume);
  if (!v) return T3D_ERR_NOT_FOUND;
  if (!v->first) return T3D_OK;
  if (depth >= 8 || l->use_light_count < v->need || (v->need && !l->use_lights)) return T3D_ERR_INVALID_ARGUMENT;
  for (uint32_t j = 0; j < l->use_light_count; j++) {
    int r = check_volume_light(c, &l->use_lights[j], depth + 1);
    if (r) return r;
  }
  return T3D_OK;
}

int t3do_render(t3do_ctx* c, const t3d_render_desc* desc, t3do_result* res) {
  memset(res, 0, sizeof *res);
  if (!desc || !desc->camera) return set_err(c, "render: missing camera");
  const t3d_camera_desc* cd = desc->camera;
  if (cd->width <= 0 || cd->height <= 0) return set_err(c, "render: camera size");
  const t3d_world_desc* world = (desc->world && desc->world->present) ? desc->world : NULL;

  ocam cam;
  cam_setup(&cam, cd);
  cam.only_visible = !desc->process_all_triangles;

  /* a LightVolume must name registered cells (as the product refuses an unknown volume handle), and the lights
   * its cells name must be in its use_lights */
  for (uint32_t i = 0; i < desc->light_count; i++) {
    int vr = check_volume_light(c, &desc->lights[i], 0);
    if (vr) return set_err(c, "render: light %u: %s", i, vr == T3D_ERR_NOT_FOUND ? "unregistered volume" : "cell lights"), vr;
  }
  for (uint32_t mi = 0; mi < desc->model_count; mi++) {
    const t3d_model_desc* md = &desc->models[mi];
    for (uint32_t k = 0; md->light_group_active && md->light_group && k < md->light_group_count; k++) {
      int vr = check_volume_light(c, &md->light_group[k], 0);
      if (vr) return set_err(c, "render: model %u light group: %s", mi, vr == T3D_ERR_NOT_FOUND ? "unregistered volume" : "cell lights"), vr;
    }
  }

  /* scene light list: camera.go:1275-1303 */
  int nl = 0;
  olight* lights = (olight*)calloc(desc->light_count + 1, sizeof(olight));
  if (world) {
    if (world->lighting_on && world->ambient_visible) {
      olight* a = &lights[nl++];
      a->type = T3D_LIGHT_AMBIENT;
      a->visible = 1;
      a->color[0] = world->ambient_color[0];
      a->color[1] = world->ambient_color[1];
      a->color[2] = world->ambient_color[2];
      a->energy = world->ambient_energy;
      light_begin_render(a, NULL);
    }
  }
  for (uint32_t i = 0; i < desc->light_count; i++) {
    const t3d_light_desc* ld = &desc->lights[i];
    if ((!world || world->lighting_on) && ld->visible) {
      light_from_desc(c, &lights[nl], ld);
      light_begin_render(&lights[nl], ld);
      nl++;
    }
  }
  const int nl_built = nl;

  /* render pairs: camera.go:1343-1457 */
  opair* pairs = (opair*)malloc(sizeof(opair) * 16);
  int np = 0, capp = 16;
  int32_t* solids = NULL;
  int32_t* transp = NULL;
  int ns = 0, nt = 0;
  int rc = T3D_OK;
  int32_t* kids = (int32_t*)malloc(sizeof(int32_t) * (desc->model_count + 1));
  for (uint32_t mi = 0; mi < desc->model_count; mi++) {
    const t3d_model_desc* model = &desc->models[mi];
    if (model->dynamic_batched) continue;                           /* DynamicBatchOwner != nil, camera.go:1346 */
    if (!model->visible || model->mesh == 0) continue;             /* camera.go:1346-1349, 1371 */
    omesh* m = find_mesh(c, model->mesh);
    if (!m) { rc = set_err(c, "render: unknown mesh handle"); free(kids); goto done; }
    rc = validate_model(c, model, m, cd);
    if (rc) { free(kids); goto done; }
    int batcher = 0;                                                /* Model.DynamicBatcher(), model.go:272-274 */
    for (uint32_t k = 0; k < desc->model_count && !batcher; k++)
      batcher = desc->models[k].dynamic_batched && desc->models[k].batch_owner == mi;
    if (batcher) {                                                  /* camera.go:1398-1453 */
      /* one `transparent` for the whole batcher: once set it stays set for its later parts, as written */
      int transparent = 0;
      for (uint32_t p = 0; p < m->d.part_count; p++) {              /* ascending part (the reference walks a map) */
        int n = batch_models(desc, mi, p, kids);
        if (!n) continue;
        const t3d_material_desc* mat = part_mat(model, p);
        if (mat && !mat->visible) continue;
        for (int i = 0; i < n; i++) {
          const t3d_model_desc* ch = &desc->models[kids[i]];
          if (!ch->visible) continue;
          if (!transparent) {
            const omesh* cm = find_mesh(c, ch->mesh);
            for (uint32_t q = 0; cm && q < cm->d.part_count; q++) {
              const t3d_material_desc* cmat = part_mat(ch, q);
              if ((!cmat || cmat->visible) && is_transparent(ch, cmat)) {
                transparent = 1;
                break;
              }
            }
          }
        }
        if (np == capp) {
          capp *= 2;
          pairs = (opair*)realloc(pairs, sizeof(opair) * capp);
        }
        opair* rp = &pairs[np];
        rp->model = mi;
        rp->part = p;
        rp->depth = (transparent || !cd->render_depth) ? v3_dist_sq(cam.pos, m4_row3(&model->transform)) : 0;
        rp->order = mat ? mat->render_order : 0;
        rp->batched = 1;
        if (transparent) {
          transp = (int32_t*)realloc(transp, sizeof(int32_t) * (nt + 1));
          transp[nt++] = np;
        } else {
          solids = (int32_t*)realloc(solids, sizeof(int32_t) * (ns + 1));
          solids[ns++] = np;
        }
        np++;
      }
      continue;
    }
    if (model->frustum_culling) {
      v3 sp;
      float sr;
      model_sphere(&model->transform, model->skinned, m->d.dimensions_min, m->d.dimensions_max, &sp, &sr);
      if (!sphere_in_frustum(&cam, sp, sr)) continue;
    }
    for (uint32_t p = 0; p < m->d.part_count; p++) {
      const t3d_material_desc* mat = part_mat(model, p);
      if (mat && !mat->visible) continue;
      if (np == capp) {
        capp *= 2;
        pairs = (opair*)realloc(pairs, sizeof(opair) * capp);
      }
      opair* rp = &pairs[np];
      rp->model = mi;
      rp->part = p;
      rp->depth = v3_dist_sq(cam.pos, m4_row3(&model->transform));
      rp->order = mat ? mat->render_order : 0;
      rp->batched = 0;
      if (is_transparent(model, mat)) {
        transp = (int32_t*)realloc(transp, sizeof(int32_t) * (nt + 1));
        transp[nt++] = np;
      } else {
        solids = (int32_t*)realloc(solids, sizeof(int32_t) * (ns + 1));
        solids[ns++] = np;
      }
      np++;
    }
  }
  free(kids);
  pair_env penv = {pairs};
  if (!cd->render_depth) go_stable_sort(solids, ns, less_solid_nodepth, &penv);  /* camera.go:1458-1467 */

  if (cd->max_light_count > 0) {                                                  /* camera.go:1474-1497 */
    light_env lenv;
    lenv.lights = lights;
    lenv.pos = cd->has_max_light_count_focus ? v3_from(cd->max_light_count_focus) : cam.pos;
    int32_t* ord = (int32_t*)malloc(sizeof(int32_t) * (nl ? nl : 1));
    for (int i = 0; i < nl; i++) ord[i] = i;
    go_stable_sort(ord, nl, less_light, &lenv);
    olight* sorted = (olight*)calloc(nl + 1, sizeof(olight));
    for (int i = 0; i < nl; i++) sorted[i] = lights[ord[i]];
    free(ord);
    free(lights);
    lights = sorted;
    if (nl > cd->max_light_count) nl = cd->max_light_count;
  }

  go_stable_sort(transp, nt, less_transparent, &penv);                            /* camera.go:2050-2055 */

  {
    uint32_t maxV = 1;
    for (int i = 0; i < c->nmeshes; i++)
      if (c->meshes[i].d.vertex_count > maxV) maxV = c->meshes[i].d.vertex_count;
    oscratch s;
    s.alt_pos = (v3*)calloc(maxV, sizeof(v3));
    s.alt_nrm = (v3*)calloc(maxV, sizeof(v3));
    s.clip = (v4*)calloc(maxV, sizeof(v4));
    s.lights = (float*)calloc((size_t)maxV * 4, sizeof(float));
    s.tnrm = cd->render_normals ? (v3*)calloc(maxV, sizeof(v3)) : NULL;
    s.udepth = (float*)calloc(maxV, sizeof(float));
    obuf ob;
    memset(&ob, 0, sizeof ob);
    ob.want_normals = cd->render_normals;
    uint32_t pair_cap = 0;
    for (int pass = 0; pass < 2 && rc == T3D_OK; pass++) {                        /* camera.go:2057-2109 */
      int32_t* list = pass == 0 ? solids : transp;
      int n = pass == 0 ? ns : nt;
      for (int i = 0; i < n && rc == T3D_OK; i++) {
        const opair* rp = &pairs[list[i]];
        const t3d_model_desc* model = &desc->models[rp->model];
        if (!model->visible || model->auto_batch_static) continue;
        if (rp->batched)
          rc = render_batch(c, desc, &cam, world, rp, lights, nl, &s, &ob, res, &pair_cap);
        else
          rc = render_model_part(c, desc, &cam, world, rp->model, rp->part, lights, nl, &s, &ob, res, &pair_cap, -1);
      }
    }
    free(s.alt_pos);
    free(s.alt_nrm);
    free(s.clip);
    free(s.lights);
    free(s.tnrm);
    free(s.udepth);
    res->vertex_count = ob.n;
    if (ob.want_normals && !ob.normals) ob.normals = (t3d_vertex*)calloc(1, sizeof(t3d_vertex));   /* present, empty */
    res->normals = ob.normals;
    res->color = ob.color;
    res->depth = ob.depth;
    res->indices = ob.idx;
    res->triangle_ids = ob.tri_ids;
    res->depth_keys = ob.keys;
  }

done:
  free(pairs);
  free(solids);
  free(transp);
  for (int i = 0; i < nl_built; i++) light_release(&lights[i]);
  free(lights);
  if (rc != T3D_OK) t3do_result_free(res);
  return rc;
}

void t3do_result_free(t3do_result* r) {
  if (!r) return;
  free(r->pairs);
  free(r->color);
  free(r->depth);
  free(r->indices);
  free(r->triangle_ids);
  free(r->depth_keys);
  free(r->normals);
  memset(r, 0, sizeof *r);
}

/* ------------------------------------------------------------------------ */
/* unit-level entry points                                                   */
/* ------------------------------------------------------------------------ */

void t3do_mat4_mult(const t3d_mat4* a, const t3d_mat4* b, t3d_mat4* out) { *out = m4_mult(a, b); }
void t3do_animate_poses(const t3d_anim_channel* ch, uint32_t nch, float time, t3d_bone_pose* poses, uint32_t n) {
  animate_poses(ch, nch, time, poses, n);
}
double t3do_go_sin(double x) { return go_sin(x); }
double t3do_go_atan2(double y, double x) { return go_atan2(y, x); }
void t3do_quat_slerp(const float a[4], const float b[4], float percent, float out[4]) {
  oquat p = {a[0], a[1], a[2], a[3]}, q = {b[0], b[1], b[2], b[3]};
  oquat r = q_slerp(p, q, percent);
  out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;
}
void t3do_quat_to_mat4(const float q[4], t3d_mat4* out) {
  oquat p = {q[0], q[1], q[2], q[3]};
  *out = q_to_m4(p);
}
int t3do_bone_palettes(const t3d_bone_pose* poses, uint32_t n, const t3d_mat4* root, t3d_mat4* out) {
  m4* world = (m4*)malloc(sizeof(m4) * (n ? n : 1));
  const int r = bone_palettes(poses, n, root, world, out);
  free(world);
  return r;
}
void t3do_mat4_inverted(const t3d_mat4* a, t3d_mat4* out) { *out = m4_inverted(a); }
void t3do_mat4_transposed(const t3d_mat4* a, t3d_mat4* out) { *out = m4_transposed(a); }
void t3do_mat4_decompose(const t3d_mat4* a, float pos[3], float scale[3], t3d_mat4* rot) {
  v3 p, s;
  m4_decompose(a, &p, &s, rot);
  pos[0] = p.x; pos[1] = p.y; pos[2] = p.z;
  scale[0] = s.x; scale[1] = s.y; scale[2] = s.z;
}
void t3do_mat4_mult_vec(const t3d_mat4* a, const float v[3], float out[3]) {
  v3 r = m4_mult_vec(a, v3_from(v));
  out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
void t3do_mat4_mult_vec_w(const t3d_mat4* a, const float v[3], float out[4]) {
  v4 r = m4_mult_vec_w(a, v3_from(v));
  out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;
}
void t3do_projection_perspective(float fovy, float near, float far, float w, float h, t3d_mat4* out) {
  *out = projection_perspective(fovy, near, far, w, h);
}
void t3do_view_matrix(const t3d_camera_desc* cam, t3d_mat4* out) { *out = view_matrix(cam); }
void t3do_vp_matrix(const t3d_camera_desc* cam, t3d_mat4* out) {
  ocam c;
  cam_setup(&c, cam);
  *out = c.vp;
}
void t3do_vec3_unit(const float v[3], float out[3]) {
  v3 r = v3_unit(v3_from(v));
  out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
int t3do_sphere_in_frustum(const t3d_camera_desc* cam, const float pos[3], float radius) {
  ocam c;
  cam_setup(&c, cam);
  return sphere_in_frustum(&c, v3_from(pos), radius);
}
void t3do_model_frustum_sphere(const t3d_mat4* transform, int skinned, const float dims_min[3],
                               const float dims_max[3], float pos[3], float* radius) {
  v3 p;
  model_sphere(transform, skinned, dims_min, dims_max, &p, radius);
  pos[0] = p.x; pos[1] = p.y; pos[2] = p.z;
}
void t3do_inverted_cam_pos(const t3d_mat4* model_transform, const float cam_pos[3], float out[3]) {
  v3 r = inverted_pos(model_transform, v3_from(cam_pos));
  out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
uint32_t t3do_bucket_bin(float depth, float min_depth, float max_depth, int32_t sort_mode) {
  return bucket_bin(depth, min_depth, max_depth, sort_mode);
}
uint32_t t3do_bucket_sort(const float* depths, uint32_t n, float min_depth, float max_depth, int32_t sort_mode,
                          uint32_t* order) {
  uint32_t* bins = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint32_t counts[NBINS + 1];
  bucket_order(depths, n, min_depth, max_depth, sort_mode, order, bins, counts);
  free(bins);
  return n;
}

typedef struct { const uint8_t* t; uint32_t n; } table_env;
static int less_table(void* env, int32_t a, int32_t b) {
  table_env* e = (table_env*)env;
  return e->t[(size_t)a * e->n + b];
}
void t3do_go_stable_sort_table(int32_t* items, uint32_t n, const uint8_t* less_table_) {
  table_env e = {less_table_, n};
  go_stable_sort(items, (int)n, less_table, &e);
}

/* ------------------------------------------------------------------------ */
/* raster sink: flush and the Kage shaders (camera.go:350-531, 1829-2048;    */
/* shaders/base3d.kage:61-153)                                               */
/* ------------------------------------------------------------------------ */
/*
 * The reference hands each flush's vertex lists to ebiten, which draws them
 * with the GPU's fixed-function rasteriser; that stage is not in the
 * reference, so its rules are fixed here (and by the HIP raster kernels, to
 * the bit):
 *  - images are RGBA8 with premultiplied alpha; a float result x is stored as
 *    floor(clamp(x, 0, 1) * 255 + 0.5);
 *  - pixel (x, y) is sampled at its centre (x + 0.5, y + 0.5); a triangle
 *    covers it when the three edge functions, computed in double precision
 *    from each edge's endpoints taken in a fixed order (smaller (y, x) first)
 *    so that a shared edge gives exactly opposite values, all have the sign of
 *    the triangle's area, a zero counting only on a top or left edge;
 *  - attributes are interpolated linearly in screen space (ebiten draws 2D
 *    triangles), with barycentric weights edge / area in double precision;
 *  - triangles of one draw are processed in index order; blending is
 *    SourceOver, out = src + dst * (1 - src.a), in float32 per channel.
 * The shader arithmetic follows the Kage sources in float32: pow(x, 2) is
 * x * x, mod(x, y) = x - y * floor(x / y), mix(a, b, t) = a * (1 - t) + b * t.
 */

int t3do_texture_register(t3do_ctx* c, uint64_t handle, uint32_t w, uint32_t h, const uint8_t* rgba8) {
  if (!w || !h || !rgba8) return set_err(c, "texture: empty");
  if (c->ntexs == c->cap_texs) {
    c->cap_texs = c->cap_texs ? 2 * c->cap_texs : 8;
    c->texs = (otexture*)realloc(c->texs, sizeof(otexture) * c->cap_texs);
  }
  otexture* t = &c->texs[c->ntexs++];
  t->handle = handle;
  t->w = w;
  t->h = h;
  t->px = (uint8_t*)dup_mem(rgba8, 4ull * w * h);
  return T3D_OK;
}

static const otexture* find_texture(t3do_ctx* c, uint64_t h) {
  for (int i = 0; i < c->ntexs; i++)
    if (c->texs[i].handle == h) return &c->texs[i];
  return NULL;
}

static const uint8_t k_white[4] = {255, 255, 255, 255};
static const otexture k_default_img = {0, 1, 1, (uint8_t*)k_white};   /* defaultImg: 1x1 white (tetra3d.go:12, 38) */

static uint8_t q8(float x) {
  if (!(x > 0.0f)) x = 0.0f;
  if (x > 1.0f) x = 1.0f;
  return (uint8_t)floorf(x * 255.0f + 0.5f);
}
static float fractf_(float x) { return x - floorf(x); }
static float modf_(float x, float y) { return x - y * floorf(x / y); }
static float mixf_(float a, float b, float t) { return a * (1.0f - t) + b * t; }

/* imageSrcNUnsafeAt(p) for an unatlased image: the texel containing p */
static void texel_at(const otexture* t, float x, float y, float out[4]) {
  float fx = floorf(x), fy = floorf(y);
  int ix = fx != fx ? 0 : (fx < 0 ? 0 : (fx >= (float)t->w ? (int)t->w - 1 : (int)fx));
  int iy = fy != fy ? 0 : (fy < 0 ? 0 : (fy >= (float)t->h ? (int)t->h - 1 : (int)fy));
  const uint8_t* p = t->px + 4 * ((size_t)iy * t->w + ix);
  for (int k = 0; k < 4; k++) out[k] = (float)p[k] / 255.0f;
}

/* bilinearFilter (base3d.kage:33-50; the clip shader's, camera.go:420-437, without the atlas origin) */
static void bilinear(const otexture* t, float x, float y, float out[4]) {
  const float w = (float)t->w, h = (float)t->h;
  const float p0x = modf_(x - 1.0f / 2.0f, w), p0y = modf_(y - 1.0f / 2.0f, h);
  const float p1x = modf_(x + 1.0f / 2.0f, w), p1y = modf_(y + 1.0f / 2.0f, h);
  float c0[4], c1[4], c2[4], c3[4];
  texel_at(t, p0x, p0y, c0);
  texel_at(t, p1x, p0y, c1);
  texel_at(t, p0x, p1y, c2);
  texel_at(t, p1x, p1y, c3);
  const float rx = fractf_(p1x), ry = fractf_(p1y);
  for (int k = 0; k < 4; k++) out[k] = mixf_(mixf_(c0[k], c1[k], rx), mixf_(c2[k], c3[k], rx), ry);
}

/* encodeDepth / decodeDepth (camera.go:354-364) */
static void encode_depth(float d, float out[4]) {
  out[0] = floorf(d * 255.0f) / 255.0f;
  out[1] = floorf(fractf_(d * 255.0f) * 255.0f) / 255.0f;
  out[2] = fractf_(d * 255.0f * 255.0f);
  out[3] = 1.0f;
}
static float decode_depth(const uint8_t px[4]) {
  return (float)px[0] / 255.0f + (((float)px[1] / 255.0f) / 255.0f) + (((float)px[2] / 255.0f) / 65025.0f);
}

static void source_over(uint8_t dst[4], const float src[4]) {
  float o[4];
  for (int k = 0; k < 4; k++) o[k] = src[k] + ((float)dst[k] / 255.0f) * (1.0f - src[3]);
  for (int k = 0; k < 4; k++) dst[k] = q8(o[k]);
}

/* The edge function of a -> b at c, evaluated from the endpoint with the
 * smaller (y, x) so that the two triangles sharing an edge get exactly
 * opposite values. */
static double edge_fn(float ax, float ay, float bx, float by, double cx, double cy) {
  const int swap = (ay > by) || (ay == by && ax > bx);
  const double px = swap ? bx : ax, py = swap ? by : ay, qx = swap ? ax : bx, qy = swap ? ay : by;
  const double e = (qx - px) * (cy - py) - (qy - py) * (cx - px);
  return swap ? -e : e;
}
/* a zero edge value counts on a top edge (horizontal, interior below) or a left edge (interior to the right, y down) */
static int top_left(float ax, float ay, float bx, float by, double s) {
  const double dx = s * ((double)bx - (double)ax), dy = s * ((double)by - (double)ay);
  return (dy == 0.0 && dx > 0.0) || dy < 0.0;
}

typedef struct {
  const t3d_vertex* v;    /* the triangle's three vertices in one list */
  double area, s;
  int x0, x1, y0, y1;     /* pixel range of its bounding box on the image */
} otri;

static int tri_setup(otri* t, const t3d_vertex* v, int W, int H) {
  t->v = v;
  t->area = edge_fn(v[0].dst_x, v[0].dst_y, v[1].dst_x, v[1].dst_y, (double)v[2].dst_x, (double)v[2].dst_y);
  if (!(t->area != 0.0) || t->area != t->area) return 0;
  t->s = t->area > 0 ? 1.0 : -1.0;
  double mnx = v[0].dst_x, mxx = v[0].dst_x, mny = v[0].dst_y, mxy = v[0].dst_y;
  for (int k = 1; k < 3; k++) {
    if (v[k].dst_x < mnx) mnx = v[k].dst_x;
    if (v[k].dst_x > mxx) mxx = v[k].dst_x;
    if (v[k].dst_y < mny) mny = v[k].dst_y;
    if (v[k].dst_y > mxy) mxy = v[k].dst_y;
  }
  /* pixels whose centre x + 0.5 lies in [mnx, mxx] */
  double fx0 = ceil(mnx - 0.5), fx1 = floor(mxx - 0.5), fy0 = ceil(mny - 0.5), fy1 = floor(mxy - 0.5);
  if (!(fx0 <= fx1) || !(fy0 <= fy1) || fx1 < 0 || fy1 < 0 || fx0 > W - 1 || fy0 > H - 1) return 0;
  t->x0 = fx0 < 0 ? 0 : (int)fx0;
  t->y0 = fy0 < 0 ? 0 : (int)fy0;
  t->x1 = fx1 > W - 1 ? W - 1 : (int)fx1;
  t->y1 = fy1 > H - 1 ? H - 1 : (int)fy1;
  return 1;
}

/* coverage of pixel (x, y) and its barycentric weights */
static int tri_cover(const otri* t, int x, int y, double l[3]) {
  const t3d_vertex* v = t->v;
  const double cx = x + 0.5, cy = y + 0.5;
  const double e12 = edge_fn(v[1].dst_x, v[1].dst_y, v[2].dst_x, v[2].dst_y, cx, cy);
  const double e20 = edge_fn(v[2].dst_x, v[2].dst_y, v[0].dst_x, v[0].dst_y, cx, cy);
  const double e01 = edge_fn(v[0].dst_x, v[0].dst_y, v[1].dst_x, v[1].dst_y, cx, cy);
  const double s = t->s;
  if (!(s * e12 > 0 || (e12 == 0 && top_left(v[1].dst_x, v[1].dst_y, v[2].dst_x, v[2].dst_y, s)))) return 0;
  if (!(s * e20 > 0 || (e20 == 0 && top_left(v[2].dst_x, v[2].dst_y, v[0].dst_x, v[0].dst_y, s)))) return 0;
  if (!(s * e01 > 0 || (e01 == 0 && top_left(v[0].dst_x, v[0].dst_y, v[1].dst_x, v[1].dst_y, s)))) return 0;
  l[0] = e12 / t->area;
  l[1] = e20 / t->area;
  l[2] = e01 / t->area;
  return 1;
}
static float interp(const double l[3], float a0, float a1, float a2) {
  return (float)(l[0] * (double)a0 + l[1] * (double)a1 + l[2] * (double)a2);
}

/* World uniforms of the colour pass (camera.go:1956-1977, world.go:96-110) */
typedef struct {
  float fog[4];           /* FogColor RGB, FogMode (-1: FogOn false) */
  float fog_range[2];
  float dither;
  int curve;
} ofog;

/* sample an image at texel position (x, y): nearest (the texel containing it) or base3d's bilinear */
static void sample(const otexture* img, int linear, float x, float y, float out[4]) {
  if (linear) bilinear(img, x, y, out);
  else texel_at(img, x, y, out);
}

static const float k_bayer[16] = {   /* camera.go:1257-1262 */
    1.0f / 17.0f,  9.0f / 17.0f, 3.0f / 17.0f,  11.0f / 17.0f, 13.0f / 17.0f, 5.0f / 17.0f, 15.0f / 17.0f, 7.0f / 17.0f,
    4.0f / 17.0f, 12.0f / 17.0f, 2.0f / 17.0f, 10.0f / 17.0f, 16.0f / 17.0f, 8.0f / 17.0f, 14.0f / 17.0f, 6.0f / 17.0f};

static float smoothstepf_(float e0, float e1, float x) {
  float t = (x - e0) / (e1 - e0);
  t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
  return t * t * (3.0f - 2.0f * t);
}

/* base3d.kage Fragment (61-153) at a pixel whose depth intermediate has alpha > 0 */
static void base3d(const otexture* img, int linear, const ofog* fg, int fogless, int pc, const uint8_t inter[4],
                   const float vc[4], float sx, float sy, float c0, float px, float py, float out[4]) {
  float tx = sx, ty = sy;
  if (pc > 0) {                                                   /* tx *= 1.0 / custom.x */
    const float r = 1.0f / c0;
    tx *= r;
    ty *= r;
  }
  tx = modf_(tx, (float)img->w);                                  /* mod(tx, srcSize) */
  ty = modf_(ty, (float)img->h);
  float tex[4];
  sample(img, linear, tx, ty, tex);
  for (int k = 0; k < 4; k++) out[k] = tex[k] * vc[k];
  for (int k = 0; k < 3; k++) out[k] *= vc[3];                    /* colorTex.rgb *= color.a */
  if (fogless) return;
  const float dd = decode_depth(inter);
  float v = dd;
  if (fg->curve == T3D_FOG_CURVE_OUT_CIRC) v = sqrtf(1.0f - (dd - 1.0f) * (dd - 1.0f));
  else if (fg->curve == T3D_FOG_CURVE_IN_CIRC) v = 1.0f - sqrtf(1.0f - dd * dd);
  else if (fg->curve != T3D_FOG_CURVE_LINEAR) v = -INFINITY;     /* no curve matches: d stays 0 */
  float d = v == -INFINITY ? 0.0f : smoothstepf_(fg->fog_range[0], fg->fog_range[1], v);
  if (fg->dither > 0) {
    const int yc = ((int)(py / fg->dither)) % 4, xc = ((int)(px / fg->dither)) % 4;
    d = (d - k_bayer[yc * 4 + xc]) < 0.0f ? 0.0f : 1.0f;          /* step(0, d - Bayer) */
  }
  const float a = out[3];
  if (fg->fog[3] == 0.0f) {
    for (int k = 0; k < 3; k++) out[k] += fg->fog[k] * d * a;
  } else if (fg->fog[3] == 1.0f) {
    for (int k = 0; k < 3; k++) out[k] -= fg->fog[k] * d * a;
  } else if (fg->fog[3] == 2.0f) {
    for (int k = 0; k < 3; k++) out[k] = mixf_(out[k], fg->fog[k], d) * a;
  }
}

/* Draws the flushes of a render into the camera's result textures: Camera.Clear (camera.go:1002-1040), then per
 * flush in order (camera.go:1829-2048) the depth pass into a cleared intermediate against the result depth (the
 * depth shader or, for AlphaClip, the clip-alpha shader), the intermediate drawn over the result depth unless the
 * pair is transparent, and the colour pass through base3d with the intermediate; without RenderDepth one plain
 * textured draw per flush.  color is W*H*4 bytes; depth (RenderDepth only) likewise, or NULL. */
int t3do_rasterize(t3do_ctx* c, const t3d_render_desc* desc, const t3do_result* res, uint8_t* color, uint8_t* depth) {
  const t3d_camera_desc* cd = desc->camera;
  const int W = cd->width, H = cd->height;
  const t3d_world_desc* world = (desc->world && desc->world->present) ? desc->world : NULL;
  const int rdepth = cd->render_depth != 0;
  if (rdepth && !depth) return set_err(c, "rasterize: a depth image is needed with RenderDepth");
  /* Camera.Clear: the world's clear colour (transparent black without a world), premultiplied */
  float clr[4] = {0, 0, 0, 0};
  if (world)
    for (int k = 0; k < 4; k++) clr[k] = world->clear_color[k];
  const uint8_t cpx[4] = {q8(clr[0] * clr[3]), q8(clr[1] * clr[3]), q8(clr[2] * clr[3]), q8(clr[3])};
  for (size_t i = 0; i < (size_t)W * H; i++) memcpy(color + 4 * i, cpx, 4);
  if (depth) memset(depth, 0, 4ull * W * H);
  ofog fg;
  memset(&fg, 0, sizeof fg);
  fg.fog_range[1] = 1.0f;
  if (world) {
    for (int k = 0; k < 3; k++) fg.fog[k] = world->fog_color[k];
    fg.fog[3] = world->fog_on ? (float)world->fog_mode : -1.0f;
    fg.fog_range[0] = world->fog_range[0];
    fg.fog_range[1] = world->fog_range[1];
    fg.dither = world->dithered_fog_size;
    fg.curve = world->fog_curve;
  }
  uint8_t* inter = rdepth ? (uint8_t*)calloc((size_t)W * H, 4) : NULL;
  int rc = T3D_OK;
  for (uint32_t f = 0; f < res->pair_count && rc == T3D_OK; f++) {
    const t3d_pair_result* pr = &res->pairs[f];
    if (!pr->survivor_count) continue;                              /* flush: vertexListIndex == 0 */
    const t3d_model_desc* model = &desc->models[pr->model_index];
    const t3d_material_desc* mat = part_mat(model, pr->part_index);
    if (mat && !mat->blend_source_over) { rc = set_err(c, "rasterize: only BlendSourceOver is drawn"); break; }
    const otexture* img = &k_default_img;
    if (mat && mat->use_texture && mat->texture) {
      img = find_texture(c, mat->texture);
      if (!img) { rc = set_err(c, "rasterize: unknown texture"); break; }
    }
    const int linear = mat && mat->texture_filter == T3D_FILTER_LINEAR;
    const int fogless = mat && mat->fogless;
    const int clip = mat && mat->transparency_mode == T3D_TRANSPARENCY_ALPHA_CLIP;
    const int pc = cd->perspective_corrected_texture_mapping ? 1 : 0;
    const t3d_vertex* cv = res->color + pr->vertex_offset;
    const t3d_vertex* dv = res->depth + pr->vertex_offset;
    const uint32_t nt = pr->survivor_count;
    if (rdepth) {
      /* depth pass: last passing fragment of the draw wins (each tests the result depth only) */
      int bx0 = W, by0 = H, bx1 = -1, by1 = -1;
      for (uint32_t k = 0; k < nt; k++) {
        otri t;
        if (!tri_setup(&t, dv + 3 * k, W, H)) continue;
        if (t.x0 < bx0) bx0 = t.x0;
        if (t.y0 < by0) by0 = t.y0;
        if (t.x1 > bx1) bx1 = t.x1;
        if (t.y1 > by1) by1 = t.y1;
        const t3d_vertex* v = t.v;
        for (int y = t.y0; y <= t.y1; y++)
          for (int x = t.x0; x <= t.x1; x++) {
            double l[3];
            if (!tri_cover(&t, x, y, l)) continue;
            const float d = interp(l, v[0].color_r, v[1].color_r, v[2].color_r);
            const uint8_t* ex = depth + 4 * ((size_t)y * W + x);
            if (!(ex[3] == 0 || decode_depth(ex) > d)) continue;
            float enc[4];
            encode_depth(d, enc);
            uint8_t* ip = inter + 4 * ((size_t)y * W + x);
            if (clip) {                                             /* camera.go:441-478 */
              float tx = interp(l, v[0].src_x, v[1].src_x, v[2].src_x), ty = interp(l, v[0].src_y, v[1].src_y, v[2].src_y);
              if (pc) {
                const float r = 1.0f / interp(l, v[0].custom0, v[1].custom0, v[2].custom0);
                tx *= r;
                ty *= r;
              }
              tx = fractf_(tx / (float)img->w) * (float)img->w;
              ty = fractf_(ty / (float)img->h) * (float)img->h;
              float tex[4];
              sample(img, linear, tx, ty, tex);
              if (tex[3] <= 0.8f) continue;
              enc[3] = tex[3];
              source_over(ip, enc);
            } else {
              source_over(ip, enc);
            }
          }
      }
      if (!is_transparent(model, mat))                              /* resultDepthTexture.DrawImage(intermediate) */
        for (int y = by0; y <= by1; y++)
          for (int x = bx0; x <= bx1; x++) {
            const uint8_t* ip = inter + 4 * ((size_t)y * W + x);
            const float src[4] = {(float)ip[0] / 255.0f, (float)ip[1] / 255.0f, (float)ip[2] / 255.0f, (float)ip[3] / 255.0f};
            source_over(depth + 4 * ((size_t)y * W + x), src);
          }
      /* colour pass through base3d.kage */
      for (uint32_t k = 0; k < nt; k++) {
        otri t;
        if (!tri_setup(&t, cv + 3 * k, W, H)) continue;
        const t3d_vertex* v = t.v;
        for (int y = t.y0; y <= t.y1; y++)
          for (int x = t.x0; x <= t.x1; x++) {
            double l[3];
            if (!tri_cover(&t, x, y, l)) continue;
            const uint8_t* ip = inter + 4 * ((size_t)y * W + x);
            if (ip[3] == 0) continue;                               /* discard() */
            const float vc[4] = {interp(l, v[0].color_r, v[1].color_r, v[2].color_r),
                                 interp(l, v[0].color_g, v[1].color_g, v[2].color_g),
                                 interp(l, v[0].color_b, v[1].color_b, v[2].color_b),
                                 interp(l, v[0].color_a, v[1].color_a, v[2].color_a)};
            float o[4];
            base3d(img, linear, &fg, fogless, pc, ip, vc, interp(l, v[0].src_x, v[1].src_x, v[2].src_x),
                   interp(l, v[0].src_y, v[1].src_y, v[2].src_y), interp(l, v[0].custom0, v[1].custom0, v[2].custom0),
                   (float)x + 0.5f, (float)y + 0.5f, o);
            source_over(color + 4 * ((size_t)y * W + x), o);
          }
      }
      for (int y = by0; y <= by1; y++) memset(inter + 4 * ((size_t)y * W + bx0), 0, bx1 >= bx0 ? 4ull * (bx1 - bx0 + 1) : 0);
    } else {
      /* DrawTriangles(colorVertexList, img): texel times the vertex colour made premultiplied */
      for (uint32_t k = 0; k < nt; k++) {
        otri t;
        if (!tri_setup(&t, cv + 3 * k, W, H)) continue;
        const t3d_vertex* v = t.v;
        for (int y = t.y0; y <= t.y1; y++)
          for (int x = t.x0; x <= t.x1; x++) {
            double l[3];
            if (!tri_cover(&t, x, y, l)) continue;
            const float a = interp(l, v[0].color_a, v[1].color_a, v[2].color_a);
            const float vc[4] = {interp(l, v[0].color_r, v[1].color_r, v[2].color_r) * a,
                                 interp(l, v[0].color_g, v[1].color_g, v[2].color_g) * a,
                                 interp(l, v[0].color_b, v[1].color_b, v[2].color_b) * a, a};
            float tx = modf_(interp(l, v[0].src_x, v[1].src_x, v[2].src_x), (float)img->w);
            float ty = modf_(interp(l, v[0].src_y, v[1].src_y, v[2].src_y), (float)img->h);
            float tex[4], o[4];
            sample(img, linear, tx, ty, tex);
            for (int q = 0; q < 4; q++) o[q] = tex[q] * vc[q];
            source_over(color + 4 * ((size_t)y * W + x), o);
          }
      }
    }
  }
  free(inter);
  return rc;
}
