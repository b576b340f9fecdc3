// bch_math.hpp -- f32 vector / quaternion / 3x3 arithmetic for the collision
// kernels.
//
// Every operation is spelled with an explicit evaluation order and the
// library is compiled with -ffp-contract=off, so each product and sum is a
// separately rounded IEEE-754 f32 operation: world AABBs (and therefore the
// broadphase pair set) come out bit-identical to a CPU evaluation of the same
// formulas, as glam evaluates them for the reference
// (crates/bullet-math/src/lib.rs:30-54).
//
//   dot(a, b)   = (a.x*b.x + a.y*b.y) + a.z*b.z
//   cross(a, b) = (a.y*b.z - b.y*a.z, a.z*b.x - b.z*a.x, a.x*b.y - b.x*a.y)
//   normalize   = v * (1 / |v|)
//   q * v       = (v (w*w - b.b) + b ((v.b) * 2)) + (b x v)(w * 2), b = q.xyz
//   M * v       = (c0 v.x + c1 v.y) + c2 v.z   (column-major)
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define BCH_HD __host__ __device__ __forceinline__
#else
#include <math.h>
#define BCH_HD inline
struct float4 {
  float x, y, z, w;
};
#endif

namespace bch {

struct V3 {
  float x, y, z;
};
BCH_HD V3 mk(float x, float y, float z) { return V3{x, y, z}; }
BCH_HD V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
BCH_HD V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
BCH_HD V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
BCH_HD V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
BCH_HD V3 operator*(float s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
BCH_HD V3 operator/(V3 a, float s) { return V3{a.x / s, a.y / s, a.z / s}; }
BCH_HD float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
BCH_HD V3 cross(V3 a, V3 b) { return V3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
BCH_HD float len2(V3 a) { return dot(a, a); }
BCH_HD float fsqrt(float x) { return sqrtf(x); }
BCH_HD float len(V3 a) { return fsqrt(dot(a, a)); }
BCH_HD V3 normalize(V3 a) { return a * (1.0f / len(a)); }
// min/max that return the first operand unless the second is strictly
// smaller/larger (the same value for every non-NaN input).
BCH_HD float mn(float a, float b) { return b < a ? b : a; }
BCH_HD float mx(float a, float b) { return b > a ? b : a; }
// NaN-ignoring max (Rust f32::max).
BCH_HD float fmax_nan(float a, float b) { return (a != a) ? b : ((b != b) ? a : (b > a ? b : a)); }
BCH_HD V3 vmin(V3 a, V3 b) { return V3{mn(a.x, b.x), mn(a.y, b.y), mn(a.z, b.z)}; }
BCH_HD V3 vmax(V3 a, V3 b) { return V3{mx(a.x, b.x), mx(a.y, b.y), mx(a.z, b.z)}; }
BCH_HD V3 vabs(V3 a) { return V3{fabsf(a.x), fabsf(a.y), fabsf(a.z)}; }
BCH_HD V3 splat(float s) { return V3{s, s, s}; }
BCH_HD float comp(V3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
// Component-wise select: scalar selects only, so structs stay in registers.
BCH_HD V3 sel(bool c, V3 a, V3 b) { return V3{c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z}; }
BCH_HD bool near_eq(V3 a, V3 b, float eps) {
  return fabsf(a.x - b.x) <= eps && fabsf(a.y - b.y) <= eps && fabsf(a.z - b.z) <= eps;
}
BCH_HD float signum(float x) { return (x != x) ? x : copysignf(1.0f, x); }
BCH_HD bool is_nan(float x) { return x != x; }

struct Q4 {
  float x, y, z, w;
};
BCH_HD V3 qrot(Q4 q, V3 v) {
  V3 b{q.x, q.y, q.z};
  float b2 = dot(b, b);
  V3 t1 = v * (q.w * q.w - b2);
  V3 t2 = b * (dot(v, b) * 2.0f);
  V3 t3 = cross(b, v) * (q.w * 2.0f);
  return (t1 + t2) + t3;
}
BCH_HD Q4 qconj(Q4 q) { return Q4{-q.x, -q.y, -q.z, q.w}; }

struct M3 {
  V3 c0, c1, c2;
};
BCH_HD V3 col(const M3& m, int i) { return sel(i == 0, m.c0, sel(i == 1, m.c1, m.c2)); }
BCH_HD M3 sel(bool c, const M3& a, const M3& b) { return M3{sel(c, a.c0, b.c0), sel(c, a.c1, b.c1), sel(c, a.c2, b.c2)}; }
BCH_HD M3 from_quat(Q4 q) {
  float x2 = q.x + q.x, y2 = q.y + q.y, z2 = q.z + q.z;
  float xx = q.x * x2, xy = q.x * y2, xz = q.x * z2;
  float yy = q.y * y2, yz = q.y * z2, zz = q.z * z2;
  float wx = q.w * x2, wy = q.w * y2, wz = q.w * z2;
  M3 m;
  m.c0 = V3{1.0f - (yy + zz), xy + wz, xz - wy};
  m.c1 = V3{xy - wz, 1.0f - (xx + zz), yz + wx};
  m.c2 = V3{xz + wy, yz - wx, 1.0f - (xx + yy)};
  return m;
}
BCH_HD V3 mmul(const M3& m, V3 v) { return (m.c0 * v.x + m.c1 * v.y) + m.c2 * v.z; }
BCH_HD M3 transpose(const M3& m) {
  M3 t;
  t.c0 = V3{m.c0.x, m.c1.x, m.c2.x};
  t.c1 = V3{m.c0.y, m.c1.y, m.c2.y};
  t.c2 = V3{m.c0.z, m.c1.z, m.c2.z};
  return t;
}
BCH_HD M3 mmulm(const M3& a, const M3& b) { return M3{mmul(a, b.c0), mmul(a, b.c1), mmul(a, b.c2)}; }

struct Box6 {
  V3 lo, hi;
};
// Closed-interval overlap (crates/bullet-math/src/util.rs:28-38).
BCH_HD bool overlaps(const Box6& a, const Box6& b) {
  return a.lo.x <= b.hi.x && a.lo.y <= b.hi.y && a.lo.z <= b.hi.z && b.lo.x <= a.hi.x && b.lo.y <= a.hi.y &&
         b.lo.z <= a.hi.z;
}

}  // namespace bch
