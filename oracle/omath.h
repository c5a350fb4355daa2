// Oracle: vector/quaternion math with the reference's operation order (packages/steel-webgpu/src/wgsl.js:47-66,
// artwgsl.js:58-95).  Test infrastructure only: the CPU restatement that the HIP kernels are checked against.
// Every expression keeps WGSL's evaluation order (left to right) and is compiled with -ffp-contract=off,
// so the float instantiation rounds exactly as the kernels do.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

namespace oracle {

template <class T> struct V3 {
  T x, y, z;
  T& operator[](uint32_t k) { return k == 0 ? x : (k == 1 ? y : z); }
  T operator[](uint32_t k) const { return k == 0 ? x : (k == 1 ? y : z); }
};
template <class T> struct V4 { T x, y, z, w; };
template <class T> struct V2 { T x, y; };

template <class T> inline V3<T> v3(T x, T y, T z) { return {x, y, z}; }
template <class T> inline V3<T> xyz(const V4<T>& a) { return {a.x, a.y, a.z}; }
template <class T> inline V4<T> v4(const V3<T>& a, T w) { return {a.x, a.y, a.z, w}; }
template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T> inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T> inline V3<T> operator/(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> inline V3<T> operator*(V3<T> a, V3<T> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class T> inline V3<T>& operator+=(V3<T>& a, V3<T> b) { a = a + b; return a; }
template <class T> inline V3<T>& operator-=(V3<T>& a, V3<T> b) { a = a - b; return a; }
template <class T> inline V4<T> operator+(V4<T> a, V4<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
template <class T> inline V4<T> operator*(V4<T> a, T s) { return {a.x * s, a.y * s, a.z * s, a.w * s}; }
template <class T> inline V4<T> operator*(T s, V4<T> a) { return {s * a.x, s * a.y, s * a.z, s * a.w}; }
template <class T> inline V2<T> operator+(V2<T> a, V2<T> b) { return {a.x + b.x, a.y + b.y}; }
template <class T> inline V2<T> operator-(V2<T> a, V2<T> b) { return {a.x - b.x, a.y - b.y}; }
template <class T> inline V2<T> operator*(V2<T> a, T s) { return {a.x * s, a.y * s}; }
template <class T> inline V2<T> operator*(T s, V2<T> a) { return {s * a.x, s * a.y}; }

template <class T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline T dot(V2<T> a, V2<T> b) { return a.x * b.x + a.y * b.y; }
template <class T> inline T dot4(V4<T> a, V4<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
template <class T> inline V3<T> cross(V3<T> a, V3<T> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T> inline T length(V3<T> a) { return std::sqrt(dot(a, a)); }
template <class T> inline T length(V2<T> a) { return std::sqrt(dot(a, a)); }
// normalize(v) = v * (1 / sqrt(dot(v, v))): the one definition used by both the oracle and the kernels.
template <class T> inline V3<T> normalize(V3<T> a) { T il = T(1) / std::sqrt(dot(a, a)); return a * il; }
template <class T> inline V2<T> normalize(V2<T> a) { T il = T(1) / std::sqrt(dot(a, a)); return a * il; }
template <class T> inline V4<T> normalize4(V4<T> a) { T il = T(1) / std::sqrt(dot4(a, a)); return a * il; }
template <class T> inline T inverse_sqrt(T x) { return T(1) / std::sqrt(x); }
template <class T> inline V3<T> vabs(V3<T> a) { return {std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)}; }
template <class T> inline T mn(T a, T b) { return b < a ? b : a; }
template <class T> inline T mx(T a, T b) { return a < b ? b : a; }
template <class T> inline T clampv(T x, T lo, T hi) { return mn(mx(x, lo), hi); }
template <class T> inline bool vne0(V3<T> a) { return a.x != T(0) || a.y != T(0) || a.z != T(0); }

// wgsl.js:48-59
template <class T> inline V3<T> qrot(V4<T> q, V3<T> v) {
  V3<T> u = xyz(q); T s = q.w;
  return (T(2) * dot(u, v)) * u + (s * s - dot(u, u)) * v + (T(2) * s) * cross(u, v);
}
template <class T> inline V4<T> qconj(V4<T> q) { return {-q.x, -q.y, -q.z, q.w}; }
template <class T> inline V3<T> qinvrot(V4<T> q, V3<T> v) { return qrot(qconj(q), v); }
template <class T> inline V4<T> qmul(V4<T> a, V4<T> b) {
  V3<T> v = a.w * xyz(b) + b.w * xyz(a) + cross(xyz(a), xyz(b));
  return {v.x, v.y, v.z, a.w * b.w - dot(xyz(a), xyz(b))};
}

// 3x3 matrix, column major (WGSL mat3x3f).
template <class T> struct M3 { V3<T> c[3]; };
template <class T> inline V3<T> mul(const M3<T>& m, V3<T> v) { return m.c[0] * v.x + m.c[1] * v.y + m.c[2] * v.z; }
template <class T> inline M3<T> transpose(const M3<T>& m) {
  M3<T> r;
  r.c[0] = {m.c[0].x, m.c[1].x, m.c[2].x};
  r.c[1] = {m.c[0].y, m.c[1].y, m.c[2].y};
  r.c[2] = {m.c[0].z, m.c[1].z, m.c[2].z};
  return r;
}
template <class T> inline M3<T> mul(const M3<T>& a, const M3<T>& b) {
  M3<T> r;
  for (int j = 0; j < 3; j++) r.c[j] = mul(a, b.c[j]);
  return r;
}
template <class T> inline M3<T> madd(const M3<T>& a, const M3<T>& b) { M3<T> r; for (int j = 0; j < 3; j++) r.c[j] = a.c[j] + b.c[j]; return r; }
template <class T> inline M3<T> msub(const M3<T>& a, const M3<T>& b) { M3<T> r; for (int j = 0; j < 3; j++) r.c[j] = a.c[j] - b.c[j]; return r; }
template <class T> inline M3<T> mscale(const M3<T>& a, T s) { M3<T> r; for (int j = 0; j < 3; j++) r.c[j] = a.c[j] * s; return r; }
template <class T> inline M3<T> mscale(T s, const M3<T>& a) { M3<T> r; for (int j = 0; j < 3; j++) r.c[j] = s * a.c[j]; return r; }
template <class T> inline M3<T> qmat(V4<T> q) {
  M3<T> m;
  m.c[0] = qrot(q, v3<T>(1, 0, 0)); m.c[1] = qrot(q, v3<T>(0, 1, 0)); m.c[2] = qrot(q, v3<T>(0, 0, 1));
  return m;
}
template <class T> inline T& comp(V3<T>& v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }
template <class T> inline T compc(const V3<T>& v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

// ---- transcendentals: Cephes (S. L. Moshier) single-precision sinf/cosf/atanf, written out so the
// float oracle and the kernels round identically; the double oracle uses libm.
inline float sinf_c(float xx);
inline float cosf_c(float xx);
inline void sincos_reduce(float xx, float& x, int& j) {
  const float DP1 = 0.78515625f, DP2 = 2.4187564849853515625e-4f, DP3 = 3.77489497744594108e-8f;
  const float FOPI = 1.27323954473516f;
  j = (int)(FOPI * xx);
  float y = (float)j;
  if (j & 1) { j += 1; y += 1.0f; }
  x = ((xx - y * DP1) - y * DP2) - y * DP3;
}
inline float sin_poly(float x) { float z = x * x; return ((-1.9515295891e-4f * z + 8.3321608736e-3f) * z - 1.6666654611e-1f) * z * x + x; }
inline float cos_poly(float x) { float z = x * x; return ((2.443315711809948e-5f * z - 1.388731625493765e-3f) * z + 4.166664568298827e-2f) * z * z - 0.5f * z + 1.0f; }
inline float sinf_c(float xx) {
  float sign = 1.0f;
  if (xx < 0.0f) { sign = -1.0f; xx = -xx; }
  float x; int j; sincos_reduce(xx, x, j);
  j &= 7;
  if (j > 3) { sign = -sign; j -= 4; }
  float y = (j == 1 || j == 2) ? cos_poly(x) : sin_poly(x);
  return sign < 0.0f ? -y : y;
}
inline float cosf_c(float xx) {
  if (xx < 0.0f) xx = -xx;
  float x; int j; sincos_reduce(xx, x, j);
  j &= 7;
  float sign = 1.0f;
  if (j > 3) { j -= 4; sign = -sign; }
  if (j > 1) sign = -sign;
  float y = (j == 1 || j == 2) ? sin_poly(x) : cos_poly(x);
  return sign < 0.0f ? -y : y;
}
inline float atanf_c(float xx) {
  float sign = 1.0f, x = xx;
  if (x < 0.0f) { sign = -1.0f; x = -x; }
  float y;
  if (x > 2.414213562373095f) { y = 1.5707963267948966f; x = -(1.0f / x); }
  else if (x > 0.4142135623730950f) { y = 0.7853981633974483f; x = (x - 1.0f) / (x + 1.0f); }
  else y = 0.0f;
  float z = x * x;
  y = y + ((((8.05374449538e-2f * z - 1.38776856032e-1f) * z + 1.99777106478e-1f) * z - 3.33329491539e-1f) * z * x + x);
  return sign < 0.0f ? -y : y;
}
inline float atan2f_c(float y, float x) {
  const float PI = 3.14159265358979f, PIO2 = 1.5707963267948966f;
  if (x == 0.0f) { if (y > 0.0f) return PIO2; if (y < 0.0f) return -PIO2; return 0.0f; }
  float z = atanf_c(y / x);
  if (x < 0.0f) z = (y < 0.0f) ? z - PI : z + PI;
  return z;
}
template <class T> inline T tsin(T x);
template <class T> inline T tcos(T x);
template <class T> inline T tatan2(T y, T x);
template <> inline float tsin<float>(float x) { return sinf_c(x); }
template <> inline float tcos<float>(float x) { return cosf_c(x); }
template <> inline float tatan2<float>(float y, float x) { return atan2f_c(y, x); }
template <> inline double tsin<double>(double x) { return std::sin(x); }
template <> inline double tcos<double>(double x) { return std::cos(x); }
template <> inline double tatan2<double>(double y, double x) { return std::atan2(y, x); }

template <class T> inline V4<T> qaxis(V3<T> axis, T angle) {
  T s = tsin<T>(T(0.5) * angle);
  return v4(axis * s, tcos<T>(T(0.5) * angle));
}

inline uint32_t hash_u32(uint32_t x) {  // wgsl.js:61-65 (lowbias32)
  x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
  return x;
}
inline uint32_t f2u(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float u2f(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
// WGSL u32(f) / i32(f) of the small non-negative integers the records hold.
template <class T> inline uint32_t tu(T f) { return f <= T(0) ? 0u : (uint32_t)f; }
template <class T> inline int32_t ti(T f) { return (int32_t)f; }
inline uint32_t ctz32(uint32_t x) { return x == 0 ? 32u : (uint32_t)__builtin_ctz(x); }

}  // namespace oracle
