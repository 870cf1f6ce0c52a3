// vec3.hpp -- 3-vector / 3x3 helpers for the gfx950 element kernels, generic over the
// scalar: double for the step itself, dd (a value with one forward-mode tangent) for the
// tangent kernels of the differentiable step.
#pragma once
#include <hip/hip_runtime.h>

#define YS_DEV __device__ __forceinline__

// ---------------------------------------------------------------------------
// dd: value v and directional derivative d.  Every operation applies the chain rule to
// d and computes v exactly as the double code does, so the value part of a dd run is
// bit-identical to the double run.
// ---------------------------------------------------------------------------
struct dd {
    double v, d;
    dd() = default;
    __host__ __device__ constexpr dd(double a) : v(a), d(0.0) {}
    __host__ __device__ constexpr dd(double a, double b) : v(a), d(b) {}
};

YS_DEV dd operator+(dd a, dd b) { return dd(a.v + b.v, a.d + b.d); }
YS_DEV dd operator-(dd a, dd b) { return dd(a.v - b.v, a.d - b.d); }
YS_DEV dd operator-(dd a) { return dd(-a.v, -a.d); }
YS_DEV dd operator*(dd a, dd b) { return dd(a.v * b.v, a.d * b.v + a.v * b.d); }
YS_DEV dd operator/(dd a, dd b) { double q = a.v / b.v; return dd(q, (a.d - q * b.d) / b.v); }
YS_DEV dd operator+(dd a, double b) { return dd(a.v + b, a.d); }
YS_DEV dd operator+(double a, dd b) { return dd(a + b.v, b.d); }
YS_DEV dd operator-(dd a, double b) { return dd(a.v - b, a.d); }
YS_DEV dd operator-(double a, dd b) { return dd(a - b.v, -b.d); }
YS_DEV dd operator*(dd a, double b) { return dd(a.v * b, a.d * b); }
YS_DEV dd operator*(double a, dd b) { return dd(a * b.v, a * b.d); }
YS_DEV dd operator/(dd a, double b) { return dd(a.v / b, a.d / b); }
YS_DEV dd operator/(double a, dd b) { double q = a / b.v; return dd(q, -q * b.d / b.v); }
YS_DEV dd& operator+=(dd& a, dd b) { a = a + b; return a; }
YS_DEV dd& operator-=(dd& a, dd b) { a = a - b; return a; }
YS_DEV bool operator<(dd a, dd b) { return a.v < b.v; }
YS_DEV bool operator>(dd a, dd b) { return a.v > b.v; }
YS_DEV bool operator<=(dd a, dd b) { return a.v <= b.v; }
YS_DEV bool operator>=(dd a, dd b) { return a.v >= b.v; }
YS_DEV bool operator==(dd a, dd b) { return a.v == b.v; }
YS_DEV bool operator!=(dd a, dd b) { return a.v != b.v; }
YS_DEV bool operator<(dd a, double b) { return a.v < b; }
YS_DEV bool operator>(dd a, double b) { return a.v > b; }
YS_DEV bool operator==(dd a, double b) { return a.v == b; }
YS_DEV bool operator!=(dd a, double b) { return a.v != b; }

YS_DEV dd sqrt(dd a) { double s = ::sqrt(a.v); return dd(s, a.d / (2.0 * s)); }
YS_DEV dd sin(dd a) { return dd(::sin(a.v), ::cos(a.v) * a.d); }
YS_DEV dd cos(dd a) { return dd(::cos(a.v), -::sin(a.v) * a.d); }
// At |a| = 1 (a straight yarn: the bend angle is 0) acos has no derivative; the angle's
// central directional derivative there is 0, which keeps the tangents finite
YS_DEV dd acos(dd a) {
    double s = ::sqrt(1.0 - a.v * a.v);
    return dd(::acos(a.v), s > 0.0 ? -a.d / s : 0.0);
}
YS_DEV dd tanh(dd a) { double t = ::tanh(a.v); return dd(t, (1.0 - t * t) * a.d); }
// a zero tangent stays zero where exp overflows (inf * 0 would be NaN)
YS_DEV dd exp(dd a) { double e = ::exp(a.v); return dd(e, a.d == 0.0 ? 0.0 : e * a.d); }
YS_DEV dd fabs(dd a) { return a.v < 0.0 ? -a : a; }
YS_DEV dd pow(dd a, double p) { double r = ::pow(a.v, p); return dd(r, p * ::pow(a.v, p - 1.0) * a.d); }
YS_DEV dd fmax(dd a, double b) { return a.v >= b ? a : dd(b); }
YS_DEV dd fmin(dd a, double b) { return a.v <= b ? a : dd(b); }

YS_DEV double val(double a) { return a; }
YS_DEV double val(dd a) { return a.v; }
YS_DEV double tan_of(double) { return 0.0; }
YS_DEV double tan_of(dd a) { return a.d; }

// a scalar argument that does not take part in template deduction
template <class T> struct nd { typedef T type; };

template <class T> struct v3 { T x, y, z; };
template <class T> struct mat3 { T m[3][3]; };
typedef v3<double> d3;
typedef mat3<double> m33;

template <class T> YS_DEV v3<T> V3(T a, T b, T c) { v3<T> r; r.x = a; r.y = b; r.z = c; return r; }
YS_DEV d3 D3(double a, double b, double c) { return V3<double>(a, b, c); }
template <class T> YS_DEV v3<T> operator+(v3<T> a, v3<T> b) { return V3<T>(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class T> YS_DEV v3<T> operator-(v3<T> a, v3<T> b) { return V3<T>(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class T> YS_DEV v3<T> operator-(v3<T> a) { return V3<T>(-a.x, -a.y, -a.z); }
template <class T> YS_DEV v3<T> operator*(typename nd<T>::type s, v3<T> a) { return V3<T>(s * a.x, s * a.y, s * a.z); }
template <class T> YS_DEV v3<T> operator/(v3<T> a, typename nd<T>::type s) { return V3<T>(a.x / s, a.y / s, a.z / s); }
template <class T> YS_DEV T dot(v3<T> a, v3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> YS_DEV T nrm(v3<T> a) { return sqrt(dot(a, a)); }
template <class T> YS_DEV v3<T> cross(v3<T> a, v3<T> b) {
    return V3<T>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <class T> YS_DEV T comp(v3<T> a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
// a double vector as a constant of scalar type T
template <class T> YS_DEV v3<T> lift(d3 a) { return V3<T>(T(a.x), T(a.y), T(a.z)); }

template <class T> YS_DEV mat3<T> Z33T() { mat3<T> r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = T(0.0); return r; }
YS_DEV m33 Z33() { return Z33T<double>(); }
template <class T> YS_DEV mat3<T> I33T(T s) { mat3<T> r = Z33T<T>(); r.m[0][0] = r.m[1][1] = r.m[2][2] = s; return r; }
YS_DEV m33 I33(double s) { return I33T<double>(s); }
YS_DEV mat3<dd> I33(dd s) { return I33T<dd>(s); }
template <class T> YS_DEV mat3<T> operator+(const mat3<T>& a, const mat3<T>& b) { mat3<T> r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][j] + b.m[i][j]; return r; }
template <class T> YS_DEV mat3<T> operator-(const mat3<T>& a, const mat3<T>& b) { mat3<T> r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][j] - b.m[i][j]; return r; }
template <class T> YS_DEV mat3<T> operator-(const mat3<T>& a) { mat3<T> r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = -a.m[i][j]; return r; }
template <class T> YS_DEV mat3<T> operator*(typename nd<T>::type s, const mat3<T>& a) { mat3<T> r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = s * a.m[i][j]; return r; }
template <class T> YS_DEV mat3<T> operator*(const mat3<T>& a, const mat3<T>& b) {
    mat3<T> r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j] + a.m[i][2] * b.m[2][j];
    return r;
}
template <class T> YS_DEV v3<T> operator*(const mat3<T>& a, v3<T> v) {
    return V3<T>(a.m[0][0] * v.x + a.m[0][1] * v.y + a.m[0][2] * v.z,
                 a.m[1][0] * v.x + a.m[1][1] * v.y + a.m[1][2] * v.z,
                 a.m[2][0] * v.x + a.m[2][1] * v.y + a.m[2][2] * v.z);
}
// v^T a
template <class T> YS_DEV v3<T> vTm(v3<T> v, const mat3<T>& a) {
    return V3<T>(v.x * a.m[0][0] + v.y * a.m[1][0] + v.z * a.m[2][0],
                 v.x * a.m[0][1] + v.y * a.m[1][1] + v.z * a.m[2][1],
                 v.x * a.m[0][2] + v.y * a.m[1][2] + v.z * a.m[2][2]);
}
template <class T> YS_DEV mat3<T> outer(v3<T> a, v3<T> b) {
    mat3<T> r;
    T av[3] = {a.x, a.y, a.z}, bv[3] = {b.x, b.y, b.z};
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = av[i] * bv[j];
    return r;
}
// a b^T P  (outer product followed by a right multiplication)
template <class T> YS_DEV mat3<T> outerP(v3<T> a, v3<T> b, const mat3<T>& P) { return outer(a, vTm(b, P)); }
template <class T> YS_DEV mat3<T> projector(v3<T> d) { return I33T<T>(T(1.0)) - outer(d, d); }
// s + A elementwise (a scalar broadcast against a matrix)
template <class T> YS_DEV mat3<T> addS(typename nd<T>::type s, const mat3<T>& a) { mat3<T> r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = s + a.m[i][j]; return r; }
