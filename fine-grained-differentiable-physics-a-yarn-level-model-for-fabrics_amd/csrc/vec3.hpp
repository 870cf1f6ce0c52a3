// vec3.hpp -- FP64 3-vector / 3x3 helpers for the gfx950 element kernels.
#pragma once
#include <hip/hip_runtime.h>

#define YS_DEV __device__ __forceinline__

struct d3 {
    double x, y, z;
};
struct m33 {
    double m[3][3];
};

YS_DEV d3 D3(double a, double b, double c) { d3 r; r.x = a; r.y = b; r.z = c; return r; }
YS_DEV d3 operator+(d3 a, d3 b) { return D3(a.x + b.x, a.y + b.y, a.z + b.z); }
YS_DEV d3 operator-(d3 a, d3 b) { return D3(a.x - b.x, a.y - b.y, a.z - b.z); }
YS_DEV d3 operator-(d3 a) { return D3(-a.x, -a.y, -a.z); }
YS_DEV d3 operator*(double s, d3 a) { return D3(s * a.x, s * a.y, s * a.z); }
YS_DEV d3 operator/(d3 a, double s) { return D3(a.x / s, a.y / s, a.z / s); }
YS_DEV double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
YS_DEV double nrm(d3 a) { return sqrt(dot(a, a)); }
YS_DEV d3 cross(d3 a, d3 b) { return D3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
YS_DEV double comp(d3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

YS_DEV m33 Z33() { m33 r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = 0.0; return r; }
YS_DEV m33 I33(double s) { m33 r = Z33(); r.m[0][0] = r.m[1][1] = r.m[2][2] = s; return r; }
YS_DEV m33 operator+(const m33& a, const m33& b) { m33 r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][j] + b.m[i][j]; return r; }
YS_DEV m33 operator-(const m33& a, const m33& b) { m33 r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][j] - b.m[i][j]; return r; }
YS_DEV m33 operator-(const m33& a) { m33 r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = -a.m[i][j]; return r; }
YS_DEV m33 operator*(double s, const m33& a) { m33 r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = s * a.m[i][j]; return r; }
YS_DEV m33 operator*(const m33& a, const m33& b) {
    m33 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j] + a.m[i][2] * b.m[2][j];
    return r;
}
YS_DEV d3 operator*(const m33& a, d3 v) {
    return D3(a.m[0][0] * v.x + a.m[0][1] * v.y + a.m[0][2] * v.z,
              a.m[1][0] * v.x + a.m[1][1] * v.y + a.m[1][2] * v.z,
              a.m[2][0] * v.x + a.m[2][1] * v.y + a.m[2][2] * v.z);
}
// v^T a
YS_DEV d3 vTm(d3 v, const m33& a) {
    return D3(v.x * a.m[0][0] + v.y * a.m[1][0] + v.z * a.m[2][0],
              v.x * a.m[0][1] + v.y * a.m[1][1] + v.z * a.m[2][1],
              v.x * a.m[0][2] + v.y * a.m[1][2] + v.z * a.m[2][2]);
}
YS_DEV m33 outer(d3 a, d3 b) {
    m33 r;
    double av[3] = {a.x, a.y, a.z}, bv[3] = {b.x, b.y, b.z};
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = av[i] * bv[j];
    return r;
}
// a b^T P  (outer product followed by a right multiplication)
YS_DEV m33 outerP(d3 a, d3 b, const m33& P) { return outer(a, vTm(b, P)); }
YS_DEV m33 projector(d3 d) { return I33(1.0) - outer(d, d); }
// s + A elementwise (a scalar broadcast against a matrix)
YS_DEV m33 addS(double s, const m33& a) { m33 r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = s + a.m[i][j]; return r; }
