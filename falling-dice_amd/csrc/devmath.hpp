// devmath.hpp — element math of the beam-lattice solver for gfx950 kernels.
//
// Templated on the element float type T (double for the FP64 path, float for
// the rotational part of the mixed path).  Every formula follows the reference
// (MC = projects/metamaterial-compression/industry-grade):
//   quaternion rotate / product / normalise   MC/include/core.hpp:40-41
//   exp / log on SO(3), J_l^{-T}               MC/include/core.hpp:51-53
//   1-D and 3-D return maps                    MC/src/sim.cpp:5-8
//   segment/segment closest points             MC/src/sim.cpp:11-16
// with the same guards (1e-12 small-angle branches, 1e-20 / 1e-30 floors).
#pragma once
#include <hip/hip_runtime.h>

namespace mmc {

#define MMC_HD __host__ __device__ __forceinline__

template <class T>
struct v3 {
  T x, y, z;
};
template <class T>
MMC_HD v3<T> mk(T x, T y, T z) {
  return v3<T>{x, y, z};
}
template <class T>
MMC_HD v3<T> operator+(v3<T> a, v3<T> b) {
  return {a.x + b.x, a.y + b.y, a.z + b.z};
}
template <class T>
MMC_HD v3<T> operator-(v3<T> a, v3<T> b) {
  return {a.x - b.x, a.y - b.y, a.z - b.z};
}
template <class T>
MMC_HD v3<T> operator*(v3<T> a, T s) {
  return {a.x * s, a.y * s, a.z * s};
}
template <class T>
MMC_HD v3<T> operator-(v3<T> a) {
  return {-a.x, -a.y, -a.z};
}
template <class T>
MMC_HD T dot(v3<T> a, v3<T> b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <class T>
MMC_HD v3<T> cross(v3<T> a, v3<T> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T>
MMC_HD T norm2(v3<T> a) {
  return dot(a, a);
}
template <class T>
MMC_HD T norm(v3<T> a) {
  return sqrt(norm2(a));
}
// v / s is evaluated as v * (1/s), as the reference's Vec3 division does.
template <class T>
MMC_HD v3<T> over(v3<T> a, T s) {
  return a * (T(1) / s);
}
template <class T>
MMC_HD v3<T> unit(v3<T> v) {
  T l2 = norm2(v);
  return (isfinite(l2) && l2 > T(1e-30)) ? over(v, sqrt(l2)) : v3<T>{T(0), T(1), T(0)};
}
template <class T, class U>
MMC_HD v3<T> cvt(v3<U> a) {
  return {T(a.x), T(a.y), T(a.z)};
}

template <class T>
struct quat {
  T w, x, y, z;
};
template <class T>
MMC_HD quat<T> qnormalize(quat<T> q) {
  T n = sqrt(q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z);
  if (n < T(1e-30)) return quat<T>{T(1), T(0), T(0), T(0)};
  return {q.w / n, q.x / n, q.y / n, q.z / n};
}
template <class T>
MMC_HD quat<T> qconj(quat<T> q) {
  return {q.w, -q.x, -q.y, -q.z};
}
template <class T>
MMC_HD quat<T> qmul(quat<T> a, quat<T> b) {
  return {a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z, a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
          a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x, a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w};
}
// R v = v + w t + u x t,  t = 2 u x v
template <class T>
MMC_HD v3<T> qrot(quat<T> q, v3<T> v) {
  v3<T> u{q.x, q.y, q.z};
  v3<T> t = cross(u, v) * T(2);
  return (v + t * q.w) + cross(u, t);
}
template <class T>
MMC_HD v3<T> qrotinv(quat<T> q, v3<T> v) {
  return qrot(qconj(q), v);
}
template <class T>
MMC_HD quat<T> qaxisangle(v3<T> axis, T angle) {
  axis = unit(axis);
  T s = sin(T(0.5) * angle);
  return qnormalize(quat<T>{cos(T(0.5) * angle), axis.x * s, axis.y * s, axis.z * s});
}
template <class T>
MMC_HD quat<T> qexp(v3<T> v) {
  T a = norm(v);
  if (a < T(1e-12)) return qnormalize(quat<T>{T(1) - T(0.125) * a * a, T(0.5) * v.x, T(0.5) * v.y, T(0.5) * v.z});
  return qaxisangle(over(v, a), a);
}
template <class T>
MMC_HD v3<T> qlog(quat<T> q) {
  q = qnormalize(q);
  if (q.w < T(0)) q = quat<T>{-q.w, -q.x, -q.y, -q.z};
  T s = sqrt(q.x * q.x + q.y * q.y + q.z * q.z);
  if (s < T(1e-12)) return {T(2) * q.x, T(2) * q.y, T(2) * q.z};
  T a = T(2) * atan2(s, fmin(T(1), fmax(T(-1), q.w)));
  return v3<T>{q.x, q.y, q.z} * (a / s);
}
// J_l^{-T}(p) v = v + 1/2 p x v + A p x (p x v),
// A = 1/a^2 - (1 + cos a) / (2 a sin a), series value 1/12 below a = 1e-7.
template <class T>
MMC_HD v3<T> jlinvt(v3<T> p, v3<T> v) {
  T a = norm(p);
  T A = (a < T(1e-7)) ? T(1.0 / 12.0) : (T(1) / (a * a) - (T(1) + cos(a)) / (T(2) * a * sin(a)));
  return (v + cross(p, v) * T(0.5)) + cross(p, cross(p, v)) * A;
}

// 1-D return map with linear isotropic hardening (MC/src/sim.cpp:6).
struct ret1 {
  double force, p, a, dg, energy, hard;
  bool yielded;
};
MMC_HD ret1 return1(double total, double pc, double ac, double k, double y, double H) {
  ret1 r;
  double trial = k * (total - pc);
  double f = fabs(trial) - (y + H * ac);
  r.p = pc;
  r.a = ac;
  r.dg = 0;
  r.yielded = false;
  if (f > 0) {
    double sgn = trial >= 0 ? 1.0 : -1.0;
    r.dg = f / fmax(1e-20, k + H);
    r.p = pc + sgn * r.dg;
    r.a = ac + r.dg;
    r.yielded = true;
  }
  double el = total - r.p;
  r.force = k * el;
  r.energy = 0.5 * k * (el * el);
  r.hard = 0.5 * H * (r.a * r.a);
  return r;
}
// 3-D radial return (MC/src/sim.cpp:8).
struct ret3 {
  v3<double> force, p;
  double a, dg, energy, hard;
  bool yielded;
};
MMC_HD ret3 return3(v3<double> t, v3<double> pc, double ac, double k, double y, double H) {
  ret3 r;
  v3<double> trial = (t - pc) * k;
  double n = norm(trial);
  double f = n - (y + H * ac);
  r.p = pc;
  r.a = ac;
  r.dg = 0;
  r.yielded = false;
  if (f > 0 && n > 1e-20) {
    v3<double> dir = over(trial, n);
    r.dg = f / fmax(1e-20, k + H);
    r.p = pc + dir * r.dg;
    r.a = ac + r.dg;
    r.yielded = true;
  }
  v3<double> el = t - r.p;
  r.force = el * k;
  r.energy = 0.5 * k * norm2(el);
  r.hard = 0.5 * H * (r.a * r.a);
  return r;
}

MMC_HD double sat(double x) { return fmax(0.0, fmin(1.0, x)); }

// Closest points of two segments (Ericson, Real-Time Collision Detection, 5.1.9).
struct closest {
  double s, t;
  v3<double> pa, pb;
};
MMC_HD closest seg_seg(v3<double> p1, v3<double> q1, v3<double> p2, v3<double> q2) {
  v3<double> d1 = q1 - p1, d2 = q2 - p2, r = p1 - p2;
  double a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r), s = 0, t = 0;
  if (a <= 1e-20 && e <= 1e-20) return {0, 0, p1, p2};
  if (a <= 1e-20) {
    s = 0;
    t = sat(f / e);
  } else {
    double c = dot(d1, r);
    if (e <= 1e-20) {
      t = 0;
      s = sat(-c / a);
    } else {
      double b = dot(d1, d2), den = a * e - b * b;
      s = fabs(den) > 1e-20 ? sat((b * f - c * e) / den) : 0;
      t = (b * s + f) / e;
      if (t < 0) {
        t = 0;
        s = sat(-c / a);
      } else if (t > 1) {
        t = 1;
        s = sat((b - c) / a);
      }
    }
  }
  return {s, t, p1 + d1 * s, p2 + d2 * t};
}

}  // namespace mmc
