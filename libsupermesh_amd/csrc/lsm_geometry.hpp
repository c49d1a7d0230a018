// lsm_geometry.hpp -- FP64 element geometry and the per-pair clippers.
//
// Every routine here runs one candidate pair per lane with its whole state in
// VGPRs: no per-lane arrays are indexed with run-time values, so nothing goes
// to scratch.  The library is compiled with -ffp-contract=off and the code
// spells out each operation in the reference's order, so every result rounds
// exactly as the reference's Fortran does (IEEE double, no FMA, correctly
// rounded division and square root).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lsm {

#define LSM_HD __host__ __device__ __forceinline__

constexpr int kTriBufSize = 22;  // TRI_BUF_SIZE, include/libsupermesh-c.h:3
constexpr int kTetBufSize = 81;  // TET_BUF_SIZE, include/libsupermesh-c.h:4
constexpr int kTriQuadBufSize = 46;  // tris_c(2, 3, 46), C_bindings.F90:59
constexpr int kTetHexBufSize = 729;  // tets_c(3, 4, 729), C_bindings.F90:47
constexpr int kQuadsBufSize = 62;    // max_n_tris_c(4, 4), Supermesh.F90:129
constexpr int kHexesBufSize = 5 * kTetHexBufSize;  // Supermesh.F90:131

// ---------------------------------------------------------------------------
// Fortran intrinsics
// ---------------------------------------------------------------------------

// SPACING(x): 2**max(e-53, -1022) with |x| = m*2**e, 0.5 <= m < 1; TINY for
// zero and subnormals.  For a biased exponent E >= 53 that is 2**(E-1075),
// whose biased exponent is E-52.
LSM_HD double fspacing(double x) {
  uint64_t bits;
  __builtin_memcpy(&bits, &x, 8);
  int64_t E = (int64_t)((bits >> 52) & 0x7ff);
  uint64_t r = E >= 53 ? (uint64_t)(E - 52) << 52 : (uint64_t)1 << 52;
  double out;
  __builtin_memcpy(&out, &r, 8);
  return out;
}

constexpr double kEpsilon = 2.220446049250313080847e-16;  // EPSILON(1.0d0)

LSM_HD double fmin_f(double a, double b) { return b < a ? b : a; }
LSM_HD double fmax_f(double a, double b) { return b > a ? b : a; }
LSM_HD double fabs_f(double a) { return a < 0.0 ? -a : (a == 0.0 ? 0.0 : a); }

// NORM2 of a 3-vector as the reference's compiler evaluates it: running
// maximum m, scaled sum of squares s, result m*sqrt(1+s)
// (Tetrahedron_Intersection.F90:676).
LSM_HD double norm2_3(double x, double y, double z) {
  double mx = 0.0, s = 0.0;
  const double v[3] = {x, y, z};
#pragma unroll
  for (int i = 0; i < 3; i++) {
    double ax = __builtin_fabs(v[i]);
    if (mx == 0.0) {
      mx = ax;
    } else if (ax > mx) {
      double t = mx / ax;
      double tsq = t * t;
      s = s * tsq;
      s = s + tsq;
      mx = ax;
    } else {
      double t = ax / mx;
      s = s + t * t;
    }
  }
  return mx * __builtin_sqrt(1.0 + s);
}

LSM_HD double dot3(double a0, double a1, double a2, double b0, double b1, double b2) {
  double s = a0 * b0;
  s = s + a1 * b1;
  s = s + a2 * b2;
  return s;
}

// ---------------------------------------------------------------------------
// Measures (Triangle_Intersection.F90:453-477, Tetrahedron_Intersection.F90:690-725)
// ---------------------------------------------------------------------------

LSM_HD double triangle_area(const double* t) {
  double ux = t[4] - t[0], uy = t[5] - t[1];
  double vx = t[2] - t[0], vy = t[3] - t[1];
  return 0.5 * __builtin_fabs(uy * vx - ux * vy);
}

LSM_HD double triangle_area6(double x0, double y0, double x1, double y1, double x2, double y2) {
  double ux = x2 - x0, uy = y2 - y0;
  double vx = x1 - x0, vy = y1 - y0;
  return 0.5 * __builtin_fabs(uy * vx - ux * vy);
}

// Edges from vertex 1, times 1/6 (the C ABI and the reference tests' Sigma).
LSM_HD double tet_volume_real(const double* t) {
  double e1x = t[3] - t[0], e1y = t[4] - t[1], e1z = t[5] - t[2];
  double e2x = t[6] - t[0], e2y = t[7] - t[1], e2z = t[8] - t[2];
  double e3x = t[9] - t[0], e3y = t[10] - t[1], e3z = t[11] - t[2];
  double d = e1x * (e2y * e3z - e2z * e3y);
  d = d + e1y * (e2z * e3x - e2x * e3z);
  d = d + e1z * (e2x * e3y - e2y * e3x);
  return (1.0 / 6.0) * __builtin_fabs(d);
}

// Edges from vertex 4, divided by 6 (the clipper's tolerance and keep test).
LSM_HD double tet_volume_tet(const double* t) {
  double ax = t[0] - t[9], ay = t[1] - t[10], az = t[2] - t[11];
  double bx = t[3] - t[9], by = t[4] - t[10], bz = t[5] - t[11];
  double cx = t[6] - t[9], cy = t[7] - t[10], cz = t[8] - t[11];
  double c0 = by * cz - bz * cy;
  double c1 = bz * cx - bx * cz;
  double c2 = bx * cy - by * cx;
  return __builtin_fabs(dot3(ax, ay, az, c0, c1, c2)) / 6.0;
}

// ---------------------------------------------------------------------------
// Triangle clipping: Sutherland-Hodgman against B's three edge lines, run as
// a pipeline (each stage keeps only its first and previous point), then a fan
// from the first surviving vertex.  Output order and rounding are those of
// Triangle_Intersection.F90:144-185,282-335.
// ---------------------------------------------------------------------------

struct Line {
  double nx, ny, px, py;
};

// get_lines_tri (Triangle_Intersection.F90:395-420).
LSM_HD void tri_lines(const double* t, Line L[3]) {
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const int k1 = (k + 1) % 3;
    L[k].nx = -(t[2 * k1 + 1] - t[2 * k + 1]);
    L[k].ny = (t[2 * k1] - t[2 * k]);
    L[k].px = t[2 * k];
    L[k].py = t[2 * k + 1];
  }
  double d = (t[2] - t[0]) * L[2].nx;
  d = d + (t[3] - t[1]) * L[2].ny;
  if (d > 0.0) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      L[k].nx = -L[k].nx;
      L[k].ny = -L[k].ny;
    }
  }
}

struct ClipStage {
  double fx, fy, fd;  // first point and its distance
  double px, py, pd;  // previous point
  int nin, nout;
};

LSM_HD void stage_init(ClipStage& s) {
  s.fx = s.fy = s.fd = s.px = s.py = s.pd = 0.0;
  s.nin = s.nout = 0;
}

// One edge (p1 -> p2) of a Sutherland-Hodgman pass: emits 0, 1 or 2 points.
template <class Emit>
LSM_HD void clip_edge(double x1, double y1, double d1, double x2, double y2, double d2,
                      int& nout, Emit&& emit) {
  if (d1 < 0.0) {
    emit(x1, y1);
    nout++;
    if (!(d2 <= 0.0)) {
      double f = __builtin_fabs(d1) / (__builtin_fabs(d1) + __builtin_fabs(d2));
      f = fmax_f(fmin_f(f, 1.0), 0.0);
      emit(x1 + f * (x2 - x1), y1 + f * (y2 - y1));
      nout++;
    }
  } else if (d1 == 0.0) {
    emit(x1, y1);
    nout++;
  } else if (d2 < 0.0) {
    double f = __builtin_fabs(d1) / (__builtin_fabs(d1) + __builtin_fabs(d2));
    f = fmax_f(fmin_f(f, 1.0), 0.0);
    emit(x1 + f * (x2 - x1), y1 + f * (y2 - y1));
    nout++;
  }
}

template <class Emit>
LSM_HD void stage_push(ClipStage& s, const Line& L, double x, double y, Emit&& emit) {
  double d = L.nx * (x - L.px);
  d = d + L.ny * (y - L.py);
  if (s.nin == 0) {
    s.fx = x;
    s.fy = y;
    s.fd = d;
  } else {
    clip_edge(s.px, s.py, s.pd, x, y, d, s.nout, emit);
  }
  s.px = x;
  s.py = y;
  s.pd = d;
  s.nin++;
}

template <class Emit>
LSM_HD void stage_close(ClipStage& s, Emit&& emit) {
  if (s.nin > 0) clip_edge(s.px, s.py, s.pd, s.fx, s.fy, s.fd, s.nout, emit);
}

// Fan triangulation of the final polygon: triangle (q1, q_{k-1}, q_k) for each
// arriving q_k, k >= 3; triangles with area < tol are dropped.
struct Fan {
  double x0, y0, x1, y1;
  int n;
};

// Pushes a point into pass S of an NL-pass pipeline (pass NL is the fan).
template <int S, int NL, class Last>
LSM_HD void pipe_push(ClipStage* st, const Line* L, double x, double y, Last& last) {
  if constexpr (S == NL) {
    last(x, y);
  } else {
    stage_push(st[S], L[S], x, y, [&](double u, double v) { pipe_push<S + 1, NL>(st, L, u, v, last); });
  }
}

template <int S, int NL, class Last>
LSM_HD void pipe_close(ClipStage* st, const Line* L, Last& last) {
  if constexpr (S < NL) {
    stage_close(st[S], [&](double u, double v) { pipe_push<S + 1, NL>(st, L, u, v, last); });
    pipe_close<S + 1, NL>(st, L, last);
  }
}

// Clips polygon A (NA points, a triangle by default) against NL lines, passes in line order, then
// fans the result (intersect_polys_lines, Triangle_Intersection.F90:216-276).
// For each kept output triangle (in the reference's order) calls
// sink(x0, y0, x1, y1, x2, y2, area).  Returns the number kept; if any pass
// ends with fewer than 3 points the reference returns 0 (and calls made to
// the sink must then be discarded, which only matters before a count is known).
template <int NL, int NA = 3, class Sink>
LSM_HD int clip_poly(const double* A, const Line* L, double tol, Sink&& sink) {
  ClipStage st[NL];
#pragma unroll
  for (int i = 0; i < NL; i++) stage_init(st[i]);
  Fan fan;
  fan.x0 = fan.y0 = fan.x1 = fan.y1 = 0.0;
  fan.n = 0;
  int kept = 0;
  auto fan_push = [&](double x, double y) {
    if (fan.n == 0) {
      fan.x0 = x;
      fan.y0 = y;
    } else if (fan.n >= 2) {
      double a = triangle_area6(fan.x0, fan.y0, fan.x1, fan.y1, x, y);
      if (!(a < tol)) {
        sink(fan.x0, fan.y0, fan.x1, fan.y1, x, y, a);
        kept++;
      }
    }
    fan.x1 = x;
    fan.y1 = y;
    fan.n++;
  };
  static_assert(NA == 3 || NA == 4, "triangles and quadrilaterals");
  // one call per vertex, written out: a loop here is inlined before it is
  // unrolled and costs the triangle kernels registers
  pipe_push<0, NL>(st, L, A[0], A[1], fan_push);
  pipe_push<0, NL>(st, L, A[2], A[3], fan_push);
  pipe_push<0, NL>(st, L, A[4], A[5], fan_push);
  if constexpr (NA == 4) pipe_push<0, NL>(st, L, A[6], A[7], fan_push);
  pipe_close<0, NL>(st, L, fan_push);
#pragma unroll
  for (int i = 0; i < NL; i++)
    if (st[i].nout < 3) return 0;
  return kept;
}

// Triangle A against triangle B's lines (intersect_tris_tri,
// Triangle_Intersection.F90:144-185).
template <class Sink>
LSM_HD int clip_tri(const double* A, const Line* L, double tol, Sink&& sink) {
  return clip_poly<3>(A, L, tol, sink);
}

// get_lines_poly for a 4-vertex polygon (Triangle_Intersection.F90:422-451):
// normals (dy, -dx) per edge, all flipped if vertex 1 is on the positive side
// of line 2.
LSM_HD void quad_lines(const double* q, Line L[4]) {
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int k1 = (k + 1) % 4;
    L[k].nx = (q[2 * k1 + 1] - q[2 * k + 1]);
    L[k].ny = -(q[2 * k1] - q[2 * k]);
    L[k].px = q[2 * k];
    L[k].py = q[2 * k + 1];
  }
  double d = (q[0] - L[1].px) * L[1].nx;
  d = d + (q[1] - L[1].py) * L[1].ny;
  if (d > 0.0) {
#pragma unroll
    for (int k = 0; k < 4; k++) {
      L[k].nx = -L[k].nx;
      L[k].ny = -L[k].ny;
    }
  }
}

LSM_HD double tri_tol(double area_a, double area_b) {
  return 10.0 * fmin_f(fspacing(area_a), fspacing(area_b));
}

// ---------------------------------------------------------------------------
// Tetrahedron clipping: B's four unit-normal planes one at a time
// (Tetrahedron_Intersection.F90:137-179,244-436), depth first.  A split tet
// yields its children in the reference's order, so the depth-first preorder
// over (plane-1 child, plane-2 child, ...) is the same lexicographic order as
// the reference's breadth-first buffer, and every output tet is produced by
// the same operations.
// ---------------------------------------------------------------------------

struct Planes {
  double n[4][3];
  double c[4];
};

LSM_HD void unit_cross(double a0, double a1, double a2, double b0, double b1, double b2,
                       double* out) {
  double c0 = a1 * b2 - a2 * b1;
  double c1 = a2 * b0 - a0 * b2;
  double c2 = a0 * b1 - a1 * b0;
  double n = norm2_3(c0, c1, c2);
  out[0] = c0 / n;
  out[1] = c1 / n;
  out[2] = c2 / n;
}

// get_planes_tet (Tetrahedron_Intersection.F90:634-667).
LSM_HD void tet_planes(const double* t, Planes& P) {
  double e10[3], e20[3], e30[3], e21[3], e31[3];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    e10[i] = t[3 + i] - t[i];
    e20[i] = t[6 + i] - t[i];
    e30[i] = t[9 + i] - t[i];
    e21[i] = t[6 + i] - t[3 + i];
    e31[i] = t[9 + i] - t[3 + i];
  }
  unit_cross(e20[0], e20[1], e20[2], e10[0], e10[1], e10[2], P.n[0]);
  unit_cross(e10[0], e10[1], e10[2], e30[0], e30[1], e30[2], P.n[1]);
  unit_cross(e30[0], e30[1], e30[2], e20[0], e20[1], e20[2], P.n[2]);
  unit_cross(e21[0], e21[1], e21[2], e31[0], e31[1], e31[2], P.n[3]);
  double det = dot3(e10[0], e10[1], e10[2], P.n[3][0], P.n[3][1], P.n[3][2]);
  if (det < 0.0) {
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int k = 0; k < 3; k++) P.n[i][k] = -P.n[i][k];
  }
#pragma unroll
  for (int i = 0; i < 4; i++)
    P.c[i] = dot3(t[3 * i], t[3 * i + 1], t[3 * i + 2], P.n[i][0], P.n[i][1], P.n[i][2]);
}

struct Tet {
  double v[12];
};

// Classification of a tet against one plane.  kind: 0 discard, 1 keep whole,
// 31 (+++-), 22 (++--), 21 (++-0), 13 (+---), 12 (+--0), 11 (+-00).
// pos/neg/zer hold ascending vertex indices, packed 2 bits each.
struct Split {
  double d[4];
  int kind, nchild;
  int pos, neg, zer;
};

LSM_HD int pick(int packed, int i) { return (packed >> (2 * i)) & 3; }

// A value the optimiser must treat as already in a register.  Selecting
// between elements of a small local array with a run-time index is otherwise
// folded into one load at a computed address, which moves the whole array to
// scratch memory.
LSM_HD double in_reg(double x) {
#ifdef __HIP_DEVICE_COMPILE__
  asm volatile("" : "+v"(x));
#endif
  return x;
}

LSM_HD int in_reg_i(int x) {
#ifdef __HIP_DEVICE_COMPILE__
  asm volatile("" : "+v"(x));
#endif
  return x;
}

LSM_HD double sel4(const double* a, int k) {
  const double a0 = in_reg(a[0]), a1 = in_reg(a[1]), a2 = in_reg(a[2]), a3 = in_reg(a[3]);
  double r = a3;
  r = k == 2 ? a2 : r;
  r = k == 1 ? a1 : r;
  r = k == 0 ? a0 : r;
  return r;
}

LSM_HD Split classify(const Tet& T, const double* n, double c) {
  Split s;
  int np = 0, nn = 0, nz = 0;
  s.pos = s.neg = s.zer = 0;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    double d = dot3(n[0], n[1], n[2], T.v[3 * i], T.v[3 * i + 1], T.v[3 * i + 2]) - c;
    s.d[i] = d;
    if (__builtin_fabs(d) < kEpsilon) {
      s.zer |= i << (2 * nz);
      nz++;
    } else if (d < 0.0) {
      s.neg |= i << (2 * nn);
      nn++;
    } else if (d > 0.0) {
      s.pos |= i << (2 * np);
      np++;
    }
  }
  if (nn == 0) {
    s.kind = 0;
    s.nchild = 0;
  } else if (np == 0) {
    s.kind = 1;
    s.nchild = 1;
  } else {
    s.kind = 10 * np + nn;
    s.nchild = (s.kind == 22 || s.kind == 13) ? 3 : (s.kind == 12 ? 2 : 1);
  }
  return s;
}

// Split point between outside vertex ip and inside vertex in:
// invdiff = 1/(d_pos - d_neg), w0 = -d_neg*invdiff, w1 = d_pos*invdiff,
// p = w0*v_pos + w1*v_neg.
LSM_HD void split_point(const Tet& T, const double* d, int ip, int in, double* p) {
  double dp = sel4(d, ip), dn = sel4(d, in);
  double invdiff = 1.0 / (dp - dn);
  double w0 = -dn * invdiff;
  double w1 = dp * invdiff;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    double vp = T.v[3 * 3 + k];
    vp = ip == 2 ? T.v[3 * 2 + k] : vp;
    vp = ip == 1 ? T.v[3 * 1 + k] : vp;
    vp = ip == 0 ? T.v[k] : vp;
    double vn = T.v[3 * 3 + k];
    vn = in == 2 ? T.v[3 * 2 + k] : vn;
    vn = in == 1 ? T.v[3 * 1 + k] : vn;
    vn = in == 0 ? T.v[k] : vn;
    p[k] = w0 * vp + w1 * vn;
  }
}

LSM_HD void vertex(const Tet& T, int i, double* p) {
#pragma unroll
  for (int k = 0; k < 3; k++) {
    double v = T.v[9 + k];
    v = i == 2 ? T.v[6 + k] : v;
    v = i == 1 ? T.v[3 + k] : v;
    v = i == 0 ? T.v[k] : v;
    p[k] = v;
  }
}

LSM_HD void set_slot(Tet& T, int s, const double* p) {
  T.v[3 * s] = p[0];
  T.v[3 * s + 1] = p[1];
  T.v[3 * s + 2] = p[2];
}

// T with vertex slot `slot` (run-time) replaced by p.
LSM_HD void replace_slot(Tet& T, int slot, const double* p) {
#pragma unroll
  for (int s = 0; s < 4; s++)
#pragma unroll
    for (int k = 0; k < 3; k++) T.v[3 * s + k] = s == slot ? p[k] : T.v[3 * s + k];
}

// Child `ci` of parent T under split s, as appended by clip_buf
// (Tetrahedron_Intersection.F90:292-434).
LSM_HD Tet make_child(const Tet& T, const Split& s, int ci) {
  Tet out = T;
  double a[3], b[3], c[3];
  switch (s.kind) {
    case 1:
      break;
    case 31: {  // +++-
      int n0 = pick(s.neg, 0);
      for (int i = 0; i < 3; i++) {
        int p = pick(s.pos, i);
        split_point(T, s.d, p, n0, a);
        replace_slot(out, p, a);
      }
      break;
    }
    case 22: {  // ++--
      int p0 = pick(s.pos, 0), p1 = pick(s.pos, 1), n0 = pick(s.neg, 0), n1 = pick(s.neg, 1);
      if (ci == 0) {
        split_point(T, s.d, p0, n1, a);  // tmp3
        split_point(T, s.d, p1, n0, b);  // tmp2
        replace_slot(out, p0, a);
        replace_slot(out, p1, b);
      } else if (ci == 1) {
        vertex(T, n1, a);
        set_slot(out, 0, a);
        split_point(T, s.d, p1, n1, a);  // tmp4
        set_slot(out, 1, a);
        split_point(T, s.d, p0, n1, a);  // tmp3
        set_slot(out, 2, a);
        split_point(T, s.d, p1, n0, a);  // tmp2
        set_slot(out, 3, a);
      } else {
        vertex(T, n0, a);
        set_slot(out, 0, a);
        split_point(T, s.d, p0, n0, a);  // tmp1
        set_slot(out, 1, a);
        split_point(T, s.d, p1, n0, a);  // tmp2
        set_slot(out, 2, a);
        split_point(T, s.d, p0, n1, a);  // tmp3
        set_slot(out, 3, a);
      }
      break;
    }
    case 21: {  // ++-0
      int n0 = pick(s.neg, 0);
      for (int i = 0; i < 2; i++) {
        int p = pick(s.pos, i);
        split_point(T, s.d, p, n0, a);
        replace_slot(out, p, a);
      }
      break;
    }
    case 13: {  // +---
      int p0 = pick(s.pos, 0), n0 = pick(s.neg, 0), n1 = pick(s.neg, 1), n2 = pick(s.neg, 2);
      if (ci == 0) {
        split_point(T, s.d, p0, n0, a);
        replace_slot(out, p0, a);
      } else if (ci == 1) {
        split_point(T, s.d, p0, n0, a);  // tmp1
        set_slot(out, 0, a);
        vertex(T, n1, b);
        set_slot(out, 1, b);
        vertex(T, n2, b);
        set_slot(out, 2, b);
        split_point(T, s.d, p0, n1, c);  // tmp2
        set_slot(out, 3, c);
      } else {
        vertex(T, n2, a);
        set_slot(out, 0, a);
        split_point(T, s.d, p0, n1, a);  // tmp2
        set_slot(out, 1, a);
        split_point(T, s.d, p0, n2, a);  // tmp3
        set_slot(out, 2, a);
        split_point(T, s.d, p0, n0, a);  // tmp1
        set_slot(out, 3, a);
      }
      break;
    }
    case 12: {  // +--0
      int p0 = pick(s.pos, 0), n0 = pick(s.neg, 0), n1 = pick(s.neg, 1), z0 = pick(s.zer, 0);
      if (ci == 0) {
        split_point(T, s.d, p0, n0, a);
        replace_slot(out, p0, a);
      } else {
        split_point(T, s.d, p0, n1, a);  // tmp2
        set_slot(out, 0, a);
        vertex(T, z0, a);
        set_slot(out, 1, a);
        vertex(T, n1, a);
        set_slot(out, 2, a);
        split_point(T, s.d, p0, n0, a);  // tmp1
        set_slot(out, 3, a);
      }
      break;
    }
    default: {  // 11: +-00
      int p0 = pick(s.pos, 0), n0 = pick(s.neg, 0);
      split_point(T, s.d, p0, n0, a);
      replace_slot(out, p0, a);
      break;
    }
  }
  return out;
}

// ---------------------------------------------------------------------------
// Flat depth-first walk: the same clip as clip_tet, written as a loop whose
// body always does the same work -- build one child, then classify it (or
// measure it, at the last plane) -- so that lanes of a wave at different
// depths, in different split cases, or on different pairs run in lockstep.
// A child is built from a recipe (up to three split points, then a source for
// each of the four output slots) that reproduces clip_buf's vertex order.
// ---------------------------------------------------------------------------

// Packed classification: bits 0-2 kind (0 discard, 1 keep, 2 +++-, 3 ++--,
// 4 ++-0, 5 +---, 6 +--0, 7 +-00), 3-4 number of children, 5-10 pos
// indices p0..p2, 11-16 neg indices n0..n2, 17-18 zero index z0.
LSM_HD int pack_class(const double* d) {
  int np = 0, nn = 0, nz = 0, pos = 0, neg = 0, zer = 0;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    if (__builtin_fabs(d[i]) < kEpsilon) {
      zer |= i << (2 * nz);
      nz++;
    } else if (d[i] < 0.0) {
      neg |= i << (2 * nn);
      nn++;
    } else if (d[i] > 0.0) {
      pos |= i << (2 * np);
      np++;
    }
  }
  int kind, nch;
  if (nn == 0) {
    kind = 0;
    nch = 0;
  } else if (np == 0) {
    kind = 1;
    nch = 1;
  } else if (np == 3) {
    kind = 2;
    nch = 1;
  } else if (np == 2) {
    kind = nn == 2 ? 3 : 4;
    nch = nn == 2 ? 3 : 1;
  } else {
    kind = nn == 3 ? 5 : (nn == 2 ? 6 : 7);
    nch = nn == 3 ? 3 : (nn == 2 ? 2 : 1);
  }
  return kind | (nch << 3) | ((pos & 63) << 5) | ((neg & 63) << 11) | ((zer & 3) << 17);
}

LSM_HD int class_nchild(int c) { return (c >> 3) & 3; }

// Recipe of child ci: split points s0..s2 as (outside, inside) vertex pairs
// and, per output slot, a source: 0-3 parent vertex, 4-6 split point.
struct Recipe {
  int sp[3], sn[3];
  int src[4];
};

LSM_HD Recipe make_recipe(int c, int ci) {
  // Written without branches: lanes of a wave are in different split cases,
  // and every case is a handful of integer selects.
  const int kind = c & 7;
  const int p0 = (c >> 5) & 3, p1 = (c >> 7) & 3, p2 = (c >> 9) & 3;
  const int n0 = (c >> 11) & 3, n1 = (c >> 13) & 3, n2 = (c >> 15) & 3;
  const int z0 = (c >> 17) & 3;
  const bool k3 = kind == 3, k5 = kind == 5, k6 = kind == 6;
  const bool k3c1 = k3 && ci == 1, k3c2 = k3 && ci == 2, k5c1 = k5 && ci == 1, k5c2 = k5 && ci == 2,
             k6c1 = k6 && ci == 1;
  Recipe r;
  // split points (outside, inside): +++- by default, (p_s, n0)
  r.sp[0] = k3c1 ? p1 : p0;
  r.sn[0] = (k3 && ci == 0) || k3c1 || k5c2 || k6c1 ? n1 : n0;
  r.sp[1] = k3c1 ? p0 : (k3c2 ? p1 : (k5c1 || k5c2 || k6c1 ? p0 : p1));
  r.sn[1] = k3c1 || k5c1 ? n1 : (k5c2 ? n2 : n0);
  r.sp[2] = k3c1 ? p1 : (k3c2 || k5c2 ? p0 : p2);
  r.sn[2] = k3c2 ? n1 : n0;
  // children of the in-place kind keep the parent's slots and move the
  // outside vertex of each split onto the plane: +++- (3 splits), ++-0 (2),
  // +-00 (1), and child 0 of ++-- (2), +--- (1), +--0 (1)
  const int nsplit = kind == 2 ? 3 : (kind == 4 || (k3 && ci == 0) ? 2 : (kind == 7 || ((k5 || k6) && ci == 0) ? 1 : 0));
#pragma unroll
  for (int j = 0; j < 4; j++) {
    int v = j;
    v = nsplit >= 3 && j == r.sp[2] ? 6 : v;
    v = nsplit >= 2 && j == r.sp[1] ? 5 : v;
    v = nsplit >= 1 && j == r.sp[0] ? 4 : v;
    r.src[j] = v;
  }
  // the other children list their four slots: ++-- (n1 | n0, 4, 5, 6),
  // +--- (4, n1, n2, 5) and (n2, 4, 5, 6), +--0 (4, z0, n1, 5)
  const bool expl = k3c1 || k3c2 || k5c1 || k5c2 || k6c1;
  const int e0 = k5c1 || k6c1 ? 4 : (k3c1 ? n1 : (k3c2 ? n0 : n2));
  const int e1 = k5c1 ? n1 : (k6c1 ? z0 : 4);
  const int e2 = k5c1 ? n2 : (k6c1 ? n1 : 5);
  const int e3 = k5c1 || k6c1 ? 5 : 6;
  r.src[0] = expl ? e0 : r.src[0];
  r.src[1] = expl ? e1 : r.src[1];
  r.src[2] = expl ? e2 : r.src[2];
  r.src[3] = expl ? e3 : r.src[3];
  return r;
}

// Child of parent T (distances d to the plane) from recipe r.  Unused split
// points are computed from vertex 0 twice and never selected.  The slots
// taken from the parent are filled first; each split point then goes into
// the slots that name it as soon as it is computed, so only one split point
// is live at a time.
LSM_HD void build_child(const double* T, const double* d, const Recipe& r, double* out) {
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int src = r.src[j];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      double v = in_reg(T[9 + k]);
      v = src == 2 ? in_reg(T[6 + k]) : v;
      v = src == 1 ? in_reg(T[3 + k]) : v;
      v = src == 0 ? in_reg(T[k]) : v;
      out[3 * j + k] = v;
    }
  }
#pragma unroll
  for (int s = 0; s < 3; s++) {
    const int ip = r.sp[s], in = r.sn[s];
    double dp = sel4(d, ip), dn = sel4(d, in);
    double invdiff = 1.0 / (dp - dn);
    double w0 = -dn * invdiff;
    double w1 = dp * invdiff;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const double t0 = in_reg(T[k]), t1 = in_reg(T[3 + k]), t2 = in_reg(T[6 + k]),
                   t3 = in_reg(T[9 + k]);
      double vp = t3;
      vp = ip == 2 ? t2 : vp;
      vp = ip == 1 ? t1 : vp;
      vp = ip == 0 ? t0 : vp;
      double vn = t3;
      vn = in == 2 ? t2 : vn;
      vn = in == 1 ? t1 : vn;
      vn = in == 0 ? t0 : vn;
      const double x = w0 * vp + w1 * vn;
#pragma unroll
      for (int j = 0; j < 4; j++) out[3 * j + k] = r.src[j] == 4 + s ? x : in_reg(out[3 * j + k]);
    }
  }
}

// Depth-first walk over the clip tree of one pair and NP planes, in which
// every step builds exactly one child: the next child is built (and, at depth
// NP, handed to the sink; otherwise pushed and classified against the next
// plane), then the walk climbs over every depth whose children are all done.
// The lanes of a wave stay in lockstep: the work of a step does not depend on
// the depth, the split case or whether the child is a leaf.  State: the class
// of the current tet at each depth 0..NP-1 against the plane below it, the
// depth of the next child to build (0 once the walk is over) and the index
// of the next child at each depth 1..NP (2 bits each).  Tets at depths
// 1..NREG live in the caller's registers R (depth d at R[(d - 1) * 12 + i]),
// and those at depths NREG+1..NP-1 in the caller's stack: depth d at
// stk[((d - NREG - 1) * 12 + i) * STRIDE], so a wave's lanes can interleave
// their stacks in LDS (STRIDE = lanes) or a host caller use a plain array
// (STRIDE = 1).  Register depths cost a select per coordinate and step; each
// takes 96 bytes per lane off the stack.
template <int NP>
struct WalkState {
  int cls[NP];
  int level;
  int digits;
};

LSM_HD void walk_distances(const double* T, const double* pl, double* d) {
#pragma unroll
  for (int v = 0; v < 4; v++) d[v] = dot3(pl[0], pl[1], pl[2], T[3 * v], T[3 * v + 1], T[3 * v + 2]) - pl[3];
}

// Climbs while the children at the current depth are all done.
template <int NP>
LSM_HD void walk_climb(WalkState<NP>& W) {
#pragma unroll
  for (int j = NP; j >= 1; j--) {
    const bool done = W.level == j && ((W.digits >> (2 * j)) & 3) >= class_nchild(W.cls[j - 1]);
    W.level = done ? j - 1 : W.level;
  }
}

// Starts the walk of tet A, whose class against the first plane is cls0.
template <int NP>
LSM_HD void walk_begin(WalkState<NP>& W, int cls0) {
  W.cls[0] = cls0;
#pragma unroll
  for (int j = 1; j < NP; j++) W.cls[j] = 0;
  W.level = 1;
  W.digits = 0;
  walk_climb(W);
}

// One step (W.level >= 1).  plane(j, out) gives plane j (0..NP-1) as
// [n0 n1 n2 c].  Calls sink(tet, vol_tet, keep) once per step: keep is true
// for the kept output tets, in the reference's order, and false otherwise
// (the tet is then not an output).
template <int NP, int STRIDE, int NREG = 0, class PlaneFn, class Sink>
LSM_HD void walk_next(WalkState<NP>& W, const double* A, double* R, double* stk, PlaneFn&& plane,
                      double tol, Sink&& sink) {
  static_assert(NREG >= 0 && NREG < NP, "register depths 1..NREG, NREG < NP");
  const int k = W.level;
  const int ci = (W.digits >> (2 * k)) & 3;
  int c = in_reg_i(W.cls[0]);
#pragma unroll
  for (int j = 1; j < NP; j++) c = k == j + 1 ? in_reg_i(W.cls[j]) : c;
  // the parent (depth k-1): A at depth 0, a register tet at depths 1..NREG,
  // else the stack's; the stack word is read in any case so that the loads
  // are issued together
  const int sd = k - 2 - NREG;
  const double* src = stk + (sd > 0 ? sd : 0) * 12 * STRIDE;
  double parent[12];
#pragma unroll
  for (int i = 0; i < 12; i++) {
    double v = in_reg(src[i * STRIDE]);
#pragma unroll
    for (int r = NREG; r >= 1; r--) v = k == r + 1 ? in_reg(R[(r - 1) * 12 + i]) : v;
    parent[i] = k == 1 ? in_reg(A[i]) : v;
  }
  double pl[4], pd[4];
  plane(k - 1, pl);
  walk_distances(parent, pl, pd);
  double child[12];
  build_child(parent, pd, make_recipe(c, ci), child);
  W.digits += 1 << (2 * k);
  const bool leaf = k == NP;
  const double vol = tet_volume_tet(child);
  sink(child, vol, leaf && vol >= tol);
  // the child's class against the next plane (computed at depth NP too, and
  // unused there)
  double cp[4], dc[4];
  plane(k < NP ? k : NP - 1, cp);
  walk_distances(child, cp, dc);
  const int cc = pack_class(dc);
#pragma unroll
  for (int r = 1; r <= NREG; r++) {
#pragma unroll
    for (int i = 0; i < 12; i++) R[(r - 1) * 12 + i] = k == r ? child[i] : R[(r - 1) * 12 + i];
  }
  if (!leaf && k > NREG) {
    double* dst = stk + (k - 1 - NREG) * 12 * STRIDE;
#pragma unroll
    for (int i = 0; i < 12; i++) dst[i * STRIDE] = child[i];
  }
#pragma unroll
  for (int j = 1; j < NP; j++) W.cls[j] = !leaf && k == j ? cc : W.cls[j];
  W.digits = leaf ? W.digits : (W.digits & ~(3 << (2 * (k + 1))));
  W.level = leaf ? k : k + 1;
  walk_climb(W);
}

// Whole walk of one pair on the host or in a single lane: sink(tet, vol_tet)
// for each kept output tet; returns how many.  planes: NP x [n0 n1 n2 c].
template <int NP, class Sink>
LSM_HD int walk_whole(const double* A, const double* planes, double tol, Sink&& sink) {
  double stk[12 * (NP - 1)];
  double d0[4];
  walk_distances(A, planes, d0);
  WalkState<NP> W;
  walk_begin(W, pack_class(d0));
  int kept = 0;
  auto plane = [&](int j, double* o) {
#pragma unroll
    for (int i = 0; i < 4; i++) o[i] = planes[4 * j + i];
  };
  while (W.level > 0)
    walk_next<NP, 1>(W, A, nullptr, stk, plane, tol, [&](const double* t, double v, bool keep) {
      if (keep) {
        sink(t, v);
        kept++;
      }
    });
  return kept;
}

// Whole clip of one pair with the walk (host checks).
template <class Sink>
LSM_HD int clip_tet_walk(const double* A, const Planes& P, double tol, Sink&& sink) {
  double planes[16];
  for (int i = 0; i < 4; i++) {
    for (int k = 0; k < 3; k++) planes[4 * i + k] = P.n[i][k];
    planes[4 * i + 3] = P.c[i];
  }
  return walk_whole<4>(A, planes, tol, sink);
}


// The six face planes of a hexahedron in Gmsh node order, normals from
// unit_cross as written in intersect_tet_hex (Supermesh.F90:269-280); no
// orientation correction.  planes: 6 x [n0 n1 n2 c].
LSM_HD void hex_planes(const double* h, double* planes) {
  // (edge a = h[a1] - h[a0], edge b = h[b1] - h[a0], plane point h[a0])
  const int a0[6] = {0, 4, 5, 1, 0, 7}, a1[6] = {1, 5, 1, 0, 4, 6}, b1[6] = {5, 6, 2, 3, 7, 2};
#pragma unroll
  for (int f = 0; f < 6; f++) {
    const double* o = h + 3 * a0[f];
    const double* p = h + 3 * a1[f];
    const double* q = h + 3 * b1[f];
    double n[3];
    unit_cross(p[0] - o[0], p[1] - o[1], p[2] - o[2], q[0] - o[0], q[1] - o[1], q[2] - o[2], n);
    planes[4 * f] = n[0];
    planes[4 * f + 1] = n[1];
    planes[4 * f + 2] = n[2];
    planes[4 * f + 3] = dot3(o[0], o[1], o[2], n[0], n[1], n[2]);
  }
}

// Split tet s of a hexahedron (intersect_hexes, Supermesh.F90:318-332):
// vertex ids 2567, 0237, 0457, 0125, 0257, one nibble each.
LSM_HD void hex_split_tet(const double* hex, int s, double* T) {
  const unsigned code = s == 0 ? 0x7652u : s == 1 ? 0x7320u : s == 2 ? 0x7540u : s == 3 ? 0x5210u : 0x7520u;
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const int hv = (code >> (4 * v)) & 15;
    T[3 * v] = hex[3 * hv];
    T[3 * v + 1] = hex[3 * hv + 1];
    T[3 * v + 2] = hex[3 * hv + 2];
  }
}

// Tet A clipped by the 6 planes of hex B, tolerance from A's volume only
// (intersect_tet_hex -> intersect_tet_planes, Tetrahedron_Intersection.F90:191-241).
template <class Sink>
LSM_HD int clip_tet_hex(const double* A, const double* hex, Sink&& sink) {
  double planes[24];
  hex_planes(hex, planes);
  const double tol = 10.0 * fspacing(tet_volume_tet(A));
  return walk_whole<6>(A, planes, tol, sink);
}

// True when tet A lies beyond one of B's face planes by a margin that no
// rounding in the clipper can bridge, so the clipper provably returns no
// tets.  Every vertex the clipper creates is a rounded convex combination of
// A's vertices: after three planes it is within ~30u*S of conv(A), S bounding
// the coordinates of A and B, and its computed distance to the plane then
// differs from the smallest distance of A's vertices by less than 1e-14*S.
// With all of A's distances above eps + 1e-12*S, every sub-tet meeting that
// plane is classified wholly outside and discarded
// (Tetrahedron_Intersection.F90:278-281).  `scale` = max(1, max |coord|).
LSM_HD bool tet_separated_by_b(const double* A, const Planes& P, double scale) {
  const double T = kEpsilon + 1e-12 * scale;
  bool sep = false;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    bool out = true;
#pragma unroll
    for (int v = 0; v < 4; v++) {
      double d = dot3(P.n[i][0], P.n[i][1], P.n[i][2], A[3 * v], A[3 * v + 1], A[3 * v + 2]) - P.c[i];
      out = out && (d > T);
    }
    sep = sep || out;
  }
  return sep;
}

LSM_HD double tet_tol(double vol_a, double vol_b) {
  return 10.0 * fmin_f(fspacing(vol_a), fspacing(vol_b));
}

// Clips tet A by the planes of B.  Calls sink(tet, vol_tet) for each kept
// output tet in the reference's order; returns how many were kept.
template <class Sink>
LSM_HD int clip_tet(const Tet& A, const Planes& P, double tol, Sink&& sink) {
  int kept = 0;
  Split s1 = classify(A, P.n[0], P.c[0]);
  for (int i1 = 0; i1 < s1.nchild; i1++) {
    Tet T1 = make_child(A, s1, i1);
    Split s2 = classify(T1, P.n[1], P.c[1]);
    for (int i2 = 0; i2 < s2.nchild; i2++) {
      Tet T2 = make_child(T1, s2, i2);
      Split s3 = classify(T2, P.n[2], P.c[2]);
      for (int i3 = 0; i3 < s3.nchild; i3++) {
        Tet T3 = make_child(T2, s3, i3);
        Split s4 = classify(T3, P.n[3], P.c[3]);
        for (int i4 = 0; i4 < s4.nchild; i4++) {
          Tet T4 = make_child(T3, s4, i4);
          double vol = tet_volume_tet(T4.v);
          if (vol >= tol) {
            sink(T4, vol);
            kept++;
          }
        }
      }
    }
  }
  return kept;
}

}  // namespace lsm
