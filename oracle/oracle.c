/*
 * oracle.c -- CPU restatement of libsupermesh's supermeshing hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Compiled with
 * -O2 -ffp-contract=off -fno-fast-math so that every operation rounds as
 * the reference's Fortran does on x86-64 (IEEE double, no FMA contraction,
 * correctly rounded / and sqrt).
 *
 * Restated reference routines (all paths under src/):
 *   orc_spacing, orc_epsilon      Fortran SPACING / EPSILON as used at
 *                                 Triangle_Intersection.F90:174,
 *                                 Tetrahedron_Intersection.F90:151,266
 *   orc_norm2_3                   Fortran NORM2 in unit_cross,
 *                                 Tetrahedron_Intersection.F90:669-677
 *   orc_triangle_area             Triangle_Intersection.F90:453-477
 *   orc_tetrahedron_volume_real   Tetrahedron_Intersection.F90:690-706
 *   orc_tetrahedron_volume_tet    Tetrahedron_Intersection.F90:708-725
 *   orc_intersect_tris            Triangle_Intersection.F90:144-185 (driver),
 *                                 :282-335 (clip_buf), :395-420 (get_lines_tri)
 *   orc_intersect_tets            Tetrahedron_Intersection.F90:137-179 (BFS
 *                                 driver), :244-436 (clip_buf),
 *                                 :634-667 (get_planes_tet), :679-688
 *   orc_intersect_tri_quad,       Supermesh.F90:211-334, get_lines_poly
 *   orc_intersect_quads,          Triangle_Intersection.F90:422-451,
 *   orc_intersect_tet_hex,        intersect_polys_lines :216-276,
 *   orc_intersect_hexes           intersect_tet_planes
 *                                 Tetrahedron_Intersection.F90:189-241
 *   orc_intersect_intervals,      Interval_Intersection.F90:45-110
 *   orc_interval_size
 *   orc_sort_finder               Intersection_Finder.F90:811-921
 *   orc_ksum_*                    Precision.F90:72-113
 *   orc_brute_force_finder        Intersection_Finder.F90:990-1050 (strict);
 *                                 closed variant = R-tree predicate,
 *                                 R-Tree_Intersection_Finder_C++.h:233-251
 *                                 with libspatialindex closed Region overlap
 *   orc_grid_finder               the brute-force set above, pruned by a
 *                                 uniform grid of cells (full-size checks)
 *   orc_tree_*                    Quadtree_Intersection_Finder.F90:174-420,
 *                                 Octree_Intersection_Finder.F90:153-411
 *   orc_mass_pairs,               test_parallel_p1_inner_product_2d.F90:59-68,
 *   orc_inner_product             205-299; ..._3d.F90:61-80,226-337
 *   orc_mesh_nelist               Graphs.F90:76-116
 *   orc_mesh_nnlist               C_bindings.F90:215-255
 *   orc_mesh_eelist               Graphs.F90:118-187 (release build)
 *   orc_transpose_sparsity        C_bindings.F90:397-430
 *   orc_advancing_front_finder    Intersection_Finder.F90:270-454
 *
 * What pins each function (tests/test_oracle_*.py):
 *   spacing / epsilon / norm2 / dot: bit-compared with the reference's own
 *     compiler (amdflang) on probe vectors when amdflang is present.
 *   tri / tet clippers: the reference tests' known answers on the reference's
 *     fixture meshes (sum area 0.5, Kahan sum volume 1/3, cube.1 x cube.2 = 1),
 *     the reference-measured candidate / non-empty / output counts on those
 *     meshes, exact rational clipping of tiny inputs, self-intersection and
 *     completeness (sum over A of |a ^ b| = |b|).
 *   finders: tree set == brute-force set; brute force == an independent numpy
 *     broadcast of the predicate; reference-measured candidate counts.
 *   mass: the reference tests' integral values (8.3333333333333398e-2 in 2D,
 *     0.12499999999999989 in 3D), 1^T M 1 = overlap measure, exactness of the
 *     degree-2 rules on linear fields.
 *   adjacency: nelist == a stable numpy argsort of the element-node list;
 *     nnlist == the sorted scipy adjacency of the node graph; eelist: the
 *     face-neighbour sets of structured grids in closed form, symmetry, and
 *     the discovery order of hand-worked meshes; transpose == scipy's CSR
 *     transpose, twice == identity.
 *   advancing front: on connected meshes every row's set equals the closed
 *     brute-force set; the row order (the front's stack order, each row
 *     grown from its clue's, a new seed per part, closed boxes) by a mesh
 *     pair worked by hand through the reference's loop.
 *   sort finder: on meshes the brute-force strict set; on intervals that
 *     overlap within a mesh (where the sweep skips pairs) cases worked by
 *     hand through the reference's sweep, and merge orders with ties.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* Operation counters (built into liboracle_count.so with -DORC_COUNT):
 * F(n) counts n FP64 add/sub/mul/div/sqrt, K(n) counts n FP64 compares,
 * abs, min or max.  Empty in the normal build. */
#ifdef ORC_COUNT
static __thread unsigned long long orc_nflop, orc_ncmp;
#define F(n) (orc_nflop += (n))
#define K(n) (orc_ncmp += (n))
#else
#define F(n) ((void)0)
#define K(n) ((void)0)
#endif

void orc_op_counts(unsigned long long *out, int reset) {
#ifdef ORC_COUNT
  out[0] = orc_nflop;
  out[1] = orc_ncmp;
  if (reset) orc_nflop = orc_ncmp = 0;
#else
  (void)reset;
  out[0] = out[1] = 0;
#endif
}

/* ------------------------------------------------------------------------ */
/* Fortran intrinsics                                                        */
/* ------------------------------------------------------------------------ */

/* SPACING(x) = 2**max(e-53, -1022) with x = m*2**e, 0.5 <= |m| < 1;
 * TINY(x) for x == 0 (Fortran 2008, 13.7.157). */
double orc_spacing(double x) {
  if (x == 0.0) return DBL_MIN;
  int e;
  frexp(x, &e);
  int k = e - 53;
  if (k < -1022) k = -1022;
  return ldexp(1.0, k);
}

double orc_epsilon(void) { return DBL_EPSILON; }

/* NORM2 as the reference's compiler evaluates it for real(8): a running
 * maximum m and a scaled sum of squares s, result m*sqrt(1+s). */
double orc_norm2_3(const double v[3]) {
  double mx = 0.0, s = 0.0;
  for (int i = 0; i < 3; i++) {
    double ax = fabs(v[i]);
    K(2);
    if (mx == 0.0) {
      mx = ax;
    } else if ((K(1), ax > mx)) {
      F(4);
      double t = mx / ax;
      double tsq = t * t;
      s = s * tsq;
      s = s + tsq;
      mx = ax;
    } else {
      F(3);
      double t = ax / mx;
      s = s + t * t;
    }
  }
  F(3);
  return mx * sqrt(1.0 + s);
}

/* DOT_PRODUCT over 3 reals: sequential accumulation. */
double orc_dot3(const double a[3], const double b[3]) {
  F(5);
  double s = a[0] * b[0];
  s = s + a[1] * b[1];
  s = s + a[2] * b[2];
  return s;
}

/* ------------------------------------------------------------------------ */
/* Measures                                                                  */
/* ------------------------------------------------------------------------ */

double orc_triangle_area(const double t[6]) {
  F(8);
  K(1);
  double ux = t[4] - t[0], uy = t[5] - t[1];
  double vx = t[2] - t[0], vy = t[3] - t[1];
  return 0.5 * fabs(uy * vx - ux * vy);
}

double orc_tetrahedron_volume_real(const double t[12]) {
  F(24);
  K(1);
  double e1[3], e2[3], e3[3];
  for (int i = 0; i < 3; i++) {
    e1[i] = t[3 + i] - t[i];
    e2[i] = t[6 + i] - t[i];
    e3[i] = t[9 + i] - t[i];
  }
  const double sixth = 1.0 / 6.0;
  double d = e1[0] * (e2[1] * e3[2] - e2[2] * e3[1]);
  d = d + e1[1] * (e2[2] * e3[0] - e2[0] * e3[2]);
  d = d + e1[2] * (e2[0] * e3[1] - e2[1] * e3[0]);
  return sixth * fabs(d);
}

double orc_tetrahedron_volume_tet(const double t[12]) {
  F(19);  /* + dot3 */
  K(1);
  double a[3], b[3], c[3], cr[3];
  for (int i = 0; i < 3; i++) {
    a[i] = t[i] - t[9 + i];
    b[i] = t[3 + i] - t[9 + i];
    c[i] = t[6 + i] - t[9 + i];
  }
  cr[0] = b[1] * c[2] - b[2] * c[1];
  cr[1] = b[2] * c[0] - b[0] * c[2];
  cr[2] = b[0] * c[1] - b[1] * c[0];
  return fabs(orc_dot3(a, cr)) / 6.0;
}

/* ------------------------------------------------------------------------ */
/* Triangle clipping (Sutherland-Hodgman, then a fan)                        */
/* ------------------------------------------------------------------------ */

typedef struct { double n[2], p[2]; } orc_line;

static void get_lines_tri(const double t[6], orc_line L[3]) {
  F(11);
  K(1);
  for (int k = 0; k < 3; k++) {
    int k1 = (k + 1) % 3;
    L[k].n[0] = -(t[2 * k1 + 1] - t[2 * k + 1]);
    L[k].n[1] = (t[2 * k1] - t[2 * k]);
    L[k].p[0] = t[2 * k];
    L[k].p[1] = t[2 * k + 1];
  }
  double d = (t[2] - t[0]) * L[2].n[0];
  d = d + (t[3] - t[1]) * L[2].n[1];
  if (d > 0.0) {
    for (int k = 0; k < 3; k++) {
      L[k].n[0] = -L[k].n[0];
      L[k].n[1] = -L[k].n[1];
    }
  }
}

#define ORC_PTS (ORC_TRI_BUF_SIZE + 2)

/* One Sutherland-Hodgman pass (initial and terminal roles swapped, as in the
 * reference); returns the new point count. */
static int clip_line(const orc_line *L, const double *pts, int n, double *out) {
  double d[ORC_PTS];
  F(5 * n);
  for (int i = 0; i < n; i++) {
    double s = L->n[0] * (pts[2 * i] - L->p[0]);
    s = s + L->n[1] * (pts[2 * i + 1] - L->p[1]);
    d[i] = s;
  }
  int m = 0;
  for (int i = 0; i < n; i++) {
    const double *p1 = pts + 2 * i;
    const double *p2 = (i == n - 1) ? pts : pts + 2 * (i + 1);
    double d1 = d[i], d2 = (i == n - 1) ? d[0] : d[i + 1];
    K(1);
    if (d1 < 0.0) {
      out[2 * m] = p1[0];
      out[2 * m + 1] = p1[1];
      m++;
      K(1);
      if (!(d2 <= 0.0)) {
        F(8);
        K(4);
        double f = fabs(d1) / (fabs(d1) + fabs(d2));
        f = fmax(fmin(f, 1.0), 0.0);
        out[2 * m] = p1[0] + f * (p2[0] - p1[0]);
        out[2 * m + 1] = p1[1] + f * (p2[1] - p1[1]);
        m++;
      }
    } else if ((K(1), d1 == 0.0)) {
      out[2 * m] = p1[0];
      out[2 * m + 1] = p1[1];
      m++;
    } else if ((K(1), d2 < 0.0)) {
      F(8);
      K(4);
      double f = fabs(d1) / (fabs(d1) + fabs(d2));
      f = fmax(fmin(f, 1.0), 0.0);
      out[2 * m] = p1[0] + f * (p2[0] - p1[0]);
      out[2 * m + 1] = p1[1] + f * (p2[1] - p1[1]);
      m++;
    }
  }
  return m;
}

int orc_intersect_tris(const double tri_a[6], const double tri_b[6], double *tris_c) {
  orc_line L[3];
  get_lines_tri(tri_b, L);
  double buf0[2 * ORC_PTS], buf1[2 * ORC_PTS];
  memcpy(buf0, tri_a, 6 * sizeof(double));
  int n = clip_line(&L[0], buf0, 3, buf1);
  if (n < 3) return 0;
  n = clip_line(&L[1], buf1, n, buf0);
  if (n < 3) return 0;
  n = clip_line(&L[2], buf0, n, buf1);
  if (n < 3) return 0;
  F(1);
  K(1);
  double tol = 10.0 * fmin(orc_spacing(orc_triangle_area(tri_a)),
                           orc_spacing(orc_triangle_area(tri_b)));
  int nc = 0;
  for (int i = 0; i < n - 2; i++) {
    double *t = tris_c + 6 * nc;
    t[0] = buf1[0];
    t[1] = buf1[1];
    t[2] = buf1[2 * (i + 1)];
    t[3] = buf1[2 * (i + 1) + 1];
    t[4] = buf1[2 * (i + 2)];
    t[5] = buf1[2 * (i + 2) + 1];
    K(1);
    if (!(orc_triangle_area(t) < tol)) nc++;
  }
  return nc;
}

/* ------------------------------------------------------------------------ */
/* Tetrahedron clipping (plane at a time, breadth first)                     */
/* ------------------------------------------------------------------------ */

static void unit_cross(const double a[3], const double b[3], double out[3]) {
  F(12);
  double c[3];
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
  double n = orc_norm2_3(c);
  out[0] = c[0] / n;
  out[1] = c[1] / n;
  out[2] = c[2] / n;
}

void orc_get_planes_tet(const double t[12], double nrm[12], double c[4]) {
  F(15);
  K(1);
  double e10[3], e20[3], e30[3], e21[3], e31[3];
  for (int i = 0; i < 3; i++) {
    e10[i] = t[3 + i] - t[i];
    e20[i] = t[6 + i] - t[i];
    e30[i] = t[9 + i] - t[i];
    e21[i] = t[6 + i] - t[3 + i];
    e31[i] = t[9 + i] - t[3 + i];
  }
  unit_cross(e20, e10, nrm + 0);
  unit_cross(e10, e30, nrm + 3);
  unit_cross(e30, e20, nrm + 6);
  unit_cross(e21, e31, nrm + 9);
  double det = orc_dot3(e10, nrm + 9);
  if (det < 0.0) {
    for (int i = 0; i < 12; i++) nrm[i] = -nrm[i];
  }
  for (int i = 0; i < 4; i++) c[i] = orc_dot3(t + 3 * i, nrm + 3 * i);
}

/* p = w0*v[ip] + w1*v[in], with the reference's weights for vertices ip
 * (outside, d > 0) and in (inside, d < 0). */
static void split_point(const double *v, const double *d, int ip, int in, double *p) {
  F(13);
  double invdiff = 1.0 / (d[ip] - d[in]);
  double w0 = -d[in] * invdiff;
  double w1 = d[ip] * invdiff;
  for (int k = 0; k < 3; k++) p[k] = w0 * v[3 * ip + k] + w1 * v[3 * in + k];
}

static void copy3(double *dst, const double *src) {
  dst[0] = src[0];
  dst[1] = src[1];
  dst[2] = src[2];
}

/* Clip tet v against (n, c) and append 0..3 tets to out; returns new count. */
static int clip_tet_plane(const double *n, double c, const double *v, double *out, int nout) {
  double d[4];
  int neg[4], pos[4], zer[4], nn = 0, np = 0, nz = 0;
  for (int i = 0; i < 4; i++) {
    d[i] = orc_dot3(n, v + 3 * i) - c;
    F(1);
    K(4);
    if (fabs(d[i]) < DBL_EPSILON) zer[nz++] = i;
    else if (d[i] < 0.0) neg[nn++] = i;
    else if (d[i] > 0.0) pos[np++] = i;
  }
  if (nn == 0) return nout;
  double *t = out + 12 * nout;
  if (np == 0) {
    memcpy(t, v, 12 * sizeof(double));
    return nout + 1;
  }
  double tmp[4][3];
  if (np == 3) {
    memcpy(t, v, 12 * sizeof(double));
    for (int i = 0; i < 3; i++) split_point(v, d, pos[i], neg[0], t + 3 * pos[i]);
    return nout + 1;
  } else if (np == 2) {
    if (nn == 2) {
      split_point(v, d, pos[0], neg[0], tmp[0]);
      split_point(v, d, pos[1], neg[0], tmp[1]);
      split_point(v, d, pos[0], neg[1], tmp[2]);
      split_point(v, d, pos[1], neg[1], tmp[3]);
      memcpy(t, v, 12 * sizeof(double));
      copy3(t + 3 * pos[0], tmp[2]);
      copy3(t + 3 * pos[1], tmp[1]);
      t += 12;
      copy3(t + 0, v + 3 * neg[1]);
      copy3(t + 3, tmp[3]);
      copy3(t + 6, tmp[2]);
      copy3(t + 9, tmp[1]);
      t += 12;
      copy3(t + 0, v + 3 * neg[0]);
      copy3(t + 3, tmp[0]);
      copy3(t + 6, tmp[1]);
      copy3(t + 9, tmp[2]);
      return nout + 3;
    } else if (nn == 1) {
      memcpy(t, v, 12 * sizeof(double));
      for (int i = 0; i < 2; i++) split_point(v, d, pos[i], neg[0], t + 3 * pos[i]);
      return nout + 1;
    }
  } else if (np == 1) {
    if (nn == 3) {
      for (int i = 0; i < 3; i++) split_point(v, d, pos[0], neg[i], tmp[i]);
      memcpy(t, v, 12 * sizeof(double));
      copy3(t + 3 * pos[0], tmp[0]);
      t += 12;
      copy3(t + 0, tmp[0]);
      copy3(t + 3, v + 3 * neg[1]);
      copy3(t + 6, v + 3 * neg[2]);
      copy3(t + 9, tmp[1]);
      t += 12;
      copy3(t + 0, v + 3 * neg[2]);
      copy3(t + 3, tmp[1]);
      copy3(t + 6, tmp[2]);
      copy3(t + 9, tmp[0]);
      return nout + 3;
    } else if (nn == 2) {
      for (int i = 0; i < 2; i++) split_point(v, d, pos[0], neg[i], tmp[i]);
      memcpy(t, v, 12 * sizeof(double));
      copy3(t + 3 * pos[0], tmp[0]);
      t += 12;
      copy3(t + 0, tmp[1]);
      copy3(t + 3, v + 3 * zer[0]);
      copy3(t + 6, v + 3 * neg[1]);
      copy3(t + 9, tmp[0]);
      return nout + 2;
    } else if (nn == 1) {
      memcpy(t, v, 12 * sizeof(double));
      split_point(v, d, pos[0], neg[0], t + 3 * pos[0]);
      return nout + 1;
    }
  }
  return nout;
}

int orc_intersect_tets(const double tet_a[12], const double tet_b[12], double *tets_c) {
  double nrm[12], c[4];
  orc_get_planes_tet(tet_b, nrm, c);
  double vol_b = orc_tetrahedron_volume_tet(tet_b);
  F(1);
  K(1);
  double tol = 10.0 * fmin(orc_spacing(orc_tetrahedron_volume_tet(tet_a)), orc_spacing(vol_b));
  double cur[ORC_TET_BUF_SIZE * 12], nxt[ORC_TET_BUF_SIZE * 12];
  memcpy(cur, tet_a, 12 * sizeof(double));
  int n = 1;
  for (int i = 0; i < 4; i++) {
    int m = 0;
    for (int j = 0; j < n; j++) m = clip_tet_plane(nrm + 3 * i, c[i], cur + 12 * j, nxt, m);
    if (i != 3) {
      memcpy(cur, nxt, (size_t)m * 12 * sizeof(double));
      n = m;
    } else {
      n = 0;
      for (int j = 0; j < m; j++) {
        K(1);
        if (orc_tetrahedron_volume_tet(nxt + 12 * j) >= tol) {
          memcpy(tets_c + 12 * n, nxt + 12 * j, 12 * sizeof(double));
          n++;
        }
      }
    }
  }
  return n;
}

/* ------------------------------------------------------------------------ */
/* Mixed element types (Supermesh.F90:211-334)                               */
/* ------------------------------------------------------------------------ */

/* get_lines_poly, Triangle_Intersection.F90:422-451: normals (dy, -dx), all
 * flipped when vertex 1 lies on the positive side of line 2. */
static void get_lines_poly(const double *q, int loc, orc_line *L) {
  for (int k = 0; k < loc; k++) {
    int k1 = (k + 1) % loc;
    L[k].n[0] = (q[2 * k1 + 1] - q[2 * k + 1]);
    L[k].n[1] = -(q[2 * k1] - q[2 * k]);
    L[k].p[0] = q[2 * k];
    L[k].p[1] = q[2 * k + 1];
  }
  double d = (q[0] - L[1].p[0]) * L[1].n[0];
  d = d + (q[1] - L[1].p[1]) * L[1].n[1];
  if (d > 0.0) {
    for (int k = 0; k < loc; k++) {
      L[k].n[0] = -L[k].n[0];
      L[k].n[1] = -L[k].n[1];
    }
  }
}

/* intersect_polys_lines, Triangle_Intersection.F90:216-276: a polygon of
 * na points clipped by nl lines, then fanned; tol as passed. */
static int polys_lines(const double *poly, int na, const orc_line *L, int nl, double tol,
                       double *tris_c) {
  double buf0[2 * 80], buf1[2 * 80];
  memcpy(buf0, poly, (size_t)na * 2 * sizeof(double));
  int n = na;
  double *cur = buf0, *nxt = buf1;
  for (int i = 0; i < nl; i++) {
    n = clip_line(&L[i], cur, n, nxt);
    if (n < 3) return 0;
    double *t = cur;
    cur = nxt;
    nxt = t;
  }
  int nc = 0;
  for (int i = 0; i < n - 2; i++) {
    double *t = tris_c + 6 * nc;
    t[0] = cur[0];
    t[1] = cur[1];
    t[2] = cur[2 * (i + 1)];
    t[3] = cur[2 * (i + 1) + 1];
    t[4] = cur[2 * (i + 2)];
    t[5] = cur[2 * (i + 2) + 1];
    if (!(orc_triangle_area(t) < tol)) nc++;
  }
  return nc;
}

/* intersect_tri_quad (Supermesh.F90:211-229) -> intersect_tri_lines with no
 * area for B: tol = 10*spacing(area(A)). */
int orc_intersect_tri_quad(const double tri_a[6], const double quad_b[8], double *tris_c) {
  orc_line L[4];
  get_lines_poly(quad_b, 4, L);
  return polys_lines(tri_a, 3, L, 4, 10.0 * orc_spacing(orc_triangle_area(tri_a)), tris_c);
}

/* intersect_quads (Supermesh.F90:231-248) -> intersect_polys_real with no
 * areas: tol = 10*tiny. */
int orc_intersect_quads(const double quad_a[8], const double quad_b[8], double *tris_c) {
  orc_line L[4];
  get_lines_poly(quad_b, 4, L);
  return polys_lines(quad_a, 4, L, 4, 10.0 * DBL_MIN, tris_c);
}

void orc_hex_planes(const double hex[24], double nrm[18], double c[6]) {
  /* plane f: unit_cross(h[a1] - h[a0], h[b1] - h[a0]), c = h[a0] . n */
  static const int a0[6] = {0, 4, 5, 1, 0, 7}, a1[6] = {1, 5, 1, 0, 4, 6},
                   b1[6] = {5, 6, 2, 3, 7, 2};
  for (int f = 0; f < 6; f++) {
    double ea[3], eb[3];
    for (int i = 0; i < 3; i++) {
      ea[i] = hex[3 * a1[f] + i] - hex[3 * a0[f] + i];
      eb[i] = hex[3 * b1[f] + i] - hex[3 * a0[f] + i];
    }
    unit_cross(ea, eb, nrm + 3 * f);
    c[f] = orc_dot3(hex + 3 * a0[f], nrm + 3 * f);
  }
}

/* intersect_tet_hex (Supermesh.F90:250-301) -> intersect_tet_planes
 * (Tetrahedron_Intersection.F90:189-241): 6 planes breadth first, keep
 * vol_tet >= 10*spacing(vol_tet(A)). */
int orc_intersect_tet_hex(const double tet_a[12], const double hex_b[24], double *tets_c) {
  double nrm[18], c[6];
  orc_hex_planes(hex_b, nrm, c);
  double tol = 10.0 * orc_spacing(orc_tetrahedron_volume_tet(tet_a));
  static __thread double cur[ORC_TET_HEX_BUF * 12], nxt[ORC_TET_HEX_BUF * 12];
  memcpy(cur, tet_a, 12 * sizeof(double));
  int n = 1;
  for (int i = 0; i < 6; i++) {
    int m = 0;
    for (int j = 0; j < n; j++) m = clip_tet_plane(nrm + 3 * i, c[i], cur + 12 * j, nxt, m);
    if (i != 5) {
      memcpy(cur, nxt, (size_t)m * 12 * sizeof(double));
      n = m;
    } else {
      n = 0;
      for (int j = 0; j < m; j++) {
        if (orc_tetrahedron_volume_tet(nxt + 12 * j) >= tol) {
          memcpy(tets_c + 12 * n, nxt + 12 * j, 12 * sizeof(double));
          n++;
        }
      }
    }
  }
  return n;
}

/* intersect_hexes (Supermesh.F90:303-334): A split into 5 tets (Gmsh
 * vertices 3678, 1348, 1568, 1236, 1368), each clipped by B's planes. */
int orc_intersect_hexes(const double hex_a[24], const double hex_b[24], double *tets_c) {
  static const int split[5][4] = {{2, 5, 6, 7}, {0, 2, 3, 7}, {0, 4, 5, 7}, {0, 1, 2, 5}, {0, 2, 5, 7}};
  int n = 0;
  for (int s = 0; s < 5; s++) {
    double t[12];
    for (int v = 0; v < 4; v++) memcpy(t + 3 * v, hex_a + 3 * split[s][v], 3 * sizeof(double));
    n += orc_intersect_tet_hex(t, hex_b, tets_c + 12 * n);
  }
  return n;
}

/* ------------------------------------------------------------------------ */
/* 1D (Interval_Intersection.F90:45-110, Intersection_Finder.F90:811-921)    */
/* ------------------------------------------------------------------------ */

double orc_interval_size(const double v[2]) { return fabs(v[1] - v[0]); }

/* Fortran MIN/MAX of two values: first argument on ties. */
static double fmin2(double a, double b) { return b < a ? b : a; }
static double fmax2(double a, double b) { return b > a ? b : a; }

int orc_intersect_intervals(const double a[2], const double b[2], double c[2]) {
  double min_a = fmin2(a[0], a[1]), max_a = fmax2(a[0], a[1]);
  double min_b = fmin2(b[0], b[1]), max_b = fmax2(b[0], b[1]);
  if (max_b <= min_a || min_b >= max_a) return 0;
  c[0] = fmax2(min_a, min_b);
  c[1] = fmin2(max_a, max_b);
  double tol = 10.0 * fmin2(orc_spacing(orc_interval_size(a)), orc_spacing(orc_interval_size(b)));
  return orc_interval_size(c) < tol ? 0 : 1;
}

/* Stable sort of left end points, ties by element number.  The reference
 * (Intersection_Finder.F90:869-917) merge-sorts with a base case for blocks
 * of at most 4 entries that gives slot i the first strict minimum of
 * v(i..n): the identity when the block is sorted, and otherwise an index list
 * with repeats, after which elements go unvisited and the reference reads an
 * undefined row count.  On every input where its result is defined, it equals
 * this stable sort. */
static void sort_merge(const double *v, int *idx, int *work, long n) {
  if (n <= 4) {
    for (long i = 0; i < n; i++) {
      long k = i;
      idx[i] = (int)i;
      while (k > 0 && v[idx[k - 1]] > v[idx[k]]) {
        int t = idx[k - 1];
        idx[k - 1] = idx[k];
        idx[k] = t;
        k--;
      }
    }
    return;
  }
  long h = n / 2;
  sort_merge(v, idx, work, h);
  sort_merge(v + h, idx + h, work + h, n - h);
  for (long i = 0; i < h; i++) work[i] = idx[i];
  for (long i = h; i < n; i++) work[i] = idx[i] + (int)h;
  long i1 = 0, i2 = h;
  for (long i = 0; i < n; i++) {
    if (i1 >= h) {
      for (; i < n; i++) idx[i] = work[i2++];
      break;
    } else if (i2 >= n) {
      for (; i < n; i++) idx[i] = work[i1++];
      break;
    } else if (v[work[i1]] <= v[work[i2]]) {
      idx[i] = work[i1++];
    } else {
      idx[i] = work[i2++];
    }
  }
}

long orc_sort_finder(long ne_a, const double *pos_a, const int *enl_a,
                     long ne_b, const double *pos_b, const int *enl_b,
                     long *indptr, int *indices, long cap) {
  double *left_a = malloc(sizeof(double) * (ne_a + 1)), *left_b = malloc(sizeof(double) * (ne_b + 1));
  int *ia = malloc(sizeof(int) * (ne_a + 1)), *ib = malloc(sizeof(int) * (ne_b + 1));
  int *work = malloc(sizeof(int) * (ne_a + ne_b + 1));
  long *rows = malloc(sizeof(long) * (ne_a + 1));
  long *first = malloc(sizeof(long) * (ne_a + 1));
  for (long e = 0; e < ne_a; e++) left_a[e] = fmin2(pos_a[enl_a[2 * e]], pos_a[enl_a[2 * e + 1]]);
  for (long e = 0; e < ne_b; e++) left_b[e] = fmin2(pos_b[enl_b[2 * e]], pos_b[enl_b[2 * e + 1]]);
  sort_merge(left_a, ia, work, ne_a);
  sort_merge(left_b, ib, work, ne_b);
  /* pass 1: per A element (in sorted order) the B ids, stored contiguously in
   * visiting order; rows[e] = count, first[e] = offset in the visit list */
  int *visit = malloc(sizeof(int) * (ne_a + 1) * 4 + sizeof(int) * (ne_b + 1));
  long nvisit = 0, vcap = (ne_a + 1) * 4 + ne_b + 1;
  long j = 0;
  for (long i = 0; i < ne_a; i++) {
    long ea = ia[i];
    double a0 = pos_a[enl_a[2 * ea]], a1 = pos_a[enl_a[2 * ea + 1]];
    double min_a = fmin2(a0, a1), max_a = fmax2(a0, a1);
    long n = 0;
    first[ea] = nvisit;
    while (j < ne_b) {
      long eb = ib[j];
      double b0 = pos_b[enl_b[2 * eb]], b1 = pos_b[enl_b[2 * eb + 1]];
      if (fmax2(b0, b1) <= min_a) {
        j++;
      } else if (fmin2(b0, b1) >= max_a) {
        break;
      } else {
        if (nvisit == vcap) {
          vcap *= 2;
          visit = realloc(visit, sizeof(int) * vcap);
        }
        visit[nvisit++] = (int)eb;
        n++;
        j++;
      }
    }
    if (n > 0) j--;
    rows[ea] = n;
  }
  indptr[0] = 0;
  for (long e = 0; e < ne_a; e++) indptr[e + 1] = indptr[e] + rows[e];
  if (indices)
    for (long e = 0; e < ne_a; e++)
      for (long k = 0; k < rows[e] && indptr[e] + k < cap; k++) indices[indptr[e] + k] = visit[first[e] + k];
  long total = indptr[ne_a];
  free(left_a); free(left_b); free(ia); free(ib); free(work); free(rows); free(first); free(visit);
  return total;
}

/* ------------------------------------------------------------------------ */
/* Kahan sum                                                                 */
/* ------------------------------------------------------------------------ */

void orc_ksum_init(orc_ksum *cs) { cs->sum = 0.0; cs->error = 0.0; }

void orc_ksum_add(orc_ksum *cs, double x) {
  double cx = cs->error + x;
  double s = cs->sum + cx;
  cs->error = cs->sum - s;
  cs->error = cs->error + cx;
  cs->sum = s;
}

double orc_ksum_value(const orc_ksum *cs) { return cs->sum + cs->error; }

/* ------------------------------------------------------------------------ */
/* Bounding boxes and finders                                                */
/* ------------------------------------------------------------------------ */

/* bbox[0..dim-1] = min, bbox[3..3+dim-1] = max (exact min/max). */
static void element_bbox(int dim, int loc, const double *pos, const int *nodes, double *bb) {
  for (int d = 0; d < dim; d++) {
    double x = pos[(long)nodes[0] * dim + d];
    bb[d] = x;
    bb[3 + d] = x;
  }
  for (int j = 1; j < loc; j++) {
    for (int d = 0; d < dim; d++) {
      double x = pos[(long)nodes[j] * dim + d];
      bb[d] = fmin(bb[d], x);
      bb[3 + d] = fmax(bb[3 + d], x);
    }
  }
}

static void coords_bbox(int dim, int loc, const double *coords, double *bb) {
  for (int d = 0; d < dim; d++) {
    bb[d] = coords[d];
    bb[3 + d] = coords[d];
  }
  for (int j = 1; j < loc; j++) {
    for (int d = 0; d < dim; d++) {
      double x = coords[j * dim + d];
      bb[d] = fmin(bb[d], x);
      bb[3 + d] = fmax(bb[3 + d], x);
    }
  }
}

static int boxes_strict(int dim, const double *a, const double *b) {
  for (int d = 0; d < dim; d++)
    if (!(b[3 + d] > a[d] && b[d] < a[3 + d])) return 0;
  return 1;
}

static int boxes_closed(int dim, const double *a, const double *b) {
  for (int d = 0; d < dim; d++)
    if (!(b[3 + d] >= a[d] && b[d] <= a[3 + d])) return 0;
  return 1;
}

long orc_brute_force_finder(int dim, int closed,
                            long ne_a, int loc_a, const double *pos_a, const int *enl_a,
                            long ne_b, int loc_b, const double *pos_b, const int *enl_b,
                            long *indptr, int *indices) {
  double *bbs = (double *)malloc(sizeof(double) * 6 * (size_t)(ne_b > 0 ? ne_b : 1));
  for (long b = 0; b < ne_b; b++) element_bbox(dim, loc_b, pos_b, enl_b + b * loc_b, bbs + 6 * b);
  long nnz = 0;
  indptr[0] = 0;
  for (long a = 0; a < ne_a; a++) {
    double ba[6];
    element_bbox(dim, loc_a, pos_a, enl_a + a * loc_a, ba);
    for (long b = 0; b < ne_b; b++) {
      int hit = closed ? boxes_closed(dim, ba, bbs + 6 * b) : boxes_strict(dim, ba, bbs + 6 * b);
      if (hit) {
        if (indices) indices[nnz] = (int)b;
        nnz++;
      }
    }
    indptr[a + 1] = nnz;
  }
  free(bbs);
  return nnz;
}

/* The same strict / closed brute-force set, found through a uniform grid of
 * cells so that it runs at the full workload sizes.  B elements are binned by
 * the cell of their box's low corner; cells are at least as wide as every
 * box of A and B, so a B box that overlaps an A box has its low corner within
 * (A.lo - E, A.hi] (E the widest B box) per axis, i.e. in a cell between
 * cell(A.lo - E) and cell(A.hi).  One cell of slack on each side covers the
 * rounding of the cell arithmetic, and every B element in those cells is
 * tested with the exact predicate of orc_brute_force_finder, so the set is
 * that of the all-pairs loop.  Rows come out ascending.  Threads take
 * contiguous ranges of A. */
typedef struct {
  int dim, closed, ncell[3];
  double lo[3], width[3];
  const double *bbs_b;
  const long *cstart;
  const int *cele;
  long a0, a1;
  int loc_a;
  const double *pos_a;
  const int *enl_a;
  double ext_b[3];
  long *counts; /* [a1 - a0] */
  int *hits;    /* grown as needed */
  long nhits, cap;
} grid_job;

static int cell_of(const grid_job *g, int d, double x) {
  double f = floor((x - g->lo[d]) / g->width[d]);
  if (f < 0) return 0;
  if (f > g->ncell[d] - 1) return g->ncell[d] - 1;
  return (int)f;
}

static int cmp_int(const void *x, const void *y) {
  int a = *(const int *)x, b = *(const int *)y;
  return (a > b) - (a < b);
}

static void *grid_worker(void *arg) {
  grid_job *g = (grid_job *)arg;
  const int dim = g->dim;
  for (long a = g->a0; a < g->a1; a++) {
    double ba[6];
    element_bbox(dim, g->loc_a, g->pos_a, g->enl_a + a * g->loc_a, ba);
    int c0[3] = {0, 0, 0}, c1[3] = {0, 0, 0};
    for (int d = 0; d < dim; d++) {
      c0[d] = cell_of(g, d, ba[d] - g->ext_b[d]) - 1;
      c1[d] = cell_of(g, d, ba[3 + d]) + 1;
      if (c0[d] < 0) c0[d] = 0;
      if (c1[d] > g->ncell[d] - 1) c1[d] = g->ncell[d] - 1;
    }
    const long start = g->nhits;
    for (int z = c0[2]; z <= c1[2]; z++)
      for (int y = c0[1]; y <= c1[1]; y++)
        for (int x = c0[0]; x <= c1[0]; x++) {
          const long c = ((long)z * g->ncell[1] + y) * g->ncell[0] + x;
          for (long k = g->cstart[c]; k < g->cstart[c + 1]; k++) {
            const int b = g->cele[k];
            const double *bb = g->bbs_b + 6 * (long)b;
            int hit = g->closed ? boxes_closed(dim, ba, bb) : boxes_strict(dim, ba, bb);
            if (hit) {
              if (g->nhits == g->cap) {
                g->cap = g->cap * 2 + 1024;
                g->hits = (int *)realloc(g->hits, sizeof(int) * (size_t)g->cap);
              }
              g->hits[g->nhits++] = b;
            }
          }
        }
    qsort(g->hits + start, (size_t)(g->nhits - start), sizeof(int), cmp_int);
    g->counts[a - g->a0] = g->nhits - start;
  }
  return NULL;
}

long orc_grid_finder(int dim, int closed, int nthreads,
                     long ne_a, int loc_a, const double *pos_a, const int *enl_a,
                     long ne_b, int loc_b, const double *pos_b, const int *enl_b,
                     long *indptr, int **indices_out) {
  double *bbs = (double *)malloc(sizeof(double) * 6 * (size_t)(ne_b > 0 ? ne_b : 1));
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0}, ext[3] = {0, 0, 0};
  for (long b = 0; b < ne_b; b++) {
    double *bb = bbs + 6 * b;
    element_bbox(dim, loc_b, pos_b, enl_b + b * loc_b, bb);
    for (int d = 0; d < dim; d++) {
      if (b == 0 || bb[d] < lo[d]) lo[d] = bb[d];
      if (b == 0 || bb[d] > hi[d]) hi[d] = bb[d];
      if (bb[3 + d] - bb[d] > ext[d]) ext[d] = bb[3 + d] - bb[d];
    }
  }
  for (long a = 0; a < ne_a; a++) {
    double ba[6];
    element_bbox(dim, loc_a, pos_a, enl_a + a * loc_a, ba);
    for (int d = 0; d < dim; d++)
      if (ba[3 + d] - ba[d] > ext[d]) ext[d] = ba[3 + d] - ba[d];
  }
  grid_job proto;
  memset(&proto, 0, sizeof(proto));
  proto.dim = dim;
  proto.closed = closed;
  const int cap = dim == 2 ? 4096 : 256;
  long ncells = 1;
  for (int d = 0; d < 3; d++) {
    proto.ncell[d] = 1;
    proto.lo[d] = lo[d];
    proto.width[d] = 1.0;
    if (d < dim) {
      double w = ext[d] > 0 ? ext[d] : 1.0;
      double span = hi[d] - lo[d];
      long n = (long)(span / w) + 1;
      if (n > cap) {
        n = cap;
        w = span / (double)(cap - 1);
        if (w < ext[d]) w = ext[d];
      }
      if (n < 1) n = 1;
      proto.ncell[d] = (int)n;
      proto.width[d] = w;
      proto.ext_b[d] = ext[d];
    }
    ncells *= proto.ncell[d];
  }
  long *cstart = (long *)calloc((size_t)ncells + 1, sizeof(long));
  int *cele = (int *)malloc(sizeof(int) * (size_t)(ne_b > 0 ? ne_b : 1));
  long *cell_of_b = (long *)malloc(sizeof(long) * (size_t)(ne_b > 0 ? ne_b : 1));
  for (long b = 0; b < ne_b; b++) {
    long c = 0;
    for (int d = dim - 1; d >= 0; d--) c = c * proto.ncell[d] + cell_of(&proto, d, bbs[6 * b + d]);
    cell_of_b[b] = c;
    cstart[c + 1]++;
  }
  for (long c = 0; c < ncells; c++) cstart[c + 1] += cstart[c];
  long *fill = (long *)malloc(sizeof(long) * (size_t)ncells);
  memcpy(fill, cstart, sizeof(long) * (size_t)ncells);
  for (long b = 0; b < ne_b; b++) cele[fill[cell_of_b[b]]++] = (int)b;
  free(fill);
  free(cell_of_b);
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  grid_job jobs[256];
  pthread_t th[256];
  for (int t = 0; t < nthreads; t++) {
    jobs[t] = proto;
    jobs[t].bbs_b = bbs;
    jobs[t].cstart = cstart;
    jobs[t].cele = cele;
    jobs[t].loc_a = loc_a;
    jobs[t].pos_a = pos_a;
    jobs[t].enl_a = enl_a;
    jobs[t].a0 = ne_a * t / nthreads;
    jobs[t].a1 = ne_a * (t + 1) / nthreads;
    jobs[t].counts = indptr + 1 + jobs[t].a0;
  }
  if (nthreads == 1) {
    grid_worker(&jobs[0]);
  } else {
    for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, grid_worker, &jobs[t]);
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
  }
  indptr[0] = 0;
  for (long a = 0; a < ne_a; a++) indptr[a + 1] += indptr[a];
  const long nnz = indptr[ne_a];
  int *out = (int *)malloc(sizeof(int) * (size_t)(nnz > 0 ? nnz : 1));
  for (int t = 0; t < nthreads; t++) {
    if (jobs[t].nhits) memcpy(out + indptr[jobs[t].a0], jobs[t].hits, sizeof(int) * (size_t)jobs[t].nhits);
    free(jobs[t].hits);
  }
  free(cstart);
  free(cele);
  free(bbs);
  *indices_out = out;
  return nnz;
}

/* Bucket quadtree / octree.  A leaf holds at most max_n elements and splits
 * into 2**dim children at the midpoint when a further element arrives.
 * Elements are inserted into every node whose box they strictly overlap. */
typedef struct {
  int n;          /* element count; INT32_MAX marks an internal node */
  int child0;     /* index of the first child (internal nodes) */
  long slot0;     /* first element slot (leaves) */
  double bb[6];
} orc_node;

struct orc_tree {
  int dim, nchild, max_n;
  orc_node *nodes;
  long nnodes, cap_nodes;
  int *slot_ele;
  double *slot_bb;
  long nslots, cap_slots;
};

static long tree_new_slots(orc_tree *t) {
  if (t->nslots + t->max_n > t->cap_slots) {
    long nc = t->cap_slots * 2 + t->max_n;
    t->slot_ele = (int *)realloc(t->slot_ele, sizeof(int) * (size_t)nc);
    t->slot_bb = (double *)realloc(t->slot_bb, sizeof(double) * 6 * (size_t)nc);
    t->cap_slots = nc;
  }
  long s = t->nslots;
  t->nslots += t->max_n;
  return s;
}

static int tree_new_nodes(orc_tree *t, int count) {
  if (t->nnodes + count > t->cap_nodes) {
    long nc = t->cap_nodes * 2 + count;
    t->nodes = (orc_node *)realloc(t->nodes, sizeof(orc_node) * (size_t)nc);
    t->cap_nodes = nc;
  }
  int first = (int)t->nnodes;
  t->nnodes += count;
  return first;
}

/* Child boxes in the reference's order (Quadtree_...F90:289-296,
 * Octree_...F90:274-289): low corner selectors per child. */
static void child_box(int dim, int k, const double *pb, const double *mid, double *cb) {
  static const int qx[4] = {0, 0, 1, 1}, qy[4] = {0, 1, 1, 0};
  static const int ox[8] = {0, 0, 1, 1, 0, 0, 1, 1}, oy[8] = {0, 1, 1, 0, 0, 1, 1, 0},
                   oz[8] = {0, 0, 0, 0, 1, 1, 1, 1};
  int sel[3];
  if (dim == 2) {
    sel[0] = qx[k];
    sel[1] = qy[k];
  } else {
    sel[0] = ox[k];
    sel[1] = oy[k];
    sel[2] = oz[k];
  }
  for (int d = 0; d < dim; d++) {
    if (sel[d]) {
      cb[d] = mid[d];
      cb[3 + d] = pb[3 + d];
    } else {
      cb[d] = pb[d];
      cb[3 + d] = mid[d];
    }
  }
}

static void tree_add(orc_tree *t, int node, int ele, const double *bb) {
  if (!boxes_strict(t->dim, t->nodes[node].bb, bb)) return;
  orc_node *nd = &t->nodes[node];
  if (nd->n <= t->max_n - 1) {
    long s = nd->slot0 + nd->n;
    t->slot_ele[s] = ele;
    memcpy(t->slot_bb + 6 * s, bb, 6 * sizeof(double));
    nd->n++;
  } else if (nd->n == t->max_n) {
    int c0 = tree_new_nodes(t, t->nchild);
    nd = &t->nodes[node];
    double mid[3];
    for (int d = 0; d < t->dim; d++) mid[d] = 0.5 * (nd->bb[d] + nd->bb[3 + d]);
    for (int k = 0; k < t->nchild; k++) {
      orc_node *ch = &t->nodes[c0 + k];
      ch->n = 0;
      ch->child0 = -1;
      memset(ch->bb, 0, sizeof(ch->bb));
      child_box(t->dim, k, t->nodes[node].bb, mid, ch->bb);
      long s = tree_new_slots(t);
      t->nodes[c0 + k].slot0 = s;
    }
    long parent_slot0 = t->nodes[node].slot0;
    t->nodes[node].child0 = c0;
    t->nodes[node].n = INT32_MAX;
    for (int i = 0; i < t->max_n; i++) {
      int e = t->slot_ele[parent_slot0 + i];
      double eb[6];
      memcpy(eb, t->slot_bb + 6 * (parent_slot0 + i), sizeof(eb));
      for (int k = 0; k < t->nchild; k++) tree_add(t, c0 + k, e, eb);
    }
    for (int k = 0; k < t->nchild; k++) tree_add(t, c0 + k, ele, bb);
  } else {
    int c0 = nd->child0;
    for (int k = 0; k < t->nchild; k++) tree_add(t, c0 + k, ele, bb);
  }
}

orc_tree *orc_tree_build(int dim, long nnodes_b, const double *pos_b,
                         long ne_b, int loc_b, const int *enl_b) {
  orc_tree *t = (orc_tree *)calloc(1, sizeof(orc_tree));
  t->dim = dim;
  t->nchild = 1 << dim;
  /* max_n = max(max node degree, 256) (Quadtree_...F90:88-108,198-204) */
  int *deg = (int *)calloc((size_t)(nnodes_b > 0 ? nnodes_b : 1), sizeof(int));
  int maxdeg = 0;
  for (long e = 0; e < ne_b; e++)
    for (int j = 0; j < loc_b; j++) {
      int v = enl_b[e * loc_b + j];
      if (++deg[v] > maxdeg) maxdeg = deg[v];
    }
  free(deg);
  t->max_n = maxdeg > 256 ? maxdeg : 256;
  int root = tree_new_nodes(t, 1);
  orc_node *r = &t->nodes[root];
  r->n = 0;
  r->child0 = -1;
  memset(r->bb, 0, sizeof(r->bb));
  for (int d = 0; d < dim; d++) {
    r->bb[d] = pos_b[d];
    r->bb[3 + d] = pos_b[d];
  }
  for (long v = 1; v < nnodes_b; v++)
    for (int d = 0; d < dim; d++) {
      r->bb[d] = fmin(r->bb[d], pos_b[v * dim + d]);
      r->bb[3 + d] = fmax(r->bb[3 + d], pos_b[v * dim + d]);
    }
  long s = tree_new_slots(t);
  t->nodes[root].slot0 = s;
  for (long e = 0; e < ne_b; e++) {
    double bb[6] = {0, 0, 0, 0, 0, 0};
    element_bbox(dim, loc_b, pos_b, enl_b + e * loc_b, bb);
    tree_add(t, root, (int)e, bb);
  }
  return t;
}

void orc_tree_free(orc_tree *t) {
  if (!t) return;
  free(t->nodes);
  free(t->slot_ele);
  free(t->slot_bb);
  free(t);
}

static void tree_query_node(const orc_tree *t, int node, const double *qb,
                            unsigned char *seen, int *out, int *nout) {
  const orc_node *nd = &t->nodes[node];
  if (nd->n == 0) return;
  if (!boxes_strict(t->dim, nd->bb, qb)) return;
  if (nd->n <= t->max_n) {
    for (int i = 0; i < nd->n; i++) {
      long s = nd->slot0 + i;
      int e = t->slot_ele[s];
      if (!seen[e] && boxes_strict(t->dim, t->slot_bb + 6 * s, qb)) {
        out[(*nout)++] = e;
        seen[e] = 1;
      }
    }
  } else {
    for (int k = 0; k < t->nchild; k++) tree_query_node(t, nd->child0 + k, qb, seen, out, nout);
  }
}

int orc_tree_query(const orc_tree *t, int loc_a, const double *coords_a,
                   unsigned char *seen, int *out) {
  double qb[6] = {0, 0, 0, 0, 0, 0};
  coords_bbox(t->dim, loc_a, coords_a, qb);
  int n = 0;
  tree_query_node(t, 0, qb, seen, out, &n);
  for (int i = 0; i < n; i++) seen[out[i]] = 0;
  return n;
}

long orc_tree_finder(int dim, long nnodes_b,
                     long ne_a, int loc_a, const double *pos_a, const int *enl_a,
                     long ne_b, int loc_b, const double *pos_b, const int *enl_b,
                     long *indptr, int *indices) {
  indptr[0] = 0;
  if (ne_a == 0 || ne_b == 0) {
    for (long a = 0; a < ne_a; a++) indptr[a + 1] = 0;
    return 0;
  }
  orc_tree *t = orc_tree_build(dim, nnodes_b, pos_b, ne_b, loc_b, enl_b);
  unsigned char *seen = (unsigned char *)calloc((size_t)ne_b, 1);
  int *buf = (int *)malloc(sizeof(int) * (size_t)ne_b);
  long nnz = 0;
  double coords[4 * 3];
  for (long a = 0; a < ne_a; a++) {
    for (int j = 0; j < loc_a; j++)
      for (int d = 0; d < dim; d++) coords[j * dim + d] = pos_a[(long)enl_a[a * loc_a + j] * dim + d];
    int n = orc_tree_query(t, loc_a, coords, seen, buf);
    if (indices) memcpy(indices + nnz, buf, sizeof(int) * (size_t)n);
    nnz += n;
    indptr[a + 1] = nnz;
  }
  free(buf);
  free(seen);
  orc_tree_free(t);
  return nnz;
}

/* ------------------------------------------------------------------------ */
/* Pair clipping, mass and drivers                                           */
/* ------------------------------------------------------------------------ */

static void gather(int dim, const double *pos, const int *enl, long e, double *out) {
  int loc = dim + 1;
  for (int j = 0; j < loc; j++)
    for (int d = 0; d < dim; d++) out[j * dim + d] = pos[(long)enl[e * loc + j] * dim + d];
}

static double measure_real(int dim, const double *el) {
  return dim == 2 ? orc_triangle_area(el) : orc_tetrahedron_volume_real(el);
}

static int clip_pair(int dim, const double *ea, const double *eb, double *buf) {
  return dim == 2 ? orc_intersect_tris(ea, eb, buf) : orc_intersect_tets(ea, eb, buf);
}

typedef struct {
  int dim;
  long p0, p1;
  const int *pair_a, *pair_b;
  const double *pos_a, *pos_b;
  const int *enl_a, *enl_b;
  int *counts;
  double *measure;
  int cap;
  double *elems;
  double *blocks;
} clip_job;

static void *clip_worker(void *arg) {
  clip_job *j = (clip_job *)arg;
  int dim = j->dim, esz = dim == 2 ? 6 : 12;
  double ea[12], eb[12];
  double buf[ORC_TET_BUF_SIZE * 12];
  for (long p = j->p0; p < j->p1; p++) {
    gather(dim, j->pos_a, j->enl_a, j->pair_a[p], ea);
    gather(dim, j->pos_b, j->enl_b, j->pair_b[p], eb);
    int n = clip_pair(dim, ea, eb, buf);
    if (j->counts) j->counts[p] = n;
    if (j->measure) {
      double s = 0.0;
      for (int k = 0; k < n; k++) s = s + measure_real(dim, buf + esz * k);
      F(n);
      j->measure[p] = s;
    }
    if (j->elems) {
      int m = n < j->cap ? n : j->cap;
      memcpy(j->elems + (size_t)p * j->cap * esz, buf, sizeof(double) * (size_t)m * esz);
    }
  }
  return NULL;
}

static void run_jobs(clip_job *base, long npairs, int nthreads, void *(*fn)(void *)) {
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  pthread_t th[256];
  clip_job jobs[256];
  for (int t = 0; t < nthreads; t++) {
    jobs[t] = *base;
    jobs[t].p0 = npairs * t / nthreads;
    jobs[t].p1 = npairs * (t + 1) / nthreads;
  }
  if (nthreads == 1) {
    fn(&jobs[0]);
    return;
  }
  for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, fn, &jobs[t]);
  for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
}

void orc_clip_pairs(int dim, long npairs, const int *pair_a, const int *pair_b,
                    const double *pos_a, const int *enl_a,
                    const double *pos_b, const int *enl_b,
                    int *counts, double *measure, int cap, double *elems,
                    int nthreads) {
  clip_job base;
  memset(&base, 0, sizeof(base));
  base.dim = dim;
  base.pair_a = pair_a;
  base.pair_b = pair_b;
  base.pos_a = pos_a;
  base.pos_b = pos_b;
  base.enl_a = enl_a;
  base.enl_b = enl_b;
  base.counts = counts;
  base.measure = measure;
  base.cap = cap;
  base.elems = elems;
  run_jobs(&base, npairs, nthreads, clip_worker);
}

/* P1 basis functions of a triangle at x: inverse-Jacobian closed form
 * (test_parallel_p1_inner_product_2d.F90:205-228). */
static void basis_p1_2d(const double *c, const double *x, double *fns) {
  F(21);
  double e1x = c[2] - c[0], e1y = c[3] - c[1];
  double e2x = c[4] - c[0], e2y = c[5] - c[1];
  double j11 = e2y, j12 = -e2x, j21 = -e1y, j22 = e1x;
  double det = j11 * j22 - j12 * j21;
  j11 = j11 / det;
  j12 = j12 / det;
  j21 = j21 / det;
  j22 = j22 / det;
  double dx = x[0] - c[0], dy = x[1] - c[1];
  double f2 = j11 * dx;
  f2 = f2 + j12 * dy;
  double f3 = j21 * dx;
  f3 = f3 + j22 * dy;
  fns[1] = f2;
  fns[2] = f3;
  fns[0] = 1.0 - f2 - f3;
}

/* P1 basis functions of a tet at x: Gaussian elimination with partial
 * pivoting on [edges | x - c1] (test_parallel_p1_inner_product_3d.F90:227-265). */
static void basis_p1_3d(const double *c, const double *x, double *fns) {
  F(75);
  K(9);
  double m[3][4];
  for (int r = 0; r < 3; r++) {
    for (int k = 0; k < 3; k++) m[r][k] = c[3 * (k + 1) + r] - c[r];
    m[r][3] = x[r] - c[r];
  }
  for (int i = 0; i < 3; i++) {
    int j = i;
    double best = fabs(m[i][i]);
    for (int r = i + 1; r < 3; r++)
      if (fabs(m[r][i]) > best) {
        best = fabs(m[r][i]);
        j = r;
      }
    if (j != i)
      for (int k = 0; k < 4; k++) {
        double tt = m[i][k];
        m[i][k] = m[j][k];
        m[j][k] = tt;
      }
    double piv = m[i][i];
    for (int k = 0; k < 4; k++) m[i][k] = m[i][k] / piv;
    for (int r = i + 1; r < 3; r++) {
      double f = m[r][i];
      for (int k = 0; k < 4; k++) m[r][k] = m[r][k] - f * m[i][k];
    }
  }
  for (int i = 2; i >= 0; i--)
    for (int r = i - 1; r >= 0; r--) {
      double f = m[r][i];
      for (int k = 0; k < 4; k++) m[r][k] = m[r][k] - f * m[i][k];
    }
  fns[1] = m[0][3];
  fns[2] = m[1][3];
  fns[3] = m[2][3];
  fns[0] = 1.0 - fns[1] - fns[2] - fns[3];
}

/* Quadrature rules of the reference tests: Dunavant degree 2 (3 points,
 * barycentric (4,1,1)/6, weights 1/3) and Stroud degree 2 (4 points). */
static const double kQ2[3][3] = {{4.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0},
                                 {1.0 / 6.0, 4.0 / 6.0, 1.0 / 6.0},
                                 {1.0 / 6.0, 1.0 / 6.0, 4.0 / 6.0}};
#define QA 5.8541019662496852e-1
#define QB 1.3819660112501050e-1
static const double kQ3[4][4] = {{QA, QB, QB, QB}, {QB, QA, QB, QB}, {QB, QB, QA, QB}, {QB, QB, QB, QA}};

static void quad_point(int dim, int q, const double *el, double *x) {
  int loc = dim + 1;
  F(2 * dim * loc);
  for (int d = 0; d < dim; d++) {
    double s = 0.0;
    for (int v = 0; v < loc; v++) {
      double w = dim == 2 ? kQ2[q][v] : kQ3[q][v];
      s = s + w * el[v * dim + d];
    }
    x[d] = s;
  }
}

static void *mass_worker(void *arg) {
  clip_job *j = (clip_job *)arg;
  int dim = j->dim, loc = dim + 1, esz = dim == 2 ? 6 : 12, nq = dim == 2 ? 3 : 4;
  double wq = dim == 2 ? 1.0 / 3.0 : 1.0 / 4.0;
  double ea[12], eb[12], buf[ORC_TET_BUF_SIZE * 12];
  for (long p = j->p0; p < j->p1; p++) {
    gather(dim, j->pos_a, j->enl_a, j->pair_a[p], ea);
    gather(dim, j->pos_b, j->enl_b, j->pair_b[p], eb);
    int n = clip_pair(dim, ea, eb, buf);
    double *blk = j->blocks + (size_t)p * loc * loc;
    for (int k = 0; k < loc * loc; k++) blk[k] = 0.0;
    for (int c = 0; c < n; c++) {
      const double *el = buf + esz * c;
      double meas = measure_real(dim, el);
      for (int q = 0; q < nq; q++) {
        double x[3], fa[4], fb[4];
        quad_point(dim, q, el, x);
        if (dim == 2) {
          basis_p1_2d(ea, x, fa);
          basis_p1_2d(eb, x, fb);
        } else {
          basis_p1_3d(ea, x, fa);
          basis_p1_3d(eb, x, fb);
        }
        double wm = wq * meas;
        F(1 + 3 * loc * loc);
        for (int a = 0; a < loc; a++)
          for (int b = 0; b < loc; b++) blk[a * loc + b] = blk[a * loc + b] + wm * fa[a] * fb[b];
      }
    }
  }
  return NULL;
}

void orc_mass_pairs(int dim, long npairs, const int *pair_a, const int *pair_b,
                    const double *pos_a, const int *enl_a,
                    const double *pos_b, const int *enl_b,
                    double *blocks, int nthreads) {
  clip_job base;
  memset(&base, 0, sizeof(base));
  base.dim = dim;
  base.pair_a = pair_a;
  base.pair_b = pair_b;
  base.pos_a = pos_a;
  base.pos_b = pos_b;
  base.enl_a = enl_a;
  base.enl_b = enl_b;
  base.blocks = blocks;
  run_jobs(&base, npairs, nthreads, mass_worker);
}

void orc_inner_product(int dim, long npairs, const int *pair_a, const int *pair_b,
                       const double *pos_a, const int *enl_a, const double *field_a,
                       const double *pos_b, const int *enl_b, const double *field_b,
                       double *measure_out, double *integral_out) {
  int loc = dim + 1, esz = dim == 2 ? 6 : 12, nq = dim == 2 ? 3 : 4;
  double wq = dim == 2 ? 1.0 / 3.0 : 1.0 / 4.0;
  double ea[12], eb[12], buf[ORC_TET_BUF_SIZE * 12];
  double area = 0.0, integral = 0.0;
  orc_ksum kv, ki;
  orc_ksum_init(&kv);
  orc_ksum_init(&ki);
  for (long p = 0; p < npairs; p++) {
    long a = pair_a[p], b = pair_b[p];
    gather(dim, pos_a, enl_a, a, ea);
    gather(dim, pos_b, enl_b, b, eb);
    int n = clip_pair(dim, ea, eb, buf);
    for (int c = 0; c < n; c++) {
      const double *el = buf + esz * c;
      double meas = measure_real(dim, el);
      if (dim == 2) area = area + meas;
      else orc_ksum_add(&kv, meas);
      for (int q = 0; q < nq; q++) {
        double x[3], fa[4], fb[4];
        quad_point(dim, q, el, x);
        if (dim == 2) {
          basis_p1_2d(eb, x, fb);
          basis_p1_2d(ea, x, fa);
        } else {
          basis_p1_3d(eb, x, fb);
          basis_p1_3d(ea, x, fa);
        }
        double ib = 0.0, ia = 0.0;
        for (int v = 0; v < loc; v++) ib = ib + fb[v] * field_b[enl_b[b * loc + v]];
        for (int v = 0; v < loc; v++) ia = ia + fa[v] * field_a[enl_a[a * loc + v]];
        double term = wq * meas * ib * ia;
        if (dim == 2) integral = integral + term;
        else orc_ksum_add(&ki, term);
      }
    }
  }
  *measure_out = dim == 2 ? area : orc_ksum_value(&kv);
  *integral_out = dim == 2 ? integral : orc_ksum_value(&ki);
}

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

typedef struct {
  int dim, finder, mass;
  long a0, a1;
  const double *pos_a, *pos_b;
  const int *enl_a, *enl_b;
  long ne_b;
  const orc_tree *tree;
  const double *bbs_b;
  long long stats[4];
  orc_ksum sum, fmg;
} sm_job;

/* The P1 x P1 mixed-mass block of one clipped pair (the reference tests'
 * quadrature, as mass_worker) and its contribution to f^T M g for the linear
 * fields f = x on A's nodes and g = y on B's. */
static void pair_mass(int dim, const double *ea, const double *eb, const double *buf, int n, orc_ksum *fmg) {
  int loc = dim + 1, esz = dim == 2 ? 6 : 12, nq = dim == 2 ? 3 : 4;
  double wq = dim == 2 ? 1.0 / 3.0 : 1.0 / 4.0;
  double blk[16];
  for (int k = 0; k < loc * loc; k++) blk[k] = 0.0;
  for (int c = 0; c < n; c++) {
    const double *el = buf + esz * c;
    double meas = measure_real(dim, el);
    for (int q = 0; q < nq; q++) {
      double x[3], fa[4], fb[4];
      quad_point(dim, q, el, x);
      if (dim == 2) {
        basis_p1_2d(ea, x, fa);
        basis_p1_2d(eb, x, fb);
      } else {
        basis_p1_3d(ea, x, fa);
        basis_p1_3d(eb, x, fb);
      }
      double wm = wq * meas;
      for (int a = 0; a < loc; a++)
        for (int b = 0; b < loc; b++) blk[a * loc + b] = blk[a * loc + b] + wm * fa[a] * fb[b];
    }
  }
  double s = 0.0;
  for (int a = 0; a < loc; a++)
    for (int b = 0; b < loc; b++) s = s + ea[a * dim] * blk[a * loc + b] * eb[b * dim + 1];
  orc_ksum_add(fmg, s);
}

static void *sm_worker(void *arg) {
  sm_job *j = (sm_job *)arg;
  int dim = j->dim, loc = dim + 1, esz = dim == 2 ? 6 : 12;
  unsigned char *seen = (unsigned char *)calloc((size_t)(j->ne_b > 0 ? j->ne_b : 1), 1);
  int *cand = (int *)malloc(sizeof(int) * (size_t)(j->ne_b > 0 ? j->ne_b : 1));
  double ea[12], eb[12], buf[ORC_TET_BUF_SIZE * 12];
  orc_ksum_init(&j->sum);
  orc_ksum_init(&j->fmg);
  memset(j->stats, 0, sizeof(j->stats));
  for (long a = j->a0; a < j->a1; a++) {
    gather(dim, j->pos_a, j->enl_a, a, ea);
    int nc = 0;
    if (j->finder == 0) {
      nc = orc_tree_query(j->tree, loc, ea, seen, cand);
    } else {
      double ba[6];
      coords_bbox(dim, loc, ea, ba);
      for (long b = 0; b < j->ne_b; b++) {
        int hit = j->finder == 2 ? boxes_closed(dim, ba, j->bbs_b + 6 * b)
                                 : boxes_strict(dim, ba, j->bbs_b + 6 * b);
        if (hit) cand[nc++] = (int)b;
      }
    }
    j->stats[0] += nc;
    for (int i = 0; i < nc; i++) {
      gather(dim, j->pos_b, j->enl_b, cand[i], eb);
      int n = clip_pair(dim, ea, eb, buf);
      if (n > 0) j->stats[1]++;
      j->stats[2] += n;
      if (n > j->stats[3]) j->stats[3] = n;
      for (int k = 0; k < n; k++) orc_ksum_add(&j->sum, measure_real(dim, buf + esz * k));
      if (j->mass && n > 0) pair_mass(dim, ea, eb, buf, n, &j->fmg);
    }
  }
  free(seen);
  free(cand);
  return NULL;
}

void orc_supermesh(int dim, int finder, int nthreads, int mass,
                   long nnodes_a, long ne_a, const double *pos_a, const int *enl_a,
                   long nnodes_b, long ne_b, const double *pos_b, const int *enl_b,
                   long long *stats, double *sum_out, double *times) {
  (void)nnodes_a;
  int loc = dim + 1;
  double t0 = now_s();
  orc_tree *tree = NULL;
  double *bbs = NULL;
  if (finder == 0 && ne_a > 0 && ne_b > 0) {
    tree = orc_tree_build(dim, nnodes_b, pos_b, ne_b, loc, enl_b);
  } else if (finder != 0) {
    bbs = (double *)malloc(sizeof(double) * 6 * (size_t)(ne_b > 0 ? ne_b : 1));
    for (long b = 0; b < ne_b; b++) element_bbox(dim, loc, pos_b, enl_b + b * loc, bbs + 6 * b);
  }
  double t1 = now_s();
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  sm_job *jobs = (sm_job *)calloc((size_t)nthreads, sizeof(sm_job));
  pthread_t th[256];
  for (int t = 0; t < nthreads; t++) {
    sm_job *j = &jobs[t];
    j->dim = dim;
    j->finder = finder;
    j->mass = mass;
    j->a0 = tree || finder != 0 ? ne_a * t / nthreads : 0;
    j->a1 = tree || finder != 0 ? ne_a * (t + 1) / nthreads : 0;
    j->pos_a = pos_a;
    j->pos_b = pos_b;
    j->enl_a = enl_a;
    j->enl_b = enl_b;
    j->ne_b = ne_b;
    j->tree = tree;
    j->bbs_b = bbs;
  }
  if (nthreads == 1) sm_worker(&jobs[0]);
  else {
    for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, sm_worker, &jobs[t]);
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
  }
  double t2 = now_s();
  memset(stats, 0, 4 * sizeof(long long));
  orc_ksum total, fmg;
  orc_ksum_init(&total);
  orc_ksum_init(&fmg);
  for (int t = 0; t < nthreads; t++) {
    for (int k = 0; k < 3; k++) stats[k] += jobs[t].stats[k];
    if (jobs[t].stats[3] > stats[3]) stats[3] = jobs[t].stats[3];
    orc_ksum_add(&total, jobs[t].sum.sum);
    orc_ksum_add(&total, jobs[t].sum.error);
    orc_ksum_add(&fmg, jobs[t].fmg.sum);
    orc_ksum_add(&fmg, jobs[t].fmg.error);
  }
  sum_out[0] = orc_ksum_value(&total);
  sum_out[1] = orc_ksum_value(&fmg);
  times[0] = t1 - t0;
  times[1] = t2 - t1;
  free(jobs);
  free(bbs);
  orc_tree_free(tree);
}

/* ------------------------------------------------------------------------ */
/* Mesh adjacency (Graphs.F90, C_bindings.F90:176-430) and the advancing    */
/* front finder (Intersection_Finder.F90:270-454).  0-based ids throughout. */
/* ------------------------------------------------------------------------ */

void orc_free(void *p) { free(p); }

/* mesh_nelist, Graphs.F90:76-116: count per node, then fill in element
 * order. */
void orc_mesh_nelist(long nnodes, long ne, int loc, const long *enl, long *indptr, long *indices) {
  long *n = calloc((size_t)nnodes + 1, sizeof(long));
  for (long i = 0; i < ne; i++)
    for (int j = 0; j < loc; j++) n[enl[i * loc + j]]++;
  indptr[0] = 0;
  for (long x = 0; x < nnodes; x++) indptr[x + 1] = indptr[x] + n[x];
  memset(n, 0, sizeof(long) * ((size_t)nnodes + 1));
  for (long i = 0; i < ne; i++)
    for (int j = 0; j < loc; j++) {
      long x = enl[i * loc + j];
      indices[indptr[x] + n[x]] = i;
      n[x]++;
    }
  free(n);
}

static int cmp_long(const void *a, const void *b) {
  long x = *(const long *)a, y = *(const long *)b;
  return x < y ? -1 : x > y;
}

/* mesh_nnlist, C_bindings.F90:215-255: per node the set (ascending, no
 * repeats) of nodes inserted by every local pair i < j of every element, in
 * both directions.  indices needs room for ne*loc*(loc-1); returns the
 * number of entries. */
long orc_mesh_nnlist(long nnodes, long ne, int loc, const long *enl, long *indptr, long *indices) {
  long *cap = calloc((size_t)nnodes + 1, sizeof(long));
  for (long e = 0; e < ne; e++)
    for (int i = 0; i < loc; i++) cap[enl[e * loc + i]] += loc - 1;
  long *start = malloc(sizeof(long) * ((size_t)nnodes + 1));
  start[0] = 0;
  for (long x = 0; x < nnodes; x++) start[x + 1] = start[x] + cap[x];
  long *buf = malloc(sizeof(long) * (size_t)(start[nnodes] > 0 ? start[nnodes] : 1));
  memset(cap, 0, sizeof(long) * ((size_t)nnodes + 1));
  for (long e = 0; e < ne; e++)
    for (int i = 0; i < loc; i++)
      for (int j = i + 1; j < loc; j++) {
        long a = enl[e * loc + i], b = enl[e * loc + j];
        buf[start[a] + cap[a]++] = b;
        buf[start[b] + cap[b]++] = a;
      }
  long m = 0;
  indptr[0] = 0;
  for (long x = 0; x < nnodes; x++) {
    long *v = buf + start[x];
    qsort(v, (size_t)cap[x], sizeof(long), cmp_long);
    for (long k = 0; k < cap[x]; k++)
      if (k == 0 || v[k] != v[k - 1]) indices[m++] = v[k];
    indptr[x + 1] = m;
  }
  free(buf);
  free(start);
  free(cap);
  return m;
}

/* mesh_eelist, Graphs.F90:118-187, as the release (NDEBUG) build runs it:
 * per element, walk its nodes' element lists, count shared nodes per other
 * element, list an element when its count reaches sloc, stop at nneigh.
 * indices needs room for ne*nneigh; returns the number of entries. */
long orc_mesh_eelist(long nnodes, long ne, int loc, long sloc, int nneigh, const long *enl, long *indptr,
                     long *indices) {
  long *nptr = malloc(sizeof(long) * ((size_t)nnodes + 1));
  long *nidx = malloc(sizeof(long) * (size_t)(ne * loc > 0 ? ne * loc : 1));
  orc_mesh_nelist(nnodes, ne, loc, enl, nptr, nidx);
  long *seen = calloc((size_t)(ne > 0 ? ne : 1), sizeof(long));
  long *seen_eles = malloc(sizeof(long) * (size_t)(ne > 0 ? ne : 1));
  long m = 0;
  indptr[0] = 0;
  for (long i = 0; i < ne; i++) {
    long nseen = 0, n = 0;
    for (int j = 0; j < loc && !(n == nneigh && n > 0); j++) {
      long x = enl[i * loc + j];
      for (long k = nptr[x]; k < nptr[x + 1]; k++) {
        long e = nidx[k];
        if (e == i) continue;
        seen[e]++;
        if (seen[e] == 1) seen_eles[nseen++] = e;
        if (seen[e] == sloc) {
          indices[m + n] = e;
          n++;
          if (n == nneigh) break;
        }
      }
      if (n == nneigh) break;
    }
    for (long k = 0; k < nseen; k++) seen[seen_eles[k]] = 0;
    m += n;
    indptr[i + 1] = m;
  }
  free(seen_eles);
  free(seen);
  free(nidx);
  free(nptr);
  return m;
}

/* transpose_sparsity, C_bindings.F90:397-430: column counts, then rows in
 * ascending order per column. */
void orc_transpose_sparsity(long nrows, long ncols, const long *indices, const long *ind_ptr, long *indices_t,
                            long *ind_ptr_t) {
  long *n = calloc((size_t)ncols + 1, sizeof(long));
  for (long r = 0; r < nrows; r++)
    for (long i = ind_ptr[r]; i < ind_ptr[r + 1]; i++) n[indices[i]]++;
  ind_ptr_t[0] = 0;
  for (long c = 0; c < ncols; c++) ind_ptr_t[c + 1] = ind_ptr_t[c] + n[c];
  memset(n, 0, sizeof(long) * ((size_t)ncols + 1));
  for (long r = 0; r < nrows; r++)
    for (long i = ind_ptr[r]; i < ind_ptr[r + 1]; i++) {
      long c = indices[i];
      indices_t[ind_ptr_t[c] + n[c]] = r;
      n[c]++;
    }
  free(n);
}

static int sloc_of(int dim, int loc) {
  if (loc == dim + 1) return loc - 1;
  if (loc == (1 << dim)) return loc / 2;
  return -1;
}

static int nneigh_of(int dim, int loc) {
  if (loc == dim + 1) return loc;
  if (loc == (1 << dim)) return 2 * dim;
  return -1;
}

/* Closed box test of the advancing front (Intersection_Finder.F90:429-450):
 * disjoint only when b.max < a.min or b.min > a.max on some axis. */
static int af_boxes_meet(int dim, const double *a, const double *b) {
  for (int d = 0; d < dim; d++)
    if (b[dim + d] < a[d] || b[d] > a[dim + d]) return 0;
  return 1;
}

static void af_bbox(int dim, int loc, const double *pos, const long *nodes, double *bb) {
  for (int d = 0; d < dim; d++) bb[d] = bb[dim + d] = pos[nodes[0] * dim + d];
  for (int v = 1; v < loc; v++)
    for (int d = 0; d < dim; d++) {
      double x = pos[nodes[v] * dim + d];
      if (x < bb[d]) bb[d] = x;
      if (x > bb[dim + d]) bb[dim + d] = x;
    }
}

/* advancing_front_intersection_finder_intersections
 * (Intersection_Finder.F90:270-454): the first A element of each connected
 * part of A is searched by brute force; every other A element, taken from a
 * stack of A's unvisited face neighbours, starts from the list of the element
 * that pushed it and advances through B's face neighbours while their boxes
 * meet.  Returns the number of pairs; *indices (malloc'd, free with orc_free)
 * holds the rows, indptr[ne_a + 1]. */
long orc_advancing_front_finder(int dim, long nn_a, long ne_a, int loc_a, const double *pos_a, const long *enl_a,
                                long nn_b, long ne_b, int loc_b, const double *pos_b, const long *enl_b,
                                long *indptr, long **indices) {
  long **rows = calloc((size_t)(ne_a > 0 ? ne_a : 1), sizeof(long *));
  long *nrow = calloc((size_t)(ne_a > 0 ? ne_a : 1), sizeof(long));
  if (ne_a > 0 && ne_b > 0) {
    int nb_b = nneigh_of(dim, loc_b), nb_a = nneigh_of(dim, loc_a);
    long *eptr_b = malloc(sizeof(long) * ((size_t)ne_b + 1));
    long *eidx_b = malloc(sizeof(long) * (size_t)(ne_b * nb_b + 1));
    long *eptr_a = malloc(sizeof(long) * ((size_t)ne_a + 1));
    long *eidx_a = malloc(sizeof(long) * (size_t)(ne_a * nb_a + 1));
    orc_mesh_eelist(nn_b, ne_b, loc_b, sloc_of(dim, loc_b), nb_b, enl_b, eptr_b, eidx_b);
    orc_mesh_eelist(nn_a, ne_a, loc_a, sloc_of(dim, loc_a), nb_a, enl_a, eptr_a, eidx_a);
    double *bb_b = malloc(sizeof(double) * 2 * dim * (size_t)ne_b);
    for (long e = 0; e < ne_b; e++) af_bbox(dim, loc_b, pos_b, enl_b + e * loc_b, bb_b + 2 * dim * e);
    char *seen_b = calloc((size_t)ne_b, 1), *seen_a = calloc((size_t)ne_a, 1);
    long *front_b = malloc(sizeof(long) * (size_t)ne_b), *ints = malloc(sizeof(long) * (size_t)ne_b);
    long *front_a = malloc(sizeof(long) * 2 * (size_t)ne_a);
    double bb_a[6];
    long seed = 0;
    while (1) {
      seen_a[seed] = 1;
      af_bbox(dim, loc_a, pos_a, enl_a + seed * loc_a, bb_a);
      long nints = 0;
      for (long b = 0; b < ne_b; b++)
        if (af_boxes_meet(dim, bb_a, bb_b + 2 * dim * b)) ints[nints++] = b;
      rows[seed] = malloc(sizeof(long) * (size_t)(nints > 0 ? nints : 1));
      memcpy(rows[seed], ints, sizeof(long) * (size_t)nints);
      nrow[seed] = nints;
      long nfa = 0;
      for (long k = eptr_a[seed]; k < eptr_a[seed + 1]; k++) {
        long na = eidx_a[k];
        if (!seen_a[na]) {
          front_a[2 * nfa] = na;
          front_a[2 * nfa + 1] = seed;
          nfa++;
          seen_a[na] = 1;
        }
      }
      while (nfa > 0) {
        nfa--;
        long a = front_a[2 * nfa], clue = front_a[2 * nfa + 1];
        af_bbox(dim, loc_a, pos_a, enl_a + a * loc_a, bb_a);
        long nfb = nrow[clue];
        memcpy(front_b, rows[clue], sizeof(long) * (size_t)nfb);
        for (long k = 0; k < nfb; k++) seen_b[front_b[k]] = 1;
        nints = 0;
        for (long i = 0; i < nfb; i++) {
          long b = front_b[i];
          if (af_boxes_meet(dim, bb_a, bb_b + 2 * dim * b)) {
            ints[nints++] = b;
            for (long k = eptr_b[b]; k < eptr_b[b + 1]; k++) {
              long nb = eidx_b[k];
              if (!seen_b[nb]) {
                front_b[nfb++] = nb;
                seen_b[nb] = 1;
              }
            }
          }
        }
        for (long k = 0; k < nfb; k++) seen_b[front_b[k]] = 0;
        rows[a] = malloc(sizeof(long) * (size_t)(nints > 0 ? nints : 1));
        memcpy(rows[a], ints, sizeof(long) * (size_t)nints);
        nrow[a] = nints;
        for (long k = eptr_a[a]; k < eptr_a[a + 1]; k++) {
          long na = eidx_a[k];
          if (!seen_a[na]) {
            front_a[2 * nfa] = na;
            front_a[2 * nfa + 1] = a;
            nfa++;
            seen_a[na] = 1;
          }
        }
      }
      while (seed < ne_a && seen_a[seed]) seed++;
      if (seed >= ne_a) break;
    }
    free(front_a);
    free(ints);
    free(front_b);
    free(seen_a);
    free(seen_b);
    free(bb_b);
    free(eidx_a);
    free(eptr_a);
    free(eidx_b);
    free(eptr_b);
  }
  indptr[0] = 0;
  for (long a = 0; a < ne_a; a++) indptr[a + 1] = indptr[a] + nrow[a];
  long total = indptr[ne_a];
  long *out = malloc(sizeof(long) * (size_t)(total > 0 ? total : 1));
  for (long a = 0; a < ne_a; a++) {
    if (nrow[a]) memcpy(out + indptr[a], rows[a], sizeof(long) * (size_t)nrow[a]);
    free(rows[a]);
  }
  free(rows);
  free(nrow);
  *indices = out;
  return total;
}
