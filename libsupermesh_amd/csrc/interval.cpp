// interval.cpp -- the 1D entry points of the C ABI, on the host.
//
// intersect_intervals / interval_size (Interval_Intersection.F90:45-110) are
// a handful of compares, and the sort finder (Intersection_Finder.F90:811-921)
// is a sequential sweep whose output depends on the sweep order; neither is
// data-parallel work, so both run here rather than as device launches.
#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/libsupermesh_gpu.h"
#include "lsm_geometry.hpp"
#include "lsm_runtime.hpp"

namespace lsm {

namespace {

// Fortran MIN / MAX of two values (the first argument on ties).
inline double min2(double a, double b) { return b < a ? b : a; }
inline double max2(double a, double b) { return b > a ? b : a; }

// Element numbers in ascending order of left end point, ties by element
// number (a stable sort).  The reference merge-sorts
// (Intersection_Finder.F90:869-917) with a base case that is the identity on
// sorted blocks of at most 4 entries; on unsorted blocks it repeats indices
// and its result is undefined.  Wherever it is defined it is this order.
std::vector<int> sorted_by_left(const std::vector<double>& left) {
  std::vector<int> idx(left.size());
  for (size_t i = 0; i < idx.size(); i++) idx[i] = (int)i;
  std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return left[x] < left[y]; });
  return idx;
}

struct Interval {
  double lo, hi;
};

std::vector<Interval> intervals(long nnodes, long ne, const double* pos, const long* enl) {
  std::vector<Interval> out((size_t)ne);
  for (long e = 0; e < ne; e++) {
    const long i0 = enl[2 * e], i1 = enl[2 * e + 1];
    LSM_REQUIRE(i0 >= 0 && i0 < nnodes && i1 >= 0 && i1 < nnodes, LIBSUPERMESH_GPU_ERROR_INVALID,
                "element node index out of range");
    out[(size_t)e] = {min2(pos[i0], pos[i1]), max2(pos[i0], pos[i1])};
  }
  return out;
}

}  // namespace

double interval_size(const double* v) { return std::fabs(v[1] - v[0]); }

int intersect_intervals(const double* a, const double* b, double* c) {
  const double min_a = min2(a[0], a[1]), max_a = max2(a[0], a[1]);
  const double min_b = min2(b[0], b[1]), max_b = max2(b[0], b[1]);
  if (max_b <= min_a || min_b >= max_a) return 0;
  c[0] = max2(min_a, min_b);
  c[1] = min2(max_a, max_b);
  const double tol = 10.0 * min2(fspacing(interval_size(a)), fspacing(interval_size(b)));
  return interval_size(c) < tol ? 0 : 1;
}

void sort_finder(long nnodes_a, long ne_a, const double* pos_a, const long* enl_a, long nnodes_b,
                 long ne_b, const double* pos_b, const long* enl_b, std::vector<long>& indptr,
                 std::vector<int>& indices) {
  LSM_REQUIRE(ne_a >= 0 && ne_b >= 0 && ne_a < (1L << 31) && ne_b < (1L << 31),
              LIBSUPERMESH_GPU_ERROR_INVALID, "invalid element count");
  const std::vector<Interval> A = intervals(nnodes_a, ne_a, pos_a, enl_a);
  const std::vector<Interval> B = intervals(nnodes_b, ne_b, pos_b, enl_b);
  std::vector<double> left_a((size_t)ne_a), left_b((size_t)ne_b);
  for (long e = 0; e < ne_a; e++) left_a[e] = A[e].lo;
  for (long e = 0; e < ne_b; e++) left_b[e] = B[e].lo;
  const std::vector<int> ia = sorted_by_left(left_a), ib = sorted_by_left(left_b);

  // Sweep A in sorted order with one cursor into sorted B; an A element that
  // found candidates hands its last one on to the next A element.
  std::vector<long> count((size_t)ne_a, 0), start((size_t)ne_a, 0);
  std::vector<int> visit;
  long j = 0;
  for (long i = 0; i < ne_a; i++) {
    const int ea = ia[i];
    const Interval a = A[ea];
    start[ea] = (long)visit.size();
    long n = 0;
    while (j < ne_b) {
      const Interval b = B[ib[j]];
      if (b.hi <= a.lo) {
        j++;
      } else if (b.lo >= a.hi) {
        break;
      } else {
        visit.push_back(ib[j]);
        n++;
        j++;
      }
    }
    if (n > 0) j--;
    count[ea] = n;
  }
  indptr.assign((size_t)ne_a + 1, 0);
  for (long e = 0; e < ne_a; e++) indptr[e + 1] = indptr[e] + count[e];
  indices.resize(visit.size());
  for (long e = 0; e < ne_a; e++)
    for (long k = 0; k < count[e]; k++) indices[indptr[e] + k] = visit[start[e] + k];
}

}  // namespace lsm
