/*
 * oracle.h -- CPU restatement of libsupermesh's supermeshing hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the GPU library is compared
 * against, and the CPU baseline timed beside it.  Only tests/, the smoke check
 * in __graft_entry__.py and the cpu_baseline leg of bench.py may load it.  The
 * product library (libsupermesh_amd) never links, loads or calls it.
 *
 * Layouts follow the reference's C ABI (include/libsupermesh-c.h):
 *   positions  double[nnodes][dim]      (Fortran dim x nnodes, column-major)
 *   enlist     int[nelements][loc]      0-based here (the C ABI passes 0-based
 *                                       long; the Fortran core uses 1-based)
 *   tri        double[3][2], tet double[4][3]
 *
 * Every routine is re-entrant: the reference keeps its scratch in module-level
 * `save` variables (src/Triangle_Intersection.F90:101-102,
 * src/Tetrahedron_Intersection.F90:102-103); here all scratch is on the stack
 * or owned by the caller, so the driver may run one thread per core.
 *
 * Parity status of each function (what pins it) is listed in oracle.c's
 * header comment.
 */
#ifndef LIBSUPERMESH_ORACLE_H
#define LIBSUPERMESH_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_TRI_BUF_SIZE 22
#define ORC_TET_BUF_SIZE 81

/* Fortran intrinsics as the reference's compiler evaluates them. */
double orc_spacing(double x);
double orc_epsilon(void);
double orc_norm2_3(const double v[3]);
double orc_dot3(const double a[3], const double b[3]);

/* Operation counts of the calling thread since the last reset:
 * out[0] FP64 add/sub/mul/div/sqrt, out[1] FP64 compare/abs/min/max.
 * Non-zero only in liboracle_count.so (built with -DORC_COUNT). */
void orc_op_counts(unsigned long long *out, int reset);

/* Element measures. */
double orc_triangle_area(const double tri[6]);
double orc_tetrahedron_volume_real(const double tet[12]);
double orc_tetrahedron_volume_tet(const double tet[12]);

/* Per-pair clippers.  Output buffers hold ORC_TRI_BUF_SIZE / ORC_TET_BUF_SIZE
 * elements.  Return the number of output elements. */
int orc_intersect_tris(const double tri_a[6], const double tri_b[6], double *tris_c);
int orc_intersect_tets(const double tet_a[12], const double tet_b[12], double *tets_c);
/* Plane set of a tet, as get_planes_tet: normals[4][3], c[4]. */
void orc_get_planes_tet(const double tet[12], double normals[12], double c[4]);

/* Mixed element types and 1D (Supermesh.F90:211-334, Interval_Intersection.F90:45-110).
 * Buffers: tri_quad 46 triangles, quads 62, tet_hex 729 tets, hexes 5*729. */
#define ORC_TRI_QUAD_BUF 46
#define ORC_QUADS_BUF 62
#define ORC_TET_HEX_BUF 729
int orc_intersect_tri_quad(const double tri_a[6], const double quad_b[8], double *tris_c);
int orc_intersect_quads(const double quad_a[8], const double quad_b[8], double *tris_c);
int orc_intersect_tet_hex(const double tet_a[12], const double hex_b[24], double *tets_c);
int orc_intersect_hexes(const double hex_a[24], const double hex_b[24], double *tets_c);
/* Hex face planes in Gmsh order, as intersect_tet_hex builds them: normals[6][3], c[6]. */
void orc_hex_planes(const double hex[24], double normals[18], double c[6]);
double orc_interval_size(const double interval[2]);
int orc_intersect_intervals(const double a[2], const double b[2], double c[2]);
/* 1D sort finder (Intersection_Finder.F90:811-921), enlists 0-based, loc 2.
 * Fills indptr[ne_a+1] and, when indices is non-NULL, indices (capacity cap);
 * returns the total count. */
long orc_sort_finder(long ne_a, const double *pos_a, const int *enl_a,
                     long ne_b, const double *pos_b, const int *enl_b,
                     long *indptr, int *indices, long cap);

/* Kahan compensated sum (src/Precision.F90:45-113). */
typedef struct { double sum, error; } orc_ksum;
void orc_ksum_init(orc_ksum *cs);
void orc_ksum_add(orc_ksum *cs, double x);
double orc_ksum_value(const orc_ksum *cs);

/* Candidate finders.  Output is CSR over A elements (indptr[ne_a+1]); when
 * indices is NULL only indptr is filled.  closed=0: strict AABB overlap (tree
 * and brute-force finders); closed=1: closed intervals (R-tree finder).
 * Brute force lists B ids ascending; the tree lists them in traversal order. */
long orc_brute_force_finder(int dim, int closed,
                            long ne_a, int loc_a, const double *pos_a, const int *enl_a,
                            long ne_b, int loc_b, const double *pos_b, const int *enl_b,
                            long *indptr, int *indices);

typedef struct orc_tree orc_tree;
/* The brute-force set through a uniform grid of cells (any size, threaded).
 * Rows ascending; *indices is allocated here (free with orc_free). */
long orc_grid_finder(int dim, int closed, int nthreads,
                     long ne_a, int loc_a, const double *pos_a, const int *enl_a,
                     long ne_b, int loc_b, const double *pos_b, const int *enl_b,
                     long *indptr, int **indices);
orc_tree *orc_tree_build(int dim, long nnodes_b, const double *pos_b,
                         long ne_b, int loc_b, const int *enl_b);
void orc_tree_free(orc_tree *t);
/* Query one A element (its vertex coordinates, loc x dim) into out (capacity
 * >= ne_b); seen is a caller-owned zeroed byte array of ne_b entries, left
 * zeroed on return.  Returns the number of B elements found. */
int orc_tree_query(const orc_tree *t, int loc_a, const double *coords_a,
                   unsigned char *seen, int *out);
long orc_tree_finder(int dim, long nnodes_b,
                     long ne_a, int loc_a, const double *pos_a, const int *enl_a,
                     long ne_b, int loc_b, const double *pos_b, const int *enl_b,
                     long *indptr, int *indices);

/* Clip a list of candidate pairs.  counts[p] = number of output elements;
 * measure[p] = sum, in output order, of the reference's Sigma measure
 * (triangle_area / tetrahedron_volume_real) over them.  Either array may be
 * NULL.  If elems is non-NULL it receives up to `cap` elements per pair
 * (cap*6 or cap*12 doubles per pair). */
void orc_clip_pairs(int dim, long npairs, const int *pair_a, const int *pair_b,
                    const double *pos_a, const int *enl_a,
                    const double *pos_b, const int *enl_b,
                    int *counts, double *measure, int cap, double *elems,
                    int nthreads);

/* P1 x P1 mixed mass: for every pair, the local block
 * block[p][i][j] = sum_c sum_q w_q |c| phi^A_i(x_q) phi^B_j(x_q)
 * (3x3 in 2D with the Dunavant 3-point rule, 4x4 in 3D with the 4-point
 * Stroud rule), with the supermesh elements c of the pair. */
void orc_mass_pairs(int dim, long npairs, const int *pair_a, const int *pair_b,
                    const double *pos_a, const int *enl_a,
                    const double *pos_b, const int *enl_b,
                    double *blocks, int nthreads);

/* f^T M g evaluated like the reference's intersection_calculation
 * callbacks, with f a nodal field on A and g on B (Kahan sum in 3D, plain in
 * 2D), plus the Sigma measure.  pairs iterate in the given order. */
void orc_inner_product(int dim, long npairs, const int *pair_a, const int *pair_b,
                       const double *pos_a, const int *enl_a, const double *field_a,
                       const double *pos_b, const int *enl_b, const double *field_b,
                       double *measure_out, double *integral_out);

/* Whole supermesh run (finder + clip + Sigma), the reference's serial caller
 * loop.  finder: 0 tree (strict), 1 brute force strict, 2 brute force closed.
 * Threads split the A elements into contiguous ranges; the tree is built once
 * (serially, as the reference does) and shared read-only.
 * With mass, each non-empty pair's P1 x P1 mixed-mass block is integrated
 * too (the reference tests' quadrature) and contracted with f = x on A's
 * nodes and g = y on B's.
 * stats[0]=candidates, [1]=non-empty pairs, [2]=output elements,
 * [3]=max per pair.  sum_out[0] = Sigma measures, [1] = f^T M g (mass).
 * times[0]=build seconds, [1]=query+clip(+mass) seconds. */
void orc_supermesh(int dim, int finder, int nthreads, int mass,
                   long nnodes_a, long ne_a, const double *pos_a, const int *enl_a,
                   long nnodes_b, long ne_b, const double *pos_b, const int *enl_b,
                   long long *stats, double *sum_out, double *times);

/* Mesh adjacency and the advancing front finder; ids 0-based, CSR from 0. */
void orc_free(void *p);
/* mesh_nelist (Graphs.F90:76-116): indptr[nnodes+1], indices[ne*loc]. */
void orc_mesh_nelist(long nnodes, long ne, int loc, const long *enl, long *indptr, long *indices);
/* mesh_nnlist (C_bindings.F90:215-255): indices needs ne*loc*(loc-1). */
long orc_mesh_nnlist(long nnodes, long ne, int loc, const long *enl, long *indptr, long *indices);
/* mesh_eelist (Graphs.F90:118-187), release build: indices needs ne*nneigh. */
long orc_mesh_eelist(long nnodes, long ne, int loc, long sloc, int nneigh, const long *enl, long *indptr,
                     long *indices);
/* transpose_sparsity (C_bindings.F90:397-430). */
void orc_transpose_sparsity(long nrows, long ncols, const long *indices, const long *ind_ptr, long *indices_t,
                            long *ind_ptr_t);
/* advancing_front_intersection_finder (Intersection_Finder.F90:270-454). */
long orc_advancing_front_finder(int dim, long nn_a, long ne_a, int loc_a, const double *pos_a, const long *enl_a,
                                long nn_b, long ne_b, int loc_b, const double *pos_b, const long *enl_b,
                                long *indptr, long **indices);

#ifdef __cplusplus
}
#endif
#endif
