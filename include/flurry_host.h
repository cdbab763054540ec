/*
 * flurry_host.h -- host-side set-up for the FR hot path (libflurry_host.so).
 *
 * This library builds what libflurry_hip consumes: the FR operator matrices,
 * a mesh with its face connectivity in FlurryPP's face order, and the metric
 * terms at solution and flux points.  It restates, in O(n), the host work that
 * FlurryPP does before its time loop:
 *   operators      src/operators.cpp:112-141,596-658,715-824,915-1001,1393-1506
 *   points/weights src/points.cpp:67-162,198-538; polynomials src/polynomials.cpp:39-151,423-493
 *   mesh           src/geo.cpp:1477-1762 (createMesh), 1021-1475 (readGmsh)
 *   connectivity   src/geo.cpp:221-570 (processConn2D/3D), 1764-1852 (periodic),
 *                  799-1019 (setupElesFaces), 2078-2225 (compareOrientation)
 *   metrics        src/solver.cpp:165-317 (setupGeometry), 1248-1313, 1587-1694
 *
 * Conventions: every function returns 0 on success and a nonzero code on
 * failure; fhost_last_error() then holds a message.  Arrays are row-major in
 * FlurryPP's layout ([point][element][dim] etc.).  Nothing here touches a GPU.
 */
#ifndef FLURRY_HOST_H
#define FLURRY_HOST_H

#ifdef __cplusplus
extern "C" {
#endif

/* Boundary-condition numbers: include/global.hpp:113-128 (enum BC_TYPE). */
enum {
  FHOST_BC_NONE = -1, FHOST_BC_PERIODIC = 0, FHOST_BC_CHAR = 1, FHOST_BC_SUP_IN = 2,
  FHOST_BC_SUP_OUT = 3, FHOST_BC_SUB_IN = 4, FHOST_BC_SUB_OUT = 5,
  FHOST_BC_SLIP_WALL = 8, FHOST_BC_ISOTHERMAL_NOSLIP = 9, FHOST_BC_ADIABATIC_NOSLIP = 10,
  FHOST_BC_OVERSET = 11, FHOST_BC_SYMMETRY = 12, FHOST_BC_ISOTHERMAL_NOSLIP_MOVING = 13
};

/* Solution-point families: src/points.cpp:163-392 (getPts1D). */
enum { FHOST_PTS_LEGENDRE = 0, FHOST_PTS_LOBATTO = 1, FHOST_PTS_EQUIDISTANT = 2 };

typedef struct fhost_mesh fhost_mesh;

const char* fhost_last_error(void);

/* Map a boundary-condition name ("periodic", "char", "slip_wall", ...) to its
 * number, as bcStr2Num does (src/global.cpp:43-57).  Returns -2 if unknown. */
int fhost_bc_from_name(const char* name);

/* ---- operators (one element type and order) ---------------------------- */

/* Sizes for nDims in {2,3} and order P: nSpts=(P+1)^D, nFpts=2D(P+1)^(D-1). */
int fhost_sizes(int nDims, int order, int* nSpts, int* nFpts, int* nFptsFace);

/* 1-D point locations (P+1 values), src/points.cpp:163-392. */
int fhost_pts1d(int ptsType, int order, double* out);
/* 1-D Gauss weights (P+1 values), src/points.cpp:425-538. */
int fhost_weights1d(int order, double* out);

/* Reference-element point locations: loc_spts [nSpts][D], loc_fpts [nFpts][D]
 * (src/points.cpp:67-162), and the parent-domain face normals tNorm [nFpts][D]
 * (src/solver.cpp:206-316). */
int fhost_ref_points(int nDims, int order, int ptsType,
                     double* loc_spts, double* loc_fpts, double* tnorm_fpts);

/* FR operators, row-major:
 *   spts_to_fpts [F][S]    opp_spts_to_fpts      (src/operators.cpp:112-141)
 *   grad_spts    [D][S][S] opp_grad_spts         (src/operators.cpp:596-658)
 *   extrap_fn    [D][F][S] opp_extrapolateFn     (src/operators.cpp:715-797)
 *   correction   [S][F]    opp_correction        (src/operators.cpp:800-824)
 *   correct_u    [D][S][F] opp_correctU          (src/operators.cpp:915-1001)
 *   w_spts       [S]       getQptWeights(P,D)    (src/points.cpp:394-423)
 *   w_face       [F/nFaces] getQptWeights(P,D-1)
 * vcjh: 0 DG, 1 SD, 2 HU, 3 C+ (src/polynomials.cpp:439-477).
 * Any output pointer may be NULL. */
int fhost_operators(int nDims, int order, int ptsType, int vcjh,
                    double* spts_to_fpts, double* grad_spts, double* extrap_fn,
                    double* correction, double* correct_u,
                    double* w_spts, double* w_face);

/* P-multigrid transfer operators of one order (oper::setupPMG,
 * src/operators.cpp:1553-1634; opp_prolong / opp_restrict,
 * include/operators.hpp:132-133), row-major:
 *   prolong     [(order+2)^nDims][nSpts]  order -> order+1
 *   restrict_op [order^nDims][nSpts]      order -> order-1 (untouched at order 0)
 * Either pointer may be NULL. */
int fhost_pmg_operators(int nDims, int order, int ptsType, double* prolong, double* restrict_op);

/* ---- meshes ------------------------------------------------------------- */

/* Cartesian mesh with createMesh's vertex, element and boundary numbering
 * (src/geo.cpp:1477-1762).  lo/hi: nDims values.  bc_names: 4 names in 2-D
 * (bottom, right, top, left) or 6 in 3-D (+ front, back).  Periodic lengths
 * are set to hi-lo, as createMesh does. */
int fhost_mesh_cartesian(int nDims, int nx, int ny, int nz,
                         const double* lo, const double* hi,
                         const char* const* bc_names, fhost_mesh** out);

/* Mesh from arrays: xv [nVerts][nDims], c2v [nEles][2^nDims] (linear quads /
 * hexes, FlurryPP node order), nBounds boundaries each with a BC number and a
 * vertex list (CSR: bnd_ptr[nBounds+1], bnd_pts).  Boundaries are kept in the
 * order given. */
int fhost_mesh_from_arrays(int nDims, int nVerts, const double* xv, int nEles,
                           const int* c2v, int nBounds, const int* bc_list,
                           const int* bnd_ptr, const int* bnd_pts,
                           fhost_mesh** out);

/* Gmsh 2.2 ASCII reader (src/geo.cpp:1021-1475): Lagrange quadrangles of
 * 4, 9, 16, 25 or 36 nodes (Gmsh types 3, 10, 36, 37, 38), Lagrange
 * hexahedra of 8, 27, 64, 125 or 216 nodes (5, 12, 92, 93, 94) and the
 * 20-node serendipity hexahedron (17), one element type per mesh, with their
 * boundary lines/quadrangles of any of those orders.  The 8-node
 * quadrangle (16) is refused, as FlurryPP's shape functions refuse it.
 * names[i] -> bcs[i] maps physical names to BC names ("fluid" marks the
 * interior). */
int fhost_mesh_read_gmsh(const char* path, int nMap, const char* const* names,
                         const char* const* bcs, fhost_mesh** out);

/* Geometry nodes of every element for curved (high-order) elements
 * (ele::nodes, src/ele.cpp:405-430): c2n [nEles][nShape] in Gmsh's node
 * order, nShape as for fhost_mesh_read_gmsh; the first 4 (8) must be the
 * element's vertices.  The metric terms then come from these nodes' shape
 * functions (src/funcs.cpp:145-468, calcTransforms src/solver.cpp:1587-1694).
 * get returns nShape and, when c2n is not NULL, the nodes (the vertices for
 * linear elements). */
int fhost_mesh_set_shape_nodes(fhost_mesh* m, int nShape, const int* c2n);
int fhost_mesh_get_shape_nodes(const fhost_mesh* m, int* nShape, int* c2n);

/* Move every vertex that is not on a boundary by a uniform offset in
 * [-amp*h_d, +amp*h_d] per coordinate d (h_d: given spacing).  The offset of
 * vertex v, coordinate d, comes from splitmix64(seed ^ (v*nDims+d)). */
int fhost_mesh_jitter(fhost_mesh* m, const double* h, double amp, unsigned long long seed);

int fhost_mesh_set_periodic(fhost_mesh* m, double dx, double dy, double dz, double tol);

/* Connectivity, periodic pairing and face list (FlurryPP order: interior
 * faces ascending by first occurrence, then matched periodic pairs, then the
 * remaining boundary faces). */
int fhost_mesh_process(fhost_mesh* m);

/* Counts after processing: nVerts, nEles, nFaces (unique), nIntFaces
 * (incl. periodic), nBndFaces. */
int fhost_mesh_counts(const fhost_mesh* m, int* nDims, int* nVerts, int* nEles,
                      int* nIntFaces, int* nBndFaces);

int fhost_mesh_get_nodes(const fhost_mesh* m, double* xv, int* c2v);

/* Face records in FlurryPP's `faces` order (size nIntFaces+nBndFaces):
 * eleL, locF_L, eleR (-1 on boundaries), locF_R (IDR), relRot (3-D; 0 in 2-D),
 * bcType (-1 interior), gid (global face id), periodic (1 if paired). */
int fhost_mesh_get_faces(const fhost_mesh* m, int* eleL, int* locL, int* eleR,
                         int* locR, int* relRot, int* bcType, int* gid, int* periodic);

/* Metric terms for order P at solution/flux points (FlurryPP layouts):
 * pos_spts [S][E][D], pos_fpts [F][E][D], jginv_spts [D][S][E][D],
 * detjac_spts [S][E], norm_fpts [F][E][D], da_fpts [F][E]. Any may be NULL.
 * Fails on a negative Jacobian (src/solver.cpp:1638). */
int fhost_mesh_geometry(const fhost_mesh* m, int order, int ptsType,
                        double* pos_spts, double* pos_fpts, double* jginv_spts,
                        double* detjac_spts, double* norm_fpts, double* da_fpts);

void fhost_mesh_destroy(fhost_mesh* m);

#ifdef __cplusplus
}
#endif
#endif
