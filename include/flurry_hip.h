/*
 * flurry_hip.h -- libflurry_hip: FlurryPP's FR residual and RK update on an
 * MI355X (gfx950), behind a C ABI.
 *
 * FlurryPP has no plugin layer; its hot path is the member-function sequence
 * of class solver (include/solver.hpp:172-194) over the solver's public
 * arrays (include/solver.hpp:94-116).  Each entry point below replaces one of
 * those members; the host program keeps mesh reading, geometry set-up and
 * everything else, and hands this library the operators, metric terms and
 * face list once.
 *
 *   fhip_create           solver::setupArrays + device copies of opers[order]
 *                         (include/solver.hpp:69), JGinv_spts/detJac_spts/
 *                         norm_fpts/dA_fpts and the face objects' index maps
 *                         (src/face.cpp:59-100, src/intFace.cpp:35-100)
 *   fhip_set_state        writes solver::U_spts            (include/solver.hpp:94)
 *   fhip_get_state        reads  solver::U_spts
 *   fhip_extrapolate_u    solver::extrapolateU             (src/solver.cpp:556-572)
 *   fhip_calc_wave_speed  ele::calcWaveSpFpts on all eles  (src/ele.cpp:1076-1125)
 *   fhip_calc_residual    solver::calcResidual(step)       (src/solver.cpp:366-445)
 *   fhip_time_step_a      solver::timeStepA(step, RKval)   (src/solver.cpp:470-502)
 *   fhip_time_step_b      solver::timeStepB(step)          (src/solver.cpp:504-536)
 *   fhip_copy_u_to_u0     solver::copyUspts_U0             (src/solver.cpp:538-545)
 *   fhip_copy_u0_to_u     solver::copyU0_Uspts             (src/solver.cpp:547-554)
 *   fhip_calc_dt          solver::calcDt                   (src/solver.cpp:447-468)
 *   fhip_update           n x { iter++; solver::update() } (src/flurry.cpp:135-138,
 *                                                           src/solver.cpp:319-363)
 *   fhip_residual_norm    ele::getNormResidual summed as in writeResidual
 *                         (src/ele.cpp:1922-1945, src/output.cpp:1035-1089)
 *   fhip_get_array        reads any of the solver arrays named in FHIP_ARR_*
 *   fhip_pmg_create       multiGrid::setup, P levels      (src/multigrid.cpp:46-143)
 *   fhip_pmg_cycle        multiGrid::cycle(Solver)         (src/multigrid.cpp:169-313)
 *
 * Layouts are FlurryPP's row-major ones: U_spts [nSpts][nEles][nFields],
 * U_fpts/Fn_fpts/disFn_fpts/dUc_fpts [nFpts][nEles][nFields],
 * F_spts/dU_spts [nDims][nSpts][nEles][nFields], dU_fpts
 * [nDims][nFpts][nEles][nFields], JGinv_spts [nDims][nSpts][nEles][nDims],
 * detJac_spts [nSpts][nEles], norm_fpts [nFpts][nEles][nDims],
 * dA_fpts [nFpts][nEles]; the wave speed is per element [nEles][nFpts] as the
 * reference's ele::waveSp_fpts vectors.
 *
 * Errors: every call returns 0 on success or a nonzero FHIP_E* code and sets a
 * thread-local message (fhip_last_error); nothing calls exit().  A context is
 * used by one host thread at a time; many contexts may share a GPU.  All
 * device work of a context runs on its own HIP stream; host-pointer calls are
 * synchronous, device-pointer calls are stream-ordered.
 */
#ifndef FLURRY_HIP_H
#define FLURRY_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  FHIP_OK = 0,
  FHIP_EINVAL = 1,   /* bad argument or unsupported configuration */
  FHIP_EDEVICE = 2,  /* HIP runtime error (allocation, launch, copy) */
  FHIP_ESTATE = 3,   /* non-finite or non-physical state detected */
  FHIP_ESIZE = 4     /* request beyond a documented maximum */
};

/* Run-time options; names and meanings follow input::readInputFile
 * (src/input.cpp:257-596).  Viscous constants are the non-dimensional ones
 * after input::nonDimensionalize (src/input.cpp:529-596). */
typedef struct fhip_params {
  int n_dims, n_fields, order;             /* nDims, nFields (4 in 2-D, 5 in 3-D), P */
  int n_eles, n_spts, n_fpts, n_fpts_face; /* E, (P+1)^D, 2D(P+1)^(D-1), (P+1)^(D-1) */
  int equation;                            /* 1: Euler/Navier-Stokes (only value supported) */
  int viscous;                             /* 0 Euler, 1 Navier-Stokes with LDG */
  int riemann_type;                        /* 0 Rusanov, 1 Roe (2-D only, as in the reference) */
  int time_type;                           /* 0, 2, 4 (RK44), 5, 6 */
  int dt_type;                             /* 0 fixed dt, 1 global CFL, 2 local (per element) CFL */
  int n_rk_steps;                          /* nRKSteps */
  int fix_vis;                             /* 1: constant viscosity, 0: Sutherland */
  int init_iter;                           /* initIter (0 unless restarting) */
  double rk_a[8], rk_b[8];                 /* RKa (nRK+1 used for time_type 5/6), RKb */
  double cfl, dt;                          /* CFL; dt used when dt_type == 0 */
  double gamma, prandtl, mu_inf, rt_inf, c_sth;
  double pen_fact, tau;                    /* LDG_penFact, LDG_tau */
  double rho_bound, u_bound, v_bound, w_bound, p_bound;
  double t_wall, r_gas, u_wall, v_wall, w_wall;
} fhip_params;

/* Everything the solver hands over once, as host pointers (copied). */
typedef struct fhip_desc {
  const fhip_params* params;
  /* opers[order] (include/operators.hpp:142-152), row-major */
  const double* opp_spts_to_fpts;   /* [F][S]    */
  const double* opp_grad_spts;      /* [D][S][S] */
  const double* opp_extrapolate_fn; /* [D][F][S] */
  const double* opp_correction;     /* [S][F]    */
  const double* opp_correct_u;      /* [D][S][F] (viscous only; may be NULL otherwise) */
  const double* w_spts;             /* getQptWeights(P, D)   [S]   */
  const double* w_face;             /* getQptWeights(P, D-1) [F/2D] */
  /* metric terms */
  const double* jginv_spts;         /* [D][S][E][D] */
  const double* detjac_spts;        /* [S][E]       */
  const double* norm_fpts;          /* [F][E][D]    */
  const double* da_fpts;            /* [F][E]       */
  /* faces, in solver::faces order: interior faces first, then boundaries
   * (src/geo.cpp:854-930).  For face i: left element and its local face
   * (eL, locF_L), right element and its local face (eR, faceID_R) or -1 on
   * boundaries, relative rotation (3-D), boundary-condition number (-1 for
   * interior faces; include/global.hpp:113-128). */
  int n_faces;
  const int* face_ele_l;
  const int* face_loc_l;
  const int* face_ele_r;
  const int* face_loc_r;
  const int* face_rel_rot;
  const int* face_bc;
  /* Batching (optional, fused paths only; all NULL/0 for one mesh).  Several
   * independent meshes of the same order and options may share a context:
   * their elements are numbered consecutively, mesh m owning elements
   * [mesh_ele_start[m], mesh_ele_start[m+1]).  Boundary faces may take their
   * free-stream state from a table instead of params: face_bnd_state[i] is
   * -1 or a row of bnd_states [n_bnd_states][5] = rho, u, v, w, p.
   * dt_type 1 then takes the CFL minimum per mesh. */
  int n_meshes;
  const int* mesh_ele_start;        /* [n_meshes+1] */
  int n_bnd_states;
  const double* bnd_states;         /* [n_bnd_states][5] */
  const int* face_bnd_state;        /* [n_faces] */
} fhip_desc;

/* Array ids for fhip_get_array / fhip_set_array / fhip_device_ptr. */
enum {
  FHIP_ARR_U_SPTS = 0, FHIP_ARR_U_FPTS = 1, FHIP_ARR_F_SPTS = 2, FHIP_ARR_DIVF_SPTS = 3,
  FHIP_ARR_DISFN_FPTS = 4, FHIP_ARR_FN_FPTS = 5, FHIP_ARR_DU_SPTS = 6, FHIP_ARR_DU_FPTS = 7,
  FHIP_ARR_DUC_FPTS = 8, FHIP_ARR_U0 = 9, FHIP_ARR_WAVESP = 10, FHIP_ARR_ELE_DT = 11,
  FHIP_ARR_FACE_FPT_L = 12, /* int32 [nFaces][nFptsFace]: fpt*nEles+ele of each left flux point */
  FHIP_ARR_FACE_FPT_R = 13  /* int32, same for the right side (-1 on boundaries) */
};

/* Execution paths for fhip_update. */
enum {
  FHIP_PATH_AUTO = 0,      /* fused kernels where available, else reference-structured */
  FHIP_PATH_STAGED = 1,    /* one kernel per reference step, dense operators */
  FHIP_PATH_FUSED = 2,     /* tensor-product (sum-factorised) fused kernels, FP64 */
  FHIP_PATH_MIXED = 3,     /* fused kernels with FP32 fluxes and face/volume
                              intermediates; FP64 state, RK accumulation,
                              metric terms, wave speeds, dt and residual */
  FHIP_PATH_MIXED_BF16 = 4,   /* FHIP_PATH_MIXED with the z contractions of the
                                 operators (gradient, divergence) of P3 hexes on
                                 the matrix cores as BF16 hi+lo products with FP32
                                 accumulation (P3 hexes only) */
  FHIP_PATH_MIXED_BF16X3 = 5  /* the same with three BF16 pieces (hi+mid+lo) */
};

typedef struct fhip_ctx fhip_ctx;

const char* fhip_last_error(void);

/* Select the GPU (hipSetDevice) for contexts created afterwards on this thread. */
int fhip_set_device(int device);

int fhip_create(const fhip_desc* desc, fhip_ctx** out);
void fhip_destroy(fhip_ctx* ctx);

/* U_spts [S][E][N] from / to host memory (synchronous). */
int fhip_set_state(fhip_ctx* ctx, const double* u_spts);
int fhip_get_state(fhip_ctx* ctx, double* u_spts);

int fhip_extrapolate_u(fhip_ctx* ctx);
int fhip_calc_wave_speed(fhip_ctx* ctx);
int fhip_calc_residual(fhip_ctx* ctx, int step);
/* src: optional multigrid source term [S][E][N] on the host (NULL for none). */
int fhip_time_step_a(fhip_ctx* ctx, int step, double rk_val, const double* src);
int fhip_time_step_b(fhip_ctx* ctx, int step, const double* src);
int fhip_copy_u_to_u0(fhip_ctx* ctx);
int fhip_copy_u0_to_u(fhip_ctx* ctx);

/* solver::calcDt: refresh != 0 recomputes every wave speed from U_fpts first
 * (the reference does so at iter == initIter+1).  dt_out may be NULL. */
int fhip_calc_dt(fhip_ctx* ctx, int refresh, double* dt_out);
int fhip_set_dt(fhip_ctx* ctx, double dt);
int fhip_get_dt(fhip_ctx* ctx, double* dt);
int fhip_get_iter(fhip_ctx* ctx, int* iter);

/* n_steps iterations of the main loop body: iter++, update().  Runs on the
 * device without host synchronisation (dt stays in device memory).
 * path: FHIP_PATH_*.  Asynchronous: call fhip_sync before reading results. */
int fhip_update(fhip_ctx* ctx, int n_steps, int path);
int fhip_sync(fhip_ctx* ctx);
/* 1 if fhip_update can take `path` for this context's configuration, else 0. */
int fhip_path_supported(fhip_ctx* ctx, int path);

/* Options: FHIP_OPT_GRAPHS (1, default on) replays the fused path's RK step
 * as one captured HIP graph. */
enum { FHIP_OPT_GRAPHS = 1 };
int fhip_set_option(fhip_ctx* ctx, int option, int value);

/* Block sizes (threads per block, 64..512) of the fused kernels, in the order
 * extrapolate, face_inviscid, volume, face_boundary (viscous LDG flux at
 * boundary points), rk_update, for path
 * FHIP_PATH_FUSED or FHIP_PATH_MIXED (no counterpart in the reference, which
 * has no kernels).  fhip_autotune times every valid size of every kernel on
 * the context's current data, keeps the fastest and returns them in
 * block_sizes (may be NULL); the state is left unchanged.  Results are cached
 * for the process per (nDims, order, precision, viscous, log2 nEles).  Block
 * shapes never change results.  Default 256 each, or the largest multiple
 * of 64 below it whose LDS slabs fit 64 KiB (P0 hexes). */
int fhip_autotune(fhip_ctx* ctx, int path, int* block_sizes);
int fhip_set_block_sizes(fhip_ctx* ctx, int path, const int* block_sizes);
int fhip_get_block_sizes(fhip_ctx* ctx, int path, int* block_sizes);

/* Residual norm over divF_spts[0] (resType 1, 2 or 3): out[nFields]. */
int fhip_residual_norm(fhip_ctx* ctx, int res_type, double* out);

/* fhip_residual_norm without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (tight
 * serving loops, harnesses that forbid hipMalloc).  Its partial sums live in
 * a static device array of 16384 doubles (128 KiB) that each GPU the library
 * runs on holds until the process ends: contexts of up to
 * (16384 / nFields - 1) * 256 elements (838,400 in 3-D, 1,048,320 in 2-D);
 * larger ones return FHIP_ESIZE.  Same result as fhip_residual_norm, bit for
 * bit.  Calls that overlap on one GPU share that array and race. */
int fhip_residual_norm_noalloc(fhip_ctx* ctx, int res_type, double* out);

/* Batched contexts: residual norm of every mesh, out[n_meshes][nFields],
 * from divF_spts[0] of the last fused step; and per-mesh activity flags
 * (0 freezes a mesh: its solution is left unchanged by fhip_update). */
int fhip_mesh_residual_norms(fhip_ctx* ctx, int res_type, double* out);
int fhip_set_mesh_active(fhip_ctx* ctx, const int* active);

/* n_steps of fhip_update on several contexts, interleaved step by step so
 * that their kernels overlap on the device (each context has its own
 * stream).  Returns after queuing; fhip_sync each context to wait. */
int fhip_update_many(fhip_ctx** ctxs, int n_ctx, int n_steps, int path);

/* ---- P-multigrid (class multiGrid, include/multigrid.hpp:40-64) ----------
 * levels[0..n_levels-1] are contexts on one mesh whose orders ascend by one;
 * levels[n_levels-1] is the fine solver, the others are the coarse pGrids
 * (the caller sets each one up and initialises it, as multiGrid::setup does,
 * with the fine level's RK coefficients).  opp_restrict[i] maps level i+1 to
 * level i: row-major [nSpts(i)][nSpts(i+1)] (opers[P].opp_restrict of level
 * i+1); opp_prolong[i] maps level i to level i+1: [nSpts(i+1)][nSpts(i)]
 * (opers[P].opp_prolong of level i); both as fhost_pmg_operators builds them.
 * smooth_steps: input key smoothSteps.  The coarse levels keep their own
 * src_spts/sol_spts/corr_spts and never refresh their wave speeds in calcDt
 * (their iteration counter stays at initIter, src/flurry.cpp:131-136).
 * fhip_pmg_cycle queues one cycle after the fine level's update, all of it on
 * the fine level's stream: fhip_sync the fine level before using a coarse
 * level directly.  Every level must be driven through the same path.  The
 * contexts must outlive mg. */
typedef struct fhip_pmg fhip_pmg;
int fhip_pmg_create(fhip_ctx* const* levels, int n_levels, const double* const* opp_restrict,
                    const double* const* opp_prolong, int smooth_steps, fhip_pmg** out);
int fhip_pmg_cycle(fhip_pmg* mg, int path);
void fhip_pmg_destroy(fhip_pmg* mg);

/* Batched P-multigrid (no counterpart in the reference: it serves callers
 * that advance many independent meshes of several orders, fhip_update_many):
 * one cycle for fine contexts of distinct orders, whose coarse levels of one
 * order are merged across the fine contexts, so that a cycle launches each
 * coarse order's kernels once for every mesh.  fine[0..n_fine-1]: orders
 * ascending, each above coarse[0]'s; coarse[j] has order coarse[0]'s + j and
 * holds, in this order, the meshes of the fine context of order + 1 (if
 * any) and then those of coarse[j+1]; the top coarse level's order is the
 * highest fine order - 1.  opp_restrict[j] / opp_prolong[j] map order
 * coarse[j] + 1 to coarse[j] and back, as for fhip_pmg_create.  Every mesh
 * goes through the operations of its own fhip_pmg_create cycle, in the same
 * order, so its result is the same bit for bit.  fhip_pmg_cycle queues the
 * cycle on fine[0]'s stream after the work queued on every fine context's
 * stream, and the fine contexts' later work after the cycle.  Fused paths
 * of one precision only. */
int fhip_pmg_create_batch(fhip_ctx* const* fine, int n_fine, fhip_ctx* const* coarse, int n_coarse,
                          const double* const* opp_restrict, const double* const* opp_prolong, int smooth_steps,
                          fhip_pmg** out);

/* Scan U_spts for NaN/Inf, rho <= 0 or p <= 0; returns FHIP_ESTATE if found. */
int fhip_check_state(fhip_ctx* ctx);

/* Number of elements (doubles, or int32 for the face maps) in an array. */
long long fhip_array_size(fhip_ctx* ctx, int which);
int fhip_get_array(fhip_ctx* ctx, int which, int step, void* host);
int fhip_set_array(fhip_ctx* ctx, int which, int step, const void* host);
/* Device address of an array (for zero-copy use on the context's stream). */
int fhip_device_ptr(fhip_ctx* ctx, int which, int step, void** dev);
/* The context's HIP stream (hipStream_t), for callers that order their own work. */
int fhip_stream(fhip_ctx* ctx, void** stream);

/* Average device time per launch of each kernel of the last fhip_update
 * calls, measured with HIP events when timing is enabled. */
int fhip_enable_timing(fhip_ctx* ctx, int on);
int fhip_kernel_times(fhip_ctx* ctx, int max, char* names, int name_len, double* avg_ms, long long* counts);

#ifdef __cplusplus
}
#endif
#endif
