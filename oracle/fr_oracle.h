/*
 * fr_oracle.h -- CPU restatement of FlurryPP's FR residual and RK update.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * or the CPU baseline; nothing in flurrypp_amd/ may call it.
 *
 * The restatement follows src/solver.cpp:319-1158, src/face.cpp:119-615,
 * src/intFace.cpp:35-161, src/boundFace.cpp:35-352, src/flux.cpp:37-210,
 * src/ele.cpp:1076-1146,1922-1945 and src/global.cpp:81-105 loop for loop,
 * over FlurryPP's array layouts, with plain row-major triple-loop GEMMs in
 * place of cblas_dgemm.  It reproduces FlurryPP's behaviour including the
 * quirks of its 3-D viscous stress (src/flux.cpp:184-185,201), its 3-D LDG
 * term (src/face.cpp:277), Rusanov on every boundary face
 * (src/face.cpp:174-177) and the left-only wave speed (src/face.cpp:372).
 */
#ifndef FR_ORACLE_H
#define FR_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_params {
  int n_dims, n_fields, order, n_eles, n_spts, n_fpts, n_fpts_face;
  int equation, viscous, riemann_type, time_type, dt_type, n_rk_steps, fix_vis, init_iter;
  double rk_a[8], rk_b[8];
  double cfl, dt;
  double gamma, prandtl, mu_inf, rt_inf, c_sth, pen_fact, tau;
  double rho_bound, u_bound, v_bound, w_bound, p_bound;
  double t_wall, r_gas, u_wall, v_wall, w_wall;
} or_params;

typedef struct or_desc {
  const or_params* params;
  const double *opp_spts_to_fpts, *opp_grad_spts, *opp_extrapolate_fn, *opp_correction, *opp_correct_u;
  const double *w_spts, *w_face;
  const double *jginv_spts, *detjac_spts, *norm_fpts, *da_fpts;
  int n_faces;
  const int *face_ele_l, *face_loc_l, *face_ele_r, *face_loc_r, *face_rel_rot, *face_bc;
} or_desc;

typedef struct or_ctx or_ctx;

/* Array ids for or_get_array / or_set_array. */
enum {
  OR_U_SPTS = 0, OR_U_FPTS = 1, OR_F_SPTS = 2, OR_DIVF_SPTS = 3, OR_DISFN_FPTS = 4, OR_FN_FPTS = 5,
  OR_DU_SPTS = 6, OR_DU_FPTS = 7, OR_DUC_FPTS = 8, OR_U0 = 9, OR_WAVESP = 10, OR_ELE_DT = 11,
  OR_FACE_FPT_L = 12, OR_FACE_FPT_R = 13
};

const char* or_last_error(void);
int or_create(const or_desc* d, or_ctx** out);
void or_destroy(or_ctx* c);
long long or_array_size(or_ctx* c, int which);
int or_get_array(or_ctx* c, int which, int step, void* host);
int or_set_array(or_ctx* c, int which, int step, const void* host);
int or_extrapolate_u(or_ctx* c);
int or_calc_wave_speed(or_ctx* c);
int or_calc_residual(or_ctx* c, int step);
int or_time_step_a(or_ctx* c, int step, double rkval, const double* src);
int or_time_step_b(or_ctx* c, int step, const double* src);
int or_copy_u_to_u0(or_ctx* c);
int or_copy_u0_to_u(or_ctx* c);
int or_calc_dt(or_ctx* c, int refresh, double* dt_out);
int or_set_dt(or_ctx* c, double dt);
int or_update(or_ctx* c, int n_steps, const double* src);
/* one P-multigrid cycle over levels lv[0..nl-1] (ascending order, lv[nl-1]
 * fine); R[i]: level i+1 -> i, Pro[i]: level i -> i+1, row-major. */
int or_pmg_cycle(or_ctx** lv, int nl, const double* const* R, const double* const* Pro, int smooth);
int or_residual_norm(or_ctx* c, int res_type, double* out);
int or_get_iter(or_ctx* c);
double or_get_dt(or_ctx* c);  /* the time step of the last update / calcDt */

/* Pointwise hooks used by the oracle's own tests (src/flux.cpp:37-210,
 * src/face.cpp:314-495,588-615, src/boundFace.cpp:65-314).  F arrays are
 * [3][5] row-major. */
void or_hook_inviscid_flux(const or_params* p, const double* U, double* F);
void or_hook_viscous_flux(const or_params* p, const double* U, const double* g, double* Fv);
double or_hook_rusanov(const or_params* p, const double* UL, const double* UR, const double* n, double* Fn);
void or_hook_roe2d(const or_params* p, const double* UL, const double* UR, const double* n, double* Fn);
void or_hook_bc(const or_params* p, int bc, const double* UL, const double* n, double* UR);
double or_hook_ldg_pen(const or_params* p, const double* n);
/* ldgSolution at one point (src/face.cpp:588-615) */
void or_hook_ldg_solution(const or_params* p, int bnd, const double* UL, const double* UR, const double* n,
                          double* UC);
/* LDG common viscous flux . n at one point (src/face.cpp:216-293); bc < 0:
 * interior point with right state UR, gR; bc >= 0: boundary point, the right
 * state comes from the boundary condition and gR is ignored.  gR_used
 * ([3][5], may be NULL) returns the right gradient the flux used. */
void or_hook_ldg_flux(const or_params* p, int bc, const double* UL, const double* UR, const double* gL,
                      const double* gR, const double* n, double* Fn, double* gR_used);

#ifdef __cplusplus
}
#endif
#endif
