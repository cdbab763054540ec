/*
 * fr_oracle.c -- CPU restatement of FlurryPP's FR residual and RK update.
 * TEST INFRASTRUCTURE ONLY (see fr_oracle.h).  Plain C99, serial.
 */
#include "fr_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static char g_err[512];

static int fail(const char* msg) {
  snprintf(g_err, sizeof g_err, "%s", msg);
  return 1;
}

const char* or_last_error(void) { return g_err; }

struct or_ctx {
  or_params p;
  int D, N, S, F, E, P1f, nRK;
  double *sf, *grad, *efn, *corr, *corru, *ws, *wf;
  double *jginv, *detj, *norm, *da;
  int nFaces;
  int *fEL, *fLL, *fER, *fLR, *fRR, *fBC;
  int* fptR; /* [nFaces][P1f] */
  double *U, *U0, *Uf, *Fs, *dU, *dUf, *dUc, *disFn, *Fn, *divF, *waveSp, *eleDt, *faceFn, *tmpF;
  double dt;
  int iter;
};

/* ---- array indexing in FlurryPP's row-major layouts ------------------- */
#define I3(a, b, c, B, C) (((size_t)(a) * (B) + (b)) * (C) + (c))
#define I4(a, b, c, d, B, C, Dd) ((((size_t)(a) * (B) + (b)) * (C) + (c)) * (Dd) + (d))

/* C[m][n] = A[m][k] B[k][n] (+ C if beta); sums ascend in k. */
static void gemm(int m, long long n, int k, const double* A, const double* B, double* C, int beta, double* tmp) {
  for (int i = 0; i < m; i++) {
    double* row = tmp;
    for (long long j = 0; j < n; j++) row[j] = 0.;
    for (int l = 0; l < k; l++) {
      double a = A[(size_t)i * k + l];
      const double* b = B + (size_t)l * n;
      for (long long j = 0; j < n; j++) row[j] += a * b[j];
    }
    double* c = C + (size_t)i * n;
    if (beta)
      for (long long j = 0; j < n; j++) c[j] = row[j] + c[j];
    else
      for (long long j = 0; j < n; j++) c[j] = row[j];
  }
}

/* ---- pointwise physics (src/flux.cpp) ---------------------------------- */

/* inviscidFlux, src/flux.cpp:37-76 */
static void inviscid_flux(const double* U, double F[3][5], const or_params* p) {
  if (p->n_dims == 2) {
    double rho = U[0], u = U[1] / rho, v = U[2] / rho, E = U[3];
    double pr = (p->gamma - 1.0) * (E - 0.5 * rho * (u * u + v * v));
    F[0][0] = rho * u;         F[1][0] = rho * v;
    F[0][1] = rho * u * u + pr; F[1][1] = rho * u * v;
    F[0][2] = rho * v * u;     F[1][2] = rho * v * v + pr;
    F[0][3] = (E + pr) * u;    F[1][3] = (E + pr) * v;
  } else {
    double rho = U[0], u = U[1] / rho, v = U[2] / rho, w = U[3] / rho, E = U[4];
    double pr = (p->gamma - 1.0) * (E - 0.5 * rho * (u * u + v * v + w * w));
    F[0][0] = rho * u;          F[1][0] = rho * v;          F[2][0] = rho * w;
    F[0][1] = rho * u * u + pr; F[1][1] = rho * u * v;      F[2][1] = rho * u * w;
    F[0][2] = rho * v * u;      F[1][2] = rho * v * v + pr; F[2][2] = rho * v * w;
    F[0][3] = rho * w * u;      F[1][3] = rho * w * v;      F[2][3] = rho * w * w + pr;
    F[0][4] = (E + pr) * u;     F[1][4] = (E + pr) * v;     F[2][4] = (E + pr) * w;
  }
}

/* viscousFlux, src/flux.cpp:79-210; g[dim][field] */
static void viscous_flux(const double* U, double g[3][5], double Fv[3][5], const or_params* p) {
  int D = p->n_dims;
  double rho = U[0], u = U[1] / rho, v = U[2] / rho;
  double e = U[D + 1] / rho - 0.5 * (u * u + v * v);
  double w = 0.;
  if (D == 3) {
    w = U[3] / rho;
    e -= 0.5 * (w * w);
  }
  double dRho_dx = g[0][0], dRhoU_dx = g[0][1], dRhoV_dx = g[0][2], dE_dx = g[0][D + 1];
  double dRho_dy = g[1][0], dRhoU_dy = g[1][1], dRhoV_dy = g[1][2], dE_dy = g[1][D + 1];
  double dRho_dz = 0, dRhoU_dz = 0, dRhoV_dz = 0, dRhoW_dx = 0, dRhoW_dy = 0, dRhoW_dz = 0, dE_dz = 0;
  if (D == 3) {
    dRho_dz = g[2][0]; dRhoU_dz = g[2][1]; dRhoV_dz = g[2][2];
    dRhoW_dx = g[0][3]; dRhoW_dy = g[1][3]; dRhoW_dz = g[2][3]; dE_dz = g[2][4];
  }
  double mu = p->mu_inf;
  if (!p->fix_vis) {
    double rt = (p->gamma - 1.0) * e / (p->rt_inf);
    mu *= pow(rt, 1.5) * (1. + (p->c_sth)) / (rt + (p->c_sth));
  }
  double du_dx = (dRhoU_dx - dRho_dx * u) / rho, du_dy = (dRhoU_dy - dRho_dy * u) / rho;
  double dv_dx = (dRhoV_dx - dRho_dx * v) / rho, dv_dy = (dRhoV_dy - dRho_dy * v) / rho;
  double du_dz = 0, dv_dz = 0, dw_dx = 0, dw_dy = 0, dw_dz = 0;
  if (D == 3) {
    du_dz = (dRhoU_dz - dRho_dz * u) / rho;
    dv_dz = (dRhoV_dz - dRho_dz * v) / rho;
    dw_dx = (dRhoW_dx - dRho_dx * w) / rho;
    dw_dy = (dRhoW_dy - dRho_dy * w) / rho;
    dw_dz = (dRhoW_dz - dRho_dz * w) / rho;
  }
  (void)dv_dz;
  double dK_dx, dK_dy, dK_dz;
  if (D == 2) {
    dK_dx = 0.5 * (u * u + v * v) * dRho_dx + rho * (u * du_dx + v * dv_dx);
    dK_dy = 0.5 * (u * u + v * v) * dRho_dy + rho * (u * du_dy + v * dv_dy);
    dK_dz = 0;
  } else {
    dK_dx = 0.5 * (u * u + v * v + w * w) * dRho_dx + rho * (u * du_dx + v * dv_dx + w * dw_dx);
    dK_dy = 0.5 * (u * u + v * v + w * w) * dRho_dy + rho * (u * du_dy + v * dv_dy + w * dw_dy);
    dK_dz = 0.5 * (u * u + v * v + w * w) * dRho_dz + rho * (u * du_dz + v * dv_dz + w * dw_dz);
  }
  double de_dx = (dE_dx - dK_dx - dRho_dx * e) / rho;
  double de_dy = (dE_dy - dK_dy - dRho_dy * e) / rho;
  double de_dz = 0;
  if (D == 3) de_dz = (dE_dz - dK_dz - dRho_dz * e) / rho;
  double diag = (du_dx + dv_dy + dw_dz) / 3.0;
  double tauxx = 2.0 * mu * (du_dx - diag);
  double tauyy = 2.0 * mu * (dv_dy - diag);
  double tauxy = mu * (du_dy + dv_dx);
  double tauxz = 0, tauyz = 0, tauzz = 0;
  if (D == 3) {
    /* FlurryPP's expressions, src/flux.cpp:184-186 */
    tauxz = mu * (du_dz + dv_dx);
    tauyz = mu * (du_dz + dv_dy);
    tauzz = 2.0 * mu * (dw_dz - diag);
  }
  double k = (mu / p->prandtl) * (p->gamma);
  Fv[0][0] = 0.0;
  Fv[0][1] = -tauxx;
  Fv[0][2] = -tauxy;
  Fv[0][D + 1] = -(u * tauxx + v * tauxy + w * tauxz + k * de_dx);
  Fv[1][0] = 0.0;
  Fv[1][1] = -tauxy;
  Fv[1][2] = -tauyy;
  Fv[1][D + 1] = -(u * tauxy + v * tauyy + w * tauyz + k * de_dy);
  if (D == 3) {
    Fv[0][3] = -tauzz; /* src/flux.cpp:201 */
    Fv[1][3] = -tauyz;
    Fv[2][0] = 0.0;
    Fv[2][1] = -tauxz;
    Fv[2][2] = -tauyz;
    Fv[2][3] = -tauzz;
    Fv[2][4] = -(u * tauxz + v * tauyz + w * tauzz + k * de_dz);
  }
}

/* getCFLLimit, src/global.cpp:81-105 */
static double cfl_limit(int order) {
  static const double t[6] = {1.393, 0.464, 0.235, 0.1453, 0.100, 0.0736};
  return (order >= 0 && order <= 5) ? t[order] : -1.;
}

/* ---- set-up ------------------------------------------------------------- */

static double* dup(const double* src, size_t n) {
  double* d = (double*)malloc(n * sizeof(double));
  if (src) memcpy(d, src, n * sizeof(double));
  else memset(d, 0, n * sizeof(double));
  return d;
}
static int* dupi(const int* src, size_t n) {
  int* d = (int*)malloc((n ? n : 1) * sizeof(int));
  if (src) memcpy(d, src, n * sizeof(int));
  return d;
}

/* intFace::setupRightState flux-point matching, src/intFace.cpp:57-95 */
static void right_fpts(int D, int order, int P1f, int idr, int relRot, int* out) {
  if (D == 2) {
    int start = idr * P1f + P1f, end = idr * P1f, f = 0;
    for (int i = start - 1; i >= end; i--) out[f++] = i;
    return;
  }
  int P1 = order + 1, s = P1f * idr, e = P1f * (idr + 1);
  for (int i = 0; i < P1f; i++) {
    int ii = i % P1, jj = i / P1;
    switch (relRot) {
      case 0: out[i] = s + ii * P1 + jj; break;
      case 1: out[i] = s + order - ii + jj * P1; break;
      case 2: out[i] = e - 1 - (ii * P1 + jj); break;
      default: out[i] = e - P1 * (jj + 1) + ii; break;
    }
  }
}

int or_create(const or_desc* d, or_ctx** out) {
  const or_params* p = d->params;
  if (p->equation != 1) return fail("oracle: only the Euler/Navier-Stokes equations are restated");
  if (p->n_dims == 3 && p->riemann_type == 1) return fail("Roe not implemented in 3D");
  or_ctx* c = (or_ctx*)calloc(1, sizeof(or_ctx));
  c->p = *p;
  c->D = p->n_dims; c->N = p->n_fields; c->S = p->n_spts; c->F = p->n_fpts; c->E = p->n_eles;
  c->P1f = p->n_fpts_face; c->nRK = p->n_rk_steps;
  int D = c->D, N = c->N, S = c->S, F = c->F, E = c->E;
  c->sf = dup(d->opp_spts_to_fpts, (size_t)F * S);
  c->grad = dup(d->opp_grad_spts, (size_t)D * S * S);
  c->efn = dup(d->opp_extrapolate_fn, (size_t)D * F * S);
  c->corr = dup(d->opp_correction, (size_t)S * F);
  c->corru = dup(d->opp_correct_u, (size_t)D * S * F);
  c->ws = dup(d->w_spts, S);
  c->wf = dup(d->w_face, c->P1f);
  c->jginv = dup(d->jginv_spts, (size_t)D * S * E * D);
  c->detj = dup(d->detjac_spts, (size_t)S * E);
  c->norm = dup(d->norm_fpts, (size_t)F * E * D);
  c->da = dup(d->da_fpts, (size_t)F * E);
  c->nFaces = d->n_faces;
  c->fEL = dupi(d->face_ele_l, d->n_faces); c->fLL = dupi(d->face_loc_l, d->n_faces);
  c->fER = dupi(d->face_ele_r, d->n_faces); c->fLR = dupi(d->face_loc_r, d->n_faces);
  c->fRR = dupi(d->face_rel_rot, d->n_faces); c->fBC = dupi(d->face_bc, d->n_faces);
  c->fptR = (int*)malloc(((size_t)d->n_faces * c->P1f + 1) * sizeof(int));
  for (int f = 0; f < d->n_faces; f++) {
    if (c->fER[f] >= 0) right_fpts(D, p->order, c->P1f, c->fLR[f], c->fRR[f], c->fptR + (size_t)f * c->P1f);
    else for (int i = 0; i < c->P1f; i++) c->fptR[(size_t)f * c->P1f + i] = -1;
  }
  size_t nS = (size_t)S * E * N, nF = (size_t)F * E * N;
  c->U = dup(0, nS); c->U0 = dup(0, nS); c->Uf = dup(0, nF);
  c->Fs = dup(0, D * nS); c->dU = dup(0, D * nS); c->dUf = dup(0, D * nF); c->dUc = dup(0, nF);
  c->disFn = dup(0, nF); c->Fn = dup(0, nF); c->divF = dup(0, c->nRK * nS);
  c->waveSp = dup(0, (size_t)E * F); c->eleDt = dup(0, E);
  c->faceFn = dup(0, (size_t)d->n_faces * c->P1f * N);
  c->tmpF = dup(0, (size_t)E * N);
  c->dt = p->dt;
  c->iter = p->init_iter;
  *out = c;
  return 0;
}

void or_destroy(or_ctx* c) {
  if (!c) return;
  double* ds[] = {c->sf, c->grad, c->efn, c->corr, c->corru, c->ws, c->wf, c->jginv, c->detj, c->norm, c->da,
                  c->U, c->U0, c->Uf, c->Fs, c->dU, c->dUf, c->dUc, c->disFn, c->Fn, c->divF, c->waveSp,
                  c->eleDt, c->faceFn, c->tmpF};
  for (size_t i = 0; i < sizeof ds / sizeof ds[0]; i++) free(ds[i]);
  free(c->fEL); free(c->fLL); free(c->fER); free(c->fLR); free(c->fRR); free(c->fBC); free(c->fptR);
  free(c);
}

static double* arr(or_ctx* c, int which, int step, size_t* n) {
  size_t nS = (size_t)c->S * c->E * c->N, nF = (size_t)c->F * c->E * c->N;
  switch (which) {
    case OR_U_SPTS: *n = nS; return c->U;
    case OR_U0: *n = nS; return c->U0;
    case OR_U_FPTS: *n = nF; return c->Uf;
    case OR_F_SPTS: *n = c->D * nS; return c->Fs;
    case OR_DIVF_SPTS: *n = nS; return (step >= 0 && step < c->nRK) ? c->divF + step * nS : NULL;
    case OR_DISFN_FPTS: *n = nF; return c->disFn;
    case OR_FN_FPTS: *n = nF; return c->Fn;
    case OR_DU_SPTS: *n = c->D * nS; return c->dU;
    case OR_DU_FPTS: *n = c->D * nF; return c->dUf;
    case OR_DUC_FPTS: *n = nF; return c->dUc;
    case OR_WAVESP: *n = (size_t)c->E * c->F; return c->waveSp;
    case OR_ELE_DT: *n = c->E; return c->eleDt;
    default: *n = 0; return NULL;
  }
}

long long or_array_size(or_ctx* c, int which) {
  if (which == OR_FACE_FPT_L || which == OR_FACE_FPT_R) return (long long)c->nFaces * c->P1f;
  size_t n;
  arr(c, which, 0, &n);
  return (long long)n;
}

int or_get_array(or_ctx* c, int which, int step, void* host) {
  if (which == OR_FACE_FPT_L || which == OR_FACE_FPT_R) {
    /* flat point index fpt*E+ele of every face flux point */
    int* o = (int*)host;
    for (int f = 0; f < c->nFaces; f++)
      for (int i = 0; i < c->P1f; i++) {
        size_t k = (size_t)f * c->P1f + i;
        if (which == OR_FACE_FPT_L) o[k] = (c->fLL[f] * c->P1f + i) * c->E + c->fEL[f];
        else o[k] = (c->fER[f] >= 0) ? c->fptR[k] * c->E + c->fER[f] : -1;
      }
    return 0;
  }
  size_t n;
  double* a = arr(c, which, step, &n);
  if (!a) return fail("bad array id");
  memcpy(host, a, n * sizeof(double));
  return 0;
}

int or_set_array(or_ctx* c, int which, int step, const void* host) {
  size_t n;
  double* a = arr(c, which, step, &n);
  if (!a) return fail("bad array id");
  memcpy(a, host, n * sizeof(double));
  return 0;
}

/* ---- residual pieces (src/solver.cpp) ---------------------------------- */

int or_extrapolate_u(or_ctx* c) { /* :556-572 */
  gemm(c->F, (long long)c->E * c->N, c->S, c->sf, c->U, c->Uf, 0, c->tmpF);
  return 0;
}

static void calc_grad_u_spts(or_ctx* c) { /* :1060-1078 */
  size_t nS = (size_t)c->S * c->E * c->N;
  for (int d = 0; d < c->D; d++)
    gemm(c->S, (long long)c->E * c->N, c->S, c->grad + (size_t)d * c->S * c->S, c->U, c->dU + d * nS, 0, c->tmpF);
}

static void calc_inviscid_flux_spts(or_ctx* c) { /* :707-736 */
  int D = c->D, N = c->N, S = c->S, E = c->E;
  double F[3][5];
  for (int s = 0; s < S; s++)
    for (int e = 0; e < E; e++) {
      inviscid_flux(&c->U[I3(s, e, 0, E, N)], F, &c->p);
      if (c->p.viscous) {
        for (int d = 0; d < D; d++)
          for (int k = 0; k < N; k++) c->Fs[I4(d, s, e, k, S, E, N)] = F[d][k];
      } else {
        for (int d1 = 0; d1 < D; d1++)
          for (int k = 0; k < N; k++) {
            double* o = &c->Fs[I4(d1, s, e, k, S, E, N)];
            *o = 0.;
            for (int d2 = 0; d2 < D; d2++) *o += c->jginv[I4(d1, s, e, d2, S, E, D)] * F[d2][k];
          }
      }
    }
}

/* boundFace::applyBCs, src/boundFace.cpp:65-314 (static grids) */
static void apply_bcs(const or_params* p, int bc, const double* UL, const double* n, double* UR) {
  int D = p->n_dims;
  double vL[3] = {0, 0, 0}, vR[3] = {0, 0, 0};
  double vB[3] = {p->u_bound, p->v_bound, p->w_bound};
  double g = p->gamma, gmo = g - 1.;
  double rhoL = UL[0], eL = UL[D + 1];
  for (int i = 0; i < D; i++) vL[i] = UL[i + 1] / UL[0];
  double vSq = 0;
  for (int i = 0; i < D; i++) vSq += vL[i] * vL[i];
  double pL = (g - 1.0) * (eL - 0.5 * rhoL * vSq);
  double rhoR = 0, pR = 0, ER = 0;
  switch (bc) {
    case 4: /* SUB_IN */
      rhoR = p->rho_bound;
      for (int i = 0; i < D; i++) vR[i] = vB[i];
      pR = pL;
      vSq = 0;
      for (int i = 0; i < D; i++) vSq += vR[i] * vR[i];
      ER = (pR / (g - 1.0)) + 0.5 * rhoR * vSq;
      break;
    case 5: /* SUB_OUT */
      rhoR = rhoL;
      for (int i = 0; i < D; i++) vR[i] = vL[i];
      pR = p->p_bound;
      vSq = 0.;
      for (int i = 0; i < D; i++) vSq += vR[i] * vR[i];
      ER = (pR / (g - 1.0)) + 0.5 * rhoR * vSq;
      break;
    case 2: /* SUP_IN */
      rhoR = p->rho_bound;
      vR[0] = p->u_bound; vR[1] = p->v_bound; vR[2] = p->w_bound;
      pR = p->p_bound;
      vSq = 0.;
      for (int i = 0; i < D; i++) vSq += vR[i] * vR[i];
      ER = (pR / (g - 1.0)) + 0.5 * rhoR * vSq;
      break;
    case 3: /* SUP_OUT */
      rhoR = rhoL;
      for (int i = 0; i < D; i++) vR[i] = vL[i];
      ER = eL;
      break;
    case 8: case 12: { /* SLIP_WALL, SYMMETRY */
      rhoR = rhoL;
      double vn = 0.;
      for (int i = 0; i < D; i++) vn += (vL[i] - 0.) * n[i];
      for (int i = 0; i < D; i++) vR[i] = vL[i] - (2.0) * vn * n[i];
      ER = eL;
      break;
    }
    case 13: /* ISOTHERMAL_NOSLIP_MOVING */
      pR = pL;
      rhoR = pR / (p->r_gas * p->t_wall);
      vR[0] = 0. + p->u_wall; vR[1] = 0. + p->v_wall;
      if (D == 3) vR[2] = 0. + p->w_wall;
      vSq = 0.;
      for (int i = 0; i < D; i++) vSq += vR[i] * vR[i];
      ER = (pR / (g - 1.0)) + 0.5 * rhoR * vSq;
      break;
    case 9: /* ISOTHERMAL_NOSLIP */
      pR = pL;
      rhoR = pR / (p->r_gas * p->t_wall);
      for (int i = 0; i < D; i++) vR[i] = 0.;
      vSq = 0.;
      for (int i = 0; i < D; i++) vSq += vR[i] * vR[i];
      ER = (pR / (g - 1.0)) + 0.5 * rhoR * vSq;
      break;
    case 10: /* ADIABATIC_NOSLIP */
      rhoR = rhoL;
      pR = pL;
      for (int i = 0; i < D; i++) vR[i] = 0.;
      vSq = 0.;
      for (int i = 0; i < D; i++) vSq += vR[i] * vR[i];
      ER = (pR / (g - 1.0)) + 0.5 * rhoR * vSq;
      break;
    case 1: { /* CHAR_INOUT */
      double vnL = 0., vnB = 0.;
      for (int i = 0; i < D; i++) vnL += vL[i] * n[i];
      for (int i = 0; i < D; i++) vnB += vB[i] * n[i];
      double cL = sqrt(g * pL / rhoL), cB = sqrt(g * p->p_bound / p->rho_bound);
      double R_L = (fabs(vnB) >= cB && vnL >= 0) ? vnB + 2. / gmo * cB : vnL + 2. / gmo * cL;
      double R_B = (fabs(vnB) >= cB && vnL < 0) ? vnL - 2. / gmo * cL : vnB - 2. / gmo * cB;
      double rhoFac = .25 * (g - 1.) * (R_L - R_B);
      double rhoBG = rhoFac * rhoFac / g;
      if (vnL < 0) rhoBG *= pow(p->rho_bound, g) / p->p_bound;
      else rhoBG *= pow(rhoL, g) / pL;
      rhoR = pow(rhoBG, 1. / gmo);
      pR = rhoFac * rhoFac * rhoR / g;
      double VB = 0.5 * (R_L + R_B);
      for (int i = 0; i < D; i++) vR[i] = (vnL < 0) ? vB[i] + (VB - vnB) * n[i] : vL[i] + (VB - vnL) * n[i];
      double v2 = 0;
      for (int i = 0; i < D; i++) v2 += vR[i] * vR[i];
      ER = pR / gmo + .5 * rhoR * v2;
      break;
    }
    default:
      rhoR = NAN; ER = NAN;
      break;
  }
  UR[0] = rhoR;
  for (int i = 0; i < D; i++) UR[i + 1] = rhoR * vR[i];
  UR[D + 1] = ER;
}

/* face::rusanovFlux, src/face.cpp:314-374; returns the wave speed */
static double rusanov(const or_params* p, const double* UL, const double* UR, const double* n, double* Fn) {
  int D = p->n_dims, N = p->n_fields;
  double FL[3][5], FR[3][5];
  inviscid_flux(UL, FL, p);
  inviscid_flux(UR, FR, p);
  double FnL[5] = {0, 0, 0, 0, 0}, FnR[5] = {0, 0, 0, 0, 0};
  double rhoL = UL[0], rhoR = UR[0];
  double uL = UL[1] / rhoL, uR = UR[1] / rhoR, vL = UL[2] / rhoL, vR = UR[2] / rhoR;
  double pL, pR, vnL = 0., vnR = 0.;
  if (D == 2) {
    pL = (p->gamma - 1.0) * (UL[3] - 0.5 * rhoL * (uL * uL + vL * vL));
    pR = (p->gamma - 1.0) * (UR[3] - 0.5 * rhoR * (uR * uR + vR * vR));
  } else {
    double wL = UL[3] / rhoL, wR = UR[3] / rhoR;
    pL = (p->gamma - 1.0) * (UL[4] - 0.5 * rhoL * (uL * uL + vL * vL + wL * wL));
    pR = (p->gamma - 1.0) * (UR[4] - 0.5 * rhoR * (uR * uR + vR * vR + wR * wR));
  }
  for (int d = 0; d < D; d++) {
    vnL += n[d] * UL[d + 1] / rhoL;
    vnR += n[d] * UR[d + 1] / rhoR;
    for (int i = 0; i < N; i++) {
      FnL[i] += n[d] * FL[d][i];
      FnR[i] += n[d] * FR[d][i];
    }
  }
  double csqL = fmax(p->gamma * pL / rhoL, 0.0), csqR = fmax(p->gamma * pR / rhoR, 0.0);
  double eigL = fabs(vnL) + sqrt(csqL), eigR = fabs(vnR) + sqrt(csqR);
  double eig = fmax(eigL, eigR);
  for (int i = 0; i < N; i++) Fn[i] = 0.5 * (FnL[i] + FnR[i] - eig * (UR[i] - UL[i]));
  return fmax(eigL, eigR);
}

/* face::roeFlux (2-D), src/face.cpp:376-495 */
static void roe2d(const or_params* p, const double* UL, const double* UR, const double* n, double* Fn) {
  double g = p->gamma;
  int D = 2;
  double vL[3], vR[3], um[3], du[5];
  for (int i = 0; i < D; i++) {
    vL[i] = UL[i + 1] / UL[0];
    vR[i] = UR[i + 1] / UR[0];
  }
  double pL = (g - 1.0) * (UL[3] - (0.5 * UL[0] * (vL[0] * vL[0] + vL[1] * vL[1])));
  double pR = (g - 1.0) * (UR[3] - (0.5 * UR[0] * (vR[0] * vR[0] + vR[1] * vR[1])));
  double hL = (UL[D + 1] + pL) / UL[0], hR = (UR[D + 1] + pR) / UR[0];
  double sq_rho = sqrt(UR[0] / UL[0]);
  double rrho = 1. / (sq_rho + 1.);
  double usq = 0., unm = 0., vgn = 0.;
  for (int i = 0; i < D; i++) {
    um[i] = rrho * (vL[i] + sq_rho * vR[i]);
    usq += 0.5 * um[i] * um[i];
    unm += um[i] * n[i];
  }
  double hm = rrho * (hL + sq_rho * hR);
  double am_sq = (g - 1.) * (hm - usq);
  double am = sqrt(am_sq);
  double rhoUnL = 0., rhoUnR = 0.;
  for (int i = 0; i < D; i++) {
    rhoUnL += UL[i + 1] * n[i];
    rhoUnR += UR[i + 1] * n[i];
  }
  Fn[0] = rhoUnL + rhoUnR;
  Fn[1] = rhoUnL * vL[0] + rhoUnR * vR[0] + (pL + pR) * n[0];
  Fn[2] = rhoUnL * vL[1] + rhoUnR * vR[1] + (pL + pR) * n[1];
  Fn[3] = rhoUnL * hL + rhoUnR * hR;
  for (int i = 0; i < 4; i++) du[i] = UR[i] - UL[i];
  double l0 = fabs(unm - vgn), lP = fabs(unm - vgn + am), lM = fabs(unm - vgn - am);
  double eps = 0.5 * (fabs(rhoUnL / UL[0] - rhoUnR / UR[0]) + fabs(sqrt(g * pL / UL[0]) - sqrt(g * pR / UR[0])));
  if (l0 < 2. * eps) l0 = 0.25 * l0 * l0 / eps + eps;
  if (lP < 2. * eps) lP = 0.25 * lP * lP / eps + eps;
  if (lM < 2. * eps) lM = 0.25 * lM * lM / eps + eps;
  double a2 = 0.5 * (lP + lM) - l0, a3 = 0.5 * (lP - lM) / am;
  double a1 = a2 * (g - 1.) / am_sq, a4 = a3 * (g - 1.);
  double a5 = usq * du[0] - um[0] * du[1] - um[1] * du[2] + du[3];
  double a6 = unm * du[0] - n[0] * du[1] - n[1] * du[2];
  double aL1 = a1 * a5 - a3 * a6, bL1 = a4 * a5 - a2 * a6;
  Fn[0] -= l0 * du[0] + aL1;
  Fn[1] -= l0 * du[1] + aL1 * um[0] + bL1 * n[0];
  Fn[2] -= l0 * du[2] + aL1 * um[1] + bL1 * n[1];
  Fn[3] -= l0 * du[3] + aL1 * hm + bL1 * unm;
  for (int i = 0; i < 4; i++) Fn[i] *= 0.5;
}

static double ldg_pen(const or_params* p, const double* n) { /* src/face.cpp:248-256, 599-609 */
  double pf = p->pen_fact;
  if (p->n_dims == 2) {
    if (n[0] + n[1] < 0) pf = -p->pen_fact;
  } else {
    if (n[0] + n[1] + sqrt(2.) * n[2] < 0) pf = -p->pen_fact;
  }
  return pf;
}

/* face::ldgSolution, src/face.cpp:588-615: the common solution UC at one
 * face flux point (bnd: plain average with the boundary state) */
static void ldg_common_u(const or_params* p, int bnd, const double* UL, const double* UR, const double* n,
                         double* UC) {
  double pf = bnd ? 0. : ldg_pen(p, n);
  for (int k = 0; k < p->n_fields; k++)
    UC[k] = bnd ? 0.5 * (UL[k] + UR[k]) : 0.5 * (UL[k] + UR[k]) - pf * (UL[k] - UR[k]);
}

/* solver::calcInviscidFlux_faces -> face::calcInviscidFlux, src/face.cpp:163-208 */
static void calc_inviscid_flux_faces(or_ctx* c) {
  int D = c->D, N = c->N, E = c->E, F = c->F, P1f = c->P1f;
  const or_params* p = &c->p;
  for (int f = 0; f < c->nFaces; f++) {
    int eL = c->fEL[f], eR = c->fER[f], bnd = (eR < 0);
    for (int i = 0; i < P1f; i++) {
      int fl = c->fLL[f] * P1f + i;
      double UL[5], UR[5], n[3], Fn[5];
      for (int k = 0; k < N; k++) UL[k] = c->Uf[I3(fl, eL, k, E, N)];
      for (int d = 0; d < D; d++) n[d] = c->norm[I3(fl, eL, d, E, D)];
      double dAL = c->da[(size_t)fl * E + eL];
      int fr = -1;
      if (bnd) apply_bcs(p, c->fBC[f], UL, n, UR);
      else {
        fr = c->fptR[(size_t)f * P1f + i];
        for (int k = 0; k < N; k++) UR[k] = c->Uf[I3(fr, eR, k, E, N)];
      }
      if (bnd || p->riemann_type == 0) c->waveSp[(size_t)eL * F + fl] = rusanov(p, UL, UR, n, Fn);
      else roe2d(p, UL, UR, n, Fn);
      for (int k = 0; k < N; k++) {
        c->faceFn[((size_t)f * P1f + i) * N + k] = Fn[k];
        c->Fn[I3(fl, eL, k, E, N)] = Fn[k] * dAL;
      }
      if (p->viscous) {
        double UC[5];
        ldg_common_u(p, bnd, UL, UR, n, UC);
        for (int k = 0; k < N; k++) c->dUc[I3(fl, eL, k, E, N)] = UC[k] - UL[k];
        if (!bnd)
          for (int k = 0; k < N; k++) c->dUc[I3(fr, eR, k, E, N)] = UC[k] - UR[k];
      }
      if (!bnd) {
        double dAR = c->da[(size_t)fr * E + eR];
        for (int k = 0; k < N; k++) c->Fn[I3(fr, eR, k, E, N)] = -Fn[k] * dAR;
      }
    }
  }
}

static void correct_grad_u(or_ctx* c) { /* :1080-1137 */
  int D = c->D, N = c->N, S = c->S, E = c->E;
  size_t nS = (size_t)S * E * N;
  for (int d = 0; d < D; d++)
    gemm(S, (long long)E * N, c->F, c->corru + (size_t)d * S * c->F, c->dUc, c->dU + d * nS, 1, c->tmpF);
  for (int s = 0; s < S; s++)
    for (int e = 0; e < E; e++) {
      double inv = 1. / c->detj[(size_t)s * E + e];
      const double* G = c->jginv;
      for (int k = 0; k < N; k++) {
        double r[3];
        for (int d = 0; d < D; d++) r[d] = c->dU[d * nS + I3(s, e, k, E, N)];
        for (int d = 0; d < D; d++) {
          double acc = r[0] * G[I4(0, s, e, d, S, E, D)] + r[1] * G[I4(1, s, e, d, S, E, D)];
          if (D == 3) acc += r[2] * G[I4(2, s, e, d, S, E, D)];
          c->dU[d * nS + I3(s, e, k, E, N)] = inv * acc;
        }
      }
    }
}

static void extrapolate_grad_u(or_ctx* c) { /* :1139-1158 */
  size_t nS = (size_t)c->S * c->E * c->N, nF = (size_t)c->F * c->E * c->N;
  for (int d = 0; d < c->D; d++)
    gemm(c->F, (long long)c->E * c->N, c->S, c->sf, c->dU + d * nS, c->dUf + d * nF, 0, c->tmpF);
}

static void calc_viscous_flux_spts(or_ctx* c) { /* :777-811 */
  int D = c->D, N = c->N, S = c->S, E = c->E;
  size_t nS = (size_t)S * E * N;
  for (int s = 0; s < S; s++)
    for (int e = 0; e < E; e++) {
      double g[3][5], Fv[3][5];
      for (int d = 0; d < D; d++)
        for (int k = 0; k < N; k++) g[d][k] = c->dU[d * nS + I3(s, e, k, E, N)];
      viscous_flux(&c->U[I3(s, e, 0, E, N)], g, Fv, &c->p);
      for (int d = 0; d < D; d++)
        for (int k = 0; k < N; k++) Fv[d][k] += c->Fs[I4(d, s, e, k, S, E, N)];
      for (int d1 = 0; d1 < D; d1++)
        for (int k = 0; k < N; k++) {
          double acc = 0.;
          for (int d2 = 0; d2 < D; d2++) acc += c->jginv[I4(d1, s, e, d2, S, E, D)] * Fv[d2][k];
          c->Fs[I4(d1, s, e, k, S, E, N)] = acc;
        }
    }
}

/* boundFace::applyViscousBCs, src/boundFace.cpp:316-352: the adiabatic wall
 * copies the interior gradient and rewrites its energy column so that the
 * temperature gradient vanishes */
static void adiabatic_grad(const or_params* p, const double* UL, const double* UR, double gL[3][5],
                           double gR[3][5]) {
  int D = p->n_dims;
  for (int d = 0; d < D; d++)
    for (int k = 0; k < p->n_fields; k++) gR[d][k] = gL[d][k];
  double rhovSq = 0.;
  for (int d = 0; d < D; d++) rhovSq += UL[d + 1] * UL[d + 1];
  double pL = (p->gamma - 1.) * (UL[D + 1] - 0.5 * rhovSq / UL[0]);
  double pR = pL;
  double e = pR / ((p->gamma - 1.) * UR[0]);
  double gv[3][3];
  for (int d1 = 0; d1 < D; d1++)
    for (int d2 = 0; d2 < D; d2++) gv[d1][d2] = (gR[d1][d2 + 1] - gR[d1][0] * UR[d2 + 1] / UR[0]) / UR[0];
  double vSq = 0.;
  for (int d = 0; d < D; d++) vSq += UR[d + 1] * UR[d + 1] / (UR[0] * UR[0]);
  for (int d1 = 0; d1 < D; d1++) {
    gR[d1][D + 1] = (e + 0.5 * vSq) * gR[d1][0];
    for (int d2 = 0; d2 < D; d2++) gR[d1][D + 1] += UR[d2 + 1] * gv[d2][d1];
  }
}

/* face::calcViscousFlux, src/face.cpp:216-293, at one face flux point: the
 * LDG common viscous flux Fc dotted with the normal, added to Fn[N].  bc < 0
 * for interior points (UR, gR the right side's values); on boundaries UR is
 * the boundary state and gR is set here (boundFace::getRightGradient,
 * applyViscousBCs, src/boundFace.cpp:56-63,316-352). */
static void ldg_point(const or_params* p, int bc, const double* UL, const double* UR, double gL[3][5],
                      double gR[3][5], const double* n, double* Fn) {
  int D = p->n_dims, N = p->n_fields, bnd = (bc >= 0);
  double FL[3][5], FR[3][5], Fc[3][5];
  if (bnd) {
    if (bc == 10) adiabatic_grad(p, UL, UR, gL, gR);
    else
      for (int d = 0; d < D; d++)
        for (int k = 0; k < N; k++) gR[d][k] = gL[d][k];
    if (bc == 10) {
      viscous_flux(UR, gR, FR, p);
      for (int d = 0; d < D; d++)
        for (int k = 0; k < N; k++) Fc[d][k] = FR[d][k] + p->tau * n[d] * (UL[k] - UR[k]);
    } else {
      viscous_flux(UL, gL, FL, p);
      for (int d = 0; d < D; d++)
        for (int k = 0; k < N; k++) Fc[d][k] = FL[d][k] + p->tau * n[d] * (UL[k] - UR[k]);
    }
  } else {
    viscous_flux(UL, gL, FL, p);
    viscous_flux(UR, gR, FR, p);
    double pf = ldg_pen(p, n);
    double nx = n[0], ny = n[1], nz = (D == 3) ? n[2] : 0;
    for (int k = 0; k < N; k++) {
      double dF0 = FL[0][k] - FR[0][k], dF1 = FL[1][k] - FR[1][k];
      double dU = UL[k] - UR[k];
      if (D == 2) {
        Fc[0][k] = 0.5 * (FL[0][k] + FR[0][k]) + pf * nx * (nx * (dF0) + ny * (dF1)) + p->tau * nx * (dU);
        Fc[1][k] = 0.5 * (FL[1][k] + FR[1][k]) + pf * ny * (nx * (dF0) + ny * (dF1)) + p->tau * ny * (dU);
      } else {
        double dF2 = FL[1][k] - FR[2][k]; /* src/face.cpp:277 */
        double s = nx * (dF0) + ny * (dF1) + nz * (dF2);
        Fc[0][k] = 0.5 * (FL[0][k] + FR[0][k]) + pf * nx * s + p->tau * nx * (dU);
        Fc[1][k] = 0.5 * (FL[1][k] + FR[1][k]) + pf * ny * s + p->tau * ny * (dU);
        Fc[2][k] = 0.5 * (FL[2][k] + FR[2][k]) + pf * nz * s + p->tau * nz * (dU);
      }
    }
  }
  for (int d = 0; d < D; d++)
    for (int k = 0; k < N; k++) Fn[k] += Fc[d][k] * n[d];
}

/* solver::calcViscousFlux_faces -> face::calcViscousFlux for every face */
static void calc_viscous_flux_faces(or_ctx* c) {
  int D = c->D, N = c->N, E = c->E, P1f = c->P1f;
  const or_params* p = &c->p;
  size_t nF = (size_t)c->F * E * N;
  for (int f = 0; f < c->nFaces; f++) {
    int eL = c->fEL[f], eR = c->fER[f], bnd = (eR < 0), bc = c->fBC[f];
    for (int i = 0; i < P1f; i++) {
      int fl = c->fLL[f] * P1f + i, fr = -1;
      double UL[5], UR[5], n[3], gL[3][5], gR[3][5];
      double* Fn = &c->faceFn[((size_t)f * P1f + i) * N];
      for (int k = 0; k < N; k++) UL[k] = c->Uf[I3(fl, eL, k, E, N)];
      for (int d = 0; d < D; d++) n[d] = c->norm[I3(fl, eL, d, E, D)];
      for (int d = 0; d < D; d++)
        for (int k = 0; k < N; k++) gL[d][k] = c->dUf[d * nF + I3(fl, eL, k, E, N)];
      if (bnd) {
        apply_bcs(p, bc, UL, n, UR);
      } else {
        fr = c->fptR[(size_t)f * P1f + i];
        for (int k = 0; k < N; k++) UR[k] = c->Uf[I3(fr, eR, k, E, N)];
        for (int d = 0; d < D; d++)
          for (int k = 0; k < N; k++) gR[d][k] = c->dUf[d * nF + I3(fr, eR, k, E, N)];
      }
      ldg_point(p, bnd ? bc : -1, UL, UR, gL, gR, n, Fn);
      double dAL = c->da[(size_t)fl * E + eL];
      for (int k = 0; k < N; k++) c->Fn[I3(fl, eL, k, E, N)] = Fn[k] * dAL;
      if (!bnd) {
        double dAR = c->da[(size_t)fr * E + eR];
        for (int k = 0; k < N; k++) c->Fn[I3(fr, eR, k, E, N)] = -Fn[k] * dAR;
      }
    }
  }
}

static void extrapolate_normal_flux(or_ctx* c) { /* :1004-1033 (static) */
  size_t nS = (size_t)c->S * c->E * c->N;
  long long n = (long long)c->E * c->N;
  gemm(c->F, n, c->S, c->efn, c->Fs, c->disFn, 0, c->tmpF);
  for (int d = 1; d < c->D; d++)
    gemm(c->F, n, c->S, c->efn + (size_t)d * c->F * c->S, c->Fs + d * nS, c->disFn, 1, c->tmpF);
}

static void calc_div_f(or_ctx* c, int step) { /* :925-954 */
  size_t nS = (size_t)c->S * c->E * c->N;
  long long n = (long long)c->E * c->N;
  double* C = c->divF + step * nS;
  gemm(c->S, n, c->S, c->grad, c->Fs, C, 0, c->tmpF);
  for (int d = 1; d < c->D; d++) gemm(c->S, n, c->S, c->grad + (size_t)d * c->S * c->S, c->Fs + d * nS, C, 1, c->tmpF);
}

static void correct_div_flux(or_ctx* c, int step) { /* :1036-1058 */
  size_t nF = (size_t)c->F * c->E * c->N, nS = (size_t)c->S * c->E * c->N;
  for (size_t i = 0; i < nF; i++) c->disFn[i] = c->Fn[i] - c->disFn[i];
  gemm(c->S, (long long)c->E * c->N, c->F, c->corr, c->disFn, c->divF + step * nS, 1, c->tmpF);
}

int or_calc_residual(or_ctx* c, int step) { /* :366-445 */
  if (step < 0 || step >= c->nRK) return fail("bad RK step");
  or_extrapolate_u(c);
  if (c->p.viscous) calc_grad_u_spts(c);
  calc_inviscid_flux_spts(c);
  calc_inviscid_flux_faces(c);
  if (c->p.viscous) {
    correct_grad_u(c);
    extrapolate_grad_u(c);
    calc_viscous_flux_spts(c);
    calc_viscous_flux_faces(c);
  }
  extrapolate_normal_flux(c);
  calc_div_f(c, step);
  correct_div_flux(c, step);
  return 0;
}

/* timeStepA / timeStepB, src/solver.cpp:470-536 */
int or_time_step_a(or_ctx* c, int step, double rk, const double* src) {
  int N = c->N, S = c->S, E = c->E;
  size_t nS = (size_t)S * E * N;
  const double* dv = c->divF + step * nS;
  for (int s = 0; s < S; s++)
    for (int e = 0; e < E; e++) {
      double dt = (c->p.dt_type != 2) ? c->dt : c->eleDt[e];
      double dj = c->detj[(size_t)s * E + e];
      for (int k = 0; k < N; k++) {
        size_t i = I3(s, e, k, E, N);
        if (src) c->U[i] = c->U0[i] - rk * (dv[i] + src[i]) / dj * dt;
        else c->U[i] = c->U0[i] - rk * dv[i] / dj * dt;
      }
    }
  return 0;
}

int or_time_step_b(or_ctx* c, int step, const double* src) {
  int N = c->N, S = c->S, E = c->E;
  size_t nS = (size_t)S * E * N;
  const double* dv = c->divF + step * nS;
  double b = c->p.rk_b[step];
  for (int s = 0; s < S; s++)
    for (int e = 0; e < E; e++) {
      double dt = (c->p.dt_type != 2) ? c->dt : c->eleDt[e];
      double dj = c->detj[(size_t)s * E + e];
      for (int k = 0; k < N; k++) {
        size_t i = I3(s, e, k, E, N);
        if (src) c->U[i] -= b * (dv[i] + src[i]) / dj * dt;
        else c->U[i] -= b * dv[i] / dj * dt;
      }
    }
  return 0;
}

int or_copy_u_to_u0(or_ctx* c) {
  memcpy(c->U0, c->U, (size_t)c->S * c->E * c->N * sizeof(double));
  return 0;
}
int or_copy_u0_to_u(or_ctx* c) {
  memcpy(c->U, c->U0, (size_t)c->S * c->E * c->N * sizeof(double));
  return 0;
}

/* ele::calcWaveSpFpts, src/ele.cpp:1076-1125 (static grids) */
int or_calc_wave_speed(or_ctx* c) {
  int D = c->D, N = c->N, E = c->E, F = c->F;
  for (int e = 0; e < E; e++)
    for (int f = 0; f < F; f++) {
      const double* U = &c->Uf[I3(f, e, 0, E, N)];
      const double* n = &c->norm[I3(f, e, 0, E, D)];
      double rho = U[0], u = U[1] / rho, v = U[2] / rho, w = 0;
      if (D == 3) w = U[3] / rho;
      double rhoVSq = rho * (u * u + v * v + w * w);
      double pr = (c->p.gamma - 1) * (U[D + 1] - 0.5 * rhoVSq);
      double vN = u * n[0] + v * n[1];
      if (D == 3) vN += w * n[2];
      double vgN = 0;
      double csq = fmax(c->p.gamma * pr / rho, 0.0);
      c->waveSp[(size_t)e * F + f] = (fabs(vN - vgN) + sqrt(csq));
    }
  return 0;
}

/* solver::calcDt + ele::calcDt, src/solver.cpp:447-468, src/ele.cpp:1127-1146 */
int or_calc_dt(or_ctx* c, int refresh, double* dt_out) {
  if (cfl_limit(c->p.order) < 0) return fail("CFL limit not available for this order!");
  if (refresh) or_calc_wave_speed(c);
  int S = c->S, F = c->F, E = c->E, P1f = c->P1f;
  double dt = INFINITY;
  for (int e = 0; e < E; e++) {
    double vol = 0;
    for (int s = 0; s < S; s++) vol += c->detj[(size_t)s * E + e] * c->ws[s];
    double iw = 0;
    for (int f = 0; f < F; f++) iw += c->waveSp[(size_t)e * F + f] * c->wf[f % P1f] * c->da[(size_t)f * E + e];
    double de = c->p.cfl * cfl_limit(c->p.order) * 2 * vol / iw;
    c->eleDt[e] = de;
    dt = fmin(dt, de);
  }
  c->dt = dt;
  if (dt_out) *dt_out = dt;
  return 0;
}

int or_set_dt(or_ctx* c, double dt) {
  c->dt = dt;
  return 0;
}

int or_get_iter(or_ctx* c) { return c->iter; }
double or_get_dt(or_ctx* c) { return c->dt; }

/* flurry main loop body (src/flurry.cpp:135-138) around solver::update
 * (src/solver.cpp:319-363), static grid */
/* solver::update (src/solver.cpp:319-363).  `count` advances the iteration
 * counter that the main loop keeps (src/flurry.cpp:131-136); the P-multigrid
 * levels below the fine one never advance theirs, so their wave speeds are
 * never refreshed by calcDt (src/solver.cpp:449-453). */
static void update_impl(or_ctx* c, int n_steps, const double* src, int count) {
  int nRK = c->nRK;
  for (int it = 0; it < n_steps; it++) {
    if (count) c->iter++;
    const int refresh = count && c->iter == c->p.init_iter + 1;
    for (int step = 0; step < nRK - 1; step++) {
      if (step == 0 && c->p.dt_type != 0) or_calc_dt(c, refresh, NULL);
      if (step == 0) or_copy_u_to_u0(c);
      or_calc_residual(c, step);
      or_time_step_a(c, step, c->p.rk_a[step + 1], src);
    }
    or_calc_residual(c, nRK - 1);
    if (c->p.time_type < 5) {
      if (nRK > 1) or_copy_u0_to_u(c);
      else if (c->p.dt_type != 0) or_calc_dt(c, refresh, NULL);
      for (int step = 0; step < nRK; step++) or_time_step_b(c, step, src);
    } else {
      or_time_step_a(c, nRK - 1, c->p.rk_a[nRK], src);
    }
  }
}

int or_update(or_ctx* c, int n_steps, const double* src) {
  update_impl(c, n_steps, src, 1);
  return 0;
}

/* ---- P-multigrid (src/multigrid.cpp) ----------------------------------- */

/* multiGrid::restrict_pmg, src/multigrid.cpp:315-345: U and divF_spts[0] */
static void pmg_restrict(or_ctx* f, or_ctx* c, const double* R, double* tmp) {
  long long n = (long long)c->E * c->N;
  gemm(c->S, n, f->S, R, f->U, c->U, 0, tmp);
  gemm(c->S, n, f->S, R, f->divF, c->divF, 0, tmp);
}

/* multiGrid::cycle, src/multigrid.cpp:169-313, P levels only (the h levels
 * are 2-D, P=0 only and off by default).  lv[0..nl-1] ascend in order by one;
 * lv[nl-1] is the fine solver.  compute_source_term (:381-398), prolong_err
 * (:347-364) and the sol/corr bookkeeping (:182-188,282-292) inline. */
int or_pmg_cycle(or_ctx** lv, int nl, const double* const* R, const double* const* Pro, int smooth) {
  if (nl < 2) return fail("P-multigrid needs at least two levels");
  for (int i = 0; i + 1 < nl; i++)
    if (lv[i]->E != lv[nl - 1]->E || lv[i]->N != lv[nl - 1]->N || lv[i]->D != lv[nl - 1]->D ||
        lv[i]->p.order + 1 != lv[i + 1]->p.order)
      return fail("P-multigrid levels must share the mesh and ascend in order by one");
  const long long n = (long long)lv[0]->E * lv[0]->N;
  double* tmp = (double*)malloc(sizeof(double) * n);
  double** sol = (double**)calloc(nl, sizeof(double*));
  double** corr = (double**)calloc(nl, sizeof(double*));
  double** src = (double**)calloc(nl, sizeof(double*));
  for (int i = 0; i + 1 < nl; i++) {
    size_t sz = sizeof(double) * (size_t)lv[i]->S * n;
    sol[i] = (double*)malloc(sz);
    corr[i] = (double*)malloc(sz);
    src[i] = (double*)malloc(sz);
  }
  or_ctx* fine = lv[nl - 1];
  or_calc_residual(fine, 0);
  pmg_restrict(fine, lv[nl - 2], R[nl - 2], tmp);
  for (int i = nl - 2; i >= 0; i--) {
    or_ctx* g = lv[i];
    const long long m = (long long)g->S * n;
    for (long long j = 0; j < m; j++) src[i][j] = g->divF[j];
    or_calc_residual(g, 0);
    for (long long j = 0; j < m; j++) src[i][j] -= g->divF[j];
    for (long long j = 0; j < m; j++) sol[i][j] = g->U[j];
    update_impl(g, smooth, src[i], 0);
    if (i > 0) {
      or_calc_residual(g, 0);
      for (long long j = 0; j < m; j++) g->divF[j] += src[i][j];
      pmg_restrict(g, lv[i - 1], R[i - 1], tmp);
    }
  }
  for (int i = 0; i + 1 < nl; i++) {
    or_ctx* g = lv[i];
    const long long m = (long long)g->S * n;
    update_impl(g, smooth, src[i], 0);
    for (long long j = 0; j < m; j++) {
      corr[i][j] = g->U[j];
      corr[i][j] -= sol[i][j];
    }
    gemm(lv[i + 1]->S, n, g->S, Pro[i], corr[i], lv[i + 1]->U, 1, tmp);
  }
  for (int i = 0; i + 1 < nl; i++) {
    free(sol[i]);
    free(corr[i]);
    free(src[i]);
  }
  free(sol);
  free(corr);
  free(src);
  free(tmp);
  return 0;
}

/* ele::getNormResidual + writeResidual, src/ele.cpp:1922-1945,
 * src/output.cpp:1011-1089 */
int or_residual_norm(or_ctx* c, int res_type, double* out) {
  int N = c->N, S = c->S, E = c->E;
  const double* dv = c->divF;
  for (int k = 0; k < N; k++) out[k] = 0;
  double tmp[5];
  for (int e = 0; e < E; e++) {
    for (int k = 0; k < N; k++) tmp[k] = 0;
    for (int s = 0; s < S; s++)
      for (int k = 0; k < N; k++) {
        double v = dv[I3(s, e, k, E, N)], dj = c->detj[(size_t)s * E + e];
        if (res_type == 1) tmp[k] += fabs(v) * c->ws[s];
        else if (res_type == 2) tmp[k] += v * v / dj * c->ws[s];
        else tmp[k] = fmax(fabs(v) / dj, tmp[k]);
      }
    for (int k = 0; k < N; k++) {
      if (res_type == 3) out[k] = fmax(out[k], tmp[k]);
      else out[k] += tmp[k];
    }
  }
  if (res_type == 2)
    for (int k = 0; k < N; k++) out[k] = sqrt(out[k]);
  return 0;
}

/* ---- pointwise hooks for the oracle's own tests ------------------------- */

void or_hook_inviscid_flux(const or_params* p, const double* U, double* F /* [3][5] */) {
  double f[3][5] = {{0}};
  inviscid_flux(U, f, p);
  memcpy(F, f, sizeof f);
}

void or_hook_viscous_flux(const or_params* p, const double* U, const double* g /* [3][5] */, double* Fv) {
  double gg[3][5], f[3][5] = {{0}};
  memcpy(gg, g, sizeof gg);
  viscous_flux(U, gg, f, p);
  memcpy(Fv, f, sizeof f);
}

double or_hook_rusanov(const or_params* p, const double* UL, const double* UR, const double* n, double* Fn) {
  return rusanov(p, UL, UR, n, Fn);
}

void or_hook_roe2d(const or_params* p, const double* UL, const double* UR, const double* n, double* Fn) {
  roe2d(p, UL, UR, n, Fn);
}

void or_hook_bc(const or_params* p, int bc, const double* UL, const double* n, double* UR) { apply_bcs(p, bc, UL, n, UR); }

double or_hook_ldg_pen(const or_params* p, const double* n) { return ldg_pen(p, n); }

void or_hook_ldg_solution(const or_params* p, int bnd, const double* UL, const double* UR, const double* n,
                          double* UC) {
  ldg_common_u(p, bnd, UL, UR, n, UC);
}

void or_hook_ldg_flux(const or_params* p, int bc, const double* UL, const double* UR, const double* gL,
                      const double* gR, const double* n, double* Fn, double* gR_used) {
  double gl[3][5], gr[3][5], ur[5];
  memcpy(gl, gL, sizeof gl);
  memcpy(gr, gR, sizeof gr);
  if (bc >= 0) apply_bcs(p, bc, UL, n, ur);
  else memcpy(ur, UR, sizeof ur);
  for (int k = 0; k < 5; k++) Fn[k] = 0.;
  ldg_point(p, bc, UL, ur, gl, gr, n, Fn);
  if (gR_used) memcpy(gR_used, gr, sizeof gr);
}
