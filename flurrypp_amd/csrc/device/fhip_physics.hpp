// Pointwise physics shared by every libflurry_hip kernel: Euler and
// Navier-Stokes fluxes, Rusanov and Roe common fluxes, boundary states and the
// LDG terms.  Templated on the dimension D (N = D+2 fields) and on the real
// type T, so the same code serves the FP64 path and the FP32 flux stage of
// the mixed-precision path.  Formulas and term order follow
// src/flux.cpp:37-210, src/face.cpp:163-615 and src/boundFace.cpp:65-352,
// except that the point fluxes multiply by one reciprocal of rho where the
// reference divides by rho (a rounding-level difference).
#pragma once

#include <hip/hip_runtime.h>

namespace fhip {

struct Phys {
  double gamma, prandtl, mu_inf, rt_inf, c_sth, pen_fact, tau;
  double rho_bound, u_bound, v_bound, w_bound, p_bound;
  double t_wall, r_gas, u_wall, v_wall, w_wall;
  int fix_vis, riemann_type, viscous;
};

template <class T>
__device__ __forceinline__ T fmax_(T a, T b) { return a > b ? a : (b != b ? b : (a != a ? a : b)); }

// std::max(a,b) semantics: returns a unless a < b.
template <class T>
__device__ __forceinline__ T smax(T a, T b) { return (a < b) ? b : a; }

// inviscidFlux, src/flux.cpp:37-76.  F[d][k].
template <int D, class T>
__device__ __forceinline__ void inviscid_flux(const T* U, T F[D][D + 2], T gamma) {
  if constexpr (D == 2) {
    T rho = U[0], ri = T(1) / rho, u = U[1] * ri, v = U[2] * ri, E = U[3];
    T p = (gamma - T(1)) * (E - T(0.5) * rho * (u * u + v * v));
    F[0][0] = rho * u;         F[1][0] = rho * v;
    F[0][1] = rho * u * u + p; F[1][1] = rho * u * v;
    F[0][2] = rho * v * u;     F[1][2] = rho * v * v + p;
    F[0][3] = (E + p) * u;     F[1][3] = (E + p) * v;
  } else {
    T rho = U[0], ri = T(1) / rho, u = U[1] * ri, v = U[2] * ri, w = U[3] * ri, E = U[4];
    T p = (gamma - T(1)) * (E - T(0.5) * rho * (u * u + v * v + w * w));
    F[0][0] = rho * u;         F[1][0] = rho * v;         F[2][0] = rho * w;
    F[0][1] = rho * u * u + p; F[1][1] = rho * u * v;     F[2][1] = rho * u * w;
    F[0][2] = rho * v * u;     F[1][2] = rho * v * v + p; F[2][2] = rho * v * w;
    F[0][3] = rho * w * u;     F[1][3] = rho * w * v;     F[2][3] = rho * w * w + p;
    F[0][4] = (E + p) * u;     F[1][4] = (E + p) * v;     F[2][4] = (E + p) * w;
  }
}

// viscousFlux, src/flux.cpp:79-210, including the reference's 3-D stress
// expressions (tauxz, tauyz and the x-flux of rho*w as written there), in
// three steps so that a caller can feed the gradient one direction at a time:
// visc_point (primitive state), visc_dir (one direction's conservative
// gradient to du, dv, [dw,] de along it), visc_from_prims (the flux).
template <int D, class T>
struct ViscPt {
  T rho, ri, u, v, w, e, q2;
};

template <int D, class T>
__device__ __forceinline__ ViscPt<D, T> visc_point(const T* U) {
  ViscPt<D, T> p;
  p.rho = U[0];
  p.ri = T(1) / p.rho;
  p.u = U[1] * p.ri;
  p.v = U[2] * p.ri;
  p.e = U[D + 1] * p.ri - T(0.5) * (p.u * p.u + p.v * p.v);
  p.w = 0;
  if constexpr (D == 3) {
    p.w = U[3] * p.ri;
    p.e -= T(0.5) * (p.w * p.w);
    p.q2 = p.u * p.u + p.v * p.v + p.w * p.w;
  } else {
    p.q2 = p.u * p.u + p.v * p.v;
  }
  return p;
}

// dp = {du, dv, [dw,] de} along one direction from that direction's gradient
// of the conserved variables gx[N].
template <int D, class T>
__device__ __forceinline__ void visc_dir(const ViscPt<D, T>& p, const T* gx, T* dp) {
  const T dRho = gx[0];
  const T du = (gx[1] - dRho * p.u) * p.ri, dv = (gx[2] - dRho * p.v) * p.ri;
  T dK;
  if constexpr (D == 3) {
    const T dw = (gx[3] - dRho * p.w) * p.ri;
    dK = T(0.5) * p.q2 * dRho + p.rho * (p.u * du + p.v * dv + p.w * dw);
    dp[2] = dw;
  } else {
    dK = T(0.5) * p.q2 * dRho + p.rho * (p.u * du + p.v * dv);
  }
  dp[0] = du;
  dp[1] = dv;
  dp[D] = (gx[D + 1] - dK - dRho * p.e) * p.ri;
}

template <int D, class T>
__device__ __forceinline__ void visc_from_prims(const ViscPt<D, T>& p, const T dp[D][D + 1], T Fv[D][D + 2],
                                                const Phys& ph) {
  const T u = p.u, v = p.v, w = p.w;
  T mu = T(ph.mu_inf);
  if (!ph.fix_vis) {
    T rt = (T(ph.gamma) - T(1)) * p.e / T(ph.rt_inf);
    mu *= T(pow(double(rt), 1.5)) * (T(1) + T(ph.c_sth)) / (rt + T(ph.c_sth));
  }
  const T kappa = mu * T(ph.gamma / ph.prandtl);
  const T du_dx = dp[0][0], dv_dx = dp[0][1], du_dy = dp[1][0], dv_dy = dp[1][1];
  if constexpr (D == 2) {
    const T de_dx = dp[0][2], de_dy = dp[1][2];
    T diag = (du_dx + dv_dy + T(0)) * T(1. / 3.);
    T tauxx = T(2) * mu * (du_dx - diag);
    T tauyy = T(2) * mu * (dv_dy - diag);
    T tauxy = mu * (du_dy + dv_dx);
    Fv[0][0] = 0; Fv[0][1] = -tauxx; Fv[0][2] = -tauxy;
    Fv[0][3] = -(u * tauxx + v * tauxy + w * T(0) + kappa * de_dx);
    Fv[1][0] = 0; Fv[1][1] = -tauxy; Fv[1][2] = -tauyy;
    Fv[1][3] = -(u * tauxy + v * tauyy + w * T(0) + kappa * de_dy);
  } else {
    const T du_dz = dp[2][0], dv_dz = dp[2][1], dw_dz = dp[2][2];
    const T de_dx = dp[0][3], de_dy = dp[1][3], de_dz = dp[2][3];
    T diag = (du_dx + dv_dy + dw_dz) * T(1. / 3.);
    T tauxx = T(2) * mu * (du_dx - diag);
    T tauyy = T(2) * mu * (dv_dy - diag);
    T tauxy = mu * (du_dy + dv_dx);
    T tauxz = mu * (du_dz + dv_dx);
    T tauyz = mu * (du_dz + dv_dy);
    T tauzz = T(2) * mu * (dw_dz - diag);
    Fv[0][0] = 0; Fv[0][1] = -tauxx; Fv[0][2] = -tauxy; Fv[0][3] = -tauzz;
    Fv[0][4] = -(u * tauxx + v * tauxy + w * tauxz + kappa * de_dx);
    Fv[1][0] = 0; Fv[1][1] = -tauxy; Fv[1][2] = -tauyy; Fv[1][3] = -tauyz;
    Fv[1][4] = -(u * tauxy + v * tauyy + w * tauyz + kappa * de_dy);
    Fv[2][0] = 0; Fv[2][1] = -tauxz; Fv[2][2] = -tauyz; Fv[2][3] = -tauzz;
    Fv[2][4] = -(u * tauxz + v * tauyz + w * tauzz + kappa * de_dz);
  }
}

template <int D, class T>
__device__ __forceinline__ void viscous_flux(const T* U, const T g[D][D + 2], T Fv[D][D + 2], const Phys& ph) {
  const ViscPt<D, T> p = visc_point<D, T>(U);
  T dp[D][D + 1];
#pragma unroll
  for (int x = 0; x < D; x++) visc_dir<D, T>(p, g[x], dp[x]);
  visc_from_prims<D, T>(p, dp, Fv, ph);
}

// Rusanov common flux (src/face.cpp:314-374).  Returns the wave speed that
// the reference stores into the left element.
template <int D, class T>
__device__ __forceinline__ T rusanov(const T* UL, const T* UR, const T* n, T* Fn, T gamma) {
  constexpr int N = D + 2;
  T FL[D][N], FR[D][N];
  inviscid_flux<D, T>(UL, FL, gamma);
  inviscid_flux<D, T>(UR, FR, gamma);
  T rhoL = UL[0], rhoR = UR[0], riL = T(1) / rhoL, riR = T(1) / rhoR;
  T uL = UL[1] * riL, uR = UR[1] * riR, vL = UL[2] * riL, vR = UR[2] * riR;
  T pL, pR;
  if constexpr (D == 2) {
    pL = (gamma - T(1)) * (UL[3] - T(0.5) * rhoL * (uL * uL + vL * vL));
    pR = (gamma - T(1)) * (UR[3] - T(0.5) * rhoR * (uR * uR + vR * vR));
  } else {
    T wL = UL[3] * riL, wR = UR[3] * riR;
    pL = (gamma - T(1)) * (UL[4] - T(0.5) * rhoL * (uL * uL + vL * vL + wL * wL));
    pR = (gamma - T(1)) * (UR[4] - T(0.5) * rhoR * (uR * uR + vR * vR + wR * wR));
  }
  T vnL = 0, vnR = 0, FnL[N], FnR[N];
#pragma unroll
  for (int i = 0; i < N; i++) FnL[i] = FnR[i] = 0;
#pragma unroll
  for (int d = 0; d < D; d++) {
    vnL += n[d] * UL[d + 1] * riL;
    vnR += n[d] * UR[d + 1] * riR;
#pragma unroll
    for (int i = 0; i < N; i++) {
      FnL[i] += n[d] * FL[d][i];
      FnR[i] += n[d] * FR[d][i];
    }
  }
  T csqL = smax(gamma * pL * riL, T(0)), csqR = smax(gamma * pR * riR, T(0));
  T eigL = fabs(vnL) + sqrt(csqL), eigR = fabs(vnR) + sqrt(csqR);
  T eig = smax(eigL, eigR);
#pragma unroll
  for (int i = 0; i < N; i++) Fn[i] = T(0.5) * (FnL[i] + FnR[i] - eig * (UR[i] - UL[i]));
  return eig;
}

// Roe common flux with entropy fix, 2-D (src/face.cpp:376-495).
template <class T>
__device__ __forceinline__ void roe2d(const T* UL, const T* UR, const T* n, T* Fn, T g) {
  T vL[2], vR[2], um[2], du[4];
  for (int i = 0; i < 2; i++) {
    vL[i] = UL[i + 1] / UL[0];
    vR[i] = UR[i + 1] / UR[0];
  }
  T pL = (g - T(1)) * (UL[3] - (T(0.5) * UL[0] * (vL[0] * vL[0] + vL[1] * vL[1])));
  T pR = (g - T(1)) * (UR[3] - (T(0.5) * UR[0] * (vR[0] * vR[0] + vR[1] * vR[1])));
  T hL = (UL[3] + pL) / UL[0], hR = (UR[3] + pR) / UR[0];
  T sq_rho = sqrt(UR[0] / UL[0]);
  T rrho = T(1) / (sq_rho + T(1));
  T usq = 0, unm = 0;
  for (int i = 0; i < 2; i++) {
    um[i] = rrho * (vL[i] + sq_rho * vR[i]);
    usq += T(0.5) * um[i] * um[i];
    unm += um[i] * n[i];
  }
  T hm = rrho * (hL + sq_rho * hR);
  T am_sq = (g - T(1)) * (hm - usq);
  T am = sqrt(am_sq);
  T rhoUnL = UL[1] * n[0] + UL[2] * n[1];
  T rhoUnR = UR[1] * n[0] + UR[2] * n[1];
  Fn[0] = rhoUnL + rhoUnR;
  Fn[1] = rhoUnL * vL[0] + rhoUnR * vR[0] + (pL + pR) * n[0];
  Fn[2] = rhoUnL * vL[1] + rhoUnR * vR[1] + (pL + pR) * n[1];
  Fn[3] = rhoUnL * hL + rhoUnR * hR;
  for (int i = 0; i < 4; i++) du[i] = UR[i] - UL[i];
  T l0 = fabs(unm), lP = fabs(unm + am), lM = fabs(unm - am);
  T eps = T(0.5) * (fabs(rhoUnL / UL[0] - rhoUnR / UR[0]) + fabs(sqrt(g * pL / UL[0]) - sqrt(g * pR / UR[0])));
  if (l0 < T(2) * eps) l0 = T(0.25) * l0 * l0 / eps + eps;
  if (lP < T(2) * eps) lP = T(0.25) * lP * lP / eps + eps;
  if (lM < T(2) * eps) lM = T(0.25) * lM * lM / eps + eps;
  T a2 = T(0.5) * (lP + lM) - l0, a3 = T(0.5) * (lP - lM) / am;
  T a1 = a2 * (g - T(1)) / am_sq, a4 = a3 * (g - T(1));
  T a5 = usq * du[0] - um[0] * du[1] - um[1] * du[2] + du[3];
  T a6 = unm * du[0] - n[0] * du[1] - n[1] * du[2];
  T aL1 = a1 * a5 - a3 * a6, bL1 = a4 * a5 - a2 * a6;
  Fn[0] -= l0 * du[0] + aL1;
  Fn[1] -= l0 * du[1] + aL1 * um[0] + bL1 * n[0];
  Fn[2] -= l0 * du[2] + aL1 * um[1] + bL1 * n[1];
  Fn[3] -= l0 * du[3] + aL1 * hm + bL1 * unm;
  for (int i = 0; i < 4; i++) Fn[i] *= T(0.5);
}

// boundFace::applyBCs for static grids (src/boundFace.cpp:65-314).
template <int D, class T>
__device__ __forceinline__ void apply_bc(int bc, const T* UL, const T* n, T* UR, const Phys& ph) {
  T vL[3] = {0, 0, 0}, vR[3] = {0, 0, 0};
  T vB[3] = {T(ph.u_bound), T(ph.v_bound), T(ph.w_bound)};
  T g = T(ph.gamma), gmo = g - T(1);
  T rhoL = UL[0], eL = UL[D + 1];
  for (int i = 0; i < D; i++) vL[i] = UL[i + 1] / UL[0];
  T vSq = 0;
  for (int i = 0; i < D; i++) vSq += vL[i] * vL[i];
  T pL = (g - T(1)) * (eL - T(0.5) * rhoL * vSq);
  T rhoR = 0, pR = 0, ER = 0;
  switch (bc) {
    case 4:  // SUB_IN
      rhoR = T(ph.rho_bound);
      for (int i = 0; i < D; i++) vR[i] = vB[i];
      pR = pL;
      vSq = 0;
      for (int i = 0; i < D; i++) vSq += vR[i] * vR[i];
      ER = (pR / (g - T(1))) + T(0.5) * rhoR * vSq;
      break;
    case 5:  // SUB_OUT
      rhoR = rhoL;
      for (int i = 0; i < D; i++) vR[i] = vL[i];
      pR = T(ph.p_bound);
      vSq = 0;
      for (int i = 0; i < D; i++) vSq += vR[i] * vR[i];
      ER = (pR / (g - T(1))) + T(0.5) * rhoR * vSq;
      break;
    case 2:  // SUP_IN
      rhoR = T(ph.rho_bound);
      vR[0] = T(ph.u_bound); vR[1] = T(ph.v_bound); vR[2] = T(ph.w_bound);
      pR = T(ph.p_bound);
      vSq = 0;
      for (int i = 0; i < D; i++) vSq += vR[i] * vR[i];
      ER = (pR / (g - T(1))) + T(0.5) * rhoR * vSq;
      break;
    case 3:  // SUP_OUT
      rhoR = rhoL;
      for (int i = 0; i < D; i++) vR[i] = vL[i];
      ER = eL;
      break;
    case 8:
    case 12: {  // SLIP_WALL, SYMMETRY
      rhoR = rhoL;
      T vn = 0;
      for (int i = 0; i < D; i++) vn += (vL[i] - T(0)) * n[i];
      for (int i = 0; i < D; i++) vR[i] = vL[i] - T(2) * vn * n[i];
      ER = eL;
      break;
    }
    case 13:  // ISOTHERMAL_NOSLIP_MOVING
      pR = pL;
      rhoR = pR / (T(ph.r_gas) * T(ph.t_wall));
      vR[0] = T(0) + T(ph.u_wall); vR[1] = T(0) + T(ph.v_wall);
      if (D == 3) vR[2] = T(0) + T(ph.w_wall);
      vSq = 0;
      for (int i = 0; i < D; i++) vSq += vR[i] * vR[i];
      ER = (pR / (g - T(1))) + T(0.5) * rhoR * vSq;
      break;
    case 9:  // ISOTHERMAL_NOSLIP
      pR = pL;
      rhoR = pR / (T(ph.r_gas) * T(ph.t_wall));
      vSq = 0;
      ER = (pR / (g - T(1))) + T(0.5) * rhoR * vSq;
      break;
    case 10:  // ADIABATIC_NOSLIP
      rhoR = rhoL;
      pR = pL;
      vSq = 0;
      ER = (pR / (g - T(1))) + T(0.5) * rhoR * vSq;
      break;
    case 1: {  // CHAR_INOUT
      T vnL = 0, vnB = 0;
      for (int i = 0; i < D; i++) vnL += vL[i] * n[i];
      for (int i = 0; i < D; i++) vnB += vB[i] * n[i];
      T cL = sqrt(g * pL / rhoL), cB = sqrt(g * T(ph.p_bound) / T(ph.rho_bound));
      T R_L = (fabs(vnB) >= cB && vnL >= T(0)) ? vnB + T(2) / gmo * cB : vnL + T(2) / gmo * cL;
      T R_B = (fabs(vnB) >= cB && vnL < T(0)) ? vnL - T(2) / gmo * cL : vnB - T(2) / gmo * cB;
      T rhoFac = T(0.25) * (g - T(1)) * (R_L - R_B);
      T rhoBG = rhoFac * rhoFac / g;
      if (vnL < T(0)) rhoBG *= T(pow(double(ph.rho_bound), double(g))) / T(ph.p_bound);
      else rhoBG *= T(pow(double(rhoL), double(g))) / pL;
      rhoR = T(pow(double(rhoBG), 1. / double(gmo)));
      pR = rhoFac * rhoFac * rhoR / g;
      T VB = T(0.5) * (R_L + R_B);
      for (int i = 0; i < D; i++) vR[i] = (vnL < T(0)) ? vB[i] + (VB - vnB) * n[i] : vL[i] + (VB - vnL) * n[i];
      T v2 = 0;
      for (int i = 0; i < D; i++) v2 += vR[i] * vR[i];
      ER = pR / gmo + T(0.5) * rhoR * v2;
      break;
    }
    default:
      rhoR = T(NAN);
      ER = T(NAN);
      break;
  }
  UR[0] = rhoR;
  for (int i = 0; i < D; i++) UR[i + 1] = rhoR * vR[i];
  UR[D + 1] = ER;
}

// Adiabatic-wall gradient state (boundFace::applyViscousBCs,
// src/boundFace.cpp:316-352): copy the left gradient, then set the energy
// gradient so that the temperature gradient vanishes.
template <int D, class T>
__device__ __forceinline__ void adiabatic_grad(const T* UL, const T* UR, T gR[D][D + 2], T gamma) {
  T rhovSq = 0;
  for (int d = 0; d < D; d++) rhovSq += UL[d + 1] * UL[d + 1];
  T pL = (gamma - T(1)) * (UL[D + 1] - T(0.5) * rhovSq / UL[0]);
  T e = pL / ((gamma - T(1)) * UR[0]);
  T gv[D][D];
  for (int d1 = 0; d1 < D; d1++)
    for (int d2 = 0; d2 < D; d2++) gv[d1][d2] = (gR[d1][d2 + 1] - gR[d1][0] * UR[d2 + 1] / UR[0]) / UR[0];
  T vSq = 0;
  for (int d = 0; d < D; d++) vSq += UR[d + 1] * UR[d + 1] / (UR[0] * UR[0]);
  for (int d1 = 0; d1 < D; d1++) {
    gR[d1][D + 1] = (e + T(0.5) * vSq) * gR[d1][0];
    for (int d2 = 0; d2 < D; d2++) gR[d1][D + 1] += UR[d2 + 1] * gv[d2][d1];
  }
}

// Sign of the LDG switch (src/face.cpp:248-256, 599-609).
template <int D, class T>
__device__ __forceinline__ T ldg_pen(const T* n, T pen) {
  if constexpr (D == 2) return (n[0] + n[1] < T(0)) ? -pen : pen;
  else return (n[0] + n[1] + T(1.4142135623730951) * n[2] < T(0)) ? -pen : pen;
}

// One side of the LDG flux: that side's viscous flux projected on n,
// nF[k] = sum_d n[d] F[d][k], and in 3-D d12[k] = F[1][k] - F[2][k], which the
// reference's jump term uses in place of the z-flux difference
// (src/face.cpp:277).
template <int D, class T>
__device__ __forceinline__ void ldg_project(const T F[D][D + 2], const T* n, T* nF, T* d12) {
  constexpr int N = D + 2;
#pragma unroll
  for (int k = 0; k < N; k++) {
    T a = n[0] * F[0][k] + n[1] * F[1][k];
    if constexpr (D == 3) {
      a += n[2] * F[2][k];
      d12[k] = F[1][k] - F[2][k];
    }
    nF[k] = a;
  }
}

template <int D, class T>
__device__ __forceinline__ void ldg_side(const T* U, const T g[D][D + 2], const T* n, T* nF, T* d12,
                                         const Phys& ph) {
  T F[D][D + 2];
  viscous_flux<D, T>(U, g, F, ph);
  ldg_project<D, T>(F, n, nF, d12);
}

// LDG common viscous flux dotted with n, added into Fn
// (face::calcViscousFlux, src/face.cpp:216-293), one side at a time.  With
// Fc[d] = (FL[d] + FR[d]) / 2 + pf n[d] s + tau n[d] (UL - UR) and
// s = n . (FL - FR) (z term FL[1] - FR[2]), s = nFL - nFR + n[2] d12L and
//   sum_d Fc[d] n[d] = [nFL / 2 + pf |n|^2 (nFL + n[2] d12L) + tau |n|^2 UL]
//                    + [nFR / 2 - pf |n|^2 nFR - tau |n|^2 UR],
// so each side adds its bracket and only N values carry over between sides.
// Adds one side's bracket for that side's viscous flux F at the point.
template <int D, class T>
__device__ __forceinline__ void ldg_add_flux(bool left, const T* U, const T F[D][D + 2], const T* n, T* Fn,
                                             const Phys& ph) {
  constexpr int N = D + 2;
  T nF[N], d12[N];
  ldg_project<D, T>(F, n, nF, d12);
  T nn = n[0] * n[0] + n[1] * n[1];
  if constexpr (D == 3) nn += n[2] * n[2];
  const T pn = ldg_pen<D, T>(n, T(ph.pen_fact)) * nn, tn = T(ph.tau) * nn;
  // one flux evaluation for both sides; only the combination differs
#pragma unroll
  for (int k = 0; k < N; k++) {
    T s = nF[k];
    if constexpr (D == 3) s += n[2] * d12[k];
    Fn[k] += left ? T(0.5) * nF[k] + pn * s + tn * U[k] : (T(0.5) - pn) * nF[k] - tn * U[k];
  }
}

template <int D, class T>
__device__ __forceinline__ void ldg_add_side(bool left, const T* U, const T g[D][D + 2], const T* n, T* Fn,
                                             const Phys& ph) {
  T F[D][D + 2];
  viscous_flux<D, T>(U, g, F, ph);
  ldg_add_flux<D, T>(left, U, F, n, Fn, ph);
}

// Boundary faces: the boundary side's flux (the right state's gradient on an
// adiabatic wall, else the left one's) plus the penalty tau n (UL - UR).
template <int D, class T>
__device__ __forceinline__ void ldg_boundary(const T* UL, const T* UR, const T* nFB, const T* n, T* Fn,
                                             const Phys& ph) {
  constexpr int N = D + 2;
  T nn = n[0] * n[0] + n[1] * n[1];
  if constexpr (D == 3) nn += n[2] * n[2];
  const T tau = T(ph.tau);
#pragma unroll
  for (int k = 0; k < N; k++) Fn[k] += nFB[k] + tau * (UL[k] - UR[k]) * nn;
}

// bnd: 0 interior, 1 Dirichlet-type boundary, 2 adiabatic wall.
template <int D, class T>
__device__ __forceinline__ void ldg_flux(int bnd, const T* UL, const T* UR, const T gL[D][D + 2],
                                         const T gR[D][D + 2], const T* n, T* Fn, const Phys& ph) {
  constexpr int N = D + 2;
  if (bnd) {
    T nFB[N], d12[N];
    ldg_side<D, T>(bnd == 2 ? UR : UL, bnd == 2 ? gR : gL, n, nFB, d12, ph);
    ldg_boundary<D, T>(UL, UR, nFB, n, Fn, ph);
  } else {
    ldg_add_side<D, T>(true, UL, gL, n, Fn, ph);
    ldg_add_side<D, T>(false, UR, gR, n, Fn, ph);
  }
}

}  // namespace fhip
