// Reference-structured ("staged") path: one kernel per step of
// solver::calcResidual (src/solver.cpp:366-445), dense FR operators applied
// as tall-skinny products over all elements, pointwise kernels over
// [point][element] and a face kernel over every face flux point.  It is the
// drop-in that mirrors the reference call for call; the fused path
// (fhip_fused.hip) is the fast one.
#include "fhip_internal.hpp"

#include <algorithm>
#include <cstdio>
#include <mutex>
#include <stdexcept>
#include <string>

namespace fhip {

#define HIPCHK(x)                                                                     \
  do {                                                                                \
    hipError_t e_ = (x);                                                              \
    if (e_ != hipSuccess) throw std::runtime_error(std::string("HIP: ") + hipGetErrorString(e_)); \
  } while (0)

static inline unsigned blocks_for(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// ---------------------------------------------------------------------------
// Dense operator product C[m][n] = A[m][k] B[k][n] (+ C when beta).  A is a
// small FR operator read through uniform (scalar) loads; each thread owns one
// column j of B/C, keeps up to KC entries of it in registers and sweeps the
// rows of A.  Columns are contiguous in memory, so every load and store is
// coalesced.  k > KC is done in chunks, accumulating into C.
template <int KC>
__global__ __launch_bounds__(256) void k_opgemm(int m, long long n, int k, const double* __restrict__ A,
                                                 const double* __restrict__ B, double* __restrict__ C, int beta) {
  long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  for (int l0 = 0; l0 < k; l0 += KC) {
    int kc = min(KC, k - l0);
    double b[KC];
#pragma unroll
    for (int l = 0; l < KC; l++) b[l] = (l < kc) ? B[(long long)(l0 + l) * n + j] : 0.0;
    bool acc = beta || l0 > 0;
    for (int i = 0; i < m; i++) {
      const double* a = A + (long long)i * k + l0;
      double s = 0.0;
#pragma unroll
      for (int l = 0; l < KC; l++)
        if (l < kc) s += a[l] * b[l];
      double* cp = C + (long long)i * n + j;
      *cp = acc ? s + *cp : s;
    }
  }
}

static void opgemm(Ctx& c, const char* name, int m, int k, const double* A, const double* B, double* C, bool beta) {
  long long n = (long long)c.z.E * c.z.N;
  Launch L(c, name);
  if (k <= 16) hipLaunchKernelGGL(k_opgemm<16>, dim3(blocks_for(n, 256)), dim3(256), 0, c.stream, m, n, k, A, B, C, beta ? 1 : 0);
  else if (k <= 32) hipLaunchKernelGGL(k_opgemm<32>, dim3(blocks_for(n, 256)), dim3(256), 0, c.stream, m, n, k, A, B, C, beta ? 1 : 0);
  else hipLaunchKernelGGL(k_opgemm<64>, dim3(blocks_for(n, 256)), dim3(256), 0, c.stream, m, n, k, A, B, C, beta ? 1 : 0);
  HIPCHK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// calcInviscidFlux_spts (src/solver.cpp:707-736): one thread per (spt, ele).
template <int D>
__global__ __launch_bounds__(256) void k_inviscid_spts(long long nPts, long long SE, const double* __restrict__ U,
                                                        const double* __restrict__ JG, double* __restrict__ Fs,
                                                        double gamma, int transform) {
  constexpr int N = D + 2;
  long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nPts) return;
  double u[N], F[D][N];
#pragma unroll
  for (int k = 0; k < N; k++) u[k] = U[p * N + k];
  inviscid_flux<D, double>(u, F, gamma);
  if (!transform) {
#pragma unroll
    for (int d = 0; d < D; d++)
#pragma unroll
      for (int k = 0; k < N; k++) Fs[(d * SE + p) * N + k] = F[d][k];
    return;
  }
  double G[D][D];
#pragma unroll
  for (int d1 = 0; d1 < D; d1++)
#pragma unroll
    for (int d2 = 0; d2 < D; d2++) G[d1][d2] = JG[(d1 * SE + p) * D + d2];
#pragma unroll
  for (int d1 = 0; d1 < D; d1++)
#pragma unroll
    for (int k = 0; k < N; k++) {
      double acc = 0.;
#pragma unroll
      for (int d2 = 0; d2 < D; d2++) acc += G[d1][d2] * F[d2][k];
      Fs[(d1 * SE + p) * N + k] = acc;
    }
}

// ---------------------------------------------------------------------------
// Inviscid face pass (face::calcInviscidFlux, src/face.cpp:163-208; right
// side src/intFace.cpp:147-161): one thread per face flux point.
template <int D>
__global__ __launch_bounds__(256) void k_face_inviscid(int nFF, const int* __restrict__ ptL, const int* __restrict__ ptR,
                                                        const int* __restrict__ bcs, const double* __restrict__ Uf,
                                                        const double* __restrict__ nrm, const double* __restrict__ dA,
                                                        double* __restrict__ Fn, double* __restrict__ faceFn,
                                                        double* __restrict__ waveSp, double* __restrict__ dUc,
                                                        Phys ph) {
  constexpr int N = D + 2;
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nFF) return;
  int pl = ptL[t], pr = ptR[t], bc = bcs[t];
  double UL[N], UR[N], n[D], fn[N];
#pragma unroll
  for (int k = 0; k < N; k++) UL[k] = Uf[(long long)pl * N + k];
#pragma unroll
  for (int d = 0; d < D; d++) n[d] = nrm[(long long)pl * D + d];
  if (pr < 0) apply_bc<D, double>(bc, UL, n, UR, ph);
  else {
#pragma unroll
    for (int k = 0; k < N; k++) UR[k] = Uf[(long long)pr * N + k];
  }
  if (pr < 0 || ph.riemann_type == 0) waveSp[pl] = rusanov<D, double>(UL, UR, n, fn, ph.gamma);
  else {
    if constexpr (D == 2) roe2d<double>(UL, UR, n, fn, ph.gamma);
  }
  double dAL = dA[pl];
#pragma unroll
  for (int k = 0; k < N; k++) {
    faceFn[(long long)t * N + k] = fn[k];
    Fn[(long long)pl * N + k] = fn[k] * dAL;
  }
  if (ph.viscous) {
    double pf = (pr < 0) ? 0.0 : ldg_pen<D, double>(n, ph.pen_fact);
#pragma unroll
    for (int k = 0; k < N; k++) {
      double uc = (pr < 0) ? 0.5 * (UL[k] + UR[k]) : 0.5 * (UL[k] + UR[k]) - pf * (UL[k] - UR[k]);
      dUc[(long long)pl * N + k] = uc - UL[k];
      if (pr >= 0) dUc[(long long)pr * N + k] = uc - UR[k];
    }
  }
  if (pr >= 0) {
    double dAR = dA[pr];
#pragma unroll
    for (int k = 0; k < N; k++) Fn[(long long)pr * N + k] = -fn[k] * dAR;
  }
}

// correctGradU physical transform (src/solver.cpp:1105-1136).
template <int D>
__global__ __launch_bounds__(256) void k_grad_transform(long long nPts, long long SE, double* __restrict__ dU,
                                                         const double* __restrict__ JG, const double* __restrict__ detj) {
  constexpr int N = D + 2;
  long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nPts) return;
  double inv = 1. / detj[p];
  double G[D][D];
#pragma unroll
  for (int r = 0; r < D; r++)
#pragma unroll
    for (int x = 0; x < D; x++) G[r][x] = JG[(r * SE + p) * D + x];
#pragma unroll
  for (int k = 0; k < N; k++) {
    double r[D];
#pragma unroll
    for (int d = 0; d < D; d++) r[d] = dU[(d * SE + p) * N + k];
#pragma unroll
    for (int x = 0; x < D; x++) {
      double acc = r[0] * G[0][x] + r[1] * G[1][x];
      if constexpr (D == 3) acc += r[2] * G[2][x];
      dU[(x * SE + p) * N + k] = inv * acc;
    }
  }
}

// calcViscousFlux_spts (src/solver.cpp:777-811).
template <int D>
__global__ __launch_bounds__(256) void k_viscous_spts(long long nPts, long long SE, const double* __restrict__ U,
                                                       const double* __restrict__ dU, const double* __restrict__ JG,
                                                       double* __restrict__ Fs, Phys ph) {
  constexpr int N = D + 2;
  long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nPts) return;
  double u[N], g[D][N], Fv[D][N];
#pragma unroll
  for (int k = 0; k < N; k++) u[k] = U[p * N + k];
#pragma unroll
  for (int d = 0; d < D; d++)
#pragma unroll
    for (int k = 0; k < N; k++) g[d][k] = dU[(d * SE + p) * N + k];
  viscous_flux<D, double>(u, g, Fv, ph);
#pragma unroll
  for (int d = 0; d < D; d++)
#pragma unroll
    for (int k = 0; k < N; k++) Fv[d][k] += Fs[(d * SE + p) * N + k];
#pragma unroll
  for (int d1 = 0; d1 < D; d1++)
#pragma unroll
    for (int k = 0; k < N; k++) {
      double acc = 0.;
#pragma unroll
      for (int d2 = 0; d2 < D; d2++) acc += JG[(d1 * SE + p) * D + d2] * Fv[d2][k];
      Fs[(d1 * SE + p) * N + k] = acc;
    }
}

// Viscous face pass (face::calcViscousFlux, src/face.cpp:210-311).
template <int D>
__global__ __launch_bounds__(256) void k_face_viscous(int nFF, const int* __restrict__ ptL, const int* __restrict__ ptR,
                                                       const int* __restrict__ bcs, const double* __restrict__ Uf,
                                                       const double* __restrict__ dUf, long long FE,
                                                       const double* __restrict__ nrm, const double* __restrict__ dA,
                                                       double* __restrict__ Fn, const double* __restrict__ faceFn,
                                                       Phys ph) {
  constexpr int N = D + 2;
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nFF) return;
  int pl = ptL[t], pr = ptR[t], bc = bcs[t];
  double UL[N], UR[N], n[D], gL[D][N], gR[D][N], fn[N];
#pragma unroll
  for (int k = 0; k < N; k++) UL[k] = Uf[(long long)pl * N + k];
#pragma unroll
  for (int d = 0; d < D; d++) n[d] = nrm[(long long)pl * D + d];
#pragma unroll
  for (int d = 0; d < D; d++)
#pragma unroll
    for (int k = 0; k < N; k++) gL[d][k] = dUf[(d * FE + pl) * N + k];
  int bnd = 0;
  if (pr < 0) {
    apply_bc<D, double>(bc, UL, n, UR, ph);
    bnd = (bc == 10) ? 2 : 1;
    if (bnd == 2) {
#pragma unroll
      for (int d = 0; d < D; d++)
#pragma unroll
        for (int k = 0; k < N; k++) gR[d][k] = gL[d][k];
      adiabatic_grad<D, double>(UL, UR, gR, ph.gamma);
    } else {
#pragma unroll
      for (int d = 0; d < D; d++)
#pragma unroll
        for (int k = 0; k < N; k++) gR[d][k] = 0.;
    }
  } else {
#pragma unroll
    for (int k = 0; k < N; k++) UR[k] = Uf[(long long)pr * N + k];
#pragma unroll
    for (int d = 0; d < D; d++)
#pragma unroll
      for (int k = 0; k < N; k++) gR[d][k] = dUf[(d * FE + pr) * N + k];
  }
#pragma unroll
  for (int k = 0; k < N; k++) fn[k] = faceFn[(long long)t * N + k];
  ldg_flux<D, double>(bnd, UL, UR, gL, gR, n, fn, ph);
  double dAL = dA[pl];
#pragma unroll
  for (int k = 0; k < N; k++) Fn[(long long)pl * N + k] = fn[k] * dAL;
  if (pr >= 0) {
    double dAR = dA[pr];
#pragma unroll
    for (int k = 0; k < N; k++) Fn[(long long)pr * N + k] = -fn[k] * dAR;
  }
}

__global__ void k_sub(long long n, const double* __restrict__ a, double* __restrict__ b) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i] - b[i];
}

__global__ void k_copy(long long n, double* __restrict__ dst, const double* __restrict__ src) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// timeStepA / timeStepB (src/solver.cpp:470-536); dt from device memory.
__global__ __launch_bounds__(256) void k_time_step(long long nPts, int N, long long E, int modeB, double rk,
                                                    double* __restrict__ U, const double* __restrict__ U0,
                                                    const double* __restrict__ divF, const double* __restrict__ src,
                                                    const double* __restrict__ detj, const double* __restrict__ dtp,
                                                    const double* __restrict__ eleDt, int local) {
  long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nPts) return;
  double dt = local ? eleDt[p % E] : *dtp;
  double dj = detj[p];
  for (int k = 0; k < N; k++) {
    long long i = p * N + k;
    double r = src ? (divF[i] + src[i]) : divF[i];
    if (modeB) U[i] -= rk * r / dj * dt;
    else U[i] = U0[i] - rk * r / dj * dt;
  }
}

// calcWaveSpFpts for every element (src/ele.cpp:1097-1124).  waveSp is
// stored [fpt][ele].
template <int D>
__global__ void k_wave_speed(long long nPts, const double* __restrict__ Uf, const double* __restrict__ nrm,
                             double* __restrict__ waveSp, double gamma) {
  constexpr int N = D + 2;
  long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nPts) return;
  const double* U = Uf + p * N;
  double rho = U[0], u = U[1] / rho, v = U[2] / rho, w = 0;
  if constexpr (D == 3) w = U[3] / rho;
  double rhoVSq = rho * (u * u + v * v + w * w);
  double pr = (gamma - 1) * (U[D + 1] - 0.5 * rhoVSq);
  double vN = u * nrm[p * D] + v * nrm[p * D + 1];
  if constexpr (D == 3) vN += w * nrm[p * D + 2];
  double csq = smax(gamma * pr / rho, 0.0);
  waveSp[p] = (fabs(vN - 0.) + sqrt(csq));
}

// ele::calcDt for every element (src/ele.cpp:1127-1146) and a per-block
// minimum; a second kernel folds the block minima into *dtp.
__global__ __launch_bounds__(256) void k_ele_dt(int E, int S, int F, int Ff, const double* __restrict__ detj,
                                                 const double* __restrict__ ws, const double* __restrict__ wf,
                                                 const double* __restrict__ waveSp, const double* __restrict__ dA,
                                                 double cflfac, double* __restrict__ eleDt, double* __restrict__ part) {
  __shared__ double red[256];
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  double dt = INFINITY;
  if (e < E) {
    double vol = 0;
    for (int s = 0; s < S; s++) vol += detj[(long long)s * E + e] * ws[s];
    double iw = 0;
    for (int f = 0; f < F; f++) iw += waveSp[(long long)f * E + e] * wf[f % Ff] * dA[(long long)f * E + e];
    dt = cflfac * 2 * vol / iw;
    eleDt[e] = dt;
  }
  red[threadIdx.x] = dt;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void k_min_final(int n, const double* __restrict__ part, double* __restrict__ out) {
  __shared__ double red[256];
  double v = INFINITY;
  for (int i = threadIdx.x; i < n; i += blockDim.x) v = fmin(v, part[i]);
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

// Residual norm (ele::getNormResidual, src/ele.cpp:1922-1945): per element
// then per block; partial sums [block][N].
__global__ __launch_bounds__(256) void k_res_norm(int E, int S, int N, int type, const double* __restrict__ divF,
                                                   const double* __restrict__ detj, const double* __restrict__ ws,
                                                   double* __restrict__ part) {
  __shared__ double red[256];
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  for (int k = 0; k < N; k++) {
    double r = 0;
    if (e < E) {
      for (int s = 0; s < S; s++) {
        double v = divF[((long long)s * E + e) * N + k], dj = detj[(long long)s * E + e];
        if (type == 1) r += fabs(v) * ws[s];
        else if (type == 2) r += v * v / dj * ws[s];
        else r = fmax(fabs(v) / dj, r);
      }
    }
    red[threadIdx.x] = r;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
      if (threadIdx.x < s)
        red[threadIdx.x] = (type == 3) ? fmax(red[threadIdx.x], red[threadIdx.x + s]) : red[threadIdx.x] + red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[(long long)blockIdx.x * N + k] = red[0];
    __syncthreads();
  }
}

// NaN / non-positive density or pressure scan.
template <int D>
__global__ void k_check(long long nPts, const double* __restrict__ U, double gamma, int* flag) {
  constexpr int N = D + 2;
  long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nPts) return;
  const double* u = U + p * N;
  double q = 0;
  for (int d = 0; d < D; d++) q += u[d + 1] * u[d + 1];
  double pr = (gamma - 1) * (u[D + 1] - 0.5 * q / u[0]);
  bool bad = !(u[0] > 0) || !(pr > 0);
  for (int k = 0; k < N; k++) bad = bad || !isfinite(u[k]);
  if (bad) atomicOr(flag, 1);
}

// ---------------------------------------------------------------------------

void op_apply(Ctx& c, const char* name, int m, int k, const double* A, const double* B, double* C, bool beta) {
  opgemm(c, name, m, k, A, B, C, beta);
}

// out may alias a
__global__ void k_sub_arr(long long n, double* out, const double* a, const double* __restrict__ b) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    double v = a[i];
    v -= b[i];
    out[i] = v;
  }
}

__global__ void k_add_arr(long long n, double* __restrict__ y, const double* __restrict__ x) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] += x[i];
}

void sub_array(Ctx& c, double* out, const double* a, const double* b, long long n) {
  Launch L(c, "pmg_sub");
  hipLaunchKernelGGL(k_sub_arr, dim3(blocks_for(n, 256)), dim3(256), 0, c.stream, n, out, a, b);
  HIPCHK(hipGetLastError());
}

void add_array(Ctx& c, double* y, const double* x, long long n) {
  Launch L(c, "pmg_add");
  hipLaunchKernelGGL(k_add_arr, dim3(blocks_for(n, 256)), dim3(256), 0, c.stream, n, y, x);
  HIPCHK(hipGetLastError());
}

void copy_array(Ctx& c, double* dst, const double* src, long long n) {
  Launch L(c, "copy");
  hipLaunchKernelGGL(k_copy, dim3(blocks_for(n, 256)), dim3(256), 0, c.stream, n, dst, src);
  HIPCHK(hipGetLastError());
}

void staged_extrapolate_u(Ctx& c) { opgemm(c, "opgemm_extrapolate_u", c.z.F, c.z.S, c.sf, c.U, c.Uf, false); }

void staged_wave_speed(Ctx& c) {
  long long n = (long long)c.z.F * c.z.E;
  Launch L(c, "wave_speed");
  if (c.z.D == 2) hipLaunchKernelGGL(k_wave_speed<2>, dim3(blocks_for(n, 256)), dim3(256), 0, c.stream, n, c.Uf, c.norm, c.waveSp, c.ph.gamma);
  else hipLaunchKernelGGL(k_wave_speed<3>, dim3(blocks_for(n, 256)), dim3(256), 0, c.stream, n, c.Uf, c.norm, c.waveSp, c.ph.gamma);
  HIPCHK(hipGetLastError());
}

void staged_residual(Ctx& c, int step) {
  const Sizes& z = c.z;
  const long long nS = z.nS(), nF = z.nF(), SE = (long long)z.S * z.E, FE = (long long)z.F * z.E;
  const bool visc = c.prm.viscous != 0;
  double* divF = c.divF + step * nS;
  unsigned gS = blocks_for(SE, 256), gFF = blocks_for(z.nFaceFpts, 256);

  staged_extrapolate_u(c);
  if (visc)
    for (int d = 0; d < z.D; d++) opgemm(c, "opgemm_grad_u", z.S, z.S, c.grad + (long long)d * z.S * z.S, c.U, c.dU + d * nS, false);
  {
    Launch L(c, "inviscid_spts");
    if (z.D == 2) hipLaunchKernelGGL(k_inviscid_spts<2>, dim3(gS), dim3(256), 0, c.stream, SE, SE, c.U, c.jginv, c.Fs, c.ph.gamma, visc ? 0 : 1);
    else hipLaunchKernelGGL(k_inviscid_spts<3>, dim3(gS), dim3(256), 0, c.stream, SE, SE, c.U, c.jginv, c.Fs, c.ph.gamma, visc ? 0 : 1);
    HIPCHK(hipGetLastError());
  }
  if (z.nFaceFpts > 0) {
    Launch L(c, "face_inviscid");
    if (z.D == 2) hipLaunchKernelGGL(k_face_inviscid<2>, dim3(gFF), dim3(256), 0, c.stream, z.nFaceFpts, c.ft.ptL, c.ft.ptR, c.ft.bc, c.Uf, c.norm, c.da, c.Fn, c.faceFn, c.waveSp, c.dUc, c.ph);
    else hipLaunchKernelGGL(k_face_inviscid<3>, dim3(gFF), dim3(256), 0, c.stream, z.nFaceFpts, c.ft.ptL, c.ft.ptR, c.ft.bc, c.Uf, c.norm, c.da, c.Fn, c.faceFn, c.waveSp, c.dUc, c.ph);
    HIPCHK(hipGetLastError());
  }
  if (visc) {
    for (int d = 0; d < z.D; d++) opgemm(c, "opgemm_correct_grad_u", z.S, z.F, c.corru + (long long)d * z.S * z.F, c.dUc, c.dU + d * nS, true);
    {
      Launch L(c, "grad_transform");
      if (z.D == 2) hipLaunchKernelGGL(k_grad_transform<2>, dim3(gS), dim3(256), 0, c.stream, SE, SE, c.dU, c.jginv, c.detj);
      else hipLaunchKernelGGL(k_grad_transform<3>, dim3(gS), dim3(256), 0, c.stream, SE, SE, c.dU, c.jginv, c.detj);
      HIPCHK(hipGetLastError());
    }
    for (int d = 0; d < z.D; d++) opgemm(c, "opgemm_extrapolate_grad_u", z.F, z.S, c.sf, c.dU + d * nS, c.dUf + d * nF, false);
    {
      Launch L(c, "viscous_spts");
      if (z.D == 2) hipLaunchKernelGGL(k_viscous_spts<2>, dim3(gS), dim3(256), 0, c.stream, SE, SE, c.U, c.dU, c.jginv, c.Fs, c.ph);
      else hipLaunchKernelGGL(k_viscous_spts<3>, dim3(gS), dim3(256), 0, c.stream, SE, SE, c.U, c.dU, c.jginv, c.Fs, c.ph);
      HIPCHK(hipGetLastError());
    }
    if (z.nFaceFpts > 0) {
      Launch L(c, "face_viscous");
      if (z.D == 2) hipLaunchKernelGGL(k_face_viscous<2>, dim3(gFF), dim3(256), 0, c.stream, z.nFaceFpts, c.ft.ptL, c.ft.ptR, c.ft.bc, c.Uf, c.dUf, FE, c.norm, c.da, c.Fn, c.faceFn, c.ph);
      else hipLaunchKernelGGL(k_face_viscous<3>, dim3(gFF), dim3(256), 0, c.stream, z.nFaceFpts, c.ft.ptL, c.ft.ptR, c.ft.bc, c.Uf, c.dUf, FE, c.norm, c.da, c.Fn, c.faceFn, c.ph);
      HIPCHK(hipGetLastError());
    }
  }
  // extrapolateNormalFlux (static grid), calcDivF_spts, correctDivFlux
  for (int d = 0; d < z.D; d++) opgemm(c, "opgemm_extrapolate_fn", z.F, z.S, c.efn + (long long)d * z.F * z.S, c.Fs + d * nS, c.disFn, d > 0);
  for (int d = 0; d < z.D; d++) opgemm(c, "opgemm_div_f", z.S, z.S, c.grad + (long long)d * z.S * z.S, c.Fs + d * nS, divF, d > 0);
  {
    Launch L(c, "sub_fn");
    hipLaunchKernelGGL(k_sub, dim3(blocks_for(nF, 256)), dim3(256), 0, c.stream, nF, c.Fn, c.disFn);
    HIPCHK(hipGetLastError());
  }
  opgemm(c, "opgemm_correction", z.S, z.F, c.corr, c.disFn, divF, true);
}

void staged_time_step_a(Ctx& c, int step, double rk, bool src) {
  long long SE = (long long)c.z.S * c.z.E;
  Launch L(c, "time_step_a");
  hipLaunchKernelGGL(k_time_step, dim3(blocks_for(SE, 256)), dim3(256), 0, c.stream, SE, c.z.N, (long long)c.z.E, 0, rk, c.U,
                     c.U0, c.divF + step * c.z.nS(), src ? c.src : nullptr, c.detj, c.d_dt, c.eleDt, c.prm.dt_type == 2 ? 1 : 0);
  HIPCHK(hipGetLastError());
}

void staged_time_step_b(Ctx& c, int step, bool src) {
  long long SE = (long long)c.z.S * c.z.E;
  Launch L(c, "time_step_b");
  hipLaunchKernelGGL(k_time_step, dim3(blocks_for(SE, 256)), dim3(256), 0, c.stream, SE, c.z.N, (long long)c.z.E, 1,
                     c.prm.rk_b[step], c.U, c.U0, c.divF + step * c.z.nS(), src ? c.src : nullptr, c.detj, c.d_dt, c.eleDt,
                     c.prm.dt_type == 2 ? 1 : 0);
  HIPCHK(hipGetLastError());
}

static double cfl_limit(int order) {  // getCFLLimit, src/global.cpp:81-105
  static const double t[6] = {1.393, 0.464, 0.235, 0.1453, 0.100, 0.0736};
  if (order < 0 || order > 5) throw std::invalid_argument("CFL limit not available for this order!");
  return t[order];
}

void calc_dt(Ctx& c, bool refresh) {
  if (refresh) {
    staged_extrapolate_u(c);
    staged_wave_speed(c);
  }
  const Sizes& z = c.z;
  unsigned nb = blocks_for(z.E, 256);
  {
    Launch L(c, "ele_dt");
    hipLaunchKernelGGL(k_ele_dt, dim3(nb), dim3(256), 0, c.stream, z.E, z.S, z.F, z.Ff, c.detj, c.ws, c.wf, c.waveSp, c.da,
                       c.prm.cfl * cfl_limit(z.P), c.eleDt, c.partial);
    HIPCHK(hipGetLastError());
  }
  Launch L(c, "dt_min");
  hipLaunchKernelGGL(k_min_final, dim3(1), dim3(256), 0, c.stream, (int)nb, c.partial, c.d_dt);
  HIPCHK(hipGetLastError());
}

void residual_norm(Ctx& c, int type, double* out) {
  const Sizes& z = c.z;
  unsigned nb = blocks_for(z.E, 256);
  hipLaunchKernelGGL(k_res_norm, dim3(nb), dim3(256), 0, c.stream, z.E, z.S, z.N, type, c.divF, c.detj, c.ws, c.partial);
  HIPCHK(hipGetLastError());
  std::vector<double> h((size_t)nb * z.N);
  HIPCHK(hipMemcpyAsync(h.data(), c.partial, h.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));
  for (int k = 0; k < z.N; k++) {
    double r = 0;
    for (unsigned b = 0; b < nb; b++) r = (type == 3) ? std::max(r, h[(size_t)b * z.N + k]) : r + h[(size_t)b * z.N + k];
    out[k] = (type == 2) ? std::sqrt(r) : r;
  }
}

// ---------------------------------------------------------------------------
// Residual norm without allocation, for callers that cannot call hipMalloc.
// The per-block partial sums and the final N values live in `scratch`, a
// file-scope device array of kNoAllocScratch doubles that every GPU the
// library runs on holds until the process ends.  Host code never passes the
// array's name to a kernel (on the host it is a stand-in, not a device
// address): the device address of each GPU's copy is looked up once, with
// that GPU current, and kept in g_scratch_ptr.  Calls on one GPU share its
// copy, so calls that overlap on one GPU race; nothing in it has to start at
// zero (every value read was written earlier in the same call).
static __device__ __align__(256) double scratch[kNoAllocScratch];
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kNoAllocMaxDevices] = {};

double* noalloc_scratch(int device) {
  if (device < 0 || device >= kNoAllocMaxDevices) throw std::invalid_argument("device ordinal out of range");
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    int cur = -1;
    HIPCHK(hipGetDevice(&cur));
    if (cur != device) HIPCHK(hipSetDevice(device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);
    HIPCHK(e);
    g_scratch_ptr[device] = static_cast<double*>(p);
  }
  return g_scratch_ptr[device];
}

// The host sum of residual_norm done on the device: one thread per field
// walks the block partials in ascending block order, as the host loop does,
// so both give the same bits.
__global__ void k_res_final(int nb, int N, int type, const double* __restrict__ part, double* __restrict__ out) {
  const int k = threadIdx.x;
  if (k >= N) return;
  double r = 0;
  for (int b = 0; b < nb; b++) r = (type == 3) ? fmax(r, part[(long long)b * N + k]) : r + part[(long long)b * N + k];
  out[k] = (type == 2) ? sqrt(r) : r;
}

void residual_norm_noalloc(Ctx& c, int type, double* out) {
  const Sizes& z = c.z;
  const unsigned nb = blocks_for(z.E, 256);
  const long long need = ((long long)nb + 1) * z.N;
  if (need > kNoAllocScratch)
    throw std::length_error("fhip_residual_norm_noalloc: " + std::to_string(z.E) + " elements need " +
                            std::to_string(need) + " doubles of scratch, more than its maximum of " +
                            std::to_string(kNoAllocScratch) + " (at most " +
                            std::to_string((kNoAllocScratch / z.N - 1) * 256) + " elements)");
  double* part = noalloc_scratch(c.device);
  double* fin = part + (long long)nb * z.N;
  hipLaunchKernelGGL(k_res_norm, dim3(nb), dim3(256), 0, c.stream, z.E, z.S, z.N, type, c.divF, c.detj, c.ws, part);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_res_final, dim3(1), dim3(64), 0, c.stream, (int)nb, z.N, type, part, fin);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, fin, (size_t)z.N * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));
}

int check_state(Ctx& c) {
  long long SE = (long long)c.z.S * c.z.E;
  int zero = 0, h = 0;
  HIPCHK(hipMemcpyAsync(c.flag, &zero, sizeof(int), hipMemcpyHostToDevice, c.stream));
  if (c.z.D == 2) hipLaunchKernelGGL(k_check<2>, dim3(blocks_for(SE, 256)), dim3(256), 0, c.stream, SE, c.U, c.ph.gamma, c.flag);
  else hipLaunchKernelGGL(k_check<3>, dim3(blocks_for(SE, 256)), dim3(256), 0, c.stream, SE, c.U, c.ph.gamma, c.flag);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(&h, c.flag, sizeof(int), hipMemcpyDeviceToHost, c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));
  return h;
}

}  // namespace fhip
