// Fused tensor-product paths: FP64 ("fused") and mixed precision ("mixed").
//
// FlurryPP applies its FR operators as dense GEMMs (src/solver.cpp:556-1158),
// but on quads and hexes every operator row is non-zero only along one line
// of solution points: the Lagrange factors of the other directions are
// evaluated at their own nodes and are exactly 0 or 1.  These paths read the
// dense operators handed to fhip_create, keep only their non-zero entries
// (at most P+1 per row for extrapolation and gradient, one per face for the
// correction) and apply them inside fused kernels, so the sums are the dense
// products minus terms that are exactly zero.
//
// Data live element-major on the device ([ele][field][point]) so that a block
// of elements is one contiguous slab.  Per RK stage:
//   face_inv   common inviscid flux times dA on both sides of every face
//              flux point (+ LDG common U, dUc)
//   volume     [gradient + correction + physical transform], inviscid (and
//              viscous) flux, transformed flux, divergence, discontinuous
//              normal flux at the faces [and each side's part of the LDG
//              viscous flux at its flux points]
//   face_bnd   LDG viscous flux at boundary flux points (viscous runs with
//              boundaries only)
//   rk         correction of the divergence (common flux = inviscid part +
//              both sides' LDG parts), RK update, U at the faces of the next
//              stage
//
// Every kernel is templated on R, the type of the face and volume
// intermediates and of the flux arithmetic: double for the FP64 path, float
// for the mixed path.  The solution, its RK stage copies, the metric terms,
// the wave speeds, dt and the residual array stay FP64 in both (the mixed
// path also keeps FP32 roundings of the geometry and of U for its kernels).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <map>
#include <mutex>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "fhip_internal.hpp"

namespace fhip {

#define FCHK(x)                                                                               \
  do {                                                                                        \
    hipError_t e_ = (x);                                                                      \
    if (e_ != hipSuccess) throw std::runtime_error(std::string("HIP: ") + hipGetErrorString(e_)); \
  } while (0)

static inline unsigned nblk(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

template <int D, int P>
struct Shape {
  static constexpr int P1 = P + 1;
  static constexpr int Ff = (D == 2) ? P1 : P1 * P1;
  static constexpr int S = (D == 2) ? P1 * P1 : P1 * P1 * P1;
  static constexpr int F = 2 * D * Ff;
  static constexpr int N = D + 2;
  static constexpr int NC = 2 * D;  // correction entries per solution point
};

// Element kernels take a runtime block of NT >= S threads holding EB = NT/S
// elements, with their LDS slabs in dynamic shared memory; the block size of
// each kernel is chosen per context (default 256, or by fused_autotune).
extern __shared__ __align__(16) unsigned char fhip_lds[];
constexpr int kMaxBlock = 512;
constexpr size_t kMaxLds = 64 * 1024;
constexpr long long kStreamBytes = 64ll << 20;
enum { kExtrap = 0, kFaceInv = 1, kVolume = 2, kFaceBnd = 3, kRk = 4 };

// Sparse operator tables (device pointers into one allocation).  Indices of
// solution points are positions in the swizzled LDS layout (spt_lds below).
struct Tabs {
  const int* ext_idx;   // [F][P1]    solution points feeding flux point f
  const double* ext_w;  // [F][P1]    opp_spts_to_fpts entries
  const int* efn_dir;   // [F]        the one direction d with opp_extrapolate_fn[d][f] != 0
  const double* efn_w;  // [F][P1]    opp_extrapolate_fn[efn_dir[f]] entries (same points)
  const int* grd_idx;   // [D][S][P1]
  const double* grd_w;  // [D][S][P1] opp_grad_spts entries
  const int* cor_idx;   // [S][NC]    flux points correcting solution point s
  const double* cor_w;  // [S][NC]    opp_correction entries
  const int* cu_idx;    // [D][S][2]  flux points with opp_correct_u[d][s][f] != 0
  const double* cu_w;   // [D][S][2]  their opp_correct_u entries
  // the same weights rounded to FP32 once, for the mixed path
  const float *ext_wf, *efn_wf, *grd_wf, *cor_wf, *cu_wf;
  // opp_grad_spts as one 1-D derivative matrix d1[i][m] (derivative of basis
  // m at point i) shared by every direction and line, FP64 and FP32; null
  // where the operator is not of that form (the fused paths then do not take
  // P3 hexes, whose FP64 volume kernel contracts with it on the matrix cores)
  const double* d1;
  const float* d1f;
};

// The 1-D contractions of FP64 P3 hexes along x and z run on the matrix
// cores (v_mfma_f64_4x4x4f64: four independent 4x4x4 products per wave).  One
// element per wave, lane L = solution point x + 4y + 16z.  The instruction
// takes A[i][k] of block b from lane 16k + 4b + i, B[k][j] from lane
// 16k + 4b + j, and returns D[i][j] = C[i][j] + sum_k A[i][k] B[k][j] in lane
// 16i + 4b + j (measured lane layout).  With b = y:
//   z: A = d1[L % 4][L / 16], B = the lane's own point: D is the z-derivative
//      at the lane's own point;
//   x: A = the point (x, z) = (L / 16, L % 4) of the lane's row y, B = the
//      same operator entry: D is the x-derivative at the lane's own point.
// C carries the terms added on the vector unit (the gradient's correction,
// the divergence's other directions).  y stays on the vector unit.
template <int D, int P, class R>
constexpr bool kMfma = D == 3 && P == 3 && sizeof(R) == 8;

__device__ __forceinline__ double mfma444(double a, double b, double c) {
  return __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ float mfma444(float, float, float c) { return c; }  // never taken (kMfma)

// The mixed path's z contractions of P3 hexes on the matrix cores in BF16
// (v_mfma_f32_16x16x32_bf16, FP32 accumulation).  Both factors are split into
// BF16 pieces: x = p0 + p1 + p2 with p0 = bf16(x), p1 = bf16(x - p0),
// p2 = bf16(x - p0 - p1).  With 2 pieces (hi+lo) the product takes the terms
// d0 v0 + d0 v1 + d1 v0, with 3 also d0 v2 + d2 v0 + d1 v1 (the rest are
// below FP32 rounding).  One element per wave, lane L = point xy + 16 z
// (xy = L % 16).  The instruction takes A[i][k] from lane 16(k/8) + i, element
// k % 8, B[k][j] from lane 16(k/8) + j, element k % 8, and returns D[i][j] in
// lane 16(i/4) + j, register i % 4 (measured lane layout).  With j = xy, the
// rows i = 4z of A hold the pieces of d1[z][m] and the column xy of B the
// pieces of v(xy, m), m = 0..3, eight K slots per lane group g = L / 16:
//   2 pieces: g0: (d0 | d0) x (v0 | v1), g1: (d1 | 0) x (v0 | 0)
//   3 pieces: g0: (d0 | d0) x (v0 | v1), g1: (d1 | d0) x (v0 | v2),
//             g2: (d2 | d1) x (v0 | v1)
// so register 0 of each lane is the z contraction at the lane's own point.
__host__ __device__ constexpr int spt_lds(int D, int P1, int s);  // the slab position of point s (below)

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void bf16_pieces(float x, __bf16 p[3]) {
  p[0] = (__bf16)x;
  const float r1 = x - float(p[0]);
  p[1] = (__bf16)r1;
  p[2] = (__bf16)(r1 - float(p[1]));
}

// Which pieces lane group g puts in its two halves of eight K slots (-1: zero).
__device__ __forceinline__ void bf16_plan(int np, int g, int& da, int& db, int& va, int& vb) {
  da = db = va = vb = -1;
  if (g == 0) da = 0, db = 0, va = 0, vb = 1;
  if (g == 1) da = 1, db = (np == 3 ? 0 : -1), va = 0, vb = (np == 3 ? 2 : -1);
  if (g == 2 && np == 3) da = 2, db = 1, va = 0, vb = 1;
}

// A operand of lane L: the pieces of row z = (L % 16) / 4 of d1 for rows
// i = L % 16 with i % 4 == 0, else zero.
__device__ __forceinline__ bf16x8 bf16_opnd_a(const float* d1, int lane, int np) {
  bf16x8 a;
  for (int e = 0; e < 8; e++) a[e] = (__bf16)0.f;
  const int i = lane % 16, g = lane / 16;
  if (i % 4 != 0) return a;
  int da, db, va, vb;
  bf16_plan(np, g, da, db, va, vb);
  for (int m = 0; m < 4; m++) {
    __bf16 p[3];
    bf16_pieces(d1[(i / 4) * 4 + m], p);
    if (da >= 0) a[m] = p[da];
    if (db >= 0) a[4 + m] = p[db];
  }
  return a;
}

// The z contraction at the lane's own point plus c: v is this field's slab
// (swizzled LDS positions), column xy = lane % 16.
template <int D, int P1>
__device__ __forceinline__ float mfma_z_bf16(const float* v, int lane, bf16x8 a, float c, int np) {
  const int xy = lane % 16, g = lane / 16;
  int da, db, va, vb;
  bf16_plan(np, g, da, db, va, vb);
  bf16x8 bv;
  for (int e = 0; e < 8; e++) bv[e] = (__bf16)0.f;
  if (va >= 0) {
    for (int m = 0; m < 4; m++) {
      __bf16 p[3];
      bf16_pieces(v[spt_lds(D, P1, xy + 16 * m)], p);
      bv[m] = p[va];
      if (vb >= 0) bv[4 + m] = p[vb];
    }
  }
  const f32x4 cc = {c, 0.f, 0.f, 0.f};
  const f32x4 d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bv, cc, 0, 0, 0);
  return d[0];
}

template <int D, int P, class R>
constexpr bool kBf16 = D == 3 && P == 3 && sizeof(R) == 4;


// Operator weights in the kernels' arithmetic type.
template <class R>
__device__ __forceinline__ const R* tw(const double* d, const float* f);
template <>
__device__ __forceinline__ const double* tw<double>(const double* d, const float*) { return d; }
template <>
__device__ __forceinline__ const float* tw<float>(const double*, const float* f) { return f; }

// Position of solution point s = i + P1*(j + P1*k) in an element's LDS slab.
// Hexes with an even P1: each (j,k) row is rotated by j+k and each k plane by
// k, so that the P1^2 points a face's flux points read at one depth fall into
// distinct LDS banks for all three face orientations (row-major order would
// give 4-way bank conflicts on the x and y faces at P=3).  With an odd P1 (rows
// of an odd number of doubles) and on quads, row-major order spreads the
// slab's reads over the banks at least as well (the rotation measured 48% of
// the P4 volume kernel's LDS cycles as bank conflicts).  A bijection on every
// plane.
__host__ __device__ constexpr int spt_lds(int D, int P1, int s) {
  if (D == 2 || P1 % 2 == 1) return s;
  const int i = s % P1, j = (s / P1) % P1, k = s / (P1 * P1);
  return (k * P1 + (j + k) % P1) * P1 + (i + k) % P1;
}

// Face and volume intermediates of one precision.
template <class R>
struct Work {
  R *divFe = nullptr, *Ufe = nullptr, *Fne = nullptr, *disFne = nullptr, *dUce = nullptr;
  // viscous runs: each flux point's side of the LDG flux [ele][N][F] (disFne
  // then holds that side's part of the flux jump, +-B dA - disFn), and the
  // gradient at the boundary flux points [D*N][nBnd]
  R *Bve = nullptr, *gBnd = nullptr;
  R* Rst = nullptr;  // mixed path: residual of each classical RK stage but the last
  // mixed path: U at the solution points rounded to FP32, the only form in
  // which the flux kernels read it; the RK stages inside a step keep their
  // states here and write the FP64 state only at the step's last stage
  R* Us = nullptr;
  // geometry in the kernels' arithmetic type: the FP64 arrays themselves, or
  // FP32 copies rounded once (the values the FP32 kernels would convert to)
  const R *JG = nullptr, *iJ = nullptr, *nrm = nullptr, *dA = nullptr, *nf = nullptr;
  void* geo_mem = nullptr;
  bool ready = false;
  void release() {
    void* ps[] = {divFe, Ufe, Fne, disFne, dUce, Bve, gBnd, Rst, Us, geo_mem};
    for (void* p : ps)
      if (p) (void)hipFree(p);
    *this = Work<R>();
  }
};

struct FusedState {
  Tabs tb{};
  void* tab_mem = nullptr;
  // FP64 element-major state, metrics and wave speeds (both paths)
  double *Ue = nullptr, *U0e = nullptr, *Uacc = nullptr, *divF0e = nullptr, *JGe = nullptr, *detJe = nullptr,
         *iJe = nullptr, *norme = nullptr, *dAe = nullptr, *waveSpe = nullptr;
  int *pL = nullptr, *pR = nullptr, *bc = nullptr;
  // per element flux point, 4 * (partner flux point e*F + f) + side (0 left,
  // 1 right) or 4 * boundary slot + 2; viscous runs: the face's normal (the
  // left element's) at every element flux point, boundary face points by slot
  int *fpart = nullptr, *bndT = nullptr;
  double* nfe = nullptr;
  int nBnd = 0;
  // batching: boundary-state table, element -> mesh map, per-mesh dt and activity
  int nMesh = 1;
  std::vector<int> h_meshStart;
  double *bst = nullptr, *meshDt = nullptr, *meshRes = nullptr;
  unsigned* dtTicket = nullptr;  // blocks of the dt kernel that have finished (zero between launches)
  double* vole = nullptr;        // element volumes, sum_s detJ w_s (calcDt's numerator)
  int *eleMesh = nullptr, *meshStart = nullptr, *active = nullptr;
  // P-multigrid source term (element-major copy of src_spts), used while use_src
  double* srce = nullptr;
  bool use_src = false;
  // the stage being formed reads its state from the stage copy Us (the
  // stages after the first of an RK step); otherwise from the state itself.
  // FP32 kernels always read the FP32 copy.
  bool read_stage = false;
  // stream once-touched arrays past the caches: set when U alone (FP64,
  // element-major) exceeds kStreamBytes, i.e. a stage's arrays are many times
  // the 256 MiB infinity cache
  bool stream = false;
  // block size of each kernel (kExtrap..kRk) for FP64 [0] and mixed [1]
  int nt[2][5] = {{256, 256, 256, 256, 256}, {256, 256, 256, 256, 256}};
  Work<double> w64;
  Work<float> w32;
  ~FusedState() {
    w64.release();
    w32.release();
    void* ps[] = {tab_mem, Ue,  U0e,     Uacc,    divF0e,    JGe,    detJe,  iJe,   norme, dAe, waveSpe, pL,
                  pR,      bc,  fpart,   bndT,    nfe,       bst,    meshDt, meshRes, eleMesh, meshStart, active, srce,
                  dtTicket, vole};
    for (void* p : ps)
      if (p) (void)hipFree(p);
  }
  template <class R>
  Work<R>& work();
};
template <>
Work<double>& FusedState::work<double>() { return w64; }
template <>
Work<float>& FusedState::work<float>() { return w32; }

// ---------------------------------------------------------------------------
// Layout conversions between FlurryPP's [pt][ele][k] and element-major
// [ele][k][pt]: both are the transpose of a matrix with one row per point,
// [P][E*K] <-> [E*K][P].  32x32 tiles through LDS so that reads and writes
// are both coalesced; tiles are numbered on a 1-D grid.
constexpr int kTile = 32;
__global__ __launch_bounds__(256) void k_transpose(long long R, long long C, const double* __restrict__ in,
                                                   double* __restrict__ out) {
  __shared__ double t[kTile][kTile + 1];
  const long long tilesC = (C + kTile - 1) / kTile;
  const long long r0 = (blockIdx.x / tilesC) * kTile, c0 = (blockIdx.x % tilesC) * kTile;
  const int tx = threadIdx.x % kTile, ty = threadIdx.x / kTile;  // ty 0..7
#pragma unroll
  for (int j = 0; j < kTile; j += 8) {
    const long long r = r0 + ty + j, c = c0 + tx;
    if (r < R && c < C) t[ty + j][tx] = in[r * C + c];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kTile; j += 8) {
    const long long c = c0 + ty + j, r = r0 + tx;
    if (r < R && c < C) out[c * R + r] = t[tx][ty + j];
  }
}

// in [P][E][K] -> out [E][K][P]
static void to_em(hipStream_t st, int P, long long E, int K, const double* in, double* out) {
  const long long R = P, C = E * K;
  const long long tiles = ((R + kTile - 1) / kTile) * ((C + kTile - 1) / kTile);
  hipLaunchKernelGGL(k_transpose, dim3((unsigned)tiles), dim3(256), 0, st, R, C, in, out);
  FCHK(hipGetLastError());
}
// in [E][K][P] -> out [P][E][K]
static void from_em(hipStream_t st, int P, long long E, int K, const double* in, double* out) {
  const long long R = E * K, C = P;
  const long long tiles = ((R + kTile - 1) / kTile) * ((C + kTile - 1) / kTile);
  hipLaunchKernelGGL(k_transpose, dim3((unsigned)tiles), dim3(256), 0, st, R, C, in, out);
  FCHK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// U at the flux points from U at the solution points (opp_spts_to_fpts rows).
template <int D, int P, class R>
__global__ __launch_bounds__(kMaxBlock) void k_extrap(int E, const double* __restrict__ Ue, R* __restrict__ Ufe,
                                                       R* __restrict__ Us, Tabs tb) {
  using Sh = Shape<D, P>;
  constexpr int S = Sh::S, F = Sh::F, N = Sh::N, P1 = Sh::P1;
  const int EB = blockDim.x / S;
  R* const sU = reinterpret_cast<R*>(fhip_lds);
  const int e0 = blockIdx.x * EB, nb = min(EB, E - e0);
  const double* src = Ue + (long long)e0 * N * S;
  // all loads first (at most N per thread), then the stores
  R v[N];
#pragma unroll
  for (int u = 0; u < N; u++) {
    const int i = threadIdx.x + u * blockDim.x;
    if (i < nb * N * S) v[u] = R(src[i]);
  }
#pragma unroll
  for (int u = 0; u < N; u++) {
    const int i = threadIdx.x + u * blockDim.x;
    if (i < nb * N * S) {
      const int r = i / S, q = i - r * S;
      sU[r * S + spt_lds(D, P1, q)] = v[u];
      if (Us) Us[(long long)e0 * N * S + i] = v[u];
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nb * F; i += blockDim.x) {
    int b = i / F, f = i - b * F;
#pragma unroll
    for (int k = 0; k < N; k++) {
      R acc = 0;
#pragma unroll
      for (int m = 0; m < P1; m++) acc += tw<R>(tb.ext_w, tb.ext_wf)[f * P1 + m] * sU[(b * N + k) * S + tb.ext_idx[f * P1 + m]];
      Ufe[((long long)(e0 + b) * N + k) * F + f] = acc;
    }
  }
}

// Loads and stores of arrays a kernel touches once per launch.  On meshes
// whose arrays are far larger than the caches (NT) they stream past them
// (nontemporal); on smaller ones they stay plain, since there the next kernel
// of the stage finds them in L2 or the infinity cache (measured both ways).
// A kernel streams only where the rows it stores are whole 64-byte blocks
// (Runner::kFaceRowsWhole, kEleRowsWhole).
template <bool NT, class T>
__device__ __forceinline__ T ld_s(const T* p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}
template <bool NT, class T>
__device__ __forceinline__ void st_s(T v, T* p) {
  if constexpr (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}
// U at a face flux point, read once per face kernel launch: streamed in FP64
// only; FP32 loads stay cached, since a 128-byte line then spans more face
// points that other waves read (measured slower when streamed).
template <bool NT, class R>
__device__ __forceinline__ R ld_face(const R* p) {
  return ld_s<NT && sizeof(R) == 8>(p);
}
// A product rounded on its own (never contracted into a following add).
__device__ __forceinline__ double mul_rn(double a, double b) { return __dmul_rn(a, b); }
__device__ __forceinline__ float mul_rn(float a, float b) { return __fmul_rn(a, b); }

// ---------------------------------------------------------------------------
// Common inviscid flux at face flux points; pL/pR are ele*F+fpt.
// Boundary state of a face flux point: the packed code holds the BC number in
// its low 4 bits and (row of the boundary-state table + 1) above them.
template <int D, class R>
__device__ __forceinline__ void bc_state(int code, const R* UL, const R* n, R* UR, const Phys& ph,
                                         const double* __restrict__ bst) {
  const int bc = code & 15, row = (code >> 4) - 1;
  if (row < 0) {
    apply_bc<D, R>(bc, UL, n, UR, ph);
  } else {
    Phys q = ph;
    q.rho_bound = bst[row * 5 + 0];
    q.u_bound = bst[row * 5 + 1];
    q.v_bound = bst[row * 5 + 2];
    q.w_bound = bst[row * 5 + 3];
    q.p_bound = bst[row * 5 + 4];
    apply_bc<D, R>(bc, UL, n, UR, q);
  }
}

template <int D, int P, class R, bool NT>
__global__ __launch_bounds__(kMaxBlock) void k_face_inv(int nFF, const int* __restrict__ pLa, const int* __restrict__ pRa,
                                                   const int* __restrict__ bca, const R* __restrict__ Ufe,
                                                   const R* __restrict__ norme,
                                                   R* __restrict__ Fne, double* __restrict__ waveSpe,
                                                   R* __restrict__ dUce, Phys ph,
                                                   const double* __restrict__ bst) {
  using Sh = Shape<D, P>;
  constexpr int F = Sh::F, N = Sh::N;
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nFF) return;
  const int pl = ld_s<NT>(&pLa[t]), pr = ld_s<NT>(&pRa[t]),
            bc = ld_s<NT>(&bca[t]);
  const long long eL = pl / F, fL = pl - eL * F;
  const long long bL = eL * N * F + fL;
  R UL[N], UR[N], n[D], fn[N];
#pragma unroll
  for (int k = 0; k < N; k++) UL[k] = ld_face<NT>(&Ufe[bL + k * F]);
#pragma unroll
  for (int d = 0; d < D; d++) n[d] = ld_s<NT>(&norme[(eL * D + d) * F + fL]);
  long long bR = 0;
  if (pr < 0) {
    bc_state<D, R>(bc, UL, n, UR, ph, bst);
  } else {
    const long long eR = pr / F, fR = pr - eR * F;
    bR = eR * N * F + fR;
#pragma unroll
    for (int k = 0; k < N; k++) UR[k] = ld_face<NT>(&Ufe[bR + k * F]);
  }
  if (pr < 0 || ph.riemann_type == 0)
    st_s<NT>(double(rusanov<D, R>(UL, UR, n, fn, R(ph.gamma))), &waveSpe[pl]);
  else {
    if constexpr (D == 2) roe2d<R>(UL, UR, n, fn, R(ph.gamma));
  }
  // the common inviscid flux once per point pair, at its left point; the
  // kernel that applies the correction scales it by each side's dA (negated
  // on the right), and viscous runs add the LDG flux there, where both
  // sides' parts meet
#pragma unroll
  for (int k = 0; k < N; k++) st_s<NT>(fn[k], &Fne[bL + k * F]);
  if (ph.viscous) {
    R pf = (pr < 0) ? R(0) : ldg_pen<D, R>(n, R(ph.pen_fact));
#pragma unroll
    for (int k = 0; k < N; k++) {
      R uc = (pr < 0) ? R(0.5) * (UL[k] + UR[k]) : R(0.5) * (UL[k] + UR[k]) - pf * (UL[k] - UR[k]);
      st_s<NT>(uc - UL[k], &dUce[bL + k * F]);
      if (pr >= 0) st_s<NT>(uc - UR[k], &dUce[bR + k * F]);
    }
  }
}

// ---------------------------------------------------------------------------
// Correction of the divergence (correctDivFlux), RK update (timeStepA/B
// folded into one pass per stage, FP64) and U at the flux points for the
// next stage.  mode: 0 first classical stage, 1 middle classical stage,
// 2 last classical stage, 3 single-stage forward Euler, 4 Jameson-form stage,
// 6 first Jameson-form stage (which also keeps U as the step's U0 where a
// separate U0 is kept); 5, 7 and 8 the residual alone (see rk_tail_p).
// Classical stages combine into U0 - sum_j b_j r_j / detJ dt in timeStepB's
// order (src/solver.cpp:504-536): either through the FP64 accumulator Uacc, or
// (mixed path, Rst given) from the stored stage residuals, which FP32 holds
// exactly there since they are sums of FP32 values; both give the same bits.
struct RkB {
  double b[8];
};
// The RK update's inputs and outputs besides the flux jump and divF: the
// stage's mode and coefficients, dt (0 global, 1 per element, 2 per mesh),
// the FP64 state and its stage copies, batching and multigrid arrays.
template <class R>
struct RkArgs {
  int mode;
  double a, bcoef;
  const double* detJe;
  const double* dtp;
  const double* eleDt;
  int local;
  double* Ue;
  double* U0e;
  double* Uacc;
  double* divF0e;
  const int* eleMesh;
  const double* meshDt;
  const int* active;
  const double* srce;
  R* Rst;
  int stage;
  RkB bs;
  R* Us;
  int fin;
  R* Ufe;
};

// correctDivFlux, timeStepA/B and U at the flux points for the next stage, for
// the block's elements: dv holds this thread's divF (one value per field),
// sD the flux jump at the block's flux points, and sU receives the new state
// (its previous contents are no longer read).  Ends with every thread of the
// block past a barrier unless mode is 5 (residual only), which returns at
// once for the whole block.
// The arrays come in as __restrict__ parameters (from RkArgs at the call
// below), so that the compiler may keep their loads ahead of the stores.
template <int D, int P, class R, bool NT>
__device__ __forceinline__ void rk_tail_p(int E, const R* dv, R* sU, const R* sD, int mode, double a, double bcoef,
                                          const double* __restrict__ detJe, const double* __restrict__ dtp,
                                          const double* __restrict__ eleDt, int local, double* __restrict__ Ue,
                                          double* __restrict__ U0e, double* __restrict__ Uacc,
                                          double* __restrict__ divF0e, const int* __restrict__ eleMesh,
                                          const double* __restrict__ meshDt, const int* __restrict__ active,
                                          const double* __restrict__ srce, R* __restrict__ Rst, int stage,
                                          const RkB& bs, R* __restrict__ Us, int fin, R* __restrict__ Ufe,
                                          const Tabs& tb) {
  using Sh = Shape<D, P>;
  constexpr int S = Sh::S, F = Sh::F, N = Sh::N, P1 = Sh::P1, NC = Sh::NC;
  const int EB = blockDim.x / S;
  const long long nS = (long long)E * N * S;
  const int e0 = blockIdx.x * EB, nb = min(EB, E - e0);
  const int b = threadIdx.x / S, s = threadIdx.x - b * S;
  if (mode == 5 || mode == 7 || mode == 8) {
    // residual only (calcResidual(0) + correctDivFlux), no update: 5 into
    // divF_spts[0]; 7 subtracted in place from the array divF0e points to
    // (the P-multigrid source); 8 plus the source into divF_spts[0]
    if (b < nb) {
      const long long e = e0 + b;
#pragma unroll
      for (int k = 0; k < N; k++) {
        R c = 0;
#pragma unroll
        for (int j = 0; j < NC; j++) c += tw<R>(tb.cor_w, tb.cor_wf)[s * NC + j] * sD[(b * N + k) * F + tb.cor_idx[s * NC + j]];
        const double r0 = double(c + dv[k]);
        const long long iu = (e * N + k) * S + s;
        if (mode == 7) {
          double v = divF0e[iu];
          v -= r0;
          divF0e[iu] = v;
        } else {
          st_s<NT>(mode == 8 ? r0 + srce[iu] : r0, &divF0e[iu]);
        }
      }
    }
    return;  // uniform over the block: no barrier follows for anyone
  }
  if (b < nb) {
    const long long e = e0 + b;
    const double dt = (local == 1) ? eleDt[e] : (local == 2 ? meshDt[eleMesh[e]] : *dtp);
    const bool frozen = active && !active[eleMesh[e]];
    const double dj = ld_s<NT>(&detJe[e * S + s]);
    // the stage's state inputs for every field first (uniform branches on the
    // mode), so that their loads are in flight together: U itself is an input
    // of the first and the single stage only (and of frozen batch members,
    // which keep it)
    double uin[N], u0in[N], accin[N];
    const bool needU = mode == 0 || mode == 3 || mode == 6 || frozen;
    const bool needU0 = mode == 1 || mode == 4 || (mode == 2 && Rst);
    const bool needAcc = !Rst && (mode == 1 || mode == 2);
#pragma unroll
    for (int k = 0; k < N; k++) uin[k] = u0in[k] = accin[k] = 0.0;
    if (needU) {
#pragma unroll
      for (int k = 0; k < N; k++) uin[k] = ld_s<NT>(&Ue[(e * N + k) * S + s]);
    }
    if (needU0) {
      // with a stage copy the state stays the step's U0 until the last stage
      // writes it, and no separate U0 is kept (U0e null)
      const double* u0 = U0e ? U0e : Ue;
#pragma unroll
      for (int k = 0; k < N; k++) u0in[k] = ld_s<NT>(&u0[(e * N + k) * S + s]);
    }
    if (needAcc) {
#pragma unroll
      for (int k = 0; k < N; k++) accin[k] = ld_s<NT>(&Uacc[(e * N + k) * S + s]);
    }
#pragma unroll
    for (int k = 0; k < N; k++) {
      const long long iu = (e * N + k) * S + s;
      R c = 0;
#pragma unroll
      for (int j = 0; j < NC; j++) c += tw<R>(tb.cor_w, tb.cor_wf)[s * NC + j] * sD[(b * N + k) * F + tb.cor_idx[s * NC + j]];
      const double r0 = double(c + dv[k]);
      if (divF0e) st_s<NT>(r0, &divF0e[iu]);
      // P-multigrid source: divF_spts[step] + src_spts (src/solver.cpp:480,514)
      const double r = srce ? r0 + srce[iu] : r0;
      const double u = uin[k];
      double un;
      if (mode == 0) {
        if (U0e) st_s<NT>(u, &U0e[iu]);
        if (Rst) st_s<NT>(R(r), &Rst[iu]);
        else st_s<NT>(u - bcoef * r / dj * dt, &Uacc[iu]);
        un = u - a * r / dj * dt;
      } else if (mode == 1) {
        if (Rst) st_s<NT>(R(r), &Rst[stage * nS + iu]);
        else st_s<NT>(accin[k] - bcoef * r / dj * dt, &Uacc[iu]);
        un = u0in[k] - a * r / dj * dt;
      } else if (mode == 2) {
        double acc;
        if (Rst) {
          acc = u0in[k];
          for (int j = 0; j < stage; j++) acc = acc - bs.b[j] * double(ld_s<NT>(&Rst[j * nS + iu])) / dj * dt;
        } else {
          acc = accin[k];
        }
        un = acc - bcoef * r / dj * dt;
      } else if (mode == 3) {
        un = u - bcoef * r / dj * dt;
      } else if (mode == 6) {
        if (U0e) st_s<NT>(u, &U0e[iu]);
        un = u - a * r / dj * dt;
      } else {
        un = u0in[k] - a * r / dj * dt;
      }
      if (frozen) un = u;
      // with a stage copy the state is written at the step's last stage
      // only (earlier stages leave it at the step's start, its U0, and write
      // the copy); the FP32 copy is written at the last stage too (the flux
      // kernels read U only through it), the FP64 one is not (the next
      // step's first stage reads the state itself)
      if (!Us || fin) st_s<NT>(un, &Ue[iu]);
      if (Us && (sizeof(R) == 4 || !fin)) st_s<NT>(R(un), &Us[iu]);
      sU[(b * N + k) * S + spt_lds(D, P1, s)] = R(un);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nb * F; i += blockDim.x) {
    int bb = i / F, f = i - bb * F;
#pragma unroll
    for (int k = 0; k < N; k++) {
      R acc = 0;
#pragma unroll
      for (int m = 0; m < P1; m++) acc += tw<R>(tb.ext_w, tb.ext_wf)[f * P1 + m] * sU[(bb * N + k) * S + tb.ext_idx[f * P1 + m]];
      st_s<NT>(acc, &Ufe[((long long)(e0 + bb) * N + k) * F + f]);
    }
  }
}

template <int D, int P, class R, bool NT>
__device__ __forceinline__ void rk_tail(int E, const R* dv, R* sU, const R* sD, const RkArgs<R>& ra, const Tabs& tb) {
  rk_tail_p<D, P, R, NT>(E, dv, sU, sD, ra.mode, ra.a, ra.bcoef, ra.detJe, ra.dtp, ra.eleDt, ra.local, ra.Ue, ra.U0e,
                         ra.Uacc, ra.divF0e, ra.eleMesh, ra.meshDt, ra.active, ra.srce, ra.Rst, ra.stage, ra.bs, ra.Us,
                         ra.fin, ra.Ufe, tb);
}

// Viscous runs, FP32: the volume kernel also folds its own side's LDG part
// into the flux jump (disFne := +-B dA - disFn), so the RK kernel reads one
// array less.  FP64 keeps disFn and B apart: its volume kernel is short of
// registers, and reading B back at the end of the element costs it more than
// the RK kernel saves (measured).
template <class R>
constexpr bool kJumpInVolume = sizeof(R) == 4;

// Shapes whose volume kernel is held to 4 waves per SIMD (128 VGPRs): quads
// and the FP32 path.  FP64 hexes need more registers and are held to 3.
template <int D, class R>
constexpr bool kHold = D == 2 || sizeof(R) == 4;

// The volume kernel's flux-point inputs and outputs besides the slabs: the
// side codes and dA (every run), and for viscous runs what it leaves at the
// flux points in place of the gradient there (see the face loop of
// volume_body).
template <class R>
struct VFace {
  const int* fpart;    // [E*F] 4 * partner flux point + side, or 4 * boundary slot + 2
  const R* nfe;        // [E][D][F] the face's normal (its left element's)
  R* Bve;              // [E][N][F] this side's part of the LDG flux . n
  R* gBnd;             // [D*N][nBnd] gradient at the boundary flux points
  const R* dA;         // [E][F] face area factor
  int nBnd;
};

// ---------------------------------------------------------------------------
// Volume kernel: [corrected gradient], flux, divergence, normal flux at faces
// [and gradient at faces].  One thread per solution point, EB elements per
// block, element data staged in LDS.
template <int D, int P, bool VISC, class R, bool NT, int BF = 0>
__device__ __forceinline__ void volume_body(int E, const double* __restrict__ Ue, const R* __restrict__ Us,
                                            const R* __restrict__ dUce,
                                            const R* __restrict__ JGe, const R* __restrict__ iJe,
                                            R* __restrict__ divFe, R* __restrict__ disFne, VFace<R> vf,
                                            Tabs tb, Phys ph, const R* __restrict__ Fne, const RkArgs<R>& ra) {
  using Sh = Shape<D, P>;
  constexpr int S = Sh::S, F = Sh::F, N = Sh::N, P1 = Sh::P1;
  const int EB = blockDim.x / S;
  // sF holds, in turn (viscous runs): dUc at the flux points, the physical
  // gradient until it has been extrapolated to the faces, then the transformed
  // flux at the same indices as the gradient.  [EB*N*S] + [EB*N*SF]; viscous
  // runs also stage the faces' normals for the face loop, [EB*D*F], inviscid
  // runs keep the flux jump of the RK update there, [EB*N*F]; then the
  // points' side codes [EB*F] ints
  constexpr int SF = (VISC && F > D * S) ? F : D * S;
  R* const sU = reinterpret_cast<R*>(fhip_lds);
  R* const sF = sU + EB * N * S;
  R* const sG = sF;
  R* const sN = sF + EB * N * SF;
  R* const sD = sN;
  int* const sP = reinterpret_cast<int*>(VISC ? sN + EB * D * F : sD + EB * N * F);
  // viscous runs with kJumpInVolume: this side's LDG part at the block's
  // flux points, kept from the first face loop for the last one [EB*N*F]
  R* const sB = reinterpret_cast<R*>(sP + EB * F);
  const int e0 = blockIdx.x * EB, nb = min(EB, E - e0);
  const int b = threadIdx.x / S, s = threadIdx.x - b * S;
  const int sw = spt_lds(D, P1, s);
  const bool act = b < nb;
  const long long e = e0 + b;
  // matrix-core contractions: this lane's operator entry d1[s % 4][s / 16]
  // and the slab position of the point whose x and z are this lane's z and x
  constexpr bool mf = kMfma<D, P, R>;
  // mixed path, P3 hexes (BF = 2 or 3): z contractions in BF16 pieces on the
  // matrix cores
  constexpr int np = kBf16<D, P, R> ? BF : 0;
  bf16x8 abf;
  if constexpr (np != 0) abf = bf16_opnd_a(tw<float>(tb.d1, tb.d1f), s, np);
  const int swx = mf ? spt_lds(D, P1, s / 16 + 4 * ((s / 4) % 4) + 16 * (s % 4)) : 0;
  const R opv = mf ? tw<R>(tb.d1, tb.d1f)[(s % 4) * P1 + s / 16] : R(0);
  // the point's metric terms, loaded first so that their latency overlaps the
  // staging of the slab (and the gradient of viscous runs)
  R G[D][D] = {};
  auto load_G = [&] {
#pragma unroll
    for (int r = 0; r < D; r++)
#pragma unroll
      for (int x = 0; x < D; x++) G[r][x] = JGe[((e * D + r) * D + x) * S + s];
  };
  R inv = 0;
  if (act) {
    load_G();
    if constexpr (VISC) inv = iJe[e * S + s];
  }
  {
    // every load of the slabs is issued before the first LDS store, so that
    // their latencies overlap: a thread stages at most N values of U (the
    // block holds at most blockDim / S elements) and KC of dUc
    constexpr int KC = VISC ? (N * F + S - 1) / S : 1, KN = VISC ? (D * F + S - 1) / S : 1,
                  KP = (F + S - 1) / S;
    const int nU = nb * N * S, nC = nb * N * F, nN = nb * D * F, nP = nb * F;
    R vu[N], vc[KC], vn[KN];
    int vp[KP];
#pragma unroll
    for (int u = 0; u < N; u++) {
      const int i = threadIdx.x + u * blockDim.x;
      if (i < nU) vu[u] = Us ? ld_s<NT>(&Us[(long long)e0 * N * S + i]) : R(ld_s<NT>(&Ue[(long long)e0 * N * S + i]));
    }
    if constexpr (VISC) {
      // dUc at the block's flux points (sF serves as sC until the gradient is written)
#pragma unroll
      for (int u = 0; u < KC; u++) {
        const int i = threadIdx.x + u * blockDim.x;
        if (i < nC) vc[u] = ld_s<NT>(&dUce[(long long)e0 * N * F + i]);
      }
#pragma unroll
      for (int u = 0; u < KN; u++) {
        const int i = threadIdx.x + u * blockDim.x;
        if (i < nN) vn[u] = ld_s<NT>(&vf.nfe[(long long)e0 * D * F + i]);
      }
    }
#pragma unroll
    for (int u = 0; u < KP; u++) {
      const int i = threadIdx.x + u * blockDim.x;
      if (i < nP) vp[u] = vf.fpart[(long long)e0 * F + i];
    }
#pragma unroll
    for (int u = 0; u < N; u++) {
      const int i = threadIdx.x + u * blockDim.x;
      if (i < nU) {
        const int r = i / S, q = i - r * S;
        sU[r * S + spt_lds(D, P1, q)] = vu[u];
      }
    }
    if constexpr (VISC) {
#pragma unroll
      for (int u = 0; u < KC; u++) {
        const int i = threadIdx.x + u * blockDim.x;
        if (i < nC) sF[i] = vc[u];
      }
#pragma unroll
      for (int u = 0; u < KN; u++) {
        const int i = threadIdx.x + u * blockDim.x;
        if (i < nN) sN[i] = vn[u];
      }
    }
#pragma unroll
    for (int u = 0; u < KP; u++) {
      const int i = threadIdx.x + u * blockDim.x;
      if (i < nP) sP[i] = vp[u];
    }
  }
  __syncthreads();
  // inviscid runs: the common flux at the points this thread finishes in the
  // last face loop (the pair's left point, the partner's on the right side)
  // and their dA, loaded now so that the loads are in flight during the flux
  // and divergence phases
  constexpr int KFP = VISC ? 1 : (F + S - 1) / S;
  R pf[KFP][N], pdA[KFP];
  int psd[KFP];
  if constexpr (!VISC) {
#pragma unroll
    for (int u = 0; u < KFP; u++) {
      const int i = threadIdx.x + u * blockDim.x;
      psd[u] = 2;
      pdA[u] = 0;
#pragma unroll
      for (int k = 0; k < N; k++) pf[u][k] = 0;
      if (i < nb * F) {
        const int code = sP[i], bb = i / F;
        psd[u] = code & 3;
        long long pb = (long long)(e0 + bb) * N * F + (i - bb * F);
        if (psd[u] == 1) {
          const long long p = code >> 2, ep = p / F;
          pb = ep * N * F + (p - ep * F);
        }
#pragma unroll
        for (int k = 0; k < N; k++) pf[u][k] = Fne[pb + k * F];
        pdA[u] = vf.dA[(long long)e0 * F + i];
      }
    }
  }
  if constexpr (VISC) {
    // phase 1: reference-space gradient (opp_grad_spts) plus correction
    // (opp_correctU), then the physical gradient (correctGradU transform,
    // src/solver.cpp:1105-1136), kept in LDS for phase 2 and the faces
    R* const sC = sF;
    R gr[D][N];
    if (act) {
#pragma unroll
      for (int d = 0; d < D; d++) {
        const int* gi = tb.grd_idx + (d * S + s) * P1;
        const R* gw = tw<R>(tb.grd_w, tb.grd_wf) + (d * S + s) * P1;
        const int* ci = tb.cu_idx + (d * S + s) * 2;
        const R* cw = tw<R>(tb.cu_w, tb.cu_wf) + (d * S + s) * 2;
#pragma unroll
        for (int k = 0; k < N; k++) {
          // the other correction entries of the row are exactly zero
          R c = 0;
#pragma unroll
          for (int j = 0; j < 2; j++) c += cw[j] * sC[(b * N + k) * F + ci[j]];
          if constexpr (np != 0) {
            if (d == 2) {
              gr[d][k] = mfma_z_bf16<D, P1>(reinterpret_cast<const float*>(sU + (b * N + k) * S), s, abf, c, np);
              continue;
            }
          }
          if (mf && d != 1) {  // compile-time: d and mf are constants here
            gr[d][k] = d == 2 ? mfma444(opv, sU[(b * N + k) * S + sw], c)
                              : mfma444(sU[(b * N + k) * S + swx], opv, c);
            continue;
          }
          R a = 0;
#pragma unroll
          for (int m = 0; m < P1; m++) a += gw[m] * sU[(b * N + k) * S + gi[m]];
          gr[d][k] = c + a;
        }
      }
      // physical gradient, in place
#pragma unroll
      for (int k = 0; k < N; k++) {
        R acc[D];
#pragma unroll
        for (int x = 0; x < D; x++) {
          acc[x] = gr[0][k] * G[0][x] + gr[1][k] * G[1][x];
          if constexpr (D == 3) acc[x] += gr[2][k] * G[2][x];
        }
#pragma unroll
        for (int x = 0; x < D; x++) gr[x][k] = inv * acc[x];
      }
    }
    __syncthreads();  // sC is overwritten by the gradient below
    if (act) {
#pragma unroll
      for (int k = 0; k < N; k++)
#pragma unroll
        for (int x = 0; x < D; x++) sG[((b * D + x) * N + k) * S + sw] = gr[x][k];
    }
    __syncthreads();
    // physical gradient at the flux points (extrapolateGradU) and, with U
    // there (extrapolateU, the same sums as the RK kernel's), this side's
    // viscous flux folded into its part of the LDG flux
    // (face::calcViscousFlux, src/face.cpp:216-293); the two sides' parts are
    // summed where the correction is applied.  Boundary points keep the
    // gradient for the boundary kernel, which needs the boundary state.
    for (int i = threadIdx.x; i < nb * F; i += blockDim.x) {
      const int bb = i / F, f = i - bb * F;
      const long long ee = e0 + bb;
      const int* xi = tb.ext_idx + f * P1;
      const R* xw = tw<R>(tb.ext_w, tb.ext_wf) + f * P1;
      // side code and normal, staged in LDS with the slab
      const int code = sP[bb * F + f], side = code & 3;
      R nv[D];
#pragma unroll
      for (int x = 0; x < D; x++) nv[x] = sN[(bb * D + x) * F + f];
      R g[D][N];
#pragma unroll
      for (int x = 0; x < D; x++)
#pragma unroll
        for (int k = 0; k < N; k++) {
          R a = 0;
#pragma unroll
          for (int m = 0; m < P1; m++) a += xw[m] * sG[((bb * D + x) * N + k) * S + xi[m]];
          g[x][k] = a;
        }
      if (side == 2) {
        const long long slot = code >> 2;
#pragma unroll
        for (int x = 0; x < D; x++)
#pragma unroll
          for (int k = 0; k < N; k++) vf.gBnd[(x * N + k) * (long long)vf.nBnd + slot] = g[x][k];
      } else {
        R u[N], bv[N];
#pragma unroll
        for (int k = 0; k < N; k++) {
          R a = 0;
#pragma unroll
          for (int m = 0; m < P1; m++) a += xw[m] * sU[(bb * N + k) * S + xi[m]];
          u[k] = a;
          bv[k] = 0;
        }
        ldg_add_side<D, R>(side == 0, u, g, nv, bv, ph);
#pragma unroll
        for (int k = 0; k < N; k++) {
          st_s<NT>(bv[k], &vf.Bve[(ee * N + k) * F + f]);
          if constexpr (kJumpInVolume<R>) sB[(bb * N + k) * F + f] = bv[k];
        }
      }
    }
    // under the register hold, the metric terms again (from cache): keeping
    // them live across the face loop above costs more registers than it allows
    if constexpr (kHold<D, R>) {
      if (act) load_G();
    }
  }
  // phase 2: flux at the point, transformed to the reference element.  It is
  // formed in registers and stored after a barrier: viscous runs store it over
  // the gradient slab, which other threads may still be reading in the face
  // loop, and threads with fewer face points go on here meanwhile.
  R Ft[D][N];
  if (act) {
    R u[N];
#pragma unroll
    for (int k = 0; k < N; k++) u[k] = sU[(b * N + k) * S + sw];
    R Fl[D][N];
    if constexpr (VISC) {
      R g[D][N];
#pragma unroll
      for (int x = 0; x < D; x++)
#pragma unroll
        for (int k = 0; k < N; k++) g[x][k] = sG[((b * D + x) * N + k) * S + sw];
      R Fi[D][N];
      viscous_flux<D, R>(u, g, Fl, ph);
      inviscid_flux<D, R>(u, Fi, R(ph.gamma));
#pragma unroll
      for (int d = 0; d < D; d++)
#pragma unroll
        for (int k = 0; k < N; k++) Fl[d][k] = Fl[d][k] + Fi[d][k];
    } else {
      inviscid_flux<D, R>(u, Fl, R(ph.gamma));
    }
#pragma unroll
    for (int d1 = 0; d1 < D; d1++)
#pragma unroll
      for (int k = 0; k < N; k++) {
        R acc = 0;
#pragma unroll
        for (int d2 = 0; d2 < D; d2++) acc += G[d1][d2] * Fl[d2][k];
        Ft[d1][k] = acc;
      }
  }
  if constexpr (VISC) __syncthreads();
  if (act) {
#pragma unroll
    for (int d1 = 0; d1 < D; d1++)
#pragma unroll
      for (int k = 0; k < N; k++) sF[((b * D + d1) * N + k) * S + sw] = Ft[d1][k];
  }
  __syncthreads();
  R dv[N] = {};  // inviscid runs: divF at the point, for the RK update below
  if (act) {
    // divergence (calcDivF_spts): sum over directions in order (matrix
    // cores: y on the vector unit, then z and x added)
#pragma unroll
    for (int k = 0; k < N; k++) {
      R tot = 0;
#pragma unroll
      for (int d = 0; d < D; d++) {
        if (mf && d != 1) continue;
        if constexpr (np != 0) {
          if (d == 2) {
            tot = mfma_z_bf16<D, P1>(reinterpret_cast<const float*>(sF + ((b * D + 2) * N + k) * S), s, abf, tot, np);
            continue;
          }
        }
        const int* gi = tb.grd_idx + (d * S + s) * P1;
        const R* gw = tw<R>(tb.grd_w, tb.grd_wf) + (d * S + s) * P1;
        R a = 0;
#pragma unroll
        for (int m = 0; m < P1; m++) a += gw[m] * sF[((b * D + d) * N + k) * S + gi[m]];
        tot = (d == 0 || mf) ? a : a + tot;
      }
      if (mf) {
        tot = mfma444(opv, sF[((b * D + 2) * N + k) * S + sw], tot);
        tot = mfma444(sF[((b * D + 0) * N + k) * S + swx], opv, tot);
      }
      if constexpr (VISC) st_s<NT>(tot, &divFe[(e * N + k) * S + s]);
      else dv[k] = tot;
    }
  }
  // faces: discontinuous normal flux.  With kJumpInVolume, viscous runs store
  // in its place this side's part of the correction's flux jump,
  // +-B dA - disFn (B as written by this thread in the face loop above, - on
  // the right side; boundary points get - disFn here and + B dA from the
  // boundary kernel).
  if constexpr (VISC) {
    for (int i = threadIdx.x; i < nb * F; i += blockDim.x) {
      const int bb = i / F, f = i - bb * F;
      const long long ee = e0 + bb;
      const int* xi = tb.ext_idx + f * P1;
      // extrapolateNormalFlux: of the sum over directions only the face's own
      // normal direction is non-zero (the other terms are exact zeros)
      const int dn = tb.efn_dir[f];
      const R* ew = tw<R>(tb.efn_w, tb.efn_wf) + f * P1;
      int side = 2;
      R dA = 0, bv[N];
      if constexpr (kJumpInVolume<R>) {
        side = sP[bb * F + f] & 3;
        dA = vf.dA[ee * F + f];
        // this thread's own LDG parts from the first face loop (the same
        // points); read for every point, unused at boundary points
#pragma unroll
        for (int k = 0; k < N; k++) bv[k] = sB[(bb * N + k) * F + f];
      }
#pragma unroll
      for (int k = 0; k < N; k++) {
        R a = 0;
#pragma unroll
        for (int m = 0; m < P1; m++) a += ew[m] * sF[((bb * D + dn) * N + k) * S + xi[m]];
        if constexpr (kJumpInVolume<R>) a = (side == 2) ? -a : (side == 0 ? bv[k] : -bv[k]) * dA - a;
        st_s<NT>(a, &disFne[(ee * N + k) * F + f]);
      }
    }
  } else {
    // inviscid runs: the flux jump, the common flux (loaded above) times
    // this side's dA, negated on the right side, less the extrapolated flux
#pragma unroll
    for (int u = 0; u < KFP; u++) {
      const int i = threadIdx.x + u * blockDim.x;
      if (i < nb * F) {
        const int bb = i / F, f = i - bb * F;
        const int* xi = tb.ext_idx + f * P1;
        const int dn = tb.efn_dir[f];
        const R* ew = tw<R>(tb.efn_w, tb.efn_wf) + f * P1;
#pragma unroll
        for (int k = 0; k < N; k++) {
          R a = 0;
#pragma unroll
          for (int m = 0; m < P1; m++) a += ew[m] * sF[((bb * D + dn) * N + k) * S + xi[m]];
          const R fo = mul_rn(pf[u][k], pdA[u]);
          sD[(bb * N + k) * F + f] = (psd[u] == 1 ? -fo : fo) - a;
        }
      }
    }
  }
  if constexpr (!VISC) {
    // inviscid runs: the correction, RK update and U at the flux points here
    // (no partner data is needed past the face kernel's common flux)
    __syncthreads();
    rk_tail<D, P, R, NT>(E, dv, sU, sD, ra, tb);
  }
}

// The volume kernel held to 4 waves per SIMD (at most 128 VGPRs) for quads
// and the FP32 path, which fit there without spilling (holding the FP32
// viscous hex kernel to 5 or 6 waves spills 26 or 41 VGPRs and was measured
// 13-16% slower), and to 3 waves (168 VGPRs) for FP64 hexes.  Viscous runs
// reload the metric terms after the face loop; inviscid runs keep them in
// registers throughout.
template <int D, int P, bool VISC, class R, bool NT>
__global__ __launch_bounds__(kMaxBlock) __attribute__((amdgpu_waves_per_eu(3))) void k_volume_w3(
    int E, const double* __restrict__ Ue, const R* __restrict__ Us, const R* __restrict__ dUce, const R* __restrict__ JGe,
    const R* __restrict__ iJe, R* __restrict__ divFe, R* __restrict__ disFne, VFace<R> vf, Tabs tb,
    Phys ph, const R* __restrict__ Fne, RkArgs<R> ra) {
  volume_body<D, P, VISC, R, NT>(E, Ue, Us, dUce, JGe, iJe, divFe, disFne, vf, tb, ph, Fne, ra);
}
template <int D, int P, bool VISC, class R, bool NT, int BF = 0>
__global__ __launch_bounds__(kMaxBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_volume_w4(
    int E, const double* __restrict__ Ue, const R* __restrict__ Us, const R* __restrict__ dUce, const R* __restrict__ JGe,
    const R* __restrict__ iJe, R* __restrict__ divFe, R* __restrict__ disFne, VFace<R> vf, Tabs tb,
    Phys ph, const R* __restrict__ Fne, RkArgs<R> ra) {
  volume_body<D, P, VISC, R, NT, BF>(E, Ue, Us, dUce, JGe, iJe, divFe, disFne, vf, tb, ph, Fne, ra);
}

// ---------------------------------------------------------------------------
// LDG flux at the boundary flux points (face::calcViscousFlux with the
// boundary state of boundFace::applyViscousBCs, src/face.cpp:216-293,
// src/boundFace.cpp:316-352): one thread per boundary point, from the
// gradient the volume kernel left there.  The result goes where the volume
// kernel puts an interior point's side of the flux, or (kJumpInVolume) times
// dA into the point's flux jump, which the volume kernel set to -disFn.
template <int D, int P, class R>
__global__ __launch_bounds__(kMaxBlock) void k_face_bnd(int nB, const int* __restrict__ bndT,
                                                   const int* __restrict__ pLa, const int* __restrict__ bca,
                                                   const R* __restrict__ Ufe, const R* __restrict__ gBnd,
                                                   const R* __restrict__ norme, const R* __restrict__ dAe,
                                                   R* __restrict__ Bve, R* __restrict__ Cne, Phys ph,
                                                   const double* __restrict__ bst) {
  using Sh = Shape<D, P>;
  constexpr int F = Sh::F, N = Sh::N;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nB) return;
  const int t = bndT[i], pl = pLa[t], bc = bca[t];
  const long long eL = pl / F, fL = pl - eL * F;
  const long long bL = eL * N * F + fL;
  R UL[N], UR[N], n[D], g[D][N], nF[N], d12[N], bv[N];
#pragma unroll
  for (int k = 0; k < N; k++) UL[k] = Ufe[bL + k * F];
#pragma unroll
  for (int d = 0; d < D; d++) n[d] = norme[(eL * D + d) * F + fL];
#pragma unroll
  for (int d = 0; d < D; d++)
#pragma unroll
    for (int k = 0; k < N; k++) g[d][k] = gBnd[(d * N + k) * (long long)nB + i];
  bc_state<D, R>(bc, UL, n, UR, ph, bst);
  const bool adiabatic = (bc & 15) == 10;
  if (adiabatic) adiabatic_grad<D, R>(UL, UR, g, R(ph.gamma));
  ldg_side<D, R>(adiabatic ? UR : UL, g, n, nF, d12, ph);
#pragma unroll
  for (int k = 0; k < N; k++) bv[k] = 0;
  ldg_boundary<D, R>(UL, UR, nF, n, bv, ph);
  if constexpr (kJumpInVolume<R>) {
    const R dA = dAe[pl];
#pragma unroll
    for (int k = 0; k < N; k++) Cne[bL + k * F] = bv[k] * dA + Cne[bL + k * F];
  } else {
#pragma unroll
    for (int k = 0; k < N; k++) Bve[bL + k * F] = bv[k];
  }
}

// Correction of the divergence, RK update and U at the flux points, after the
// viscous volume kernel.  With NT, the arrays this kernel reads or writes
// exactly once per launch go through nontemporal loads and stores (they
// stream past the caches); arrays read twice (the LDG side parts and the
// inviscid common flux, read by both sides of a face) stay cached.
template <int D, int P, class R, bool NT>
__global__ __launch_bounds__(kMaxBlock) void k_rk(int E, const R* __restrict__ divFe, const R* __restrict__ Fne,
                                             const R* __restrict__ disFne, RkArgs<R> ra, Tabs tb,
                                             const int* __restrict__ fpart, const R* __restrict__ dAe,
                                             const R* __restrict__ Bve) {
  using Sh = Shape<D, P>;
  constexpr int S = Sh::S, F = Sh::F, N = Sh::N;
  const int EB = blockDim.x / S;
  R* const sU = reinterpret_cast<R*>(fhip_lds);  // [EB*N*S]
  R* const sD = sU + EB * N * S;                  // [EB*N*F] Fn - disFn at the block's flux points
  const int e0 = blockIdx.x * EB, nb = min(EB, E - e0);
  const int b = threadIdx.x / S, s = threadIdx.x - b * S;
  // The flux points' loads are unconditional (a boundary point reads its own
  // row where an interior point reads its partner's, and selects) and, for
  // the points a thread stages, issued before their LDS stores, so that their
  // latencies overlap.  A thread stages at most KF = ceil(F / S) points.
  constexpr int KF = (F + S - 1) / S;
  if constexpr (kJumpInVolume<R>) {
    // viscous: disFne holds this side's jump part +-B dA - disFn; add the
    // inviscid part (Fne at the pair's left point, times dA) and the partner
    // side's B times dA, negated on the right side
    int cd[KF];
#pragma unroll
    for (int u = 0; u < KF; u++) {
      const int i = threadIdx.x + u * blockDim.x;
      cd[u] = (i < nb * F) ? fpart[(long long)e0 * F + i] : 2;
    }
#pragma unroll
    for (int u = 0; u < KF; u++) {
      const int i = threadIdx.x + u * blockDim.x;
      if (i >= nb * F) break;
      const int bb = i / F, f = i - bb * F;
      const long long ef = (long long)(e0 + bb) * F + f, ob = (long long)(e0 + bb) * N * F + f;
      const int code = cd[u], side = code & 3;
      const R dA = dAe[ef];
      long long pb = ob;
      if (side != 2) {
        const long long p = code >> 2, ep = p / F;
        pb = ep * N * F + (p - ep * F);
      }
#pragma unroll
      for (int k = 0; k < N; k++) {
        const R bp = Bve[pb + k * F];
        // the pair's inviscid flux (at its left point) times this side's dA
        const R fo = mul_rn(Fne[(side == 1 ? pb : ob) + k * F], dA);
        const R base = (side == 1 ? -fo : fo) + ld_s<NT>(&disFne[ob + k * F]);
        sD[(bb * N + k) * F + f] = (side == 2) ? base : base + (side == 0 ? bp : -bp) * dA;
      }
    }
  } else {
    // viscous: common flux = inviscid part (Fne at the pair's left point)
    // plus both sides' LDG parts, times dA, negated on the right side
    int cd[KF];
#pragma unroll
    for (int u = 0; u < KF; u++) {
      const int i = threadIdx.x + u * blockDim.x;
      cd[u] = (i < nb * F) ? fpart[(long long)e0 * F + i] : 2;
    }
#pragma unroll
    for (int u = 0; u < KF; u++) {
      const int i = threadIdx.x + u * blockDim.x;
      if (i >= nb * F) break;
      const int bb = i / F, f = i - bb * F;
      const long long ef = (long long)(e0 + bb) * F + f, ob = (long long)(e0 + bb) * N * F + f;
      const int code = cd[u], side = code & 3;
      const R dA = dAe[ef];
      long long pb = ob;
      if (side != 2) {
        const long long p = code >> 2, ep = p / F;
        pb = ep * N * F + (p - ep * F);
      }
#pragma unroll
      for (int k = 0; k < N; k++) {
        const R bo = Bve[ob + k * F], bp = Bve[pb + k * F];
        R bv = (side != 2) ? bo + bp : bo;
        if (side == 1) bv = -bv;
        // the pair's inviscid flux (at its left point) times this side's dA
        const R fo = mul_rn(Fne[(side == 1 ? pb : ob) + k * F], dA);
        sD[(bb * N + k) * F + f] = ((side == 1 ? -fo : fo) + bv * dA) - ld_s<NT>(&disFne[ob + k * F]);
      }
    }
  }
  __syncthreads();
  R dv[N];
  if (b < nb) {
#pragma unroll
    for (int k = 0; k < N; k++) dv[k] = ld_s<NT>(&divFe[((e0 + b) * (long long)N + k) * S + s]);
  }
  rk_tail<D, P, R, NT>(E, dv, sU, sD, ra, tb);
}

__global__ void k_copy_em(long long n, double* __restrict__ dst, const double* __restrict__ src) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// P-multigrid transfers in element-major layout ([ele][field][point] rows of
// S values), the products the staged operator GEMM forms in FlurryPP's layout
// with the same sums (chunks of kc columns of the operator, added in turn).
// multiGrid::restrict_pmg (src/multigrid.cpp:315-345) and the start of
// compute_source_term (:381-398): the coarse level's U = R U_fine, its FAS
// source = R divF_fine (the residual of U is subtracted later) and sol = U.
__global__ void k_restrict_em(long long rows, int Sc, int Sf, int kc, const double* __restrict__ R,
                              const double* __restrict__ Uf, const double* __restrict__ divFf,
                              double* __restrict__ Uc, double* __restrict__ srcc, double* __restrict__ solc) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= rows * Sc) return;
  const long long r = t / Sc;
  const int i = (int)(t - r * Sc);
  const double* a = R + (long long)i * Sf;
  const double* u = Uf + r * Sf;
  const double* d = divFf + r * Sf;
  double cu = 0, cd = 0;
  for (int l0 = 0; l0 < Sf; l0 += kc) {
    const int n = min(kc, Sf - l0);
    double su = 0, sd = 0;
    for (int l = 0; l < n; l++) {
      su += a[l0 + l] * u[l0 + l];
      sd += a[l0 + l] * d[l0 + l];
    }
    cu = l0 ? su + cu : su;
    cd = l0 ? sd + cd : sd;
  }
  Uc[t] = cu;
  solc[t] = cu;
  srcc[t] = cd;
}

// prolong_err (src/multigrid.cpp:347-364): U_up += P (U - sol) of the level
// below.
__global__ void k_prolong_em(long long rows, int Su, int Sc, int kc, const double* __restrict__ Pm,
                             const double* __restrict__ Uc, const double* __restrict__ solc, double* __restrict__ Uu) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= rows * Su) return;
  const long long r = t / Su;
  const int i = (int)(t - r * Su);
  const double* a = Pm + (long long)i * Sc;
  const double* u = Uc + r * Sc;
  const double* o = solc + r * Sc;
  double c = Uu[t];
  for (int l0 = 0; l0 < Sc; l0 += kc) {
    const int n = min(kc, Sc - l0);
    double s = 0;
    for (int l = 0; l < n; l++) {
      double v = u[l0 + l];
      v -= o[l0 + l];
      s += a[l0 + l] * v;
    }
    c = s + c;
  }
  Uu[t] = c;
}

// The same transfers with the operator (transposed, so that the lanes of a
// wave read consecutive entries) and the block's rows staged in LDS: nr rows
// per block, one output per thread, the sums in the same order as above.
// Used where the operator and the rows fit in 64 KiB.
__global__ void k_restrict_em_lds(long long rows, int Sc, int Sf, int kc, int nr, const double* __restrict__ R,
                                  const double* __restrict__ Uf, const double* __restrict__ divFf,
                                  double* __restrict__ Uc, double* __restrict__ srcc, double* __restrict__ solc) {
  double* const sA = reinterpret_cast<double*>(fhip_lds);  // [Sf][Sc]
  double* const sU = sA + Sc * Sf;                          // [nr][Sf]
  double* const sV = sU + nr * Sf;                          // [nr][Sf]
  const long long r0 = (long long)blockIdx.x * nr;
  const int nb = (int)min((long long)nr, rows - r0);
  for (int j = threadIdx.x; j < Sc * Sf; j += blockDim.x) {
    const int i = j / Sf;
    sA[(j - i * Sf) * Sc + i] = R[j];
  }
  for (int j = threadIdx.x; j < nb * Sf; j += blockDim.x) {
    sU[j] = Uf[r0 * Sf + j];
    sV[j] = divFf[r0 * Sf + j];
  }
  __syncthreads();
  const int rr = threadIdx.x / Sc, i = threadIdx.x - rr * Sc;
  if (rr >= nb) return;
  const double* u = sU + rr * Sf;
  const double* d = sV + rr * Sf;
  double cu = 0, cd = 0;
  for (int l0 = 0; l0 < Sf; l0 += kc) {
    const int n = min(kc, Sf - l0);
    double su = 0, sd = 0;
    for (int l = 0; l < n; l++) {
      const double a = sA[(l0 + l) * Sc + i];
      su += a * u[l0 + l];
      sd += a * d[l0 + l];
    }
    cu = l0 ? su + cu : su;
    cd = l0 ? sd + cd : sd;
  }
  const long long t = r0 * Sc + threadIdx.x;
  Uc[t] = cu;
  solc[t] = cu;
  srcc[t] = cd;
}

__global__ void k_prolong_em_lds(long long rows, int Su, int Sc, int kc, int nr, const double* __restrict__ Pm,
                                 const double* __restrict__ Uc, const double* __restrict__ solc,
                                 double* __restrict__ Uu) {
  double* const sA = reinterpret_cast<double*>(fhip_lds);  // [Sc][Su]
  double* const sV = sA + Su * Sc;                          // [nr][Sc]: U - sol
  const long long r0 = (long long)blockIdx.x * nr;
  const int nb = (int)min((long long)nr, rows - r0);
  for (int j = threadIdx.x; j < Su * Sc; j += blockDim.x) {
    const int i = j / Sc;
    sA[(j - i * Sc) * Su + i] = Pm[j];
  }
  for (int j = threadIdx.x; j < nb * Sc; j += blockDim.x) {
    double v = Uc[r0 * Sc + j];
    v -= solc[r0 * Sc + j];
    sV[j] = v;
  }
  __syncthreads();
  const int rr = threadIdx.x / Su, i = threadIdx.x - rr * Su;
  if (rr >= nb) return;
  const double* v = sV + rr * Sc;
  const long long t = r0 * Su + threadIdx.x;
  double c = Uu[t];
  for (int l0 = 0; l0 < Sc; l0 += kc) {
    const int n = min(kc, Sc - l0);
    double s = 0;
    for (int l = 0; l < n; l++) s += sA[(l0 + l) * Su + i] * v[l0 + l];
    c = s + c;
  }
  Uu[t] = c;
}

// Wave speeds from U at the flux points (ele::calcWaveSpFpts), element-major.
template <int D, int P, class R>
__global__ void k_wave_em(int E, const R* __restrict__ Ufe, const double* __restrict__ norme,
                          double* __restrict__ waveSpe, double gamma) {
  using Sh = Shape<D, P>;
  constexpr int F = Sh::F, N = Sh::N;
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)E * F) return;
  long long e = i / F, f = i - e * F;
  double U[N];
#pragma unroll
  for (int k = 0; k < N; k++) U[k] = double(Ufe[(e * N + k) * F + f]);
  double rho = U[0], u = U[1] / rho, v = U[2] / rho, w = 0;
  if constexpr (D == 3) w = U[3] / rho;
  double rhoVSq = rho * (u * u + v * v + w * w);
  double pr = (gamma - 1) * (U[D + 1] - 0.5 * rhoVSq);
  double vN = u * norme[(e * D + 0) * F + f] + v * norme[(e * D + 1) * F + f];
  if constexpr (D == 3) vN += w * norme[(e * D + 2) * F + f];
  double csq = smax(gamma * pr / rho, 0.0);
  waveSpe[i] = (fabs(vN - 0.) + sqrt(csq));
}

// Minimum over the block (a multiple of 64 threads, at most 256), returned to
// every thread: across each wave by lane exchanges, then across the waves
// through red.  min is exact, so the order of the comparisons does not matter.
__device__ __forceinline__ double block_min(double v, double* red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o));
  __syncthreads();  // red may still be read by an earlier call
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  v = red[0];
  for (int w = 1; w < (int)(blockDim.x >> 6); w++) v = fmin(v, red[w]);
  return v;
}

// The element volumes calcDt divides by, once per context (the geometry is
// fixed): the same sum as the reference's per call.
template <int D, int P>
__global__ __launch_bounds__(256) void k_ele_vol(int E, const double* __restrict__ detJe, const double* __restrict__ ws,
                                                  double* __restrict__ vole) {
  constexpr int S = Shape<D, P>::S;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  double vol = 0;
  for (int s = 0; s < S; s++) vol += detJe[(long long)e * S + s] * ws[s];
  vole[e] = vol;
}

// ele::calcDt per element (element-major inputs) and block minima.  With a
// ticket counter, the block that finishes last also reduces the block minima
// to the global dt (solver::calcDt's minimum, src/solver.cpp:447-468) and
// resets the counter to zero for the next launch (a single block's minimum is
// the global dt at once).
template <int D, int P>
__global__ __launch_bounds__(256) void k_ele_dt_em(int E, const double* __restrict__ vole, const double* __restrict__ wf,
                                                    const double* __restrict__ waveSpe, const double* __restrict__ dAe,
                                                    double cflfac, double* __restrict__ eleDt, double* part,
                                                    unsigned* ticket, double* __restrict__ out) {
  using Sh = Shape<D, P>;
  constexpr int F = Sh::F, Ff = Sh::Ff;
  __shared__ double red[4];
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  double dt = INFINITY;
  if (e < E) {
    const double vol = vole[e];
    double iw = 0;
    for (int f = 0; f < F; f++) iw += waveSpe[(long long)e * F + f] * wf[f % Ff] * dAe[(long long)e * F + f];
    dt = cflfac * 2 * vol / iw;
    eleDt[e] = dt;
  }
  const double bmin = block_min(dt, red);
  if (!ticket) return;  // merged meshes: per-mesh minima of eleDt follow
  if (gridDim.x == 1) {  // one block: its minimum is the global one
    if (threadIdx.x == 0) *out = bmin;
    return;
  }
  if (threadIdx.x == 0) part[blockIdx.x] = bmin;
  __shared__ bool last;
  if (threadIdx.x == 0) {
    // the block minimum above is released with the ticket; the last taker
    // acquires every other block's
    const unsigned t = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    last = t == gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return;
  double v = INFINITY;
  for (int i = threadIdx.x; i < (int)gridDim.x; i += blockDim.x) v = fmin(v, part[i]);
  v = block_min(v, red);
  if (threadIdx.x == 0) {
    *out = v;
    __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Per-mesh minimum of the element time steps (batched contexts, dt_type 1).
__global__ void k_mesh_dt_min(const int* __restrict__ start, const double* __restrict__ eleDt,
                              double* __restrict__ meshDt) {
  __shared__ double red[256];
  const int m = blockIdx.x;
  double v = INFINITY;
  for (int e = start[m] + threadIdx.x; e < start[m + 1]; e += blockDim.x) v = fmin(v, eleDt[e]);
  red[threadIdx.x] = v;
  __syncthreads();
  for (int st = blockDim.x / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + st]);
    __syncthreads();
  }
  if (threadIdx.x == 0) meshDt[m] = red[0];
}

// Per-mesh residual norm over divF_spts[0] (element-major), as
// ele::getNormResidual summed over the mesh's elements (src/ele.cpp:1922-1945).
__global__ void k_mesh_res(int S, int N, int type, const int* __restrict__ start, const double* __restrict__ divF0e,
                           const double* __restrict__ detJe, const double* __restrict__ ws, double* __restrict__ out) {
  __shared__ double red[256];
  const int m = blockIdx.x;
  for (int k = 0; k < N; k++) {
    double r = 0;
    for (long long i = (long long)start[m] * S + threadIdx.x; i < (long long)start[m + 1] * S; i += blockDim.x) {
      long long e = i / S;
      int s = (int)(i - e * S);
      double v = divF0e[(e * N + k) * S + s], dj = detJe[i];
      if (type == 1) r += fabs(v) * ws[s];
      else if (type == 2) r += v * v / dj * ws[s];
      else r = fmax(r, fabs(v) / dj);
    }
    red[threadIdx.x] = r;
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
      if (threadIdx.x < st)
        red[threadIdx.x] = (type == 3) ? fmax(red[threadIdx.x], red[threadIdx.x + st]) : red[threadIdx.x] + red[threadIdx.x + st];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[(long long)m * N + k] = (type == 2) ? sqrt(red[0]) : red[0];
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Host side.

static FusedState& fs(Ctx& c) { return *reinterpret_cast<FusedState*>(c.fused); }

template <class T>
static T* dal(size_t n) {
  void* p = nullptr;
  FCHK(hipMalloc(&p, (n ? n : 1) * sizeof(T)));
  poison_fill(p, (n ? n : 1) * sizeof(T));
  return (T*)p;
}

// Extract the non-zero structure of the dense operators.  Returns false if an
// operator row has more non-zeros than the tensor-product structure allows.
// opp_grad_spts as one 1-D derivative matrix (Tabs::d1): true when every
// entry equals its tensor-product counterpart d1[coordinate of q along d]
// [coordinate of r along d] on q's d-line and is zero off it (the dense
// operator of a tensor-product element on one 1-D point set is: its factors
// in the other directions are exactly 0 or 1).
static bool build_d1(const Ctx& c, const fhip_desc* dsc, std::vector<double>& d1) {
  const int D = c.z.D, S = c.z.S, P1 = c.z.P + 1;
  auto coord = [&](int q, int dd) { return dd == 0 ? q % P1 : (dd == 1 ? (q / P1) % P1 : q / (P1 * P1)); };
  auto tangent = [&](int q, int dd) {  // the point's coordinates in the other directions
    return dd == 0 ? q / P1 : (dd == 1 ? q % P1 + P1 * (q / (P1 * P1)) : q % (P1 * P1));
  };
  d1.assign(P1 * P1, 0.);
  for (int i = 0; i < P1; i++)
    for (int m = 0; m < P1; m++) d1[i * P1 + m] = dsc->opp_grad_spts[(size_t)i * S + m];  // d = 0, line 0
  for (int dd = 0; dd < D; dd++)
    for (int q = 0; q < S; q++)
      for (int r = 0; r < S; r++) {
        const double v = dsc->opp_grad_spts[((size_t)dd * S + q) * S + r];
        const bool on = tangent(r, dd) == tangent(q, dd);
        if (!(v == (on ? d1[coord(q, dd) * P1 + coord(r, dd)] : 0.))) return false;  // +0 == -0
      }
  return true;
}

static bool build_tables(const Ctx& c, const fhip_desc* d, std::vector<char>& blob, size_t off[17]) {
  const int D = c.z.D, S = c.z.S, F = c.z.F, P1 = c.z.P + 1, NC = 2 * D;
  std::vector<int> ext_idx((size_t)F * P1, 0), efn_dir(F, 0), grd_idx((size_t)D * S * P1, 0),
      cor_idx((size_t)S * NC, 0), cu_idx((size_t)D * S * 2, 0);
  std::vector<double> ext_w((size_t)F * P1, 0.), efn_w((size_t)F * P1, 0.), grd_w((size_t)D * S * P1, 0.),
      cor_w((size_t)S * NC, 0.), cu_w((size_t)D * S * 2, 0.);
  for (int f = 0; f < F; f++) {
    // the face's normal direction: the only d with non-zero extrapolate_fn entries
    int dn = -1;
    for (int dd = 0; dd < D; dd++)
      for (int q = 0; q < S; q++)
        if (d->opp_extrapolate_fn[((size_t)dd * F + f) * S + q] != 0.) {
          if (dn >= 0 && dn != dd) return false;
          dn = dd;
        }
    efn_dir[f] = dn < 0 ? 0 : dn;
    int n = 0;
    for (int q = 0; q < S; q++) {
      double v = d->opp_spts_to_fpts[(size_t)f * S + q];
      double w = dn < 0 ? 0. : d->opp_extrapolate_fn[((size_t)dn * F + f) * S + q];
      if (v == 0. && w == 0.) continue;
      if (n >= P1) return false;
      ext_idx[(size_t)f * P1 + n] = spt_lds(D, P1, q);
      ext_w[(size_t)f * P1 + n] = v;
      efn_w[(size_t)f * P1 + n] = w;
      n++;
    }
    for (int m = n; m < P1; m++) ext_idx[(size_t)f * P1 + m] = ext_idx[(size_t)f * P1];
  }
  for (int dd = 0; dd < D; dd++)
    for (int q0 = 0; q0 < S; q0++) {
      int n = 0;
      for (int q = 0; q < S; q++) {
        double v = d->opp_grad_spts[((size_t)dd * S + q0) * S + q];
        if (v == 0.) continue;
        if (n >= P1) return false;
        grd_idx[((size_t)dd * S + q0) * P1 + n] = spt_lds(D, P1, q);
        grd_w[((size_t)dd * S + q0) * P1 + n] = v;
        n++;
      }
      for (int m = n; m < P1; m++) grd_idx[((size_t)dd * S + q0) * P1 + m] = spt_lds(D, P1, q0);
    }
  for (int q = 0; q < S; q++) {
    int n = 0;
    for (int f = 0; f < F; f++) {
      double v = d->opp_correction[(size_t)q * F + f];
      if (v == 0.) continue;
      if (n >= NC) return false;
      cor_idx[(size_t)q * NC + n] = f;
      cor_w[(size_t)q * NC + n] = v;
      n++;
    }
    if (c.prm.viscous)
      for (int dd = 0; dd < D; dd++) {
        int m = 0;
        for (int f = 0; f < F; f++) {
          double v = d->opp_correct_u[((size_t)dd * S + q) * F + f];
          if (v == 0.) continue;
          if (m >= 2) return false;
          cu_idx[((size_t)dd * S + q) * 2 + m] = f;
          cu_w[((size_t)dd * S + q) * 2 + m] = v;
          m++;
        }
      }
  }
  auto put = [&](const void* p, size_t bytes, size_t& o) {
    o = (blob.size() + 255) / 256 * 256;
    blob.resize(o + bytes);
    std::memcpy(blob.data() + o, p, bytes);
  };
  blob.clear();
  put(ext_idx.data(), ext_idx.size() * 4, off[0]);
  put(ext_w.data(), ext_w.size() * 8, off[1]);
  put(efn_dir.data(), efn_dir.size() * 4, off[2]);
  put(efn_w.data(), efn_w.size() * 8, off[3]);
  put(grd_idx.data(), grd_idx.size() * 4, off[4]);
  put(grd_w.data(), grd_w.size() * 8, off[5]);
  put(cor_idx.data(), cor_idx.size() * 4, off[6]);
  put(cor_w.data(), cor_w.size() * 8, off[7]);
  put(cu_idx.data(), cu_idx.size() * 4, off[8]);
  put(cu_w.data(), cu_w.size() * 8, off[9]);
  auto putf = [&](const std::vector<double>& v, size_t& o) {
    std::vector<float> f(v.begin(), v.end());
    put(f.data(), f.size() * 4, o);
  };
  putf(ext_w, off[10]);
  putf(efn_w, off[11]);
  putf(grd_w, off[12]);
  putf(cor_w, off[13]);
  putf(cu_w, off[14]);
  // the matrix-core path of FP64 P3 hexes needs the 1-D derivative matrix
  std::vector<double> d1;
  off[15] = off[16] = 0;
  if (build_d1(c, d, d1)) {
    put(d1.data(), d1.size() * 8, off[15]);
    putf(d1, off[16]);
  } else if (D == 3 && P1 == 4) {
    return false;
  }
  return true;
}

__global__ void k_round_f32(long long n, const double* __restrict__ in, float* __restrict__ out) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    out[i] = float(in[i]);
}

// Geometry views of one precision (Work::JG .. nf).
static void set_geometry(Ctx& c, Work<double>& w) {
  FusedState& st = fs(c);
  w.JG = st.JGe;
  w.iJ = st.iJe;
  w.nrm = st.norme;
  w.dA = st.dAe;
  w.nf = st.nfe;
}
static void set_geometry(Ctx& c, Work<float>& w) {
  FusedState& st = fs(c);
  const long long E = c.z.E, S = c.z.S, F = c.z.F, D = c.z.D;
  const long long n[5] = {E * D * D * S, E * S, E * D * F, E * F, st.nfe ? E * D * F : 0};
  const double* src[5] = {st.JGe, st.iJe, st.norme, st.dAe, st.nfe};
  long long tot = 0;
  for (long long v : n) tot += (v + 63) / 64 * 64;
  float* m = dal<float>(tot);
  w.geo_mem = m;
  const float** dst[5] = {&w.JG, &w.iJ, &w.nrm, &w.dA, &w.nf};
  for (int i = 0; i < 5; i++) {
    *dst[i] = n[i] ? m : nullptr;
    if (n[i]) {
      hipLaunchKernelGGL(k_round_f32, dim3(1024), dim3(256), 0, c.stream, n[i], src[i], m);
      FCHK(hipGetLastError());
    }
    m += (n[i] + 63) / 64 * 64;
  }
}

template <class R>
static void ensure_work(Ctx& c) {
  Work<R>& w = fs(c).work<R>();
  if (w.ready) return;
  const long long E = c.z.E, S = c.z.S, F = c.z.F, N = c.z.N, D = c.z.D;
  w.Ufe = dal<R>(E * N * F);
  w.Fne = dal<R>(E * N * F);
  if (c.prm.viscous) {
    // inviscid runs keep divF and the discontinuous normal flux in the volume kernel
    w.divFe = dal<R>(E * N * S);
    w.disFne = dal<R>(E * N * F);
    w.dUce = dal<R>(E * N * F);
    w.Bve = dal<R>(E * N * F);
    w.gBnd = dal<R>((long long)D * N * fs(c).nBnd);
  }
  if (sizeof(R) == 4 && c.z.nRK > 1 && c.prm.time_type < 5) w.Rst = dal<R>((long long)(c.z.nRK - 1) * E * N * S);
  w.Us = dal<R>(E * N * S);
  set_geometry(c, w);
  w.ready = true;
}

bool fused_setup(Ctx& c, const fhip_desc* d) {
  const Sizes& z = c.z;
  if (z.P < 0 || z.P > 5) return false;
  std::vector<char> blob;
  size_t off[17];
  if (!build_tables(c, d, blob, off)) return false;
  auto* st = new FusedState();
  c.fused = st;
  st->tab_mem = dal<char>(blob.size());
  FCHK(hipMemcpy(st->tab_mem, blob.data(), blob.size(), hipMemcpyHostToDevice));
  char* base = (char*)st->tab_mem;
  st->tb = Tabs{(const int*)(base + off[0]),    (const double*)(base + off[1]), (const int*)(base + off[2]),
                (const double*)(base + off[3]), (const int*)(base + off[4]),    (const double*)(base + off[5]),
                (const int*)(base + off[6]),    (const double*)(base + off[7]), (const int*)(base + off[8]),
                (const double*)(base + off[9]), (const float*)(base + off[10]), (const float*)(base + off[11]),
                (const float*)(base + off[12]), (const float*)(base + off[13]), (const float*)(base + off[14]),
                off[15] ? (const double*)(base + off[15]) : nullptr, off[16] ? (const float*)(base + off[16]) : nullptr};
  const long long E = z.E, S = z.S, F = z.F, N = z.N, D = z.D;
  st->stream = E * N * S * (long long)sizeof(double) > kStreamBytes;
  st->Ue = dal<double>(E * N * S);
  st->Uacc = dal<double>(E * N * S);
  st->divF0e = dal<double>(E * N * S);
  st->waveSpe = dal<double>(E * F);
  st->dtTicket = dal<unsigned>(1);
  FCHK(hipMemset(st->dtTicket, 0, sizeof(unsigned)));
  st->JGe = dal<double>(E * D * D * S);
  st->detJe = dal<double>(E * S);
  st->iJe = dal<double>(E * S);
  st->norme = dal<double>(E * D * F);
  st->dAe = dal<double>(E * F);
  {
    std::vector<double> jg((size_t)E * D * D * S), dj((size_t)E * S), nr((size_t)E * D * F), da((size_t)E * F);
    for (long long r = 0; r < D; r++)
      for (long long s = 0; s < S; s++)
        for (long long e = 0; e < E; e++)
          for (long long x = 0; x < D; x++)
            jg[((e * D + r) * D + x) * S + s] = d->jginv_spts[((r * S + s) * E + e) * D + x];
    for (long long s = 0; s < S; s++)
      for (long long e = 0; e < E; e++) dj[e * S + s] = d->detjac_spts[s * E + e];
    for (long long f = 0; f < F; f++)
      for (long long e = 0; e < E; e++) {
        da[e * F + f] = d->da_fpts[f * E + e];
        for (long long x = 0; x < D; x++) nr[(e * D + x) * F + f] = d->norm_fpts[(f * E + e) * D + x];
      }
    FCHK(hipMemcpy(st->JGe, jg.data(), jg.size() * 8, hipMemcpyHostToDevice));
    FCHK(hipMemcpy(st->detJe, dj.data(), dj.size() * 8, hipMemcpyHostToDevice));
    for (double& v : dj) v = 1. / v;
    FCHK(hipMemcpy(st->iJe, dj.data(), dj.size() * 8, hipMemcpyHostToDevice));
    FCHK(hipMemcpy(st->norme, nr.data(), nr.size() * 8, hipMemcpyHostToDevice));
    FCHK(hipMemcpy(st->dAe, da.data(), da.size() * 8, hipMemcpyHostToDevice));
  }
  // batching: meshes, boundary-state table
  if (d->n_meshes > 1 || d->mesh_ele_start) {
    int nm = std::max(1, d->n_meshes);
    if (!d->mesh_ele_start) throw std::invalid_argument("n_meshes > 1 needs mesh_ele_start");
    st->h_meshStart.assign(d->mesh_ele_start, d->mesh_ele_start + nm + 1);
    if (st->h_meshStart[0] != 0 || st->h_meshStart[nm] != z.E) throw std::invalid_argument("mesh_ele_start must span [0, nEles]");
    std::vector<int> em(z.E);
    for (int m = 0; m < nm; m++) {
      if (st->h_meshStart[m + 1] <= st->h_meshStart[m]) throw std::invalid_argument("mesh_ele_start must increase");
      for (int e = st->h_meshStart[m]; e < st->h_meshStart[m + 1]; e++) em[e] = m;
    }
    st->nMesh = nm;
    st->eleMesh = dal<int>(z.E);
    st->meshStart = dal<int>(nm + 1);
    st->meshDt = dal<double>(nm);
    st->meshRes = dal<double>((size_t)nm * z.N);
    FCHK(hipMemcpy(st->eleMesh, em.data(), em.size() * 4, hipMemcpyHostToDevice));
    FCHK(hipMemcpy(st->meshStart, st->h_meshStart.data(), (nm + 1) * 4, hipMemcpyHostToDevice));
  } else {
    st->h_meshStart = {0, z.E};
  }
  if (d->n_bnd_states > 0) {
    if (!d->bnd_states || !d->face_bnd_state) throw std::invalid_argument("boundary-state table needs bnd_states and face_bnd_state");
    st->bst = dal<double>((size_t)d->n_bnd_states * 5);
    FCHK(hipMemcpy(st->bst, d->bnd_states, (size_t)d->n_bnd_states * 5 * 8, hipMemcpyHostToDevice));
  }
  {
    std::vector<int> pl(z.nFaceFpts), pr(z.nFaceFpts), bc(z.nFaceFpts);
    for (int t = 0; t < z.nFaceFpts; t++) {
      int a = c.h_ptL[t];  // fpt*E + ele
      pl[t] = (a % z.E) * z.F + a / z.E;
      int r = c.h_ptR[t];
      pr[t] = (r < 0) ? -1 : (r % z.E) * z.F + r / z.E;
      int row = -1;
      if (r < 0 && d->n_bnd_states > 0) {
        row = d->face_bnd_state[t / z.Ff];
        if (row < -1 || row >= d->n_bnd_states) throw std::invalid_argument("face_bnd_state out of range");
      }
      bc[t] = (r < 0) ? (d->face_bc[t / z.Ff] | ((row + 1) << 4)) : -1;
    }
    st->pL = dal<int>(pl.size());
    st->pR = dal<int>(pr.size());
    st->bc = dal<int>(bc.size());
    FCHK(hipMemcpy(st->pL, pl.data(), pl.size() * 4, hipMemcpyHostToDevice));
    FCHK(hipMemcpy(st->pR, pr.data(), pr.size() * 4, hipMemcpyHostToDevice));
    FCHK(hipMemcpy(st->bc, bc.data(), bc.size() * 4, hipMemcpyHostToDevice));
    {
      // every element flux point must lie on exactly one face point pair;
      // its code: 4 * partner flux point + side, or 4 * boundary slot + 2
      // (viscous runs also keep the face's normal at every point)
      const bool visc = c.prm.viscous != 0;
      if (E * F >= (1LL << 29)) throw std::invalid_argument("fused path: more than 2^29 element flux points");
      std::vector<int> part((size_t)E * F, -1), bt;
      std::vector<double> nl(visc ? (size_t)E * D * F : 0);
      // the left element's normal at flux point q (element-major q = e*F + f)
      auto put_normal = [&](int q, int qL) {
        if (!visc) return;
        const long long e = q / F, f = q - e * F, eL = qL / F, fL = qL - eL * F;
        for (long long x = 0; x < D; x++) nl[(e * D + x) * F + f] = d->norm_fpts[(fL * E + eL) * D + x];
      };
      for (int t = 0; t < z.nFaceFpts; t++) {
        if (part[pl[t]] != -1 || (pr[t] >= 0 && part[pr[t]] != -1)) {
          fused_release(c);
          return false;
        }
        put_normal(pl[t], pl[t]);
        if (pr[t] < 0) {
          bt.push_back(t);
          part[pl[t]] = 4 * (int)(bt.size() - 1) + 2;
        } else {
          part[pl[t]] = 4 * pr[t];
          part[pr[t]] = 4 * pl[t] + 1;
          put_normal(pr[t], pl[t]);
        }
      }
      for (int q : part)
        if (q == -1) {
          fused_release(c);
          return false;
        }
      st->fpart = dal<int>(part.size());
      FCHK(hipMemcpy(st->fpart, part.data(), part.size() * 4, hipMemcpyHostToDevice));
      if (visc) {
        st->nBnd = (int)bt.size();
        st->bndT = dal<int>(bt.size());
        st->nfe = dal<double>(nl.size());
        FCHK(hipMemcpy(st->nfe, nl.data(), nl.size() * 8, hipMemcpyHostToDevice));
        if (!bt.empty()) FCHK(hipMemcpy(st->bndT, bt.data(), bt.size() * 4, hipMemcpyHostToDevice));
      }
    }
  }
  return true;
}

void fused_mesh_residuals(Ctx& c, int type, double* out) {
  FusedState& st = fs(c);
  const int nm = st.nMesh;
  int* start = st.meshStart;
  double* res = st.meshRes;
  int* tmpStart = nullptr;
  double* tmpRes = nullptr;
  if (nm == 1) {  // single mesh: a two-entry range
    tmpStart = dal<int>(2);
    tmpRes = dal<double>(c.z.N);
    int h[2] = {0, c.z.E};
    FCHK(hipMemcpyAsync(tmpStart, h, 8, hipMemcpyHostToDevice, c.stream));
    start = tmpStart;
    res = tmpRes;
  }
  hipLaunchKernelGGL(k_mesh_res, dim3(nm), dim3(256), 0, c.stream, c.z.S, c.z.N, type, start, st.divF0e, st.detJe, c.ws, res);
  FCHK(hipGetLastError());
  FCHK(hipMemcpyAsync(out, res, (size_t)nm * c.z.N * 8, hipMemcpyDeviceToHost, c.stream));
  FCHK(hipStreamSynchronize(c.stream));
  if (tmpStart) (void)hipFree(tmpStart);
  if (tmpRes) (void)hipFree(tmpRes);
}

int fused_n_meshes(Ctx& c) { return fs(c).nMesh; }

void fused_set_active(Ctx& c, const int* active) {
  FusedState& st = fs(c);
  if (!st.active || !st.eleMesh) c.drop_graphs();  // captured steps hold null flag pointers
  if (!st.active) st.active = dal<int>(st.nMesh);
  if (!st.eleMesh) {  // single mesh: every element belongs to mesh 0
    std::vector<int> em(c.z.E, 0);
    st.eleMesh = dal<int>(c.z.E);
    FCHK(hipMemcpy(st.eleMesh, em.data(), em.size() * 4, hipMemcpyHostToDevice));
  }
  FCHK(hipMemcpyAsync(st.active, active, (size_t)st.nMesh * 4, hipMemcpyHostToDevice, c.stream));
  FCHK(hipStreamSynchronize(c.stream));
}

void fused_release(Ctx& c) {
  if (c.fused) delete reinterpret_cast<FusedState*>(c.fused);
  c.fused = nullptr;
}

// ---- per-order / per-precision dispatch ------------------------------------

template <int D, int P, class R>
struct Runner {
  using Sh = Shape<D, P>;
  static constexpr int S = Sh::S, F = Sh::F, N = Sh::N;
  static const char* nm(const char* f64, const char* f32) { return sizeof(R) == 8 ? f64 : f32; }
  // The face kernel stores each side's flux into that element's rows of one
  // face (Ff values per field).  Streamed (nontemporal) stores pay only when
  // such a row is whole 64-byte blocks, the memory's write granule: a partial
  // block written past the caches costs a read-modify-write (measured: FP64
  // P2/P4/P5 face kernels 1.7-1.4x faster without streaming, P3 rows of 128
  // and 64 bytes faster with it).
  static constexpr bool kFaceRowsWhole = (Sh::Ff * sizeof(R)) % 64 == 0;
  // The element kernels store rows of S or F values per field; a wave's
  // streamed store covers whole blocks only when those rows are whole blocks.
  static constexpr bool kEleRowsWhole = (S * sizeof(R)) % 64 == 0 && (F * sizeof(R)) % 64 == 0;
  static int& blk(Ctx& c, int k) { return fs(c).nt[sizeof(R) == 4 ? 1 : 0][k]; }
  // LDS bytes of an element kernel at block size nt (0 for the face kernels)
  static size_t lds(int k, int nt, bool visc) {
    const size_t EB = nt / S;
    const size_t SF = (visc && F > D * S) ? F : D * S;
    if (k == kExtrap) return EB * N * S * sizeof(R);
    if (k == kVolume)
      return EB * N * (S + SF) * sizeof(R) +
             (visc ? EB * (D * F * sizeof(R) + F * sizeof(int) + (kJumpInVolume<R> ? N * F * sizeof(R) : 0))
                   : EB * (N * F * sizeof(R) + F * sizeof(int)));
    if (k == kRk) return EB * N * (S + F) * sizeof(R);
    return 0;
  }
  static bool valid(int k, int nt, bool visc) {
    if (nt < 64 || nt > kMaxBlock) return false;
    if (k == kFaceInv || k == kFaceBnd) return true;
    return nt >= S && lds(k, nt, visc) <= kMaxLds;
  }
  static void launch_extrap(Ctx& c, int nt) {
    FusedState& st = fs(c);
    Work<R>& w = st.work<R>();
    // the FP32 stage copy starts from the state; the FP64 one is first
    // written by a step's first stage (which reads the state itself)
    hipLaunchKernelGGL((k_extrap<D, P, R>), dim3(nblk(c.z.E, nt / S)), dim3(nt), lds(kExtrap, nt, false), c.stream,
                       c.z.E, st.Ue, w.Ufe, sizeof(R) == 4 ? w.Us : nullptr, st.tb);
    FCHK(hipGetLastError());
  }
  static void launch_face_inv(Ctx& c, int nt) {
    FusedState& st = fs(c);
    Work<R>& w = st.work<R>();
    const int nFF = c.z.nFaceFpts;
    if (nFF == 0) return;
    if (st.stream && kFaceRowsWhole)
      hipLaunchKernelGGL((k_face_inv<D, P, R, true>), dim3(nblk(nFF, nt)), dim3(nt), 0, c.stream, nFF, st.pL, st.pR,
                         st.bc, w.Ufe, w.nrm, w.Fne, st.waveSpe, w.dUce, c.ph, st.bst);
    else
      hipLaunchKernelGGL((k_face_inv<D, P, R, false>), dim3(nblk(nFF, nt)), dim3(nt), 0, c.stream, nFF, st.pL, st.pR,
                         st.bc, w.Ufe, w.nrm, w.Fne, st.waveSpe, w.dUce, c.ph, st.bst);
    FCHK(hipGetLastError());
  }
  template <bool VISC, bool NT>
  static void launch_volume_t(Ctx& c, int nt, const RkArgs<R>& ra) {
    FusedState& st = fs(c);
    Work<R>& w = st.work<R>();
    const VFace<R> vf{st.fpart, w.nf, w.Bve, w.gBnd, w.dA, st.nBnd};
    const R* us = (sizeof(R) == 4 || st.read_stage) ? w.Us : nullptr;
    const Tabs& tb = st.tb;
    // quads and FP32 are held to 4 waves per SIMD (128 VGPRs); FP64 hexes,
    // which spill under that hold, to 3 (168 VGPRs: P3 without spills, where
    // the unheld compiler would take 211 and leave 2 waves)
    if constexpr (kBf16<D, P, R>) {
      if (c.bf16 == 2 || c.bf16 == 3) {
        auto kern = c.bf16 == 2 ? k_volume_w4<D, P, VISC, R, NT, 2> : k_volume_w4<D, P, VISC, R, NT, 3>;
        hipLaunchKernelGGL(kern, dim3(nblk(c.z.E, nt / S)), dim3(nt), lds(kVolume, nt, VISC), c.stream, c.z.E, st.Ue,
                           us, w.dUce, w.JG, w.iJ, w.divFe, w.disFne, vf, tb, c.ph, w.Fne, ra);
        FCHK(hipGetLastError());
        return;
      }
    }
    if constexpr (kHold<D, R>)
      hipLaunchKernelGGL((k_volume_w4<D, P, VISC, R, NT>), dim3(nblk(c.z.E, nt / S)), dim3(nt),
                         lds(kVolume, nt, VISC), c.stream, c.z.E, st.Ue, us, w.dUce, w.JG, w.iJ, w.divFe, w.disFne,
                         vf, tb, c.ph, w.Fne, ra);
    else
      hipLaunchKernelGGL((k_volume_w3<D, P, VISC, R, NT>), dim3(nblk(c.z.E, nt / S)), dim3(nt), lds(kVolume, nt, VISC),
                         c.stream, c.z.E, st.Ue, us, w.dUce, w.JG, w.iJ, w.divFe, w.disFne, vf, tb, c.ph, w.Fne, ra);
    FCHK(hipGetLastError());
  }
  // viscous runs: the volume kernel (ra unused); inviscid runs: the volume
  // kernel with the RK update of stage ra
  static void launch_volume(Ctx& c, int nt, const RkArgs<R>& ra) {
    const bool sm = fs(c).stream && kEleRowsWhole;
    if (c.prm.viscous) sm ? launch_volume_t<true, true>(c, nt, ra) : launch_volume_t<true, false>(c, nt, ra);
    else sm ? launch_volume_t<false, true>(c, nt, ra) : launch_volume_t<false, false>(c, nt, ra);
  }
  static void launch_face_bnd(Ctx& c, int nt) {
    FusedState& st = fs(c);
    Work<R>& w = st.work<R>();
    if (st.nBnd == 0) return;
    hipLaunchKernelGGL((k_face_bnd<D, P, R>), dim3(nblk(st.nBnd, nt)), dim3(nt), 0, c.stream, st.nBnd, st.bndT, st.pL,
                       st.bc, w.Ufe, w.gBnd, w.nrm, w.dA, w.Bve, w.disFne, c.ph, st.bst);
    FCHK(hipGetLastError());
  }
  static RkArgs<R> rk_args(Ctx& c, int mode, double a, double b, bool write0, int stage) {
    FusedState& st = fs(c);
    Work<R>& w = st.work<R>();
    RkB bs{};
    for (int j = 0; j < c.z.nRK && j < 8; j++) bs.b[j] = c.prm.rk_b[j];
    // residual modes 7 and 8: the source is the array written, or read
    double* d0 = mode == 7 ? st.srce : (write0 ? st.divF0e : nullptr);
    const double* src = mode == 8 ? st.srce : (st.use_src && mode != 7 ? st.srce : nullptr);
    return RkArgs<R>{mode, a, b, st.detJe, c.d_dt, c.eleDt,
                     c.prm.dt_type == 2 ? 1 : (c.prm.dt_type == 1 && st.nMesh > 1 ? 2 : 0), st.Ue,
                     w.Us ? nullptr : st.U0e, st.Uacc,
                     d0, st.eleMesh, st.meshDt, st.active, src,
                     (w.Rst && !st.use_src) ? w.Rst : nullptr, stage, bs, w.Us,
                     (mode == 2 || mode == 3 || (mode == 4 && stage == c.z.nRK - 1)) ? 1 : 0, w.Ufe};
  }
  // the RK kernel of viscous runs; inviscid runs do its work at the end of
  // their volume kernel
  static void launch_rk(Ctx& c, int nt, int mode, double a, double b, bool write0, int stage = 0) {
    FusedState& st = fs(c);
    Work<R>& w = st.work<R>();
    auto* kern = (st.stream && kEleRowsWhole) ? k_rk<D, P, R, true> : k_rk<D, P, R, false>;
    hipLaunchKernelGGL(kern, dim3(nblk(c.z.E, nt / S)), dim3(nt), lds(kRk, nt, false), c.stream, c.z.E, w.divFe, w.Fne,
                       w.disFne, rk_args(c, mode, a, b, write0, stage), st.tb, st.fpart, w.dA, w.Bve);
    FCHK(hipGetLastError());
  }
  static void extrap(Ctx& c) {
    Launch L(c, nm("fused_extrapolate", "mixed_extrapolate"));
    launch_extrap(c, blk(c, kExtrap));
  }
  static void wave(Ctx& c) {
    FusedState& st = fs(c);
    long long n = (long long)c.z.E * Sh::F;
    Launch L(c, nm("fused_wave_speed", "mixed_wave_speed"));
    hipLaunchKernelGGL((k_wave_em<D, P, R>), dim3(nblk(n, 256)), dim3(256), 0, c.stream, c.z.E, st.work<R>().Ufe, st.norme,
                       st.waveSpe, c.ph.gamma);
    FCHK(hipGetLastError());
  }
  static void dt(Ctx& c, double cflfac) {
    FusedState& st = fs(c);
    unsigned nb = nblk(c.z.E, 256);
    {
      // one mesh: the element kernel's last block forms the global minimum
      Launch L(c, nm("fused_ele_dt", "mixed_ele_dt"));
      const bool one = st.nMesh == 1;
      hipLaunchKernelGGL((k_ele_dt_em<D, P>), dim3(nb), dim3(256), 0, c.stream, c.z.E, st.vole, c.wf, st.waveSpe,
                         st.dAe, cflfac, c.eleDt, c.partial, one ? st.dtTicket : nullptr, c.d_dt);
      FCHK(hipGetLastError());
    }
    // merged meshes: per-mesh minima for a global CFL dt per mesh (a local
    // CFL dt uses the element values alone)
    if (st.nMesh > 1 && c.prm.dt_type == 1) {
      Launch L(c, nm("fused_dt_min", "mixed_dt_min"));
      hipLaunchKernelGGL(k_mesh_dt_min, dim3(st.nMesh), dim3(256), 0, c.stream, st.meshStart, c.eleDt, st.meshDt);
      FCHK(hipGetLastError());
    }
  }
  // calcResidual up to what the RK update needs: viscous runs the face,
  // volume and boundary kernels; inviscid runs the face kernel only (their
  // volume kernel runs with the update, in rk below)
  static void residual(Ctx& c) {
    const bool visc = c.prm.viscous != 0;
    {
      Launch L(c, nm("fused_face_inviscid", "mixed_face_inviscid"));
      launch_face_inv(c, blk(c, kFaceInv));
    }
    if (!visc) return;
    {
      Launch L(c, nm("fused_volume_viscous", "mixed_volume_viscous"));
      launch_volume(c, blk(c, kVolume), RkArgs<R>{});
    }
    if (fs(c).nBnd > 0) {
      Launch L(c, nm("fused_face_boundary", "mixed_face_boundary"));
      launch_face_bnd(c, blk(c, kFaceBnd));
    }
  }
  static void rk(Ctx& c, int mode, double a, double b, bool write0, int stage) {
    if (!c.prm.viscous) {
      Launch L(c, nm("fused_volume_rk", "mixed_volume_rk"));
      launch_volume(c, blk(c, kVolume), rk_args(c, mode, a, b, write0, stage));
      return;
    }
    Launch L(c, nm("fused_rk_update", "mixed_rk_update"));
    launch_rk(c, blk(c, kRk), mode, a, b, write0, stage);
  }
  // Time every valid block size of every kernel on this context's data and
  // keep the fastest.  The RK kernel (and the volume kernel of inviscid runs,
  // which includes the update) runs in its middle-stage mode, which rewrites
  // U and the RK accumulator: both are saved and restored, as are the wave
  // speeds the face kernel writes; block shapes never change results.
  static void tune(Ctx& c) {
    FusedState& st = fs(c);
    const bool visc = c.prm.viscous != 0;
    const long long nS = (long long)c.z.E * N * S, nW = (long long)c.z.E * F;
    double *bU = nullptr, *bA = nullptr, *bW = nullptr;
    FCHK(hipMalloc(&bU, nS * 8));
    FCHK(hipMalloc(&bA, nS * 8));
    FCHK(hipMalloc(&bW, nW * 8));
    FCHK(hipMemcpyAsync(bU, st.Ue, nS * 8, hipMemcpyDeviceToDevice, c.stream));
    FCHK(hipMemcpyAsync(bA, st.Uacc, nS * 8, hipMemcpyDeviceToDevice, c.stream));
    FCHK(hipMemcpyAsync(bW, st.waveSpe, nW * 8, hipMemcpyDeviceToDevice, c.stream));
    hipEvent_t e0, e1;
    FCHK(hipEventCreate(&e0));
    FCHK(hipEventCreate(&e1));
    // multiples of the 64-wide wavefront; element kernels hold floor(NT/S)
    // elements, so sizes between the powers of two pay off where S divides them
    const int cand[7] = {64, 128, 192, 256, 320, 384, 512};
    for (int k = kExtrap; k <= kRk; k++) {
      if ((k == kFaceBnd && (!visc || st.nBnd == 0)) || (k == kRk && !visc)) continue;
      auto run = [&](int nt) {
        switch (k) {
          case kExtrap: launch_extrap(c, nt); break;
          case kFaceInv: launch_face_inv(c, nt); break;
          case kVolume: launch_volume(c, nt, visc ? RkArgs<R>{} : rk_args(c, 1, 0.5, 1. / 3., false, 0)); break;
          case kFaceBnd: launch_face_bnd(c, nt); break;
          default: launch_rk(c, nt, 1, 0.5, 1. / 3., false);
        }
      };
      int best = blk(c, k);
      float bt = 1e30f;
      for (int nt : cand) {
        if (!valid(k, nt, visc)) continue;
        run(nt);
        FCHK(hipEventRecord(e0, c.stream));
        for (int r = 0; r < 3; r++) run(nt);
        FCHK(hipEventRecord(e1, c.stream));
        FCHK(hipEventSynchronize(e1));
        float ms = 0;
        FCHK(hipEventElapsedTime(&ms, e0, e1));
        if (ms < bt) bt = ms, best = nt;
      }
      blk(c, k) = best;
      run(best);  // leave this kernel's outputs for the next one's inputs
    }
    FCHK(hipMemcpyAsync(st.Ue, bU, nS * 8, hipMemcpyDeviceToDevice, c.stream));
    FCHK(hipMemcpyAsync(st.Uacc, bA, nS * 8, hipMemcpyDeviceToDevice, c.stream));
    FCHK(hipMemcpyAsync(st.waveSpe, bW, nW * 8, hipMemcpyDeviceToDevice, c.stream));
    FCHK(hipStreamSynchronize(c.stream));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(bU);
    (void)hipFree(bA);
    (void)hipFree(bW);
    launch_extrap(c, blk(c, kExtrap));  // U at the faces of the restored state
  }
  static bool check(Ctx& c, const int* nt) {
    const bool visc = c.prm.viscous != 0;
    for (int k = kExtrap; k <= kRk; k++)
      if (!valid(k, nt[k], visc)) return false;
    return true;
  }
  // Each block size brought within its kernel's limits: the default of 256
  // threads exceeds kMaxLds for the RK kernel and the viscous volume kernel
  // of P0 hexes (256 elements of 7 FP64 N-vectors per block)
  static void fit(Ctx& c) {
    const bool visc = c.prm.viscous != 0;
    for (int k = kExtrap; k <= kRk; k++)
      while (!valid(k, blk(c, k), visc) && blk(c, k) > 64) blk(c, k) -= 64;
  }
  static void ensure(Ctx& c) {
    ensure_work<R>(c);
    fit(c);
    FusedState& st = fs(c);
    if (!st.vole) {
      st.vole = dal<double>(c.z.E);
      hipLaunchKernelGGL((k_ele_vol<D, P>), dim3(nblk(c.z.E, 256)), dim3(256), 0, c.stream, c.z.E, st.detJe, c.ws,
                         st.vole);
      FCHK(hipGetLastError());
    }
  }
};

struct Ops {
  void (*extrap)(Ctx&);
  void (*wave)(Ctx&);
  void (*dt)(Ctx&, double);
  void (*residual)(Ctx&);
  void (*rk)(Ctx&, int, double, double, bool, int);
  void (*ensure)(Ctx&);
  void (*tune)(Ctx&);
  bool (*check)(Ctx&, const int*);
  void (*fit)(Ctx&);
};

template <int D, int P, class R>
static Ops ops_for() {
  using Rn = Runner<D, P, R>;
  return Ops{Rn::extrap, Rn::wave, Rn::dt, Rn::residual, Rn::rk, Rn::ensure, Rn::tune, Rn::check, Rn::fit};
}

template <class R>
static Ops get_ops_r(const Ctx& c) {
  const int D = c.z.D, P = c.z.P;
  if (D == 2) {
    switch (P) {
      case 0: return ops_for<2, 0, R>();
      case 1: return ops_for<2, 1, R>();
      case 2: return ops_for<2, 2, R>();
      case 3: return ops_for<2, 3, R>();
      case 4: return ops_for<2, 4, R>();
      default: return ops_for<2, 5, R>();
    }
  }
  switch (P) {
    case 0: return ops_for<3, 0, R>();
    case 1: return ops_for<3, 1, R>();
    case 2: return ops_for<3, 2, R>();
    case 3: return ops_for<3, 3, R>();
    case 4: return ops_for<3, 4, R>();
    default: return ops_for<3, 5, R>();
  }
}

static Ops get_ops(const Ctx& c, bool mixed) { return mixed ? get_ops_r<float>(c) : get_ops_r<double>(c); }

// Tuned block sizes, shared by the contexts of one process: keyed by
// (nDims, order, precision, viscous, log2 of the element count).
static std::mutex g_tune_mu;
static std::map<std::array<int, 5>, std::array<int, 5>> g_tuned;

void fused_autotune(Ctx& c, bool mixed, int* out) {
  FusedState& st = fs(c);
  Ops o = get_ops(c, mixed);
  const int m = mixed ? 1 : 0;
  int lg = 0;
  for (long long e = c.z.E; e > 1; e >>= 1) lg++;
  const std::array<int, 5> key{c.z.D, c.z.P, m, c.prm.viscous ? 1 : 0, lg};
  bool hit = false;
  {
    std::lock_guard<std::mutex> g(g_tune_mu);
    auto it = g_tuned.find(key);
    if (it != g_tuned.end()) {
      std::copy(it->second.begin(), it->second.end(), st.nt[m]);
      hit = true;
    }
  }
  if (!hit) {
    c.drop_graphs();
    fused_begin(c, mixed, false);  // reads U and the wave speeds; both are left as they were
    o.tune(c);
    std::array<int, 5> v;
    std::copy(st.nt[m], st.nt[m] + 5, v.begin());
    std::lock_guard<std::mutex> g(g_tune_mu);
    g_tuned[key] = v;
  }
  c.drop_graphs();
  if (out) std::copy(st.nt[m], st.nt[m] + 5, out);
}

void fused_set_blocks(Ctx& c, bool mixed, const int* nt) {
  Ops o = get_ops(c, mixed);
  if (!o.check(c, nt)) throw std::invalid_argument("block size not valid for this kernel and order");
  c.drop_graphs();
  std::copy(nt, nt + 5, fs(c).nt[mixed ? 1 : 0]);
}

void fused_get_blocks(Ctx& c, bool mixed, int* out) {
  get_ops(c, mixed).fit(c);
  const int* nt = fs(c).nt[mixed ? 1 : 0];
  std::copy(nt, nt + 5, out);
}

static double cfl_limit_f(int order) {  // getCFLLimit, src/global.cpp:81-105
  static const double t[6] = {1.393, 0.464, 0.235, 0.1453, 0.100, 0.0736};
  if (order < 0 || order > 5) throw std::invalid_argument("CFL limit not available for this order!");
  return t[order];
}

void fused_begin(Ctx& c, bool mixed, bool src) {
  FusedState& st = fs(c);
  const Sizes& z = c.z;
  Ops o = get_ops(c, mixed);
  o.ensure(c);
  long long nS = (long long)z.E * z.N * z.S;
  to_em(c.stream, z.S, z.E, z.N, c.U, st.Ue);
  st.use_src = src;
  if (src) {
    if (!c.src) throw std::invalid_argument("no P-multigrid source term on this context");
    if (!st.srce) {
      c.drop_graphs();
      FCHK(hipMalloc(&st.srce, nS * sizeof(double)));
      poison_fill(st.srce, nS * sizeof(double));
    }
    to_em(c.stream, z.S, z.E, z.N, c.src, st.srce);
  }
  to_em(c.stream, z.F, z.E, 1, c.waveSp, st.waveSpe);
  FCHK(hipGetLastError());
  o.extrap(c);
}

// P-multigrid cycle with every level kept element-major (the transfers above;
// a level's FlurryPP-layout arrays are written back by the caller at the end).
static int opgemm_chunk(int k) { return k <= 16 ? 16 : (k <= 32 ? 32 : 64); }
// LDS the staged transfers may take (operator plus the block's rows)
constexpr size_t kTransferLds = 64 * 1024;

void fused_em_prepare(Ctx& c, bool mixed) {
  FusedState& st = fs(c);
  get_ops(c, mixed).ensure(c);
  if (!st.srce) {
    c.drop_graphs();
    FCHK(hipMalloc(&st.srce, (size_t)c.z.E * c.z.N * c.z.S * sizeof(double)));
    poison_fill(st.srce, (size_t)c.z.E * c.z.N * c.z.S * sizeof(double));
  }
}

void fused_em_extrap(Ctx& c, bool mixed) { get_ops(c, mixed).extrap(c); }

// The wave speeds from FlurryPP's layout: the face kernel rewrites only the
// entries of left and boundary flux points (the reference's calcWaveSp
// stores into the left element); the others keep what the level holds.
void fused_em_import_wave(Ctx& c) {
  to_em(c.stream, c.z.F, c.z.E, 1, c.waveSp, fs(c).waveSpe);
  FCHK(hipGetLastError());
}

void fused_em_residual(Ctx& c, bool mixed, int how) {
  Ops o = get_ops(c, mixed);
  fs(c).use_src = false;
  o.residual(c);
  o.rk(c, how == 1 ? 7 : (how == 2 ? 8 : 5), 0., 0., true, 0);
}

void fused_em_set_src(Ctx& c, bool on) { fs(c).use_src = on; }

double* fused_em_array(Ctx& c, int which) {
  FusedState& st = fs(c);
  return which == 0 ? st.Ue : (which == 1 ? st.divF0e : st.srce);
}

void fused_em_export_u(Ctx& c) {
  FusedState& st = fs(c);
  st.use_src = false;
  from_em(c.stream, c.z.S, c.z.E, c.z.N, st.Ue, c.U);
  FCHK(hipGetLastError());
}

void pmg_restrict_em(Ctx& c, long long e0, long long ne, int Sf, const double* R, const double* Uf,
                     const double* divFf, double* solc) {
  FusedState& st = fs(c);
  if (e0 < 0 || ne < 0 || e0 + ne > c.z.E) throw std::invalid_argument("restriction rows out of range");
  if (ne == 0) return;
  const long long rows = ne * c.z.N;
  const int Sc = c.z.S, nr = 256 / Sc;
  const long long off = e0 * c.z.N * Sc;  // the first coarse element's rows
  Launch L(c, "pmg_restrict");
  const size_t lds = ((size_t)Sc * Sf + 2 * (size_t)nr * Sf) * sizeof(double);
  if (nr >= 1 && lds <= kTransferLds)
    hipLaunchKernelGGL(k_restrict_em_lds, dim3((unsigned)((rows + nr - 1) / nr)), dim3(nr * Sc), lds, c.stream,
                       rows, Sc, Sf, opgemm_chunk(Sf), nr, R, Uf, divFf, st.Ue + off, st.srce + off, solc + off);
  else
    hipLaunchKernelGGL(k_restrict_em, dim3(nblk(rows * Sc, 256)), dim3(256), 0, c.stream, rows, Sc, Sf,
                       opgemm_chunk(Sf), R, Uf, divFf, st.Ue + off, st.srce + off, solc + off);
  FCHK(hipGetLastError());
}

void pmg_prolong_em(Ctx& up, int Sc, const double* Pm, const double* Uc, const double* solc) {
  FusedState& st = fs(up);
  const long long rows = (long long)up.z.E * up.z.N;
  Launch L(up, "pmg_prolong");
  const int Su = up.z.S, nr = 256 / Su;
  const size_t lds = ((size_t)Su * Sc + (size_t)nr * Sc) * sizeof(double);
  if (nr >= 1 && lds <= kTransferLds)
    hipLaunchKernelGGL(k_prolong_em_lds, dim3((unsigned)((rows + nr - 1) / nr)), dim3(nr * Su), lds, up.stream,
                       rows, Su, Sc, opgemm_chunk(Sc), nr, Pm, Uc, solc, st.Ue);
  else
    hipLaunchKernelGGL(k_prolong_em, dim3(nblk(rows * Su, 256)), dim3(256), 0, up.stream, rows, Su, Sc,
                       opgemm_chunk(Sc), Pm, Uc, solc, st.Ue);
  FCHK(hipGetLastError());
}

void fused_end(Ctx& c) {
  FusedState& st = fs(c);
  st.use_src = false;
  const Sizes& z = c.z;
  from_em(c.stream, z.S, z.E, z.N, st.Ue, c.U);
  from_em(c.stream, z.S, z.E, z.N, st.divF0e, c.divF);
  from_em(c.stream, z.F, z.E, 1, st.waveSpe, c.waveSp);
  FCHK(hipGetLastError());
}

// calcResidual(0) at the current state (src/solver.cpp:366-445) on a fused
// path: divF_spts[0] and the face wave speeds in FlurryPP's layouts.
void fused_residual(Ctx& c, bool mixed) {
  FusedState& st = fs(c);
  const Sizes& z = c.z;
  Ops o = get_ops(c, mixed);
  fused_begin(c, mixed, false);
  o.residual(c);
  o.rk(c, 5, 0., 0., true, 0);
  from_em(c.stream, z.S, z.E, z.N, st.divF0e, c.divF);
  from_em(c.stream, z.F, z.E, 1, st.waveSpe, c.waveSp);
  FCHK(hipGetLastError());
}

// One iteration of solver::update (src/solver.cpp:319-363) on a fused path.
void fused_rk_step(Ctx& c, bool refresh, bool mixed, bool keep) {
  Ops o = get_ops(c, mixed);
  FusedState& st = fs(c);
  const int nRK = c.z.nRK;
  const double* A = c.prm.rk_a;
  const double* B = c.prm.rk_b;
  const bool cfl = c.prm.dt_type != 0;
  const double cflfac = cfl ? c.prm.cfl * cfl_limit_f(c.z.P) : 0.;
  const bool jameson = c.prm.time_type >= 5;
  if (nRK == 1) {
    o.residual(c);
    if (cfl) {
      if (refresh) o.wave(c);
      o.dt(c, cflfac);
    }
    o.rk(c, jameson ? 4 : 3, A[1], B[0], keep, 0);
    return;
  }
  if (cfl) {
    if (refresh) o.wave(c);
    o.dt(c, cflfac);
  }
  for (int step = 0; step < nRK; step++) {
    st.read_stage = step > 0;
    o.residual(c);
    // the Jameson form's first stage reads the state itself (the step's U0)
    int mode;
    if (jameson) mode = step == 0 ? 6 : 4;
    else mode = (step == 0) ? 0 : (step == nRK - 1 ? 2 : 1);
    o.rk(c, mode, A[step + 1], B[step], keep && step == 0, step);
  }
  st.read_stage = false;
}

}  // namespace fhip
