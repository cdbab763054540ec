// Internal declarations of libflurry_hip: the context and the launchers the
// C-ABI layer (fhip_api.hip) drives.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdlib>

#include <map>
#include <string>
#include <vector>

#include "fhip_physics.hpp"
#include "flurry_hip.h"

namespace fhip {

struct Sizes {
  int D, N, P, S, F, Ff, E, nRK, nFaces, nFaceFpts;
  long long nS() const { return (long long)S * E * N; }  // doubles in one [S][E][N] array
  long long nF() const { return (long long)F * E * N; }
};

// Face flux points grouped for coalesced access: every face flux point has
// a left point index (fpt*E+ele), a right one (or -1) and a kind.
struct FaceTables {
  int* ptL = nullptr;   // [nFaceFpts]
  int* ptR = nullptr;   // [nFaceFpts], -1 on boundaries
  int* bc = nullptr;    // [nFaceFpts], -1 interior
};

// Per-kernel timing with HIP events recorded on the context's stream around
// each launch; nothing synchronises until the times are read (or the event
// pool is full), so timing runs inside the timed region itself.
struct Timer {
  std::map<std::string, std::pair<double, long long>> acc;  // total ms, launches
  std::vector<hipEvent_t> pool;
  std::vector<std::pair<const char*, size_t>> pending;       // name, index of start event
  size_t used = 0;
};

// Debug aid: with FHIP_POISON_ALLOC=1 in the environment every device
// allocation of the library is filled with 0xFF bytes (NaN in FP64/FP32, -1 in
// int32), so that a read of memory no kernel has written shows up as NaN.
inline bool poison_alloc() {
  static const bool on = [] {
    const char* v = std::getenv("FHIP_POISON_ALLOC");
    return v && v[0] == '1';
  }();
  return on;
}
inline void poison_fill(void* p, size_t bytes) {
  if (!(poison_alloc() && p && bytes)) return;
  (void)hipMemset(p, 0xFF, bytes);
  (void)hipDeviceSynchronize();
}

struct Ctx {
  Sizes z;
  fhip_params prm;
  Phys ph;
  int device = 0;
  hipStream_t stream = nullptr;

  // operators
  double *sf = nullptr, *grad = nullptr, *efn = nullptr, *corr = nullptr, *corru = nullptr, *ws = nullptr,
         *wf = nullptr;
  // 1-D tensor-product factors for the fused path (see fhip_fused.hip)
  double* tp = nullptr;
  // metrics
  double *jginv = nullptr, *detj = nullptr, *norm = nullptr, *da = nullptr;
  FaceTables ft;
  // face-fpt order used by the staged kernels (natural) and original face ids
  std::vector<int> h_ptL, h_ptR;
  // state and work arrays (FlurryPP layouts)
  double *U = nullptr, *U0 = nullptr, *Uf = nullptr, *Fs = nullptr, *dU = nullptr, *dUf = nullptr,
         *dUc = nullptr, *disFn = nullptr, *Fn = nullptr, *divF = nullptr, *waveSp = nullptr, *eleDt = nullptr,
         *faceFn = nullptr, *src = nullptr, *Uacc = nullptr;
  double* d_dt = nullptr;       // device copy of params->dt
  double* partial = nullptr;    // reduction scratch
  int* flag = nullptr;          // state-check flag
  double h_dt = 0;
  int iter = 0;
  bool timing = false;
  Timer timer;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // captured RK steps (fused path): [with P-multigrid source][keeping
  // divF_spts[0]][FP64, mixed, mixed with BF16 hi+lo, mixed with BF16
  // hi+mid+lo].  A graph holds the device pointers of the moment it was
  // captured: drop them whenever an array they use is (re)allocated.
  hipGraphExec_t graph[2][2][4] = {};
  // mixed path: BF16 pieces of the matrix-core z contractions (0: none)
  int bf16 = 0;
  unsigned long long graph_gen = 0;  // bumped by drop_graphs: graphs over several contexts check it
  void drop_graphs() {
    graph_gen++;
    bool any = false;
    for (auto& plane : graph)
      for (auto& row : plane)
        for (auto& g : row) any = any || g != nullptr;
    if (any) (void)hipStreamSynchronize(stream);  // no replay still in flight
    for (auto& plane : graph)
      for (auto& row : plane)
        for (auto& g : row) {
          if (g) (void)hipGraphExecDestroy(g);
          g = nullptr;
        }
  }
  bool fused_ok = false;
  bool use_graphs = true;
  bool batched = false;  // several meshes or a boundary-state table
  void* fused = nullptr;  // FusedState (fhip_fused.hip)
};

void timer_resolve(Ctx& c);

// Launch helper: brackets one kernel launch with timing events when enabled.
struct Launch {
  Ctx& c;
  const char* name;
  size_t slot = 0;
  Launch(Ctx& c_, const char* n) : c(c_), name(n) {
    if (!c.timing) return;
    Timer& t = c.timer;
    if (t.used + 2 > 8192) timer_resolve(c);
    while (t.pool.size() < t.used + 2) {
      hipEvent_t e;
      (void)hipEventCreate(&e);
      t.pool.push_back(e);
    }
    slot = t.used;
    t.used += 2;
    (void)hipEventRecord(t.pool[slot], c.stream);
  }
  ~Launch() {
    if (!c.timing) return;
    (void)hipEventRecord(c.timer.pool[slot + 1], c.stream);
    c.timer.pending.push_back({name, slot});
  }
};

// staged path (fhip_staged.hip)
void staged_extrapolate_u(Ctx& c);
void staged_wave_speed(Ctx& c);
void staged_residual(Ctx& c, int step);
void staged_time_step_a(Ctx& c, int step, double rk, bool src);
void staged_time_step_b(Ctx& c, int step, bool src);
void copy_array(Ctx& c, double* dst, const double* src, long long n);
void calc_dt(Ctx& c, bool refresh);
// C[m][E*N] = A[m][k] B[k][E*N] (+ C when beta): an operator applied to every
// element and field of a FlurryPP-layout array
void op_apply(Ctx& c, const char* name, int m, int k, const double* A, const double* B, double* C, bool beta);
void sub_array(Ctx& c, double* out, const double* a, const double* b, long long n);  // out = a; out -= b
void add_array(Ctx& c, double* y, const double* x, long long n);                     // y += x
void residual_norm(Ctx& c, int res_type, double* out_host);
// The same without allocating: partial sums in a static device array of
// kNoAllocScratch doubles per GPU (std::length_error beyond it).
constexpr long long kNoAllocScratch = 16384;  // 128 KiB per GPU
constexpr int kNoAllocMaxDevices = 64;
void residual_norm_noalloc(Ctx& c, int res_type, double* out_host);
double* noalloc_scratch(int device);  // device address of that GPU's scratch array
int check_state(Ctx& c);

// fused path (fhip_fused.hip)
bool fused_setup(Ctx& c, const fhip_desc* d);  // false: configuration not covered
void fused_release(Ctx& c);
// import U, wave speeds [and the P-multigrid source], extrapolate U to the faces
void fused_begin(Ctx& c, bool mixed, bool src = false);
void fused_end(Ctx& c);                // export U, divF_spts[0] and wave speeds
// (keep: write divF_spts[0] at the first stage; steps whose residual nothing
// reads skip it)
void fused_rk_step(Ctx& c, bool refresh, bool mixed, bool keep = true);
// P-multigrid cycle on element-major levels: work arrays and the FAS source
// allocated; U at the flux points of the current U; the residual (no source)
// and the wave speeds, the residual written (how) 0: into divF_spts[0],
// 1: subtracted from the source in place, 2: plus the source into
// divF_spts[0]; the source on or off for the RK
// update; U, divF_spts[0] or the source (which 0, 1, 2); U written back to
// FlurryPP's layout; the transfers into the coarse level c (U, source, sol)
// and into the upper level up (U += P (U_c - sol_c)).
void fused_em_prepare(Ctx& c, bool mixed);
void fused_em_extrap(Ctx& c, bool mixed);
void fused_em_import_wave(Ctx& c);
void fused_em_residual(Ctx& c, bool mixed, int how = 0);
void fused_em_set_src(Ctx& c, bool on);
double* fused_em_array(Ctx& c, int which);
void fused_em_export_u(Ctx& c);
// (restriction into the coarse elements [e0, e0 + ne) from the fine rows Uf,
// divFf of those elements)
void pmg_restrict_em(Ctx& c, long long e0, long long ne, int Sf, const double* R, const double* Uf,
                     const double* divFf, double* solc);
void pmg_prolong_em(Ctx& up, int Sc, const double* Pm, const double* Uc, const double* solc);
void fused_residual(Ctx& c, bool mixed);  // divF_spts[0] and wave speeds of the current U
// block sizes of the fused kernels (extrapolate, face_inviscid, volume, face_boundary, rk)
void fused_autotune(Ctx& c, bool mixed, int* out);
void fused_set_blocks(Ctx& c, bool mixed, const int* nt);
void fused_get_blocks(Ctx& c, bool mixed, int* out);
void fused_mesh_residuals(Ctx& c, int type, double* out);  // out [nMesh][N]
int fused_n_meshes(Ctx& c);
void fused_set_active(Ctx& c, const int* active);

}  // namespace fhip

struct fhip_ctx {
  fhip::Ctx c;
};
