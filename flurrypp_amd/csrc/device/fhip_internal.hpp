// Internal declarations of libflurry_hip: the context and the launchers the
// C-ABI layer (fhip_api.hip) drives.
#pragma once

#include <hip/hip_runtime.h>

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

struct Timer {
  std::map<std::string, std::pair<double, long long>> acc;  // total ms, launches
};

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
  // captured RK steps (fused path): index 0 without, 1 with wave-speed refresh
  hipGraphExec_t graph[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
  bool fused_ok = false;
  void* fused = nullptr;  // FusedState (fhip_fused.hip)
};

// Launch helper: records per-kernel time with events when timing is on.
struct Launch {
  Ctx& c;
  const char* name;
  Launch(Ctx& c_, const char* n) : c(c_), name(n) {
    if (c.timing) (void)hipEventRecord(c.ev0, c.stream);
  }
  ~Launch() {
    if (c.timing) {
      (void)hipEventRecord(c.ev1, c.stream);
      (void)hipEventSynchronize(c.ev1);
      float ms = 0;
      (void)hipEventElapsedTime(&ms, c.ev0, c.ev1);
      auto& a = c.timer.acc[name];
      a.first += ms;
      a.second += 1;
    }
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
void residual_norm(Ctx& c, int res_type, double* out_host);
int check_state(Ctx& c);

// fused path (fhip_fused.hip)
bool fused_setup(Ctx& c, const fhip_desc* d);  // false: configuration not covered
void fused_release(Ctx& c);
void fused_begin(Ctx& c);  // import U and wave speeds, extrapolate U to the faces
void fused_end(Ctx& c);    // export U, divF_spts[0] and wave speeds
void fused_rk_step(Ctx& c, bool refresh);

}  // namespace fhip

struct fhip_ctx {
  fhip::Ctx c;
};
