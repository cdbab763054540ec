// C ABI of libflurry_hip (declared in include/flurry_hip.h).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "fhip_internal.hpp"

namespace {

thread_local std::string g_err;

struct ApiError : std::runtime_error {
  int code;
  ApiError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HCHK(x)                                                                           \
  do {                                                                                    \
    hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess) throw ApiError(FHIP_EDEVICE, std::string("HIP: ") + hipGetErrorString(e_)); \
  } while (0)

template <class F>
int guard(F&& f) {
  try {
    return f();
  } catch (const ApiError& e) {
    g_err = e.what();
    return e.code;
  } catch (const std::invalid_argument& e) {
    g_err = e.what();
    return FHIP_EINVAL;
  } catch (const std::exception& e) {
    g_err = e.what();
    return FHIP_EDEVICE;
  }
}

void require(bool ok, const std::string& msg) {
  if (!ok) throw ApiError(FHIP_EINVAL, msg);
}

template <class T>
T* dalloc(size_t n) {
  void* p = nullptr;
  if (n == 0) n = 1;
  HCHK(hipMalloc(&p, n * sizeof(T)));
  fhip::poison_fill(p, n * sizeof(T));
  return (T*)p;
}

template <class T>
T* dcopy(const T* h, size_t n, hipStream_t s) {
  T* d = dalloc<T>(n);
  if (h && n) HCHK(hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, s));
  else if (n) HCHK(hipMemsetAsync(d, 0, n * sizeof(T), s));
  return d;
}

// Right-element flux point of left flux point i (src/intFace.cpp:57-95).
int right_fpt(int D, int P, int Ff, int idr, int relRot, int i) {
  if (D == 2) return idr * Ff + Ff - 1 - i;
  int P1 = P + 1, s = Ff * idr, e = Ff * (idr + 1), ii = i % P1, jj = i / P1;
  switch (relRot) {
    case 0: return s + ii * P1 + jj;
    case 1: return s + P - ii + jj * P1;
    case 2: return e - 1 - (ii * P1 + jj);
    default: return e - P1 * (jj + 1) + ii;
  }
}

void destroy_ctx(fhip::Ctx& c) {
  if (c.device >= 0) (void)hipSetDevice(c.device);
  fhip::fused_release(c);
  for (auto& plane : c.graph)
    for (auto& row : plane)
      for (auto& g : row)
        if (g) (void)hipGraphExecDestroy(g);
  void* ptrs[] = {c.sf, c.grad, c.efn, c.corr, c.corru, c.ws, c.wf, c.tp, c.jginv, c.detj, c.norm, c.da,
                  c.ft.ptL, c.ft.ptR, c.ft.bc, c.U, c.U0, c.Uf, c.Fs, c.dU, c.dUf, c.dUc, c.disFn, c.Fn,
                  c.divF, c.waveSp, c.eleDt, c.faceFn, c.src, c.Uacc, c.d_dt, c.partial, c.flag};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (hipEvent_t e : c.timer.pool) (void)hipEventDestroy(e);
  if (c.ev0) (void)hipEventDestroy(c.ev0);
  if (c.ev1) (void)hipEventDestroy(c.ev1);
  if (c.stream) (void)hipStreamDestroy(c.stream);
}

double* array_ptr(fhip::Ctx& c, int which, int step, long long* n) {
  const fhip::Sizes& z = c.z;
  switch (which) {
    case FHIP_ARR_U_SPTS: *n = z.nS(); return c.U;
    case FHIP_ARR_U0: *n = z.nS(); return c.U0;
    case FHIP_ARR_U_FPTS: *n = z.nF(); return c.Uf;
    case FHIP_ARR_F_SPTS: *n = z.D * z.nS(); return c.Fs;
    case FHIP_ARR_DIVF_SPTS:
      require(step >= 0 && step < z.nRK, "RK step out of range");
      *n = z.nS();
      return c.divF + step * z.nS();
    case FHIP_ARR_DISFN_FPTS: *n = z.nF(); return c.disFn;
    case FHIP_ARR_FN_FPTS: *n = z.nF(); return c.Fn;
    case FHIP_ARR_DU_SPTS: *n = c.dU ? z.D * z.nS() : 0; return c.dU;
    case FHIP_ARR_DU_FPTS: *n = c.dUf ? z.D * z.nF() : 0; return c.dUf;
    case FHIP_ARR_DUC_FPTS: *n = c.dUc ? z.nF() : 0; return c.dUc;
    case FHIP_ARR_WAVESP: *n = (long long)z.F * z.E; return c.waveSp;
    case FHIP_ARR_ELE_DT: *n = z.E; return c.eleDt;
    default: throw ApiError(FHIP_EINVAL, "unknown array id");
  }
}

}  // namespace

namespace fhip {
void timer_resolve(Ctx& c) {
  Timer& t = c.timer;
  if (t.pending.empty()) {
    t.used = 0;
    return;
  }
  (void)hipStreamSynchronize(c.stream);
  for (auto& pr : t.pending) {
    float ms = 0;
    (void)hipEventElapsedTime(&ms, t.pool[pr.second], t.pool[pr.second + 1]);
    auto& a = t.acc[pr.first];
    a.first += ms;
    a.second += 1;
  }
  t.pending.clear();
  t.used = 0;
}
}  // namespace fhip

extern "C" {

const char* fhip_last_error(void) { return g_err.c_str(); }

int fhip_set_device(int device) {
  return guard([&] {
    HCHK(hipSetDevice(device));
    return 0;
  });
}

int fhip_create(const fhip_desc* d, fhip_ctx** out) {
  return guard([&] {
    require(d && d->params && out, "null descriptor");
    const fhip_params& p = *d->params;
    require(p.equation == 1, "only the Euler/Navier-Stokes equations (equation 1) are supported");
    require(p.n_dims == 2 || p.n_dims == 3, "nDims must be 2 or 3");
    require(p.n_fields == p.n_dims + 2, "nFields must be nDims+2");
    require(p.order >= 0 && p.order <= 10, "order out of range");
    int P1 = p.order + 1;
    int Ff = (p.n_dims == 2) ? P1 : P1 * P1;
    require(p.n_fpts_face == Ff && p.n_spts == (p.n_dims == 2 ? P1 * P1 : P1 * P1 * P1) && p.n_fpts == 2 * p.n_dims * Ff,
            "point counts do not match nDims and order");
    require(p.n_eles > 0, "nEles must be positive");
    require(p.n_rk_steps >= 1 && p.n_rk_steps <= 7, "nRKSteps out of range");
    require(!(p.n_dims == 3 && p.riemann_type == 1), "Roe not implemented in 3D");
    require(p.riemann_type == 0 || p.riemann_type == 1, "riemannType must be 0 or 1");
    require(p.dt_type >= 0 && p.dt_type <= 2, "dtType must be 0, 1 or 2");
    require((long long)p.n_fpts * p.n_eles < (1ll << 31), "mesh too large for 32-bit point indices");
    require(d->opp_spts_to_fpts && d->opp_grad_spts && d->opp_extrapolate_fn && d->opp_correction && d->w_spts &&
                d->w_face && d->jginv_spts && d->detjac_spts && d->norm_fpts && d->da_fpts,
            "missing operator or metric array");
    require(!p.viscous || d->opp_correct_u, "viscous runs need opp_correct_u");
    require(d->n_faces >= 0 && (d->n_faces == 0 || (d->face_ele_l && d->face_loc_l && d->face_ele_r && d->face_loc_r &&
                                                   d->face_rel_rot && d->face_bc)),
            "missing face arrays");

    auto* h = new fhip_ctx();
    fhip::Ctx& c = h->c;
    try {
      HCHK(hipGetDevice(&c.device));
      c.prm = p;
      fhip::Sizes& z = c.z;
      z.D = p.n_dims; z.N = p.n_fields; z.P = p.order; z.S = p.n_spts; z.F = p.n_fpts; z.Ff = Ff; z.E = p.n_eles;
      z.nRK = p.n_rk_steps; z.nFaces = d->n_faces; z.nFaceFpts = d->n_faces * Ff;
      c.ph = fhip::Phys{p.gamma, p.prandtl, p.mu_inf, p.rt_inf, p.c_sth, p.pen_fact, p.tau,
                        p.rho_bound, p.u_bound, p.v_bound, p.w_bound, p.p_bound,
                        p.t_wall, p.r_gas, p.u_wall, p.v_wall, p.w_wall, p.fix_vis, p.riemann_type, p.viscous};
      HCHK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
      HCHK(hipEventCreate(&c.ev0));
      HCHK(hipEventCreate(&c.ev1));
      const int D = z.D, S = z.S, F = z.F, E = z.E;
      c.sf = dcopy(d->opp_spts_to_fpts, (size_t)F * S, c.stream);
      c.grad = dcopy(d->opp_grad_spts, (size_t)D * S * S, c.stream);
      c.efn = dcopy(d->opp_extrapolate_fn, (size_t)D * F * S, c.stream);
      c.corr = dcopy(d->opp_correction, (size_t)S * F, c.stream);
      if (p.viscous) c.corru = dcopy(d->opp_correct_u, (size_t)D * S * F, c.stream);
      c.ws = dcopy(d->w_spts, S, c.stream);
      c.wf = dcopy(d->w_face, Ff, c.stream);
      c.jginv = dcopy(d->jginv_spts, (size_t)D * S * E * D, c.stream);
      c.detj = dcopy(d->detjac_spts, (size_t)S * E, c.stream);
      c.norm = dcopy(d->norm_fpts, (size_t)F * E * D, c.stream);
      c.da = dcopy(d->da_fpts, (size_t)F * E, c.stream);

      // Face flux-point tables.
      std::vector<int> bc((size_t)z.nFaceFpts);
      c.h_ptL.resize(z.nFaceFpts);
      c.h_ptR.resize(z.nFaceFpts);
      std::vector<char> owned((size_t)F * E, 0);
      for (int f = 0; f < z.nFaces; f++) {
        int eL = d->face_ele_l[f], lL = d->face_loc_l[f], eR = d->face_ele_r[f];
        require(eL >= 0 && eL < E && lL >= 0 && lL < 2 * D, "face " + std::to_string(f) + ": bad left element/face");
        require(eR < E, "face " + std::to_string(f) + ": bad right element");
        if (eR >= 0) {
          int lR = d->face_loc_r[f], rr = d->face_rel_rot[f];
          require(lR >= 0 && lR < 2 * D && (D == 2 || (rr >= 0 && rr < 4)), "face " + std::to_string(f) + ": bad right face/rotation");
        } else {
          int b = d->face_bc[f];
          require(b == 1 || b == 2 || b == 3 || b == 4 || b == 5 || b == 8 || b == 9 || b == 10 || b == 12 || b == 13,
                  "face " + std::to_string(f) + ": unsupported boundary condition " + std::to_string(b));
        }
        for (int i = 0; i < Ff; i++) {
          size_t t = (size_t)f * Ff + i;
          int fl = lL * Ff + i;
          c.h_ptL[t] = fl * E + eL;
          if (eR >= 0) {
            int fr = right_fpt(D, z.P, Ff, d->face_loc_r[f], d->face_rel_rot[f], i);
            c.h_ptR[t] = fr * E + eR;
          } else {
            c.h_ptR[t] = -1;
          }
          bc[t] = (eR >= 0) ? -1 : d->face_bc[f];
          for (int side = 0; side < 2; side++) {
            int pt = side ? c.h_ptR[t] : c.h_ptL[t];
            if (pt < 0) continue;
            require(!owned[pt], "flux point owned by two faces (face " + std::to_string(f) + ")");
            owned[pt] = 1;
          }
        }
      }
      c.ft.ptL = dcopy(c.h_ptL.data(), c.h_ptL.size(), c.stream);
      c.ft.ptR = dcopy(c.h_ptR.data(), c.h_ptR.size(), c.stream);
      c.ft.bc = dcopy(bc.data(), bc.size(), c.stream);

      const size_t nS = z.nS(), nF = z.nF();
      c.U = dcopy<double>(nullptr, nS, c.stream);
      c.U0 = dcopy<double>(nullptr, nS, c.stream);
      c.Uf = dcopy<double>(nullptr, nF, c.stream);
      c.Fs = dcopy<double>(nullptr, (size_t)D * nS, c.stream);
      if (p.viscous) {
        c.dU = dcopy<double>(nullptr, (size_t)D * nS, c.stream);
        c.dUf = dcopy<double>(nullptr, (size_t)D * nF, c.stream);
        c.dUc = dcopy<double>(nullptr, nF, c.stream);
      }
      c.disFn = dcopy<double>(nullptr, nF, c.stream);
      c.Fn = dcopy<double>(nullptr, nF, c.stream);
      c.divF = dcopy<double>(nullptr, (size_t)z.nRK * nS, c.stream);
      c.waveSp = dcopy<double>(nullptr, (size_t)F * E, c.stream);
      c.eleDt = dcopy<double>(nullptr, E, c.stream);
      c.faceFn = dcopy<double>(nullptr, (size_t)z.nFaceFpts * z.N, c.stream);
      c.d_dt = dcopy<double>(&p.dt, 1, c.stream);
      c.partial = dcopy<double>(nullptr, ((size_t)E / 256 + 2) * z.N + 1024, c.stream);
      c.flag = dcopy<int>(nullptr, 1, c.stream);
      c.h_dt = p.dt;
      c.iter = p.init_iter;
      c.batched = d->n_meshes > 1 || d->n_bnd_states > 0;
      c.fused_ok = fhip::fused_setup(c, d);
      require(c.fused_ok || !c.batched, "batched contexts need the fused path (quads/hexes, P 0..5)");
      HCHK(hipStreamSynchronize(c.stream));
    } catch (...) {
      destroy_ctx(c);
      delete h;
      throw;
    }
    *out = h;
    return 0;
  });
}

void fhip_destroy(fhip_ctx* h) {
  if (!h) return;
  destroy_ctx(h->c);
  delete h;
}

int fhip_set_state(fhip_ctx* h, const double* u) {
  return guard([&] {
    require(h && u, "null argument");
    fhip::Ctx& c = h->c;
    HCHK(hipMemcpyAsync(c.U, u, c.z.nS() * sizeof(double), hipMemcpyHostToDevice, c.stream));
    HCHK(hipStreamSynchronize(c.stream));
    return 0;
  });
}

int fhip_get_state(fhip_ctx* h, double* u) {
  return guard([&] {
    require(h && u, "null argument");
    fhip::Ctx& c = h->c;
    HCHK(hipMemcpyAsync(u, c.U, c.z.nS() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    HCHK(hipStreamSynchronize(c.stream));
    return 0;
  });
}

int fhip_extrapolate_u(fhip_ctx* h) {
  return guard([&] {
    require(h, "null context");
    fhip::staged_extrapolate_u(h->c);
    return 0;
  });
}

int fhip_calc_wave_speed(fhip_ctx* h) {
  return guard([&] {
    require(h, "null context");
    fhip::staged_wave_speed(h->c);
    return 0;
  });
}

int fhip_calc_residual(fhip_ctx* h, int step) {
  return guard([&] {
    require(h, "null context");
    require(step >= 0 && step < h->c.z.nRK, "RK step out of range");
    require(!h->c.batched, "batched contexts run through fhip_update");
    fhip::staged_residual(h->c, step);
    return 0;
  });
}

static void upload_src(fhip::Ctx& c, const double* src) {
  if (!c.src) c.src = dalloc<double>(c.z.nS());
  HCHK(hipMemcpyAsync(c.src, src, c.z.nS() * sizeof(double), hipMemcpyHostToDevice, c.stream));
}

int fhip_time_step_a(fhip_ctx* h, int step, double rk, const double* src) {
  return guard([&] {
    require(h, "null context");
    require(step >= 0 && step < h->c.z.nRK, "RK step out of range");
    if (src) upload_src(h->c, src);
    fhip::staged_time_step_a(h->c, step, rk, src != nullptr);
    HCHK(hipStreamSynchronize(h->c.stream));
    return 0;
  });
}

int fhip_time_step_b(fhip_ctx* h, int step, const double* src) {
  return guard([&] {
    require(h, "null context");
    require(step >= 0 && step < h->c.z.nRK, "RK step out of range");
    if (src) upload_src(h->c, src);
    fhip::staged_time_step_b(h->c, step, src != nullptr);
    HCHK(hipStreamSynchronize(h->c.stream));
    return 0;
  });
}

int fhip_copy_u_to_u0(fhip_ctx* h) {
  return guard([&] {
    require(h, "null context");
    fhip::copy_array(h->c, h->c.U0, h->c.U, h->c.z.nS());
    return 0;
  });
}

int fhip_copy_u0_to_u(fhip_ctx* h) {
  return guard([&] {
    require(h, "null context");
    fhip::copy_array(h->c, h->c.U, h->c.U0, h->c.z.nS());
    return 0;
  });
}

int fhip_calc_dt(fhip_ctx* h, int refresh, double* dt_out) {
  return guard([&] {
    require(h, "null context");
    fhip::Ctx& c = h->c;
    fhip::calc_dt(c, refresh != 0);
    HCHK(hipMemcpyAsync(&c.h_dt, c.d_dt, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    HCHK(hipStreamSynchronize(c.stream));
    if (dt_out) *dt_out = c.h_dt;
    return 0;
  });
}

int fhip_set_dt(fhip_ctx* h, double dt) {
  return guard([&] {
    require(h, "null context");
    fhip::Ctx& c = h->c;
    c.h_dt = dt;
    HCHK(hipMemcpyAsync(c.d_dt, &c.h_dt, sizeof(double), hipMemcpyHostToDevice, c.stream));
    HCHK(hipStreamSynchronize(c.stream));
    return 0;
  });
}

int fhip_get_dt(fhip_ctx* h, double* dt) {
  return guard([&] {
    require(h && dt, "null argument");
    fhip::Ctx& c = h->c;
    HCHK(hipMemcpyAsync(&c.h_dt, c.d_dt, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    HCHK(hipStreamSynchronize(c.stream));
    *dt = c.h_dt;
    return 0;
  });
}

int fhip_get_iter(fhip_ctx* h, int* iter) {
  if (!h || !iter) return FHIP_EINVAL;
  *iter = h->c.iter;
  return 0;
}

static bool resolve_path(fhip::Ctx& c, int path, bool& mixed) {
  require(path >= FHIP_PATH_AUTO && path <= FHIP_PATH_MIXED_BF16X3, "unknown path");
  const bool bf = path == FHIP_PATH_MIXED_BF16 || path == FHIP_PATH_MIXED_BF16X3;
  require(!bf || (c.z.D == 3 && c.z.P == 3), "the BF16 matrix-core forms cover P3 hexes only");
  c.bf16 = path == FHIP_PATH_MIXED_BF16 ? 2 : (path == FHIP_PATH_MIXED_BF16X3 ? 3 : 0);
  mixed = path == FHIP_PATH_MIXED || bf;
  bool fused = mixed || (path == FHIP_PATH_FUSED) || (path == FHIP_PATH_AUTO && c.fused_ok);
  require(!fused || c.fused_ok, "the fused paths do not support this configuration");
  require(fused || !c.batched, "batched contexts (several meshes or a boundary-state table) need a fused path");
  return fused;
}

// Set while a P-multigrid cycle is captured into one graph: the steps inside
// it are captured kernel by kernel rather than as graph launches.
static thread_local bool t_capturing_cycle = false;

// The launches of body() on stream st captured into an instantiated graph.
// On any failure the capture is ended, nothing is kept and the error is
// thrown: no half-built graph or executable is ever stored or launched.
extern "C++" template <class F>
static hipGraphExec_t capture(hipStream_t st, F&& body) {
  hipGraph_t g = nullptr;
  HCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  try {
    body();
  } catch (...) {
    (void)hipStreamEndCapture(st, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  hipError_t e = hipStreamEndCapture(st, &g);
  if (e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    HCHK(e);
  }
  hipGraphExec_t ge = nullptr;
  e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    if (ge) (void)hipGraphExecDestroy(ge);
    HCHK(e);
  }
  return ge;
}

// One iteration of the main loop body: iter++, solver::update(src).  The
// coarse P-multigrid levels run update(true) without advancing their counter.
// keep = false (fused paths): the step's divF_spts[0] is not stored, for
// steps that another step or a residual follows before anything reads it.
static void update_step(fhip::Ctx& c, bool fused, bool mixed, bool src = false, bool count = true,
                        bool keep = true) {
  const int nRK = c.z.nRK;
  if (count) c.iter++;
  const bool refresh = count && (c.iter == c.prm.init_iter + 1);
  if (fused) {
    // Steady-state steps replay one captured graph; the wave-speed refresh
    // step and timed runs launch kernel by kernel.
    if (!refresh && !c.timing && c.use_graphs && !t_capturing_cycle) {
      hipGraphExec_t& ge = c.graph[src ? 1 : 0][keep ? 1 : 0][mixed ? 1 + (c.bf16 ? c.bf16 - 1 : 0) : 0];
      if (!ge) ge = capture(c.stream, [&] { fhip::fused_rk_step(c, false, mixed, keep); });
      HCHK(hipGraphLaunch(ge, c.stream));
    } else {
      fhip::fused_rk_step(c, refresh, mixed, keep);
    }
    return;
  }
  // solver::update (src/solver.cpp:319-363), static grid
  for (int step = 0; step < nRK - 1; step++) {
    if (step == 0 && c.prm.dt_type != 0) fhip::calc_dt(c, refresh);
    if (step == 0) fhip::copy_array(c, c.U0, c.U, c.z.nS());
    fhip::staged_residual(c, step);
    fhip::staged_time_step_a(c, step, c.prm.rk_a[step + 1], src);
  }
  fhip::staged_residual(c, nRK - 1);
  if (c.prm.time_type < 5) {
    if (nRK > 1) fhip::copy_array(c, c.U, c.U0, c.z.nS());
    else if (c.prm.dt_type != 0) fhip::calc_dt(c, refresh);
    for (int step = 0; step < nRK; step++) fhip::staged_time_step_b(c, step, src);
  } else {
    fhip::staged_time_step_a(c, nRK - 1, c.prm.rk_a[nRK], src);
  }
}

int fhip_update(fhip_ctx* h, int n_steps, int path) {
  return guard([&] {
    require(h, "null context");
    require(n_steps >= 0, "n_steps must be >= 0");
    fhip::Ctx& c = h->c;
    bool mixed;
    const bool fused = resolve_path(c, path, mixed);
    if (fused && n_steps > 0) fhip::fused_begin(c, mixed);
    // divF_spts[0] of the last step only: the earlier ones are overwritten
    for (int it = 0; it < n_steps; it++) update_step(c, fused, mixed, false, true, it == n_steps - 1);
    if (fused && n_steps > 0) fhip::fused_end(c);
    return 0;
  });
}

int fhip_update_many(fhip_ctx** hs, int n, int n_steps, int path) {
  return guard([&] {
    require(hs && n >= 0 && n_steps >= 0, "bad arguments");
    std::vector<bool> fused(n), mixed(n);
    for (int i = 0; i < n; i++) {
      require(hs[i] != nullptr, "null context");
      bool m;
      fused[i] = resolve_path(hs[i]->c, path, m);
      mixed[i] = m;
    }
    if (n_steps == 0) return 0;
    for (int i = 0; i < n; i++)
      if (fused[i]) fhip::fused_begin(hs[i]->c, mixed[i]);
    for (int it = 0; it < n_steps; it++)
      for (int i = 0; i < n; i++) update_step(hs[i]->c, fused[i], mixed[i], false, true, it == n_steps - 1);
    for (int i = 0; i < n; i++)
      if (fused[i]) fhip::fused_end(hs[i]->c);
    return 0;
  });
}

// ---- P-multigrid -------------------------------------------------------------

struct fhip_pmg {
  std::vector<fhip_ctx*> lv;
  std::vector<double*> R, Pro, sol, corr;  // device; R[i]: i+1 -> i, Pro[i]: i -> i+1
  int smooth = 1;
  // A cycle's levels run one after the other, so every launch of a cycle goes
  // to the fine level's stream (the coarse levels' own streams are swapped
  // for it while the cycle is queued); work a caller queued on a coarse
  // level's own stream is finished before a cycle is queued.
  // one captured cycle per path, replayed on the fine level's stream; valid
  // while no level has re-allocated arrays since (graph_gen of every level)
  hipGraphExec_t graph[6] = {};  // by path
  std::vector<unsigned long long> gen[6];
  bool warm[6] = {};
  // batched form (fhip_pmg_create_batch): lv holds the merged coarse levels,
  // fine the fine contexts; coarse level j takes its first splitE[j]
  // elements from fine context fineOf[j] (-1: none) and the rest from level
  // j + 1.  ev[k], k >= 1: fine context k's queued work; ev[0]: a cycle's end.
  std::vector<fhip_ctx*> fine;
  std::vector<int> fineOf;
  std::vector<long long> splitE;
  std::vector<hipEvent_t> ev;
  // the stream a cycle runs on
  fhip_ctx* owner() const { return fine.empty() ? lv.back() : fine[0]; }
};

static void pmg_free(fhip_pmg* mg) {
  if (!mg) return;
  if (!mg->lv.empty()) (void)hipSetDevice(mg->owner()->c.device);
  if (!mg->lv.empty()) (void)hipStreamSynchronize(mg->owner()->c.stream);
  for (fhip_ctx* h : mg->fine) (void)hipStreamSynchronize(h->c.stream);  // their waits on ev[0]
  for (hipGraphExec_t g : mg->graph)
    if (g) (void)hipGraphExecDestroy(g);
  for (hipEvent_t e : mg->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto* v : {&mg->R, &mg->Pro, &mg->sol, &mg->corr})
    for (double* p : *v)
      if (p) (void)hipFree(p);
  delete mg;
}

int fhip_pmg_create(fhip_ctx* const* levels, int n_levels, const double* const* opp_restrict,
                    const double* const* opp_prolong, int smooth_steps, fhip_pmg** out) {
  fhip_pmg* mg = nullptr;
  int rc = guard([&] {
    require(levels && opp_restrict && opp_prolong && out, "null argument");
    require(n_levels >= 2, "P-multigrid needs at least two levels");
    require(smooth_steps >= 0, "smoothSteps must be >= 0");
    const fhip::Ctx& f = levels[n_levels - 1]->c;
    for (int i = 0; i < n_levels; i++) {
      require(levels[i] != nullptr, "null level");
      const fhip::Ctx& c = levels[i]->c;
      require(c.z.E == f.z.E && c.z.N == f.z.N && c.z.D == f.z.D && c.device == f.device,
              "P-multigrid levels must share the mesh, the equation and the device");
      if (i + 1 < n_levels) {
        require(c.z.P + 1 == levels[i + 1]->c.z.P, "P-multigrid level orders must ascend by one");
        require(opp_restrict[i] && opp_prolong[i], "null transfer operator");
      }
    }
    HCHK(hipSetDevice(f.device));
    mg = new fhip_pmg();
    mg->lv.assign(levels, levels + n_levels);
    mg->smooth = smooth_steps;
    for (int i = 0; i + 1 < n_levels; i++) {
      fhip::Ctx& c = levels[i]->c;
      const size_t Sc = c.z.S, Sf = levels[i + 1]->c.z.S, nS = c.z.nS();
      mg->R.push_back(dalloc<double>(Sc * Sf));
      HCHK(hipMemcpy(mg->R.back(), opp_restrict[i], Sc * Sf * sizeof(double), hipMemcpyHostToDevice));
      mg->Pro.push_back(dalloc<double>(Sc * Sf));
      HCHK(hipMemcpy(mg->Pro.back(), opp_prolong[i], Sc * Sf * sizeof(double), hipMemcpyHostToDevice));
      mg->sol.push_back(dalloc<double>(nS));
      mg->corr.push_back(dalloc<double>(nS));
      if (!c.src) c.src = dalloc<double>(nS);
    }
    *out = mg;
    return 0;
  });
  if (rc != 0) {
    pmg_free(mg);
    if (out) *out = nullptr;
  }
  return rc;
}

void fhip_pmg_destroy(fhip_pmg* mg) { pmg_free(mg); }

int fhip_pmg_create_batch(fhip_ctx* const* fine, int n_fine, fhip_ctx* const* coarse, int n_coarse,
                          const double* const* opp_restrict, const double* const* opp_prolong, int smooth_steps,
                          fhip_pmg** out) {
  fhip_pmg* mg = nullptr;
  int rc = guard([&] {
    require(fine && coarse && opp_restrict && opp_prolong && out, "null argument");
    require(n_fine >= 1 && n_coarse >= 1, "a batched P-multigrid needs fine and coarse levels");
    require(smooth_steps >= 0, "smoothSteps must be >= 0");
    for (int k = 0; k < n_fine; k++) require(fine[k] != nullptr, "null fine level");
    for (int j = 0; j < n_coarse; j++) require(coarse[j] != nullptr, "null coarse level");
    const fhip::Ctx& c0 = coarse[0]->c;
    const int low = c0.z.P;
    for (int k = 0; k < n_fine; k++) {
      const fhip::Ctx& f = fine[k]->c;
      require(f.z.N == c0.z.N && f.z.D == c0.z.D && f.device == c0.device,
              "batched P-multigrid levels must share the equation, the dimension and the device");
      require(f.z.P > low && f.z.P <= low + n_coarse, "fine orders must lie above the coarse levels'");
      require(k == 0 || f.z.P > fine[k - 1]->c.z.P, "fine orders must ascend");
    }
    require(fine[n_fine - 1]->c.z.P == low + n_coarse, "the top coarse level must be one below the highest fine order");
    std::vector<int> of(n_coarse, -1);
    for (int k = 0; k < n_fine; k++) of[fine[k]->c.z.P - low - 1] = k;
    for (int j = 0; j < n_coarse; j++) {
      const fhip::Ctx& c = coarse[j]->c;
      require(c.z.N == c0.z.N && c.z.D == c0.z.D && c.device == c0.device,
              "batched P-multigrid levels must share the equation, the dimension and the device");
      require(c.z.P == low + j, "coarse level orders must ascend by one");
      require(opp_restrict[j] && opp_prolong[j], "null transfer operator");
      const long long eA = of[j] >= 0 ? fine[of[j]]->c.z.E : 0;
      const long long eB = j + 1 < n_coarse ? coarse[j + 1]->c.z.E : 0;
      require(c.z.E == eA + eB, "a coarse level must hold the meshes of the fine level one order up and of the "
                                "coarse level above it");
    }
    HCHK(hipSetDevice(c0.device));
    mg = new fhip_pmg();
    mg->lv.assign(coarse, coarse + n_coarse);
    mg->fine.assign(fine, fine + n_fine);
    mg->fineOf = of;
    mg->smooth = smooth_steps;
    for (int j = 0; j < n_coarse; j++) {
      fhip::Ctx& c = coarse[j]->c;
      // the level one order up: the coarse level above, or at the top the fine one
      const size_t Sc = c.z.S, nS = c.z.nS();
      const size_t Sf = j + 1 < n_coarse ? coarse[j + 1]->c.z.S : fine[of[j]]->c.z.S;
      mg->splitE.push_back(of[j] >= 0 ? fine[of[j]]->c.z.E : 0);
      mg->R.push_back(dalloc<double>(Sc * Sf));
      HCHK(hipMemcpy(mg->R.back(), opp_restrict[j], Sc * Sf * sizeof(double), hipMemcpyHostToDevice));
      mg->Pro.push_back(dalloc<double>(Sc * Sf));
      HCHK(hipMemcpy(mg->Pro.back(), opp_prolong[j], Sc * Sf * sizeof(double), hipMemcpyHostToDevice));
      mg->sol.push_back(dalloc<double>(nS));
      if (!c.src) c.src = dalloc<double>(nS);
    }
    mg->ev.assign(n_fine, nullptr);
    for (int k = 0; k < n_fine; k++) HCHK(hipEventCreateWithFlags(&mg->ev[k], hipEventDisableTiming));
    *out = mg;
    return 0;
  });
  if (rc != 0) {
    pmg_free(mg);
    if (out) *out = nullptr;
  }
  return rc;
}

// While a cycle is queued every level launches on the fine level's stream
// (batched form: fine context 0's); the others get their own streams back at
// scope exit.
struct CycleStream {
  fhip_pmg* mg;
  std::vector<fhip_ctx*> all;
  std::vector<hipStream_t> own;
  explicit CycleStream(fhip_pmg* m) : mg(m) {
    hipStream_t st = mg->owner()->c.stream;
    all = mg->lv;
    all.insert(all.end(), mg->fine.begin(), mg->fine.end());
    for (fhip_ctx* h : all) {
      own.push_back(h->c.stream);
      h->c.stream = st;
    }
  }
  ~CycleStream() {
    for (size_t i = 0; i < own.size(); i++) all[i]->c.stream = own[i];
  }
};

// Work a caller queued on a coarse level's own stream (set-up, a residual)
// is finished on the host before a cycle is queued: cycles leave nothing
// there, so this is rare, and no device-side wait ties the fine stream to the
// coarse ones (a wait can hold back other contexts' streams that share its
// hardware queue).
static void pmg_settle_coarse(fhip_pmg* mg) {
  const int n = (int)mg->lv.size();
  for (int i = 0; i + (mg->fine.empty() ? 1 : 0) < n; i++) {
    hipStream_t coarse = mg->lv[i]->c.stream;
    const hipError_t q = hipStreamQuery(coarse);
    if (q == hipErrorNotReady) HCHK(hipStreamSynchronize(coarse));
    else HCHK(q);
  }
}

static void pmg_residual(fhip::Ctx& c, bool fused, bool mixed) {
  if (fused) fhip::fused_residual(c, mixed);
  else fhip::staged_residual(c, 0);
}

static void pmg_smooth(fhip::Ctx& c, int n, bool fused, bool mixed) {
  if (n <= 0) return;
  if (fused) fhip::fused_begin(c, mixed, true);
  for (int i = 0; i < n; i++) update_step(c, fused, mixed, true, false);
  if (fused) fhip::fused_end(c);
}

// multiGrid::restrict_pmg (src/multigrid.cpp:315-345), on the coarse stream
static void pmg_restrict(fhip_pmg* mg, int i) {
  fhip::Ctx& f = mg->lv[i + 1]->c;
  fhip::Ctx& c = mg->lv[i]->c;
  fhip::op_apply(c, "pmg_restrict", c.z.S, f.z.S, mg->R[i], f.U, c.U, false);
  fhip::op_apply(c, "pmg_restrict", c.z.S, f.z.S, mg->R[i], f.divF, c.divF, false);
}

// multiGrid::cycle (src/multigrid.cpp:169-313), P levels: the launches of one
// cycle, all on the fine level's stream (CycleStream).
static void pmg_cycle_body(fhip_pmg* mg, const std::vector<char>& fused, const std::vector<char>& mixed) {
  CycleStream one(mg);
  const int n = (int)mg->lv.size();
  fhip::Ctx& fine = mg->lv[n - 1]->c;
  {
    pmg_residual(fine, fused[n - 1], mixed[n - 1]);
    pmg_restrict(mg, n - 2);
    for (int i = n - 2; i >= 0; i--) {
      fhip::Ctx& g = mg->lv[i]->c;
      const long long nS = g.z.nS();
      // compute_source_term (src/multigrid.cpp:381-398)
      fhip::copy_array(g, g.src, g.divF, nS);
      pmg_residual(g, fused[i], mixed[i]);
      fhip::sub_array(g, g.src, g.src, g.divF, nS);
      fhip::copy_array(g, mg->sol[i], g.U, nS);
      pmg_smooth(g, mg->smooth, fused[i], mixed[i]);
      if (i > 0) {
        pmg_residual(g, fused[i], mixed[i]);
        fhip::add_array(g, g.divF, g.src, nS);
        pmg_restrict(mg, i - 1);
      }
    }
    // upward: smooth again, corr = U - sol, prolong_err (src/multigrid.cpp:347-364)
    for (int i = 0; i + 1 < n; i++) {
      fhip::Ctx& g = mg->lv[i]->c;
      fhip::Ctx& up = mg->lv[i + 1]->c;
      pmg_smooth(g, mg->smooth, fused[i], mixed[i]);
      fhip::sub_array(g, mg->corr[i], g.U, mg->sol[i], g.z.nS());
      fhip::op_apply(up, "pmg_prolong", up.z.S, g.z.S, mg->Pro[i], mg->corr[i], up.U, true);
    }
  }
}

// The same cycle when every level runs a fused path of one precision: the
// levels stay element-major from the fine level's import to the end (no
// layout conversions around each residual and smoothing pass); the
// transfers act on the element-major arrays.  The coarse levels' wave speeds
// come in once (the face kernel rewrites only the left and boundary entries,
// the dt sums all of them).  At the end every level's U, and the fine level's
// divF_spts[0] and wave speeds, are written back to FlurryPP's layouts (the
// coarse levels' residual and the wave speeds the cycle computed stay
// element-major: the next cycle computes them again before any use).
static void pmg_cycle_body_em(fhip_pmg* mg, bool mixed) {
  CycleStream one(mg);
  const int n = (int)mg->lv.size();
  fhip::Ctx& fine = mg->lv[n - 1]->c;
  for (int i = 0; i + 1 < n; i++) {
    fhip::fused_em_prepare(mg->lv[i]->c, mixed);
    fhip::fused_em_import_wave(mg->lv[i]->c);
  }
  fhip::fused_begin(fine, mixed, false);
  fhip::fused_em_residual(fine, mixed);
  auto smooth = [&](fhip::Ctx& g) {
    fhip::fused_em_set_src(g, true);
    for (int k = 0; k < mg->smooth; k++) update_step(g, true, mixed, true, false, false);
    fhip::fused_em_set_src(g, false);
  };
  for (int i = n - 2; i >= 0; i--) {
    fhip::Ctx& g = mg->lv[i]->c;
    fhip::Ctx& f = mg->lv[i + 1]->c;
    // restriction of U and divF_spts[0]; compute_source_term (src/multigrid.cpp:381-398)
    fhip::pmg_restrict_em(g, 0, g.z.E, f.z.S, mg->R[i], fhip::fused_em_array(f, 0), fhip::fused_em_array(f, 1),
                          mg->sol[i]);
    fhip::fused_em_extrap(g, mixed);
    fhip::fused_em_residual(g, mixed, 1);  // source -= residual
    smooth(g);
    if (i > 0) fhip::fused_em_residual(g, mixed, 2);  // residual + source, restricted below
  }
  // upward: smooth again, prolong_err (src/multigrid.cpp:347-364)
  for (int i = 0; i + 1 < n; i++) {
    fhip::Ctx& g = mg->lv[i]->c;
    fhip::Ctx& up = mg->lv[i + 1]->c;
    smooth(g);
    fhip::pmg_prolong_em(up, g.z.S, mg->Pro[i], fhip::fused_em_array(g, 0), mg->sol[i]);
    if (i + 2 < n) fhip::fused_em_extrap(up, mixed);  // the next level up smooths from here
    fhip::fused_em_export_u(g);
  }
  fhip::fused_end(fine);
}

// The batched cycle: every mesh goes through the steps of
// pmg_cycle_body_em in the same order; a merged coarse level's steps run once
// for all of its meshes.  Down the levels a coarse level restricts its first
// elements from the fine context one order up and the rest from the coarse
// level above; up the levels it prolongs into the same two.
static void pmg_cycle_body_batch(fhip_pmg* mg, bool mixed) {
  CycleStream one(mg);
  const int nc = (int)mg->lv.size();
  for (int j = 0; j < nc; j++) {
    fhip::fused_em_prepare(mg->lv[j]->c, mixed);
    fhip::fused_em_import_wave(mg->lv[j]->c);
  }
  for (fhip_ctx* h : mg->fine) {
    fhip::fused_begin(h->c, mixed, false);
    fhip::fused_em_residual(h->c, mixed);
  }
  auto smooth = [&](fhip::Ctx& g) {
    fhip::fused_em_set_src(g, true);
    for (int k = 0; k < mg->smooth; k++) update_step(g, true, mixed, true, false, false);
    fhip::fused_em_set_src(g, false);
  };
  for (int j = nc - 1; j >= 0; j--) {
    fhip::Ctx& g = mg->lv[j]->c;
    const long long eA = mg->splitE[j];
    if (mg->fineOf[j] >= 0) {
      fhip::Ctx& f = mg->fine[mg->fineOf[j]]->c;
      fhip::pmg_restrict_em(g, 0, eA, f.z.S, mg->R[j], fhip::fused_em_array(f, 0), fhip::fused_em_array(f, 1),
                            mg->sol[j]);
    }
    if (j + 1 < nc) {
      fhip::Ctx& f = mg->lv[j + 1]->c;
      fhip::pmg_restrict_em(g, eA, g.z.E - eA, f.z.S, mg->R[j], fhip::fused_em_array(f, 0),
                            fhip::fused_em_array(f, 1), mg->sol[j]);
    }
    fhip::fused_em_extrap(g, mixed);
    fhip::fused_em_residual(g, mixed, 1);  // source -= residual
    smooth(g);
    if (j > 0) fhip::fused_em_residual(g, mixed, 2);  // residual + source, restricted below
  }
  for (int j = 0; j < nc; j++) {
    fhip::Ctx& g = mg->lv[j]->c;
    const long long offA = mg->splitE[j] * g.z.N * g.z.S;
    smooth(g);
    if (mg->fineOf[j] >= 0)
      fhip::pmg_prolong_em(mg->fine[mg->fineOf[j]]->c, g.z.S, mg->Pro[j], fhip::fused_em_array(g, 0), mg->sol[j]);
    if (j + 1 < nc) {
      fhip::Ctx& up = mg->lv[j + 1]->c;
      fhip::pmg_prolong_em(up, g.z.S, mg->Pro[j], fhip::fused_em_array(g, 0) + offA, mg->sol[j] + offA);
      fhip::fused_em_extrap(up, mixed);  // the next level up smooths from here
    }
    fhip::fused_em_export_u(g);
  }
  for (fhip_ctx* h : mg->fine) fhip::fused_end(h->c);
}

// The first cycle on a path runs launch by launch (it also allocates what the
// fused paths allocate on first use); the second is captured into one graph
// on the fine level's stream, which later cycles replay there.  Timed levels,
// disabled graphs or re-allocated arrays fall back to launch-by-launch cycles
// (and a new capture).
int fhip_pmg_cycle(fhip_pmg* mg, int path) {
  return guard([&] {
    require(mg, "null multigrid");
    std::vector<fhip_ctx*> all = mg->lv;
    all.insert(all.end(), mg->fine.begin(), mg->fine.end());
    const int n = (int)all.size();
    std::vector<char> fused(n), mixed(n);
    bool graphs = true;
    std::vector<unsigned long long> gen(n);
    for (int i = 0; i < n; i++) {
      bool m;
      fhip::Ctx& c = all[i]->c;
      fused[i] = resolve_path(c, path, m);
      mixed[i] = m;
      graphs = graphs && c.use_graphs && !c.timing;
      gen[i] = c.graph_gen;
    }
    fhip::Ctx& own = mg->owner()->c;
    HCHK(hipSetDevice(own.device));
    hipGraphExec_t& ge = mg->graph[path];
    if (ge && (!graphs || mg->gen[path] != gen)) {
      HCHK(hipStreamSynchronize(own.stream));
      (void)hipGraphExecDestroy(ge);
      ge = nullptr;
    }
    pmg_settle_coarse(mg);
    bool em = true;
    for (int i = 0; i < n; i++) em = em && fused[i] && mixed[i] == mixed[0];
    const bool batch = !mg->fine.empty();
    require(!batch || em, "a batched P-multigrid runs on a fused path of one precision");
    auto body = [&] {
      if (batch) pmg_cycle_body_batch(mg, mixed[0] != 0);
      else if (em) pmg_cycle_body_em(mg, mixed[0] != 0);
      else pmg_cycle_body(mg, fused, mixed);
    };
    // batched form: the cycle follows the work queued on every fine
    // context's stream, and their later work follows the cycle
    if (batch) {
      for (size_t k = 1; k < mg->fine.size(); k++) {
        HCHK(hipEventRecord(mg->ev[k], mg->fine[k]->c.stream));
        HCHK(hipStreamWaitEvent(own.stream, mg->ev[k], 0));
      }
    }
    auto join = [&] {
      if (!batch || mg->fine.size() < 2) return;
      HCHK(hipEventRecord(mg->ev[0], own.stream));
      for (size_t k = 1; k < mg->fine.size(); k++) HCHK(hipStreamWaitEvent(mg->fine[k]->c.stream, mg->ev[0], 0));
    };
    if (!graphs || !mg->warm[path]) {
      body();
      join();
      mg->warm[path] = true;
      return 0;
    }
    if (!ge) {
      struct Flag {  // t_capturing_cycle for the capture's duration, however it ends
        Flag() { t_capturing_cycle = true; }
        ~Flag() { t_capturing_cycle = false; }
      } flag;
      ge = capture(own.stream, [&] { body(); });
      for (int i = 0; i < n; i++) gen[i] = all[i]->c.graph_gen;
      mg->gen[path] = gen;
    }
    HCHK(hipGraphLaunch(ge, own.stream));
    join();
    return 0;
  });
}

int fhip_mesh_residual_norms(fhip_ctx* h, int res_type, double* out) {
  return guard([&] {
    require(h && out, "null argument");
    require(res_type >= 1 && res_type <= 3, "resType must be 1, 2 or 3");
    require(h->c.fused_ok, "per-mesh residuals need the fused path");
    fhip::fused_mesh_residuals(h->c, res_type, out);
    return 0;
  });
}

int fhip_set_mesh_active(fhip_ctx* h, const int* active) {
  return guard([&] {
    require(h && active, "null argument");
    require(h->c.fused_ok, "mesh activity flags need the fused path");
    fhip::fused_set_active(h->c, active);
    return 0;
  });
}

int fhip_path_supported(fhip_ctx* h, int path) {
  if (!h) return 0;
  if (path == FHIP_PATH_STAGED || path == FHIP_PATH_AUTO) return 1;
  if (path == FHIP_PATH_FUSED || path == FHIP_PATH_MIXED) return h->c.fused_ok ? 1 : 0;
  if (path == FHIP_PATH_MIXED_BF16 || path == FHIP_PATH_MIXED_BF16X3)
    return h->c.fused_ok && h->c.z.D == 3 && h->c.z.P == 3 ? 1 : 0;
  return 0;
}

static bool tuning_path(fhip::Ctx& c, int path) {
  bool mixed;
  const bool fused = resolve_path(c, path, mixed);
  require(fused, "block sizes apply to the fused paths");
  return mixed;
}

int fhip_autotune(fhip_ctx* h, int path, int* block_sizes) {
  return guard([&] {
    require(h, "null context");
    fhip::Ctx& c = h->c;
    const bool mixed = tuning_path(c, path);
    fhip::fused_autotune(c, mixed, block_sizes);
    return 0;
  });
}

int fhip_set_block_sizes(fhip_ctx* h, int path, const int* block_sizes) {
  return guard([&] {
    require(h && block_sizes, "null argument");
    fhip::Ctx& c = h->c;
    const bool mixed = tuning_path(c, path);
    fhip::fused_set_blocks(c, mixed, block_sizes);
    return 0;
  });
}

int fhip_get_block_sizes(fhip_ctx* h, int path, int* block_sizes) {
  return guard([&] {
    require(h && block_sizes, "null argument");
    fhip::Ctx& c = h->c;
    const bool mixed = tuning_path(c, path);
    fhip::fused_get_blocks(c, mixed, block_sizes);
    return 0;
  });
}

int fhip_set_option(fhip_ctx* h, int option, int value) {
  return guard([&] {
    require(h, "null context");
    if (option == FHIP_OPT_GRAPHS) h->c.use_graphs = value != 0;
    else throw ApiError(FHIP_EINVAL, "unknown option");
    return 0;
  });
}

int fhip_sync(fhip_ctx* h) {
  return guard([&] {
    require(h, "null context");
    HCHK(hipStreamSynchronize(h->c.stream));
    return 0;
  });
}

int fhip_residual_norm(fhip_ctx* h, int res_type, double* out) {
  return guard([&] {
    require(h && out, "null argument");
    require(res_type >= 1 && res_type <= 3, "resType must be 1, 2 or 3");
    fhip::residual_norm(h->c, res_type, out);
    return 0;
  });
}

int fhip_residual_norm_noalloc(fhip_ctx* h, int res_type, double* out) {
  return guard([&] {
    require(h && out, "null argument");
    require(res_type >= 1 && res_type <= 3, "resType must be 1, 2 or 3");
    try {
      fhip::residual_norm_noalloc(h->c, res_type, out);
    } catch (const std::length_error& e) {
      throw ApiError(FHIP_ESIZE, e.what());
    }
    return 0;
  });
}

int fhip_check_state(fhip_ctx* h) {
  return guard([&] {
    require(h, "null context");
    if (fhip::check_state(h->c)) throw ApiError(FHIP_ESTATE, "non-finite or non-physical state (NaN/Inf, rho<=0 or p<=0)");
    return 0;
  });
}

long long fhip_array_size(fhip_ctx* h, int which) {
  if (!h) return -1;
  if (which == FHIP_ARR_FACE_FPT_L || which == FHIP_ARR_FACE_FPT_R) return h->c.z.nFaceFpts;
  long long n = -1;
  int rc = guard([&] {
    array_ptr(h->c, which, 0, &n);
    return 0;
  });
  return rc ? -1 : n;
}

int fhip_get_array(fhip_ctx* h, int which, int step, void* host) {
  return guard([&] {
    require(h && host, "null argument");
    fhip::Ctx& c = h->c;
    if (which == FHIP_ARR_FACE_FPT_L || which == FHIP_ARR_FACE_FPT_R) {
      const std::vector<int>& v = (which == FHIP_ARR_FACE_FPT_L) ? c.h_ptL : c.h_ptR;
      std::memcpy(host, v.data(), v.size() * sizeof(int));
      return 0;
    }
    long long n;
    double* d = array_ptr(c, which, step, &n);
    require(d != nullptr && n > 0, "array not allocated in this configuration");
    HCHK(hipStreamSynchronize(c.stream));
    if (which == FHIP_ARR_WAVESP) {
      // device [fpt][ele] -> reference per-element vectors [ele][fpt]
      std::vector<double> tmp(n);
      HCHK(hipMemcpy(tmp.data(), d, n * sizeof(double), hipMemcpyDeviceToHost));
      double* o = (double*)host;
      for (int f = 0; f < c.z.F; f++)
        for (int e = 0; e < c.z.E; e++) o[(size_t)e * c.z.F + f] = tmp[(size_t)f * c.z.E + e];
      return 0;
    }
    HCHK(hipMemcpy(host, d, n * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
  });
}

int fhip_set_array(fhip_ctx* h, int which, int step, const void* host) {
  return guard([&] {
    require(h && host, "null argument");
    fhip::Ctx& c = h->c;
    require(which != FHIP_ARR_FACE_FPT_L && which != FHIP_ARR_FACE_FPT_R, "face maps are read-only");
    long long n;
    double* d = array_ptr(c, which, step, &n);
    require(d != nullptr && n > 0, "array not allocated in this configuration");
    HCHK(hipStreamSynchronize(c.stream));
    if (which == FHIP_ARR_WAVESP) {
      std::vector<double> tmp(n);
      const double* in = (const double*)host;
      for (int f = 0; f < c.z.F; f++)
        for (int e = 0; e < c.z.E; e++) tmp[(size_t)f * c.z.E + e] = in[(size_t)e * c.z.F + f];
      HCHK(hipMemcpy(d, tmp.data(), n * sizeof(double), hipMemcpyHostToDevice));
      return 0;
    }
    HCHK(hipMemcpy(d, host, n * sizeof(double), hipMemcpyHostToDevice));
    return 0;
  });
}

int fhip_device_ptr(fhip_ctx* h, int which, int step, void** dev) {
  return guard([&] {
    require(h && dev, "null argument");
    long long n;
    *dev = array_ptr(h->c, which, step, &n);
    return 0;
  });
}

int fhip_stream(fhip_ctx* h, void** stream) {
  if (!h || !stream) return FHIP_EINVAL;
  *stream = (void*)h->c.stream;
  return 0;
}

int fhip_enable_timing(fhip_ctx* h, int on) {
  if (!h) return FHIP_EINVAL;
  fhip::timer_resolve(h->c);
  h->c.timing = on != 0;
  h->c.timer.acc.clear();
  return 0;
}

int fhip_kernel_times(fhip_ctx* h, int max, char* names, int name_len, double* avg_ms, long long* counts) {
  if (!h) return -1;
  fhip::timer_resolve(h->c);
  int i = 0;
  for (auto& kv : h->c.timer.acc) {
    if (i >= max) break;
    if (names) {
      std::strncpy(names + (size_t)i * name_len, kv.first.c_str(), name_len - 1);
      names[(size_t)i * name_len + name_len - 1] = 0;
    }
    if (avg_ms) avg_ms[i] = kv.second.second ? kv.second.first / kv.second.second : 0;
    if (counts) counts[i] = kv.second.second;
    i++;
  }
  return i;
}

}  // extern "C"
