// Exercises libflurry_host's C API (operators, P-multigrid transfer operators,
// Cartesian and array meshes, periodic pairing, face lists, metric terms) so
// that the sources can run under AddressSanitizer and UBSan on the CPU.
// Exit status 0 when every call succeeds.
#include <cstdio>
#include <vector>

#include "flurry_host.h"

#define CHK(x)                                                                 \
  do {                                                                         \
    if ((x) != 0) {                                                            \
      std::fprintf(stderr, "%s failed: %s\n", #x, fhost_last_error());         \
      return 1;                                                                \
    }                                                                          \
  } while (0)

static int operators_all() {
  for (int D = 2; D <= 3; D++)
    for (int P = 0; P <= 5; P++)
      for (int pt = 0; pt <= 1; pt++) {
        int S, F, Ff;
        CHK(fhost_sizes(D, P, &S, &F, &Ff));
        std::vector<double> sf((size_t)F * S), g((size_t)D * S * S), ef((size_t)D * F * S), c((size_t)S * F),
            cu((size_t)D * S * F), ws(S), wf(Ff);
        CHK(fhost_operators(D, P, pt, 0, sf.data(), g.data(), ef.data(), c.data(), cu.data(), ws.data(), wf.data()));
        int up = 1, dn = 1;
        for (int d = 0; d < D; d++) up *= P + 2, dn *= P;
        std::vector<double> pro((size_t)up * S), res((size_t)(dn ? dn : 1) * S);
        CHK(fhost_pmg_operators(D, P, pt, pro.data(), res.data()));
      }
  return 0;
}

static int mesh_case(int D, int n, const char* const* bcs, double jitter, int order) {
  const double lo[3] = {0, 0, 0}, hi[3] = {1, 2, 3};
  fhost_mesh* m = nullptr;
  CHK(fhost_mesh_cartesian(D, n, n + 1, D == 3 ? n : 1, lo, hi, bcs, &m));
  if (jitter > 0) {
    const double h[3] = {1.0 / n, 2.0 / (n + 1), 3.0 / n};
    CHK(fhost_mesh_jitter(m, h, jitter, 7));
  }
  CHK(fhost_mesh_process(m));
  int nd, nv, ne, ni, nb;
  CHK(fhost_mesh_counts(m, &nd, &nv, &ne, &ni, &nb));
  const int nc = D == 2 ? 4 : 8, nf = ni + nb;
  std::vector<double> xv((size_t)nv * D);
  std::vector<int> c2v((size_t)ne * nc);
  CHK(fhost_mesh_get_nodes(m, xv.data(), c2v.data()));
  std::vector<int> a(nf), b(nf), c(nf), d(nf), e(nf), f(nf), g(nf), h(nf);
  CHK(fhost_mesh_get_faces(m, a.data(), b.data(), c.data(), d.data(), e.data(), f.data(), g.data(), h.data()));
  int S, F, Ff;
  CHK(fhost_sizes(D, order, &S, &F, &Ff));
  std::vector<double> ps((size_t)S * ne * D), pf((size_t)F * ne * D), jg((size_t)D * S * ne * D), dj((size_t)S * ne),
      nr((size_t)F * ne * D), da((size_t)F * ne);
  CHK(fhost_mesh_geometry(m, order, 0, ps.data(), pf.data(), jg.data(), dj.data(), nr.data(), da.data()));
  fhost_mesh_destroy(m);
  // the same nodes and elements through the array interface, one boundary of all boundary vertices
  std::vector<int> bnd;
  for (int v = 0; v < nv; v++) {
    bool on = false;
    for (int k = 0; k < D; k++) on = on || xv[(size_t)v * D + k] == lo[k] || xv[(size_t)v * D + k] == hi[k];
    if (on) bnd.push_back(v);
  }
  const int bc_list[1] = {fhost_bc_from_name("slip_wall")}, ptr[2] = {0, (int)bnd.size()};
  CHK(fhost_mesh_from_arrays(D, nv, xv.data(), ne, c2v.data(), 1, bc_list, ptr, bnd.data(), &m));
  CHK(fhost_mesh_process(m));
  CHK(fhost_mesh_geometry(m, order, 1, nullptr, nullptr, jg.data(), dj.data(), nr.data(), da.data()));
  fhost_mesh_destroy(m);
  return 0;
}

int main() {
  if (operators_all()) return 1;
  const char* per2[4] = {"periodic", "periodic", "periodic", "periodic"};
  const char* mix2[4] = {"slip_wall", "char", "sup_out", "sup_in"};
  const char* per3[6] = {"periodic", "periodic", "periodic", "periodic", "periodic", "periodic"};
  const char* mix3[6] = {"slip_wall", "periodic", "char", "periodic", "adiabatic_noslip", "isothermal_noslip"};
  for (int P = 0; P <= 4; P++) {
    if (mesh_case(2, 5, per2, 0.1, P) || mesh_case(2, 4, mix2, 0.0, P)) return 1;
    if (mesh_case(3, 3, per3, 0.1, P) || mesh_case(3, 3, mix3, 0.0, P)) return 1;
  }
  if (fhost_bc_from_name("no_such_bc") != -2) return 1;
  std::printf("host sanitize driver: ok\n");
  return 0;
}
