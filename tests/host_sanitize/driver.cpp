// Exercises libflurry_host's C API (operators, P-multigrid transfer operators,
// Cartesian and array meshes, periodic pairing, face lists, metric terms) so
// that the sources can run under AddressSanitizer and UBSan on the CPU.
// Exit status 0 when every call succeeds.
#include <cstdio>
#include <string>
#include <cmath>
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

// A 2x1 quad mesh written as Gmsh 2.2 and read back, plus a file whose
// element names a node that does not exist (must fail cleanly).
static int gmsh_case(const char* dir) {
  const std::string good = std::string(dir) + "/fhost_san_good.msh", bad = std::string(dir) + "/fhost_san_bad.msh";
  const char* body =
      "$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$PhysicalNames\n2\n1 1 \"Wall\"\n2 2 \"FLUID\"\n"
      "$EndPhysicalNames\n$Nodes\n6\n1 0 0 0\n2 1 0 0\n3 2 0 0\n4 0 1 0\n5 1 1 0\n6 2 1 0\n$EndNodes\n"
      "$Elements\n8\n1 1 2 1 1 1 2\n2 1 2 1 1 2 3\n3 1 2 1 1 3 6\n4 1 2 1 1 6 5\n5 1 2 1 1 5 4\n"
      "6 1 2 1 1 4 1\n7 3 2 2 2 1 2 5 4\n8 3 2 2 2 2 3 6 %d\n$EndElements\n";
  const char* names[2] = {"wall", "fluid"};
  const char* bcs[2] = {"slip_wall", "fluid"};
  for (int pass = 0; pass < 2; pass++) {
    const std::string& path = pass ? bad : good;
    FILE* fp = std::fopen(path.c_str(), "w");
    if (!fp) return 1;
    std::fprintf(fp, body, pass ? 9 : 5);
    std::fclose(fp);
    fhost_mesh* m = nullptr;
    const int rc = fhost_mesh_read_gmsh(path.c_str(), 2, names, bcs, &m);
    std::remove(path.c_str());
    if (pass == 1) {
      if (rc == 0) return 1;  // node 9 of 6 must be rejected
      continue;
    }
    CHK(rc);
    CHK(fhost_mesh_process(m));
    std::vector<double> dj((size_t)4 * 2);
    CHK(fhost_mesh_geometry(m, 1, 0, nullptr, nullptr, nullptr, dj.data(), nullptr, nullptr));
    fhost_mesh_destroy(m);
  }
  return 0;
}

// Curved elements: a 9-node quadrangle and a 27-node hexahedron at their
// reference node positions (Gmsh order), metric terms at orders 0-5, the
// shape nodes read back, and shape-node counts that must be refused.
static int curved_case() {
  const double q9[9][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}, {0, -1}, {1, 0}, {0, 1}, {-1, 0}, {0, 0}};
  const int ce[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1}, {-1, -1, 1}, {1, -1, 1}, {1, 1, 1}, {-1, 1, 1}};
  const int ed[12][2] = {{0, 1}, {0, 3}, {0, 4}, {1, 2}, {1, 5}, {2, 3}, {2, 6}, {3, 7}, {4, 5}, {4, 7}, {5, 6}, {6, 7}};
  const int fa[6][4] = {{0, 3, 2, 1}, {0, 1, 5, 4}, {0, 4, 7, 3}, {1, 2, 6, 5}, {2, 3, 7, 6}, {4, 5, 6, 7}};
  std::vector<double> h27;
  for (auto& c : ce)
    for (int d = 0; d < 3; d++) h27.push_back(c[d]);
  for (auto& e : ed)
    for (int d = 0; d < 3; d++) h27.push_back(0.5 * (ce[e[0]][d] + ce[e[1]][d]));
  for (auto& f : fa)
    for (int d = 0; d < 3; d++) h27.push_back(0.25 * (ce[f[0]][d] + ce[f[1]][d] + ce[f[2]][d] + ce[f[3]][d]));
  for (int d = 0; d < 3; d++) h27.push_back(0.0);
  for (int D = 2; D <= 3; D++) {
    const int nn = D == 2 ? 9 : 27, nc = D == 2 ? 4 : 8;
    std::vector<int> c2v(nc), c2n(nn);
    for (int i = 0; i < nc; i++) c2v[i] = i;
    for (int i = 0; i < nn; i++) c2n[i] = i;
    const int ptr[1] = {0};
    fhost_mesh* m = nullptr;
    CHK(fhost_mesh_from_arrays(D, nn, D == 2 ? &q9[0][0] : h27.data(), 1, c2v.data(), 0, nullptr, ptr, nullptr, &m));
    if (fhost_mesh_set_shape_nodes(m, nn - 1, c2n.data()) == 0) return 1;  // not a Lagrange element
    CHK(fhost_mesh_set_shape_nodes(m, nn, c2n.data()));
    int ns = 0;
    std::vector<int> back(nn);
    CHK(fhost_mesh_get_shape_nodes(m, &ns, back.data()));
    if (ns != nn || back != c2n) return 1;
    for (int order = 0; order <= 5; order++) {
      int S, F, Ff;
      CHK(fhost_sizes(D, order, &S, &F, &Ff));
      std::vector<double> ps((size_t)S * D), pf((size_t)F * D), jg((size_t)D * S * D), dj(S), nr((size_t)F * D), da(F);
      CHK(fhost_mesh_geometry(m, order, 0, ps.data(), pf.data(), jg.data(), dj.data(), nr.data(), da.data()));
      for (double v : dj)
        if (std::fabs(v - 1.0) > 1e-12) return 1;
    }
    fhost_mesh_destroy(m);
  }
  return 0;
}

int main(int argc, char** argv) {
  if (operators_all()) return 1;
  if (curved_case()) return 1;
  if (gmsh_case(argc > 1 ? argv[1] : "/tmp")) return 1;
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
