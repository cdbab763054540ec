// Mesh, face connectivity and metric terms for quads and hexes (linear or
// curved: Lagrange elements up to quintic, the 20-node serendipity hex).
//
// FlurryPP builds its face list with an O(n^2) search (matrix::unique,
// src/matrix.cpp:668-695; findEq, src/geo.cpp:451; the periodic pairing loop,
// src/geo.cpp:1783-1834).  This file produces the same faces in the same
// order in O(n): unique faces come from a hash map in first-occurrence order,
// and periodic partners from a spatial hash whose hits are then accepted with
// FlurryPP's own matching predicate, lowest boundary index first.
#include "flurry_host.h"
#include "fr_host_internal.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <set>
#include <sstream>
#include <unordered_map>

using fr::Error;

namespace {

thread_local std::string g_err;

struct FaceRec {
  int eL, locL, eR, locR, relRot, bcType, gid, periodic;
};

struct ArrHash {
  size_t operator()(const std::array<int, 4>& a) const {
    uint64_t h = 1469598103934665603ull;
    for (int v : a) {
      h ^= (uint64_t)(uint32_t)v;
      h *= 1099511628211ull;
    }
    return (size_t)h;
  }
};

// Vertex lists of the six hex faces, src/geo.cpp:365-370.
const int kHexFaceVerts[6][4] = {{0, 1, 2, 3}, {4, 5, 6, 7}, {3, 0, 4, 7}, {2, 1, 5, 6}, {1, 0, 4, 5}, {3, 2, 6, 7}};
// Oriented vertex lists used for the relative rotation, src/geo.cpp:2089-2135.
const int kHexOrientVerts[6][4] = {{0, 1, 2, 3}, {5, 4, 7, 6}, {0, 3, 7, 4}, {2, 1, 5, 6}, {1, 0, 4, 5}, {3, 2, 6, 7}};

}  // namespace

struct fhost_mesh {
  int nDims = 0, nVerts = 0, nEles = 0, nNodes = 0, nFacesPerEle = 0, nvFace = 0;
  std::vector<double> xv;
  std::vector<int> c2v;
  std::vector<int> bcList;
  std::vector<std::vector<int>> bndPts;  // sorted, unique
  double pdx = INFINITY, pdy = INFINITY, pdz = INFINITY, ptol = 1e-6;
  bool processed = false;

  int nFaces = 0;
  std::vector<std::array<int, 4>> f2v;  // sorted vertex ids (2-D uses 2)
  std::vector<int> c2f, f2c, faceType;
  std::vector<int> intFaces, bndFaces, bndBc;
  std::vector<FaceRec> faces;
  int nIntFaces = 0, nBndFaces = 0;

  // Geometry nodes (fhost_mesh_set_shape_nodes): nShape per element in
  // Gmsh's node order, the first 4 (8) being the element's vertices; empty
  // for linear elements, whose vertices are their nodes.  lattice[n] is node
  // n's position (i, j, k) on the element's grid of nSide equidistant points
  // per direction (Lagrange elements); the 20-node hexahedron is serendipity.
  int nShape = 0, nSide = 2;
  bool serendipity = false;
  std::vector<int> c2n;
  std::vector<std::array<int, 3>> lattice;

  const double* X(int v) const { return &xv[(size_t)v * nDims]; }
  double x(int v, int d) const { return xv[(size_t)v * nDims + d]; }
  int node(size_t e, int n) const { return nShape ? c2n[e * nShape + n] : c2v[e * nNodes + n]; }
};

namespace {

void set_shape(fhost_mesh* m) {
  if (m->nDims == 2) {
    m->nNodes = 4;
    m->nFacesPerEle = 4;
    m->nvFace = 2;
  } else {
    m->nNodes = 8;
    m->nFacesPerEle = 6;
    m->nvFace = 4;
  }
}

// ---- periodic predicates -------------------------------------------------

// checkPeriodicFaces, src/geo.cpp:1870-1898
bool periodic2d(const fhost_mesh* m, const int* e1, const int* e2) {
  double x11 = m->x(e1[0], 0), y11 = m->x(e1[0], 1);
  double x12 = m->x(e1[1], 0), y12 = m->x(e1[1], 1);
  double x21 = m->x(e2[0], 0), y21 = m->x(e2[0], 1);
  double x22 = m->x(e2[1], 0), y22 = m->x(e2[1], 1);
  double tol = m->ptol, dx = m->pdx, dy = m->pdy;
  using std::abs;
  if (abs(abs(x21 - x11) - dx) < tol && abs(abs(x22 - x12) - dx) < tol && abs(y21 - y11) < tol && abs(y22 - y12) < tol)
    return true;
  if (abs(abs(x22 - x11) - dx) < tol && abs(abs(x21 - x12) - dx) < tol && abs(y22 - y11) < tol && abs(y21 - y12) < tol)
    return true;
  if (abs(abs(y21 - y11) - dy) < tol && abs(abs(y22 - y12) - dy) < tol && abs(x21 - x11) < tol && abs(x22 - x12) < tol)
    return true;
  if (abs(abs(y22 - y11) - dy) < tol && abs(abs(y21 - y12) - dy) < tol && abs(x22 - x11) < tol && abs(x21 - x12) < tol)
    return true;
  return false;
}

void face_normal_centroid3(const fhost_mesh* m, const int* f, double* n, double* c) {
  double v1[3], v2[3];
  for (int d = 0; d < 3; d++) {
    v1[d] = m->x(f[1], d) - m->x(f[0], d);
    v2[d] = m->x(f[2], d) - m->x(f[0], d);
  }
  c[0] = c[1] = c[2] = 0;
  for (int j = 0; j < 4; j++)
    for (int d = 0; d < 3; d++) c[d] += m->x(f[j], d);
  for (int d = 0; d < 3; d++) c[d] /= 4;
  n[0] = v1[1] * v2[2] - v1[2] * v2[1];
  n[1] = v1[2] * v2[0] - v1[0] * v2[2];
  n[2] = v1[0] * v2[1] - v1[1] * v2[0];
  double mag = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  for (int d = 0; d < 3; d++) n[d] /= mag;
}

// checkPeriodicFaces3D, src/geo.cpp:1900-1960
bool periodic3d(const fhost_mesh* m, const int* f1, const int* f2) {
  double n1[3], c1[3], n2[3], c2[3];
  face_normal_centroid3(m, f1, n1, c1);
  face_normal_centroid3(m, f2, n2, c2);
  double tol = m->ptol;
  double dot = n1[0] * n2[0] + n1[1] * n2[1] + n1[2] * n2[2];
  if (std::abs(1 - std::abs(dot)) > tol) return false;
  double nd[3] = {std::abs(n1[0] * m->pdx), std::abs(n1[1] * m->pdy), std::abs(n1[2] * m->pdz)};
  for (int d = 0; d < 3; d++) {
    double off = std::abs(c2[d] - c1[d]);
    double diff = std::abs(off - nd[d]);
    if (diff > tol) return false;  // NaN (0*inf) passes, as in the reference
  }
  return true;
}

bool periodic_match(const fhost_mesh* m, int fa, int fb) {
  if (m->nDims == 2) return periodic2d(m, m->f2v[fa].data(), m->f2v[fb].data());
  return periodic3d(m, m->f2v[fa].data(), m->f2v[fb].data());
}

// Bucket key for a periodic face: the dimension its normal points along and
// its tangential centroid coordinates quantized by the periodic tolerance.
struct PKey {
  int dir;
  long long t[2];
  bool operator==(const PKey& o) const { return dir == o.dir && t[0] == o.t[0] && t[1] == o.t[1]; }
};
struct PKeyHash {
  size_t operator()(const PKey& k) const {
    uint64_t h = (uint64_t)k.dir * 0x9E3779B97F4A7C15ull;
    h ^= (uint64_t)k.t[0] + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
    h ^= (uint64_t)k.t[1] + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
    return (size_t)h;
  }
};

void periodic_key(const fhost_mesh* m, int f, int& dir, double* tang) {
  const int* v = m->f2v[f].data();
  int D = m->nDims;
  double c[3] = {0, 0, 0}, n[3] = {0, 0, 0};
  if (D == 2) {
    for (int d = 0; d < 2; d++) c[d] = 0.5 * (m->x(v[0], d) + m->x(v[1], d));
    n[0] = m->x(v[1], 1) - m->x(v[0], 1);
    n[1] = -(m->x(v[1], 0) - m->x(v[0], 0));
  } else {
    face_normal_centroid3(m, v, n, c);
  }
  dir = 0;
  for (int d = 1; d < D; d++)
    if (std::abs(n[d]) > std::abs(n[dir])) dir = d;
  int k = 0;
  for (int d = 0; d < D; d++)
    if (d != dir) tang[k++] = c[d];
  if (D == 2) tang[1] = 0;
}

// compareOrientation, src/geo.cpp:2078-2225 (3-D only; 2-D returns 1 there).
int compare_orientation(const fhost_mesh* m, int ic1, int f1, int ic2, int f2) {
  if (m->nDims == 2) return 0;
  int t1[4], t2[4];
  for (int i = 0; i < 4; i++) {
    t1[i] = m->c2v[(size_t)ic1 * 8 + kHexOrientVerts[f1][i]];
    t2[i] = m->c2v[(size_t)ic2 * 8 + kHexOrientVerts[f2][i]];
  }
  for (int i = 0; i < 4; i++)
    if (t1[i] == t2[0]) return i;
  if (periodic3d(m, t1, t2)) {
    double c1[3] = {0, 0, 0}, c2[3] = {0, 0, 0};
    for (int i = 0; i < 4; i++)
      for (int d = 0; d < 3; d++) {
        c1[d] += m->x(t1[i], d);
        c2[d] += m->x(t2[i], d);
      }
    double fd[3], mag = 0;
    for (int d = 0; d < 3; d++) {
      c1[d] /= 4;
      c2[d] /= 4;
      fd[d] = c2[d] - c1[d];
      mag += fd[d] * fd[d];
    }
    mag = std::sqrt(mag);
    for (int d = 0; d < 3; d++) fd[d] /= mag;
    for (int i = 0; i < 4; i++) {
      double pd[3], pm = 0;
      for (int d = 0; d < 3; d++) {
        pd[d] = m->x(t2[0], d) - m->x(t1[i], d);
        pm += pd[d] * pd[d];
      }
      pm = std::sqrt(pm);
      double dot = 0;
      for (int d = 0; d < 3; d++) dot += fd[d] * (pd[d] / pm);
      if (std::abs(1 - std::abs(dot)) < m->ptol) return i;
    }
    throw Error("Unable to orient periodic faces using simple algorithm.");
  }
  throw Error("Internal faces improperly matched.");
}

int find_local_face(const fhost_mesh* m, int ic, int ff) {
  for (int j = 0; j < m->nFacesPerEle; j++)
    if (m->c2f[(size_t)ic * m->nFacesPerEle + j] == ff) return j;
  return -1;
}

void process(fhost_mesh* m) {
  const int D = m->nDims, NF = m->nFacesPerEle, NV = m->nvFace;
  const size_t nRows = (size_t)m->nEles * NF;
  std::vector<int> iF(nRows);
  m->f2v.clear();

  // processConn2D/3D: unique faces in first-occurrence order.
  if (D == 2) {
    std::unordered_map<uint64_t, int> map;
    map.reserve(nRows);
    for (int e = 0; e < m->nEles; e++)
      for (int ie = 0; ie < 4; ie++) {
        int a = m->c2v[(size_t)e * 4 + ie], b = m->c2v[(size_t)e * 4 + (ie + 1) % 4];
        if (a == b) throw Error("collapsed quad edges are not supported");
        if (a > b) std::swap(a, b);
        uint64_t key = (uint64_t)a * (uint64_t)m->nVerts + (uint64_t)b;
        auto it = map.find(key);
        int id;
        if (it == map.end()) {
          id = (int)m->f2v.size();
          map.emplace(key, id);
          m->f2v.push_back({a, b, -1, -1});
        } else {
          id = it->second;
        }
        iF[(size_t)e * 4 + ie] = id;
      }
  } else {
    std::unordered_map<std::array<int, 4>, int, ArrHash> map;
    map.reserve(nRows);
    for (int e = 0; e < m->nEles; e++)
      for (int f = 0; f < 6; f++) {
        std::array<int, 4> fv;
        for (int i = 0; i < 4; i++) fv[i] = m->c2v[(size_t)e * 8 + kHexFaceVerts[f][i]];
        std::sort(fv.begin(), fv.end());
        if (fv[0] == fv[1] || fv[1] == fv[2] || fv[2] == fv[3]) throw Error("collapsed hex faces are not supported");
        auto it = map.find(fv);
        int id;
        if (it == map.end()) {
          id = (int)m->f2v.size();
          map.emplace(fv, id);
          m->f2v.push_back(fv);
        } else {
          id = it->second;
        }
        iF[(size_t)e * 6 + f] = id;
      }
  }
  m->nFaces = (int)m->f2v.size();

  std::vector<int> count(m->nFaces, 0);
  for (size_t r = 0; r < nRows; r++) count[iF[r]]++;
  m->faceType.assign(m->nFaces, 0);
  m->intFaces.clear();
  m->bndFaces.clear();
  for (int f = 0; f < m->nFaces; f++) {
    if (count[f] > 2) throw Error("More than 2 cells for face " + std::to_string(f));
    if (count[f] == 2) m->intFaces.push_back(f);
    else {
      m->bndFaces.push_back(f);
      m->faceType[f] = 1;
    }
  }

  // Boundary-condition matching: first boundary (in bcList order) that
  // holds every vertex of the face (src/geo.cpp:275-290, 477-498).
  const int nB = (int)m->bcList.size();
  m->bndBc.assign(m->bndFaces.size(), -1);
  for (size_t i = 0; i < m->bndFaces.size(); i++) {
    const int* fv = m->f2v[m->bndFaces[i]].data();
    for (int b = 0; b < nB; b++) {
      const std::vector<int>& bp = m->bndPts[b];
      bool on = true;
      for (int j = 0; j < NV && on; j++) on = std::binary_search(bp.begin(), bp.end(), fv[j]);
      if (on) {
        m->bndBc[i] = m->bcList[b];
        break;
      }
    }
  }

  // Cell-to-face / face-to-cell.
  m->c2f.assign(nRows, -1);
  m->f2c.assign((size_t)m->nFaces * 2, -1);
  for (int ic = 0; ic < m->nEles; ic++)
    for (int j = 0; j < NF; j++) {
      int ff = iF[(size_t)ic * NF + j];
      m->c2f[(size_t)ic * NF + j] = ff;
      if (m->f2c[(size_t)ff * 2] == -1) m->f2c[(size_t)ff * 2] = ic;
      else m->f2c[(size_t)ff * 2 + 1] = ic;
    }

  // processPeriodicBoundaries (src/geo.cpp:1764-1852).
  std::vector<int> iPer;
  for (size_t i = 0; i < m->bndFaces.size(); i++)
    if (m->bndBc[i] == fr::BC_PERIODIC) iPer.push_back((int)i);
  std::vector<char> gone(m->bndFaces.size(), 0);
  std::vector<char> isPeriodicFace(m->nFaces, 0);
  if (!iPer.empty()) {
    if (iPer.size() % 2 != 0) throw Error("Expecting even number of periodic faces; have odd number.");
    const double q = (m->ptol > 0) ? m->ptol : 1e-12;
    std::unordered_map<PKey, std::vector<int>, PKeyHash> buckets;
    std::vector<PKey> keys(iPer.size());
    for (size_t a = 0; a < iPer.size(); a++) {
      int dir;
      double t[2];
      periodic_key(m, m->bndFaces[iPer[a]], dir, t);
      PKey k{dir, {(long long)std::floor(t[0] / q), (long long)std::floor(t[1] / q)}};
      keys[a] = k;
      buckets[k].push_back((int)a);
    }
    const int span1 = (D == 3) ? 1 : 0;
    int nUnmatched = 0;
    for (size_t a = 0; a < iPer.size(); a++) {
      int i = iPer[a];
      if (gone[i]) continue;
      int best = -1;
      for (int d0 = -1; d0 <= 1; d0++)
        for (int d1 = -span1; d1 <= span1; d1++) {
          PKey k = keys[a];
          k.t[0] += d0;
          k.t[1] += d1;
          auto it = buckets.find(k);
          if (it == buckets.end()) continue;
          for (int b : it->second) {
            int j = iPer[b];
            if (j == i || gone[j]) continue;
            if (best != -1 && j >= best) continue;
            if (periodic_match(m, m->bndFaces[i], m->bndFaces[j])) best = j;
          }
        }
      if (best == -1) {
        // Faces whose partner is not axis-shifted: scan as the reference does.
        for (int j : iPer) {
          if (j == i || gone[j]) continue;
          if (periodic_match(m, m->bndFaces[i], m->bndFaces[j])) {
            best = j;
            break;
          }
        }
      }
      if (best == -1) {
        nUnmatched++;
        continue;
      }
      int j = best;
      if (i > j) throw Error("periodic pairing out of order");
      int bi = m->bndFaces[i], bj = m->bndFaces[j];
      m->intFaces.push_back(bi);
      m->faceType[bi] = 0;
      m->faceType[bj] = 0;
      m->f2c[(size_t)bi * 2 + 1] = m->f2c[(size_t)bj * 2];
      m->f2c[(size_t)bj * 2] = -1;
      int ic = m->f2c[(size_t)bi * 2 + 1];
      int fID = find_local_face(m, ic, bj);
      m->c2f[(size_t)ic * NF + fID] = bi;
      gone[i] = gone[j] = 1;
      isPeriodicFace[bi] = 1;
    }
    if (nUnmatched > 0) throw Error("Unmatched periodic faces exist.");
  }
  std::vector<int> bnd2, bc2;
  for (size_t i = 0; i < m->bndFaces.size(); i++)
    if (!gone[i]) {
      bnd2.push_back(m->bndFaces[i]);
      bc2.push_back(m->bndBc[i]);
    }
  m->bndFaces.swap(bnd2);
  m->bndBc.swap(bc2);
  for (size_t i = 0; i < m->bndBc.size(); i++)
    if (m->bndBc[i] < 0) throw Error("boundary face " + std::to_string(m->bndFaces[i]) + " matches no boundary");

  // setupElesFaces (src/geo.cpp:854-930): interior faces, then boundaries.
  m->faces.clear();
  for (int ff : m->intFaces) {
    int ic1 = m->f2c[(size_t)ff * 2], ic2 = m->f2c[(size_t)ff * 2 + 1];
    if (ic2 == -1) throw Error("Interior face does not have a right element assigned.");
    int fid1 = find_local_face(m, ic1, ff), fid2 = find_local_face(m, ic2, ff);
    int rr = compare_orientation(m, ic1, fid1, ic2, fid2);
    m->faces.push_back(FaceRec{ic1, fid1, ic2, fid2, rr, -1, ff, (int)isPeriodicFace[ff]});
  }
  for (size_t i = 0; i < m->bndFaces.size(); i++) {
    int ff = m->bndFaces[i];
    int ic = m->f2c[(size_t)ff * 2];
    if (m->f2c[(size_t)ff * 2 + 1] != -1) throw Error("Boundary face has a right element assigned.");
    int fid1 = find_local_face(m, ic, ff);
    m->faces.push_back(FaceRec{ic, fid1, -1, -1, 0, m->bndBc[i], ff, 0});
  }
  m->nIntFaces = (int)m->intFaces.size();
  m->nBndFaces = (int)m->bndFaces.size();
  m->processed = true;
}

// ---- high-order geometry nodes ----------------------------------------------
//
// Gmsh numbers the nodes of a Lagrange quadrangle or hexahedron of n points
// per edge as: its corners; the n-2 nodes inside each edge, walked from the
// edge's first corner (hexahedron edges as corner pairs {0,1} {0,3} {0,4}
// {1,2} {1,5} {2,3} {2,6} {3,7} {4,5} {4,7} {5,6} {6,7}); the nodes inside
// each face, numbered as a quadrangle with n-2 points per edge whose corners
// follow the face's corner list ({0,3,2,1} {0,1,5,4} {0,4,7,3} {1,2,6,5}
// {2,3,7,6} {4,5,6,7}); then the interior, a hexahedron with n-2 points per
// edge numbered the same way.  FlurryPP's shape functions use this order
// (src/funcs.cpp:145-468).  Positions are lattice points (i, j, k), corner 0
// at the origin and corner 6 at (n-1, n-1, n-1).
using Lat = std::array<int, 3>;

Lat lat_step(const Lat& a, const Lat& b, int num, int den) {  // a + (b - a) * num / den
  Lat r;
  for (int d = 0; d < 3; d++) r[d] = a[d] + (b[d] - a[d]) * num / den;
  return r;
}

// Quadrangle with corners c[0..3] (lattice points) and n points per edge.
void gmsh_quad(const Lat c[4], int n, std::vector<Lat>& out) {
  if (n <= 0) return;
  if (n == 1) {
    out.push_back(c[0]);
    return;
  }
  for (int k = 0; k < 4; k++) out.push_back(c[k]);
  for (int k = 0; k < 4; k++)
    for (int j = 1; j < n - 1; j++) out.push_back(lat_step(c[k], c[(k + 1) % 4], j, n - 1));
  // the inner quadrangle: every corner one step along both of its edges
  Lat in[4];
  for (int k = 0; k < 4; k++) {
    const Lat a = lat_step(c[k], c[(k + 1) % 4], 1, n - 1), b = lat_step(c[k], c[(k + 3) % 4], 1, n - 1);
    for (int d = 0; d < 3; d++) in[k][d] = a[d] + b[d] - c[k][d];
  }
  gmsh_quad(in, n - 2, out);
}

void gmsh_hex(const Lat c[8], int n, std::vector<Lat>& out) {
  static const int edges[12][2] = {{0, 1}, {0, 3}, {0, 4}, {1, 2}, {1, 5}, {2, 3},
                                   {2, 6}, {3, 7}, {4, 5}, {4, 7}, {5, 6}, {6, 7}};
  static const int faces[6][4] = {{0, 3, 2, 1}, {0, 1, 5, 4}, {0, 4, 7, 3}, {1, 2, 6, 5}, {2, 3, 7, 6}, {4, 5, 6, 7}};
  if (n <= 0) return;
  if (n == 1) {
    out.push_back(c[0]);
    return;
  }
  for (int k = 0; k < 8; k++) out.push_back(c[k]);
  for (auto& e : edges)
    for (int j = 1; j < n - 1; j++) out.push_back(lat_step(c[e[0]], c[e[1]], j, n - 1));
  if (n > 2) {
    for (auto& f : faces) {
      Lat fc[4], in[4];
      for (int k = 0; k < 4; k++) fc[k] = c[f[k]];
      for (int k = 0; k < 4; k++) {
        const Lat a = lat_step(fc[k], fc[(k + 1) % 4], 1, n - 1), b = lat_step(fc[k], fc[(k + 3) % 4], 1, n - 1);
        for (int d = 0; d < 3; d++) in[k][d] = a[d] + b[d] - fc[k][d];
      }
      gmsh_quad(in, n - 2, out);
    }
    // the inner hexahedron: every corner one step inwards along its three edges
    Lat in[8];
    for (int k = 0; k < 8; k++)
      for (int d = 0; d < 3; d++) in[k][d] = c[k][d] + (2 * c[k][d] < c[0][d] + c[6][d] ? 1 : -1);
    gmsh_hex(in, n - 2, out);
  }
}

std::vector<Lat> gmsh_lattice(int D, int n) {
  std::vector<Lat> out;
  const int h = n - 1;
  if (D == 2) {
    const Lat c[4] = {{0, 0, 0}, {h, 0, 0}, {h, h, 0}, {0, h, 0}};
    gmsh_quad(c, n, out);
  } else {
    const Lat c[8] = {{0, 0, 0}, {h, 0, 0}, {h, h, 0}, {0, h, 0}, {0, 0, h}, {h, 0, h}, {h, h, h}, {0, h, h}};
    gmsh_hex(c, n, out);
  }
  return out;
}

// Shape functions and their reference derivatives ds[D][nShape] of the
// mesh's elements at p: Lagrange products on the equidistant grid
// (src/funcs.cpp:145-468) or the 20-node serendipity hexahedron
// (src/funcs.cpp:205-240; nodes in Gmsh order: corners, then the mid-points
// of the edges in the edge order above).
void shape_eval(const fhost_mesh* m, const fr::Pt& p, double* s, double* ds) {
  const int D = m->nDims, NS = m->nShape ? m->nShape : m->nNodes;
  const double q[3] = {p.x, p.y, p.z};
  if (m->serendipity) {
    static const int sg[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                 {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
    static const int edges[12][2] = {{0, 1}, {0, 3}, {0, 4}, {1, 2}, {1, 5}, {2, 3},
                                     {2, 6}, {3, 7}, {4, 5}, {4, 7}, {5, 6}, {6, 7}};
    for (int n = 0; n < 20; n++) {
      double c[3];
      for (int d = 0; d < 3; d++)
        c[d] = n < 8 ? sg[n][d] : 0.5 * (sg[edges[n - 8][0]][d] + sg[edges[n - 8][1]][d]);
      if (n < 8) {
        // (1 + x c)(1 + y c)(1 + z c)(x cx + y cy + z cz - 2) / 8
        double f[3], sum = -2;
        for (int d = 0; d < 3; d++) f[d] = 1 + q[d] * c[d], sum += q[d] * c[d];
        if (s) s[n] = 0.125 * f[0] * f[1] * f[2] * sum;
        if (ds)
          for (int d = 0; d < 3; d++) {
            const int a = (d + 1) % 3, b = (d + 2) % 3;
            ds[d * 20 + n] = 0.125 * f[a] * f[b] * (c[d] * sum + f[d] * c[d]);
          }
      } else {
        // the coordinate with c = 0: (1 - x^2)(1 + y cy)(1 + z cz) / 4
        const int z0 = c[0] == 0 ? 0 : (c[1] == 0 ? 1 : 2), a = (z0 + 1) % 3, b = (z0 + 2) % 3;
        const double g = 1 - q[z0] * q[z0], fa = 1 + q[a] * c[a], fb = 1 + q[b] * c[b];
        if (s) s[n] = 0.25 * g * fa * fb;
        if (ds) {
          ds[z0 * 20 + n] = 0.25 * (-2 * q[z0]) * fa * fb;
          ds[a * 20 + n] = 0.25 * g * c[a] * fb;
          ds[b * 20 + n] = 0.25 * g * fa * c[b];
        }
      }
    }
    return;
  }
  const int n1 = m->nShape ? m->nSide : 2;
  std::vector<double> xl(n1);
  for (int i = 0; i < n1; i++) xl[i] = -1. + i * (2. / (n1 - 1));
  double L[3][16], dL[3][16];
  for (int d = 0; d < D; d++)
    for (int i = 0; i < n1; i++) L[d][i] = fr::lagrange(xl, q[d], i), dL[d][i] = fr::dlagrange(xl, q[d], i);
  static const int lin[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  for (int n = 0; n < NS; n++) {
    const int* ijk = m->nShape ? m->lattice[n].data() : lin[n];
    double v = 1;
    for (int d = 0; d < D; d++) v *= L[d][ijk[d]];
    if (s) s[n] = v;
    if (ds)
      for (int r = 0; r < D; r++) {
        double g = 1;
        for (int d = 0; d < D; d++) g *= (d == r) ? dL[d][ijk[d]] : L[d][ijk[d]];
        ds[r * NS + n] = g;
      }
  }
}

// Attach geometry nodes (Gmsh order) to a mesh: nShape per element.
void set_shape_nodes(fhost_mesh* m, int nShape, const int* c2n) {
  const int D = m->nDims;
  int side = 0;
  bool ser = false;
  if (D == 3 && nShape == 20) {
    ser = true;
  } else {
    for (int n = 2; n <= 6; n++)
      if ((D == 2 ? n * n : n * n * n) == nShape) side = n;
    if (!side)
      throw Error(D == 2 ? "For Lagrange quad of order N, must have (N+1)^2 shape points."
                         : "For Lagrange hex of order N, must have (N+1)^3 shape points (or the 20-node serendipity hex).");
  }
  for (size_t e = 0; e < (size_t)m->nEles; e++)
    for (int n = 0; n < nShape; n++) {
      const int v = c2n[e * nShape + n];
      if (v < 0 || v >= m->nVerts) throw Error("shape node out of range");
      if (n < m->nNodes && v != m->c2v[e * m->nNodes + n])
        throw Error("the first shape nodes of an element must be its vertices");
    }
  m->nShape = nShape;
  m->serendipity = ser;
  m->nSide = ser ? 3 : side;
  m->lattice = ser ? std::vector<Lat>() : gmsh_lattice(D, side);
  m->c2n.assign(c2n, c2n + (size_t)m->nEles * nShape);
}

uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

template <class F>
int guard(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

}  // namespace

extern "C" {

const char* fhost_last_error(void) { return g_err.c_str(); }

int fhost_bc_from_name(const char* name) {
  static const std::map<std::string, int> m = {
      {"none", -1}, {"fluid", -1}, {"periodic", 0}, {"char", 1}, {"sup_in", 2}, {"sup_out", 3},
      {"sub_in", 4}, {"sub_out", 5}, {"slip_wall", 8}, {"isothermal_noslip", 9},
      {"isothermal_noslip_moving", 13}, {"adiabatic_noslip", 10}, {"overset", 11}, {"symmetry", 12}};
  std::string s(name ? name : "");
  std::transform(s.begin(), s.end(), s.begin(), ::tolower);
  auto it = m.find(s);
  return it == m.end() ? -2 : it->second;
}

int fhost_sizes(int nDims, int order, int* nSpts, int* nFpts, int* nFptsFace) {
  return guard([&] {
    int S, F, Ff;
    fr::sizes(nDims, order, S, F, Ff);
    if (nSpts) *nSpts = S;
    if (nFpts) *nFpts = F;
    if (nFptsFace) *nFptsFace = Ff;
  });
}

int fhost_pts1d(int ptsType, int order, double* out) {
  return guard([&] {
    auto p = fr::pts1d(ptsType, order);
    std::copy(p.begin(), p.end(), out);
  });
}

int fhost_weights1d(int order, double* out) {
  return guard([&] {
    auto w = fr::weights1d(order);
    std::copy(w.begin(), w.end(), out);
  });
}

int fhost_ref_points(int nDims, int order, int ptsType, double* loc_spts, double* loc_fpts, double* tnorm) {
  return guard([&] {
    std::vector<fr::Pt> s, f;
    fr::ref_points(nDims, order, ptsType, s, f);
    int S, F, Ff;
    fr::sizes(nDims, order, S, F, Ff);
    for (int i = 0; i < S && loc_spts; i++) {
      double v[3] = {s[i].x, s[i].y, s[i].z};
      for (int d = 0; d < nDims; d++) loc_spts[i * nDims + d] = v[d];
    }
    for (int i = 0; i < F; i++) {
      double v[3] = {f[i].x, f[i].y, f[i].z};
      if (loc_fpts)
        for (int d = 0; d < nDims; d++) loc_fpts[i * nDims + d] = v[d];
      if (tnorm) fr::face_tnorm(nDims, i / Ff, &tnorm[i * nDims]);
    }
  });
}

int fhost_operators(int nDims, int order, int ptsType, int vcjh, double* spts_to_fpts, double* grad_spts,
                    double* extrap_fn, double* correction, double* correct_u, double* w_spts, double* w_face) {
  return guard([&] {
    fr::Operators op;
    fr::operators(nDims, order, ptsType, vcjh, op);
    auto cp = [](const std::vector<double>& v, double* o) {
      if (o) std::copy(v.begin(), v.end(), o);
    };
    cp(op.spts_to_fpts, spts_to_fpts);
    cp(op.grad_spts, grad_spts);
    cp(op.extrap_fn, extrap_fn);
    cp(op.correction, correction);
    cp(op.correct_u, correct_u);
    cp(op.w_spts, w_spts);
    cp(op.w_face, w_face);
  });
}

int fhost_pmg_operators(int nDims, int order, int ptsType, double* prolong, double* restrict_op) {
  return guard([&] {
    if (nDims != 2 && nDims != 3) throw Error("nDims must be 2 or 3");
    if (order < 0 || order > 9) throw Error("P-multigrid operators need 0 <= order <= 9");
    std::vector<double> pro, res;
    fr::pmg_operators(nDims, order, ptsType, pro, res);
    if (prolong) std::copy(pro.begin(), pro.end(), prolong);
    if (restrict_op) std::copy(res.begin(), res.end(), restrict_op);
  });
}

int fhost_mesh_cartesian(int nDims, int nx, int ny, int nz, const double* lo, const double* hi,
                         const char* const* bc_names, fhost_mesh** out) {
  return guard([&] {
    if (nDims != 2 && nDims != 3) throw Error("nDims must be 2 or 3");
    if (nx < 1 || ny < 1 || (nDims == 3 && nz < 1)) throw Error("mesh dimensions must be positive");
    if (nDims == 2) nz = 1;
    auto m = new fhost_mesh();
    m->nDims = nDims;
    set_shape(m);
    double xmin = lo[0], xmax = hi[0], ymin = lo[1], ymax = hi[1];
    double zmin = nDims == 3 ? lo[2] : 0, zmax = nDims == 3 ? hi[2] : 0;
    double dx = (xmax - xmin) / nx, dy = (ymax - ymin) / ny, dz = (zmax - zmin) / nz;
    m->pdx = xmax - xmin;
    m->pdy = ymax - ymin;
    m->pdz = zmax - zmin;
    m->nEles = nx * ny * nz;
    if (nDims == 2) {
      m->nVerts = (nx + 1) * (ny + 1);
      m->xv.resize((size_t)m->nVerts * 2);
      int nv = 0;
      for (int i = 0; i < ny + 1; i++)
        for (int j = 0; j < nx + 1; j++) {
          m->xv[nv * 2 + 0] = xmin + j * dx;
          m->xv[nv * 2 + 1] = ymin + i * dy;
          nv++;
        }
      for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
          int c[4] = {j * (nx + 1) + i, j * (nx + 1) + i + 1, (j + 1) * (nx + 1) + i + 1, (j + 1) * (nx + 1) + i};
          m->c2v.insert(m->c2v.end(), c, c + 4);
        }
    } else {
      m->nVerts = (nx + 1) * (ny + 1) * (nz + 1);
      m->xv.resize((size_t)m->nVerts * 3);
      size_t nv = 0;
      for (int k = 0; k < nz + 1; k++)
        for (int j = 0; j < ny + 1; j++)
          for (int i = 0; i < nx + 1; i++) {
            m->xv[nv * 3 + 0] = xmin + i * dx;
            m->xv[nv * 3 + 1] = ymin + j * dy;
            m->xv[nv * 3 + 2] = zmin + k * dz;
            nv++;
          }
      m->c2v.reserve((size_t)m->nEles * 8);
      for (int k = 0; k < nz; k++)
        for (int i = 0; i < nx; i++)
          for (int j = 0; j < ny; j++) {
            int c[8] = {i + (nx + 1) * (j + (ny + 1) * k),       i + (nx + 1) * (j + (ny + 1) * k) + 1,
                        i + (nx + 1) * (j + 1 + (ny + 1) * k) + 1, i + (nx + 1) * (j + 1 + (ny + 1) * k),
                        i + (nx + 1) * (j + (ny + 1) * (k + 1)),   i + (nx + 1) * (j + (ny + 1) * (k + 1)) + 1,
                        i + (nx + 1) * (j + 1 + (ny + 1) * (k + 1)) + 1, i + (nx + 1) * (j + 1 + (ny + 1) * (k + 1))};
            m->c2v.insert(m->c2v.end(), c, c + 8);
          }
    }
    // Boundaries (src/geo.cpp:1588-1760): sorted unique BC numbers; each
    // boundary's vertex list also holds vertex 0, since FlurryPP sorts a
    // zero-initialised buffer that is larger than the list.
    int nSide = (nDims == 2) ? 4 : 6;
    int bcs[6];
    for (int s = 0; s < nSide; s++) {
      bcs[s] = fhost_bc_from_name(bc_names[s]);
      if (bcs[s] < -1) throw Error(std::string("Unrecognized boundary condition: ") + bc_names[s]);
    }
    std::vector<int> list(bcs, bcs + nSide);
    std::sort(list.begin(), list.end());
    list.erase(std::unique(list.begin(), list.end()), list.end());
    m->bcList = list;
    std::vector<std::set<int>> pts(list.size());
    auto slot = [&](int bc) { return (int)(std::lower_bound(list.begin(), list.end(), bc) - list.begin()); };
    for (auto& s : pts) s.insert(0);
    if (nDims == 2) {
      // bottom, top, left, right (names: 0 bottom, 1 right, 2 top, 3 left)
      for (int ix = 0; ix < nx; ix++) {
        pts[slot(bcs[0])].insert(ix);
        pts[slot(bcs[0])].insert(ix + 1);
      }
      for (int ix = 0; ix < nx; ix++) {
        pts[slot(bcs[2])].insert((nx + 1) * ny + ix + 1);
        pts[slot(bcs[2])].insert((nx + 1) * ny + ix);
      }
      for (int iy = 0; iy < ny; iy++) {
        pts[slot(bcs[3])].insert((iy + 1) * (nx + 1));
        pts[slot(bcs[3])].insert(iy * (nx + 1));
      }
      for (int iy = 0; iy < ny; iy++) {
        pts[slot(bcs[1])].insert(iy * (nx + 1) + nx);
        pts[slot(bcs[1])].insert((iy + 1) * (nx + 1) + nx);
      }
    } else {
      int NX = nx + 1, NXY = (nx + 1) * (ny + 1);
      auto add4 = [&](int s, int a, int b, int c, int d) {
        pts[s].insert(a);
        pts[s].insert(b);
        pts[s].insert(c);
        pts[s].insert(d);
      };
      for (int ix = 0; ix < nx; ix++)
        for (int iy = 0; iy < ny; iy++)
          add4(slot(bcs[0]), iy * NX + ix, iy * NX + ix + 1, (iy + 1) * NX + ix + 1, (iy + 1) * NX + ix);
      for (int ix = 0; ix < nx; ix++)
        for (int iy = 0; iy < ny; iy++)
          add4(slot(bcs[2]), NXY * nz + iy * NX + ix, NXY * nz + iy * NX + ix + 1, NXY * nz + (iy + 1) * NX + ix + 1,
               NXY * nz + (iy + 1) * NX + ix);
      for (int iz = 0; iz < nz; iz++)
        for (int iy = 0; iy < ny; iy++)
          add4(slot(bcs[3]), iz * NXY + iy * NX, (iz + 1) * NXY + iy * NX, (iz + 1) * NXY + (iy + 1) * NX,
               iz * NXY + (iy + 1) * NX);
      for (int iz = 0; iz < nz; iz++)
        for (int iy = 0; iy < ny; iy++)
          add4(slot(bcs[1]), iz * NXY + iy * NX + nx, (iz + 1) * NXY + iy * NX + nx, (iz + 1) * NXY + (iy + 1) * NX + nx,
               iz * NXY + (iy + 1) * NX + nx);
      for (int iz = 0; iz < nz; iz++)
        for (int ix = 0; ix < nx; ix++)
          add4(slot(bcs[5]), iz * NXY + ix, (iz + 1) * NXY + ix, (iz + 1) * NXY + ix + 1, iz * NXY + ix + 1);
      for (int iz = 0; iz < nz; iz++)
        for (int ix = 0; ix < nx; ix++)
          add4(slot(bcs[4]), iz * NXY + ny * NX + ix, (iz + 1) * NXY + ny * NX + ix, (iz + 1) * NXY + ny * NX + ix + 1,
               iz * NXY + ny * NX + ix + 1);
    }
    for (auto& s : pts) m->bndPts.emplace_back(s.begin(), s.end());
    *out = m;
  });
}

int fhost_mesh_from_arrays(int nDims, int nVerts, const double* xv, int nEles, const int* c2v, int nBounds,
                           const int* bc_list, const int* bnd_ptr, const int* bnd_pts, fhost_mesh** out) {
  return guard([&] {
    if (nDims != 2 && nDims != 3) throw Error("nDims must be 2 or 3");
    auto m = new fhost_mesh();
    m->nDims = nDims;
    set_shape(m);
    m->nVerts = nVerts;
    m->nEles = nEles;
    m->xv.assign(xv, xv + (size_t)nVerts * nDims);
    m->c2v.assign(c2v, c2v + (size_t)nEles * m->nNodes);
    for (size_t i = 0; i < m->c2v.size(); i++)
      if (m->c2v[i] < 0 || m->c2v[i] >= nVerts) {
        delete m;
        throw Error("c2v entry out of range");
      }
    for (int b = 0; b < nBounds; b++) {
      m->bcList.push_back(bc_list[b]);
      std::vector<int> p(bnd_pts + bnd_ptr[b], bnd_pts + bnd_ptr[b + 1]);
      std::sort(p.begin(), p.end());
      p.erase(std::unique(p.begin(), p.end()), p.end());
      m->bndPts.push_back(p);
    }
    *out = m;
  });
}

int fhost_mesh_set_shape_nodes(fhost_mesh* m, int nShape, const int* c2n) {
  return guard([&] {
    if (!m || !c2n) throw Error("null argument");
    set_shape_nodes(m, nShape, c2n);
  });
}

int fhost_mesh_get_shape_nodes(const fhost_mesh* m, int* nShape, int* c2n) {
  return guard([&] {
    if (!m) throw Error("null mesh");
    const int ns = m->nShape ? m->nShape : m->nNodes;
    if (nShape) *nShape = ns;
    if (c2n)
      for (size_t e = 0; e < (size_t)m->nEles; e++)
        for (int n = 0; n < ns; n++) c2n[e * ns + n] = m->node(e, n);
  });
}

int fhost_mesh_read_gmsh(const char* path, int nMap, const char* const* names, const char* const* bcs,
                         fhost_mesh** out) {
  return guard([&] {
    std::ifstream f(path);
    if (!f.is_open()) throw Error("Unable to open mesh file.");
    std::map<std::string, std::string> bmap;
    for (int i = 0; i < nMap; i++) {
      std::string a(names[i]), b(bcs[i]);
      std::transform(a.begin(), a.end(), a.begin(), ::tolower);
      std::transform(b.begin(), b.end(), b.begin(), ::tolower);
      bmap[a] = b;
    }
    std::string line;
    auto seek = [&](const char* tag) {
      f.clear();
      f.seekg(0);
      while (std::getline(f, line))
        if (line.find(tag) != std::string::npos) return;
      throw Error(std::string(tag) + " tag not found in Gmsh file!");
    };
    // $PhysicalNames (src/geo.cpp:1039-1095)
    seek("$PhysicalNames");
    int nG;
    f >> nG;
    std::getline(f, line);
    int nDims = 0;
    std::map<int, int> bcIdMap;
    std::vector<int> bcList;
    for (int i = 0; i < nG; i++) {
      std::getline(f, line);
      std::stringstream ss(line);
      int dim, id;
      std::string s;
      ss >> dim >> id >> s;
      s.erase(std::remove(s.begin(), s.end(), '"'), s.end());
      std::transform(s.begin(), s.end(), s.begin(), ::tolower);
      if (!bmap.count(s)) throw Error("Unrecognized mesh boundary: \"" + s + "\"");
      std::string bc = bmap[s];
      int num = fhost_bc_from_name(bc.c_str());
      if (num < -1) throw Error("Unrecognized boundary condition: \"" + bc + "\"");
      if (bc == "fluid") {
        nDims = dim;
        bcIdMap[id] = -1;
      } else {
        bcIdMap[id] = (int)bcList.size();
        bcList.push_back(num);
      }
    }
    if (nDims != 2 && nDims != 3) throw Error("no 'fluid' physical name in the Gmsh file");
    seek("$Nodes");
    int nV = -1;
    f >> nV;
    if (!f || nV < 0) throw Error("malformed $Nodes count");
    std::getline(f, line);
    std::vector<double> xv((size_t)nV * nDims);
    for (int i = 0; i < nV; i++) {
      int iv;
      double z;
      f >> iv >> xv[(size_t)i * nDims] >> xv[(size_t)i * nDims + 1];
      if (nDims == 3) f >> xv[(size_t)i * nDims + 2];
      else f >> z;
      if (!f) throw Error("malformed $Nodes entry " + std::to_string(i + 1));
      std::getline(f, line);
    }
    seek("$Elements");
    int nE = -1;
    f >> nE;
    if (!f || nE < 0) throw Error("malformed $Elements count");
    std::getline(f, line);
    auto vertex = [&](int k) {
      int v = 0;
      f >> v;
      if (!f) throw Error("malformed $Elements entry " + std::to_string(k + 1));
      if (v < 1 || v > nV) throw Error("$Elements entry " + std::to_string(k + 1) + " names node " +
                                       std::to_string(v) + ", outside 1.." + std::to_string(nV));
      return v - 1;
    };
    // element types (src/geo.cpp:1163-1344): Lagrange quadrangles of 4, 9,
    // 16, 25, 36 nodes (Gmsh 3, 10, 36, 37, 38) and hexahedra of 8, 27, 64,
    // 125, 216 nodes (5, 12, 92, 93, 94), the 20-node serendipity hexahedron
    // (17); the 8-node quadrangle (16) is read, as FlurryPP reads it, and then
    // refused, as its shape functions refuse it.  Boundary faces: lines of 2
    // to 6 nodes (1, 8, 26, 27, 28) and quadrangles (3, 16, 10, 36, 37, 38);
    // their corner vertices mark the boundary.
    auto ele_nodes = [&](int type) {
      switch (type) {
        case 3: return nDims == 2 ? 4 : 0;
        case 16: return nDims == 2 ? 8 : 0;
        case 10: return nDims == 2 ? 9 : 0;
        case 36: return nDims == 2 ? 16 : 0;
        case 37: return nDims == 2 ? 25 : 0;
        case 38: return nDims == 2 ? 36 : 0;
        case 5: return nDims == 3 ? 8 : 0;
        case 17: return nDims == 3 ? 20 : 0;
        case 12: return nDims == 3 ? 27 : 0;
        case 92: return nDims == 3 ? 64 : 0;
        case 93: return nDims == 3 ? 125 : 0;
        case 94: return nDims == 3 ? 216 : 0;
        default: return 0;
      }
    };
    auto face_nodes = [&](int type, int& corners) {
      corners = nDims == 2 ? 2 : 4;
      switch (type) {
        case 1: return nDims == 2 ? 2 : 0;
        case 8: return nDims == 2 ? 3 : 0;
        case 26: return nDims == 2 ? 4 : 0;
        case 27: return nDims == 2 ? 5 : 0;
        case 28: return nDims == 2 ? 6 : 0;
        case 3: return nDims == 3 ? 4 : 0;
        case 16: return nDims == 3 ? 8 : 0;
        case 10: return nDims == 3 ? 9 : 0;
        case 36: return nDims == 3 ? 16 : 0;
        case 37: return nDims == 3 ? 25 : 0;
        case 38: return nDims == 3 ? 36 : 0;
        default: return 0;
      }
    };
    const int nv = (nDims == 2) ? 4 : 8;
    std::vector<int> c2v, c2n;
    int nShape = -1;
    std::vector<std::set<int>> bp(bcList.size());
    for (int k = 0; k < nE; k++) {
      int id, type, nTags, tag, tmp;
      f >> id >> type >> nTags >> tag;
      for (int t = 0; t < nTags - 1; t++) f >> tmp;
      if (!f) throw Error("malformed $Elements entry " + std::to_string(k + 1));
      int b = bcIdMap.count(tag) ? bcIdMap[tag] : -2;
      if (b == -1) {
        const int nn = ele_nodes(type);
        if (!nn) throw Error("Gmsh element type " + std::to_string(type) + " not supported (quadrangles and hexahedra)");
        if (nShape >= 0 && nn != nShape) throw Error("elements of different orders in one mesh are not supported");
        nShape = nn;
        for (int i = 0; i < nn; i++) {
          const int v = vertex(k);
          c2n.push_back(v);
          if (i < nv) c2v.push_back(v);
        }
      } else if (b >= 0) {
        int corners = 0;
        const int np = face_nodes(type, corners);
        if (!np) throw Error("Boundary Element (Face) Type Not Recognized!");
        for (int i = 0; i < np; i++) {
          const int v = vertex(k);
          if (i < corners) bp[b].insert(v);
        }
      } else {
        throw Error("element tagged with an unknown physical id");
      }
      std::getline(f, line);
    }
    std::vector<int> ptr(1, 0), pts;
    for (auto& s : bp) {
      pts.insert(pts.end(), s.begin(), s.end());
      ptr.push_back((int)pts.size());
    }
    if (fhost_mesh_from_arrays(nDims, nV, xv.data(), (int)(c2v.size() / nv), c2v.data(), (int)bcList.size(),
                               bcList.data(), ptr.data(), pts.data(), out))
      throw Error(g_err);
    if (nShape > nv) {
      try {
        set_shape_nodes(*out, nShape, c2n.data());
      } catch (...) {
        fhost_mesh_destroy(*out);
        *out = nullptr;
        throw;
      }
    }
  });
}

int fhost_mesh_jitter(fhost_mesh* m, const double* h, double amp, unsigned long long seed) {
  return guard([&] {
    if (m->nShape) throw Error("jitter moves the vertices of linear meshes only (this mesh has curved elements)");
    // A vertex is interior if no boundary face (before periodic pairing)
    // touches it; find those faces with a count over all element faces.
    const int NF = m->nFacesPerEle;
    std::unordered_map<std::array<int, 4>, int, ArrHash> cnt;
    cnt.reserve((size_t)m->nEles * NF);
    for (int e = 0; e < m->nEles; e++)
      for (int f = 0; f < NF; f++) {
        std::array<int, 4> fv = {-1, -1, -1, -1};
        if (m->nDims == 2) {
          fv[0] = m->c2v[(size_t)e * 4 + f];
          fv[1] = m->c2v[(size_t)e * 4 + (f + 1) % 4];
          std::sort(fv.begin(), fv.begin() + 2);
        } else {
          for (int i = 0; i < 4; i++) fv[i] = m->c2v[(size_t)e * 8 + kHexFaceVerts[f][i]];
          std::sort(fv.begin(), fv.end());
        }
        cnt[fv]++;
      }
    std::vector<char> onB(m->nVerts, 0);
    for (auto& kv : cnt)
      if (kv.second == 1)
        for (int i = 0; i < m->nvFace; i++) onB[kv.first[i]] = 1;
    for (int v = 0; v < m->nVerts; v++) {
      if (onB[v]) continue;
      for (int d = 0; d < m->nDims; d++) {
        uint64_t r = splitmix64((uint64_t)seed ^ (uint64_t)((uint64_t)v * m->nDims + d));
        double u = (double)(r >> 11) * (1.0 / 9007199254740992.0);
        m->xv[(size_t)v * m->nDims + d] += (2.0 * u - 1.0) * amp * h[d];
      }
    }
  });
}

int fhost_mesh_set_periodic(fhost_mesh* m, double dx, double dy, double dz, double tol) {
  return guard([&] {
    m->pdx = dx;
    m->pdy = dy;
    m->pdz = dz;
    m->ptol = tol;
  });
}

int fhost_mesh_process(fhost_mesh* m) { return guard([&] { process(m); }); }

int fhost_mesh_counts(const fhost_mesh* m, int* nDims, int* nVerts, int* nEles, int* nIntFaces, int* nBndFaces) {
  if (nDims) *nDims = m->nDims;
  if (nVerts) *nVerts = m->nVerts;
  if (nEles) *nEles = m->nEles;
  if (nIntFaces) *nIntFaces = m->nIntFaces;
  if (nBndFaces) *nBndFaces = m->nBndFaces;
  return 0;
}

int fhost_mesh_get_nodes(const fhost_mesh* m, double* xv, int* c2v) {
  if (xv) std::copy(m->xv.begin(), m->xv.end(), xv);
  if (c2v) std::copy(m->c2v.begin(), m->c2v.end(), c2v);
  return 0;
}

int fhost_mesh_get_faces(const fhost_mesh* m, int* eleL, int* locL, int* eleR, int* locR, int* relRot, int* bcType,
                         int* gid, int* periodic) {
  if (!m->processed) {
    g_err = "mesh not processed";
    return 1;
  }
  for (size_t i = 0; i < m->faces.size(); i++) {
    const FaceRec& f = m->faces[i];
    if (eleL) eleL[i] = f.eL;
    if (locL) locL[i] = f.locL;
    if (eleR) eleR[i] = f.eR;
    if (locR) locR[i] = f.locR;
    if (relRot) relRot[i] = f.relRot;
    if (bcType) bcType[i] = f.bcType;
    if (gid) gid[i] = f.gid;
    if (periodic) periodic[i] = f.periodic;
  }
  return 0;
}

int fhost_mesh_geometry(const fhost_mesh* m, int order, int ptsType, double* pos_spts, double* pos_fpts,
                        double* jginv_spts, double* detjac_spts, double* norm_fpts, double* da_fpts) {
  return guard([&] {
    const int D = m->nDims, NN = m->nShape ? m->nShape : m->nNodes;
    const size_t E = (size_t)m->nEles;
    int S, F, Ff;
    fr::sizes(D, order, S, F, Ff);
    std::vector<fr::Pt> ls, lf;
    fr::ref_points(D, order, ptsType, ls, lf);
    std::vector<double> sh_s((size_t)S * NN), sh_f((size_t)F * NN), ds_s((size_t)D * S * NN), ds_f((size_t)D * F * NN);
    std::vector<double> tmp((size_t)D * NN);
    for (int s = 0; s < S; s++) {
      shape_eval(m, ls[s], &sh_s[(size_t)s * NN], tmp.data());
      for (int d = 0; d < D; d++)
        for (int n = 0; n < NN; n++) ds_s[((size_t)d * S + s) * NN + n] = tmp[d * NN + n];
    }
    for (int f = 0; f < F; f++) {
      shape_eval(m, lf[f], &sh_f[(size_t)f * NN], tmp.data());
      for (int d = 0; d < D; d++)
        for (int n = 0; n < NN; n++) ds_f[((size_t)d * F + f) * NN + n] = tmp[d * NN + n];
    }
    std::vector<double> tn((size_t)F * D);
    for (int f = 0; f < F; f++) fr::face_tnorm(D, f / Ff, &tn[(size_t)f * D]);

    std::string err;
#pragma omp parallel for schedule(static)
    for (long long e = 0; e < (long long)E; e++) {
      std::vector<std::array<double, 3>> nodes(NN);
      for (int n = 0; n < NN; n++)
        for (int d = 0; d < D; d++) nodes[n][d] = m->x(m->node((size_t)e, n), d);
      // positions: shape * nodes (setPosSptsFpts, src/solver.cpp:1248-1313)
      for (int s = 0; s < S && pos_spts; s++)
        for (int d = 0; d < D; d++) {
          double acc = 0;
          for (int n = 0; n < NN; n++) acc += sh_s[(size_t)s * NN + n] * nodes[n][d];
          pos_spts[((size_t)s * E + e) * D + d] = acc;
        }
      for (int f = 0; f < F && pos_fpts; f++)
        for (int d = 0; d < D; d++) {
          double acc = 0;
          for (int n = 0; n < NN; n++) acc += sh_f[(size_t)f * NN + n] * nodes[n][d];
          pos_fpts[((size_t)f * E + e) * D + d] = acc;
        }
      // Jacobian J(r,x) = dx_x/dxi_r and its adjugate (calcTransforms,
      // src/solver.cpp:1587-1694).
      auto jac = [&](const std::vector<double>& ds, int npts, int p, double J[3][3]) {
        for (int r = 0; r < D; r++)
          for (int x = 0; x < D; x++) {
            double acc = 0;
            for (int n = 0; n < NN; n++) acc += ds[((size_t)r * npts + p) * NN + n] * nodes[n][x];
            J[r][x] = acc;
          }
      };
      auto adj = [&](double J[3][3], double& det, double G[3][3]) {
        if (D == 2) {
          det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
          G[0][0] = J[1][1];
          G[0][1] = -J[1][0];
          G[1][0] = -J[0][1];
          G[1][1] = J[0][0];
        } else {
          double xr = J[0][0], xs = J[1][0], xt = J[2][0];
          double yr = J[0][1], ys = J[1][1], yt = J[2][1];
          double zr = J[0][2], zs = J[1][2], zt = J[2][2];
          det = xr * (ys * zt - yt * zs) - xs * (yr * zt - yt * zr) + xt * (yr * zs - ys * zr);
          G[0][0] = ys * zt - yt * zs;
          G[0][1] = xt * zs - xs * zt;
          G[0][2] = xs * yt - xt * ys;
          G[1][0] = yt * zr - yr * zt;
          G[1][1] = xr * zt - xt * zr;
          G[1][2] = xt * yr - xr * yt;
          G[2][0] = yr * zs - ys * zr;
          G[2][1] = xs * zr - xr * zs;
          G[2][2] = xr * ys - xs * yr;
        }
      };
      for (int s = 0; s < S; s++) {
        double J[3][3] = {}, G[3][3] = {}, det;
        jac(ds_s, S, s, J);
        adj(J, det, G);
        if (det < 0) {
#pragma omp critical
          err = "Negative Jacobian at solution points.";
        }
        if (detjac_spts) detjac_spts[(size_t)s * E + e] = det;
        if (jginv_spts)
          for (int r = 0; r < D; r++)
            for (int x = 0; x < D; x++) jginv_spts[(((size_t)r * S + s) * E + e) * D + x] = G[r][x];
      }
      for (int f = 0; f < F; f++) {
        double J[3][3] = {}, G[3][3] = {}, det;
        jac(ds_f, F, f, J);
        adj(J, det, G);
        double nv[3];
        for (int d1 = 0; d1 < D; d1++) {
          nv[d1] = 0.;
          for (int d2 = 0; d2 < D; d2++) nv[d1] += G[d2][d1] * tn[(size_t)f * D + d2];
        }
        double dA = 0;
        for (int d = 0; d < D; d++) dA += nv[d] * nv[d];
        dA = std::sqrt(dA);
        if (std::fabs(dA) < 1e-10) {
          dA = 0.;
          for (int d = 0; d < D; d++) nv[d] = 0;
        } else {
          for (int d = 0; d < D; d++) nv[d] /= dA;
        }
        if (da_fpts) da_fpts[(size_t)f * E + e] = dA;
        if (norm_fpts)
          for (int d = 0; d < D; d++) norm_fpts[((size_t)f * E + e) * D + d] = nv[d];
      }
    }
    if (!err.empty()) throw Error(err);
  });
}

void fhost_mesh_destroy(fhost_mesh* m) { delete m; }

}  // extern "C"
