// FR basis: 1-D point sets, quadrature weights, Lagrange/Legendre/VCJH
// polynomials and the dense operator matrices libflurry_hip applies.
//
// The abscissa and weight tables are data that has to equal FlurryPP's to the
// last bit (its Gauss abscissae are truncated to 15 decimals, with a few
// asymmetric roundings, src/points.cpp:198-392, 425-538), so they are listed
// here verbatim as numbers.  Every formula below evaluates its terms in the
// same order as FlurryPP so that the operators are bit-identical.
#include "fr_host_internal.hpp"

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

namespace fr {

namespace {

// Gauss-Legendre abscissae, orders 0..10 (src/points.cpp:198-297).
const double kLegendre[11][11] = {
  {0.0},
  {-0.577350269189626, 0.577350269189626},
  {-0.774596669241483, 0.000000000000000, 0.774596669241483},
  {-0.861136311594053, -0.339981043584856, 0.339981043584856, 0.861136311594053},
  {-0.906179845938664, -0.538469310105683, 0.000000000000000, 0.538469310105683, 0.906179845938664},
  {-0.932469514203152, -0.661209386466264, -0.238619186083197, 0.238619186083197, 0.661209386466265,
   0.932469514203152},
  {-0.949107912342758, -0.741531185599395, -0.405845151377397, 0.000000000000000, 0.405845151377397,
   0.741531185599394, 0.949107912342758},
  {-0.960289856497536, -0.796666477413627, -0.525532409916329, -0.183434642495650, 0.183434642495650,
   0.525532409916329, 0.796666477413627, 0.960289856497536},
  {-0.968160239507626, -0.836031107326636, -0.613371432700591, -0.324253423403809, 0.000000000000000,
   0.324253423403809, 0.613371432700591, 0.836031107326636, 0.968160239507626},
  {-0.973906528517172, -0.865063366688985, -0.679409568299024, -0.433395394129247, -0.148874338981631,
   0.148874338981632, 0.433395394129247, 0.679409568299024, 0.865063366688984, 0.973906528517171},
  {-0.978228658146057, -0.887062599768095, -0.730152005574049, -0.519096129206812, -0.269543155952345,
   0.000000000000000, 0.269543155952345, 0.519096129206812, 0.730152005574049, 0.887062599768095,
   0.978228658146057}};

// Gauss-Lobatto abscissae, orders 0..10 (src/points.cpp:299-385).
const double kLobatto[11][11] = {
  {0.0},
  {-1.000000000000000, 1.000000000000000},
  {-1.000000000000000, -0.000000000000000, 1.000000000000000},
  {-1.000000000000000, -0.447213595499958, 0.447213595499958, 1.000000000000000},
  {-1.000000000000000, -0.654653670707977, -0.000000000000000, 0.654653670707977, 1.000000000000000},
  {-1.000000000000000, -0.765055323929465, -0.285231516480645, 0.285231516480645, 0.765055323929465,
   1.000000000000000},
  {-1.000000000000000, -0.830223896278567, -0.468848793470714, 0.000000000000000, 0.468848793470714,
   0.830223896278567, 1.000000000000000},
  {-1.000000000000000, -0.871740148509607, -0.591700181433142, -0.209299217902479, 0.209299217902479,
   0.591700181433142, 0.871740148509607, 1.000000000000000},
  {-1.000000000000000, -0.899757995411460, -0.677186279510738, -0.363117463826178, -0.000000000000000,
   0.363117463826178, 0.677186279510738, 0.899757995411460, 1.000000000000000},
  {-1.000000000000000, -0.919533908166459, -0.738773865105505, -0.477924949810445, -0.165278957666387,
   0.165278957666387, 0.477924949810444, 0.738773865105505, 0.919533908166459, 1.000000000000000},
  {-1.000000000000000, -0.934001430408059, -0.784483473663144, -0.565235326996205, -0.295758135586939,
   0.000000000000000, 0.295758135586939, 0.565235326996205, 0.784483473663144, 0.934001430408059,
   1.000000000000000}};

// Gauss quadrature weights, orders 0..10 (src/points.cpp:425-538).
const double kWeights[11][11] = {
  {2.0},
  {1.0, 1.0},
  {0.5555555555555556, 0.8888888888888888, 0.5555555555555556},
  {0.3478548451374538, 0.6521451548625461, 0.6521451548625461, 0.3478548451374538},
  {0.2369268850561891, 0.4786286704993665, 0.5688888888888889, 0.4786286704993665, 0.2369268850561891},
  {0.1713244923791704, 0.3607615730481386, 0.4679139345726910, 0.4679139345726910, 0.3607615730481386,
   0.1713244923791704},
  {0.1294849661688697, 0.2797053914892766, 0.3818300505051189, 0.4179591836734694, 0.3818300505051189,
   0.2797053914892766, 0.1294849661688697},
  {0.1012285362903763, 0.2223810344533745, 0.3137066458778873, 0.3626837833783620, 0.3626837833783620,
   0.3137066458778873, 0.2223810344533745, 0.1012285362903763},
  {0.0812743883615744, 0.1806481606948574, 0.2606106964029354, 0.3123470770400029, 0.3302393550012598,
   0.3123470770400029, 0.2606106964029354, 0.1806481606948574, 0.0812743883615744},
  {0.0666713443086881, 0.1494513491505806, 0.2190863625159820, 0.2692667193099963, 0.2955242247147529,
   0.2955242247147529, 0.2692667193099963, 0.2190863625159820, 0.1494513491505806, 0.0666713443086881},
  {0.0556685671161737, 0.1255803694649046, 0.1862902109277343, 0.2331937645919905, 0.2628045445102467,
   0.2729250867779006, 0.2628045445102467, 0.2331937645919905, 0.1862902109277343, 0.1255803694649046,
   0.0556685671161737}};

int factorial(int n) { return (n <= 1) ? 1 : factorial(n - 1) * n; }

}  // namespace

std::vector<double> pts1d(int ptsType, int order) {
  if (order < 0 || order > 10) throw Error("point locations not available for order " + std::to_string(order));
  std::vector<double> out(order + 1);
  if (ptsType == PTS_EQUIDISTANT) {
    double dxi = 2. / order;
    for (int i = 0; i < order + 1; i++) out[i] = -1 + dxi * i;
  } else if (ptsType == PTS_LEGENDRE) {
    for (int i = 0; i <= order; i++) out[i] = kLegendre[order][i];
  } else if (ptsType == PTS_LOBATTO) {
    for (int i = 0; i <= order; i++) out[i] = kLobatto[order][i];
  } else {
    throw Error("points type not recognized");
  }
  return out;
}

std::vector<double> weights1d(int order) {
  if (order < 0 || order > 10) throw Error("Gauss weights not available for order " + std::to_string(order));
  return std::vector<double>(kWeights[order], kWeights[order] + order + 1);
}

// getQptWeights (src/points.cpp:394-423): tensor products, slowest index first.
std::vector<double> weightsNd(int order, int nDims) {
  std::vector<double> w = weights1d(order);
  if (nDims == 1) return w;
  int P1 = order + 1;
  std::vector<double> out;
  if (nDims == 2) {
    out.resize(P1 * P1);
    for (int i = 0; i < P1; i++)
      for (int j = 0; j < P1; j++) out[j + i * P1] = w[i] * w[j];
  } else {
    out.resize(P1 * P1 * P1);
    for (int i = 0; i < P1; i++)
      for (int j = 0; j < P1; j++)
        for (int k = 0; k < P1; k++) out[k + P1 * (j + i * P1)] = w[i] * w[j] * w[k];
  }
  return out;
}

// src/polynomials.cpp:39-50
double lagrange(const std::vector<double>& x, double y, int mode) {
  double lag = 1.0;
  for (int i = 0; i < (int)x.size(); i++)
    if (i != mode) lag = lag * ((y - x[i]) / (x[mode] - x[i]));
  return lag;
}

// src/polynomials.cpp:53-80
double dlagrange(const std::vector<double>& x, double y, int mode) {
  double d = 0.0;
  int n = (int)x.size();
  for (int i = 0; i < n; i++) {
    if (i == mode) continue;
    double num = 1.0, den = 1.0;
    for (int j = 0; j < n; j++) {
      if (j != mode && j != i) num = num * (y - x[j]);
      if (j != mode) den = den * (x[mode] - x[j]);
    }
    d = d + (num / den);
  }
  return d;
}

// src/polynomials.cpp:119-129
double legendre(double r, int mode) {
  if (mode < 0) return 0.;
  if (mode == 0) return 1.0;
  if (mode == 1) return r;
  return ((2 * mode - 1) * r * legendre(r, mode - 1) - (mode - 1) * legendre(r, mode - 2)) / mode;
}

// src/polynomials.cpp:131-151
double dlegendre(double r, int mode) {
  double d = 0.;
  if (mode == 0) return 0.;
  if (r > -1.0 && r < 1.0) {
    d = mode * ((r * legendre(r, mode)) - legendre(r, mode - 1)) / (r * r - 1.0);
  } else {
    if (r == -1.0) d = std::pow(-1.0, mode - 1.0) * 0.5 * mode * (mode + 1.0);
    if (r == 1.0) d = 0.5 * mode * (mode + 1.0);
  }
  return d;
}

// src/polynomials.cpp:439-477
double vcjh_eta(int vcjh, int order) {
  if (order == 0 && vcjh != VCJH_DG) throw Error("P=0 only compatible with DG");
  switch (vcjh) {
    case VCJH_DG: return 0.0;
    case VCJH_SD: return (1.0 * order) / (1.0 * (order + 1));
    case VCJH_HU: return (1.0 * (order + 1)) / (1.0 * order);
    case VCJH_CPLUS: {
      double c;
      if (order == 2) c = 0.206;
      else if (order == 3) c = 3.80e-3;
      else if (order == 4) c = 4.67e-5;
      else if (order == 5) c = 4.28e-7;
      else throw Error("C_plus scheme not implemented for this order");
      double ap = 1. / std::pow(2.0, order) * factorial(2 * order) / (factorial(order) * factorial(order));
      return c * (2 * order + 1) / 2 * (factorial(order) * ap) * (factorial(order) * ap);
    }
    default: throw Error("invalid VCJH scheme");
  }
}

// Derivative of the left (mode 0) / right (mode 1) VCJH correction function,
// src/polynomials.cpp:490-508.
double dvcjh1d(double r, int mode, int order, double eta) {
  if (mode == 0) {
    if (order == 0) return 0.5 * std::pow(-1.0, order) * (dlegendre(r, order) - ((dlegendre(r, order + 1)) / (1.0 + eta)));
    return 0.5 * std::pow(-1.0, order) *
           (dlegendre(r, order) - (((eta * dlegendre(r, order - 1)) + dlegendre(r, order + 1)) / (1.0 + eta)));
  }
  if (order == 0) return 0.5 * (dlegendre(r, order) + ((dlegendre(r, order + 1)) / (1.0 + eta)));
  return 0.5 * (dlegendre(r, order) + (((eta * dlegendre(r, order - 1)) + dlegendre(r, order + 1)) / (1.0 + eta)));
}

void sizes(int nDims, int order, int& nSpts, int& nFpts, int& nFptsFace) {
  if (nDims != 2 && nDims != 3) throw Error("nDims must be 2 or 3");
  if (order < 0 || order > 10) throw Error("order out of range");
  int P1 = order + 1;
  nFptsFace = (nDims == 2) ? P1 : P1 * P1;
  nSpts = (nDims == 2) ? P1 * P1 : P1 * P1 * P1;
  nFpts = 2 * nDims * nFptsFace;
}

// Parent-domain outward normal of face iFace (src/solver.cpp:229-246, 266-297).
void face_tnorm(int nDims, int iFace, double* t) {
  static const double q[4][2] = {{0, -1}, {1, 0}, {0, 1}, {-1, 0}};
  static const double h[6][3] = {{0, 0, -1}, {0, 0, 1}, {-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}};
  for (int d = 0; d < nDims; d++) t[d] = (nDims == 2) ? q[iFace][d] : h[iFace][d];
}

// getLocSpts / getLocFpts (src/points.cpp:67-162).
void ref_points(int nDims, int order, int ptsType, std::vector<Pt>& spts, std::vector<Pt>& fpts) {
  std::vector<double> p = pts1d(ptsType, order);
  int P1 = order + 1;
  spts.clear();
  fpts.clear();
  if (nDims == 2) {
    spts.resize(P1 * P1);
    for (int i = 0; i < P1; i++)
      for (int j = 0; j < P1; j++) spts[j + i * P1] = Pt{p[j], p[i], 0.};
    fpts.resize(4 * P1);
    for (int i = 0; i < P1; i++) {
      fpts[i] = Pt{p[i], -1., 0.};
      fpts[i + P1] = Pt{1., p[i], 0.};
      fpts[i + 2 * P1] = Pt{p[order - i], 1., 0.};
      fpts[i + 3 * P1] = Pt{-1., p[order - i], 0.};
    }
  } else {
    int P12 = P1 * P1;
    spts.resize(P12 * P1);
    for (int k = 0; k < P1; k++)
      for (int j = 0; j < P1; j++)
        for (int i = 0; i < P1; i++) spts[i + P1 * (j + P1 * k)] = Pt{p[i], p[j], p[k]};
    fpts.resize(6 * P12);
    for (int i = 0; i < P1; i++) {
      for (int j = 0; j < P1; j++) {
        int ind = i + j * P1;
        fpts[ind] = Pt{p[i], p[j], -1.};
        fpts[ind + P12] = Pt{p[order - i], p[j], 1.};
        fpts[ind + 2 * P12] = Pt{-1., p[i], p[j]};
        fpts[ind + 3 * P12] = Pt{1., p[order - i], p[j]};
        fpts[ind + 4 * P12] = Pt{p[order - i], -1., p[j]};
        fpts[ind + 5 * P12] = Pt{p[i], 1., p[j]};
      }
    }
  }
}

// divVCJH_quad / divVCJH_hex (src/operators.cpp:1448-1506).
static double div_vcjh(int nDims, int fpt, const Pt& loc, const std::vector<double>& s1, int order, double eta) {
  int P1 = order + 1;
  if (nDims == 2) {
    int i = fpt / P1, j = fpt - P1 * i;
    if (i == 0) return -lagrange(s1, loc.x, j) * dvcjh1d(loc.y, 0, order, eta);
    if (i == 1) return lagrange(s1, loc.y, j) * dvcjh1d(loc.x, 1, order, eta);
    if (i == 2) return lagrange(s1, loc.x, order - j) * dvcjh1d(loc.y, 1, order, eta);
    return -lagrange(s1, loc.y, order - j) * dvcjh1d(loc.x, 0, order, eta);
  }
  int P12 = P1 * P1;
  int f = fpt / P12, j = (fpt - P12 * f) / P1, i = fpt - P12 * f - P1 * j;
  switch (f) {
    case 0: return lagrange(s1, loc.x, i) * lagrange(s1, loc.y, j) * -dvcjh1d(loc.z, 0, order, eta);
    case 1: return lagrange(s1, loc.x, order - i) * lagrange(s1, loc.y, j) * dvcjh1d(loc.z, 1, order, eta);
    case 2: return lagrange(s1, loc.y, i) * lagrange(s1, loc.z, j) * -dvcjh1d(loc.x, 0, order, eta);
    case 3: return lagrange(s1, loc.y, order - i) * lagrange(s1, loc.z, j) * dvcjh1d(loc.x, 1, order, eta);
    case 4: return lagrange(s1, loc.x, order - i) * lagrange(s1, loc.z, j) * -dvcjh1d(loc.y, 0, order, eta);
    default: return lagrange(s1, loc.x, i) * lagrange(s1, loc.z, j) * dvcjh1d(loc.y, 1, order, eta);
  }
}

void operators(int nDims, int order, int ptsType, int vcjh, Operators& op) {
  int S, F, Ff;
  sizes(nDims, order, S, F, Ff);
  op.nDims = nDims;
  op.order = order;
  op.nSpts = S;
  op.nFpts = F;
  std::vector<double> s1 = pts1d(ptsType, order);
  std::vector<Pt> ls, lf;
  ref_points(nDims, order, ptsType, ls, lf);
  int P1 = order + 1;

  // opp_spts_to_fpts (src/operators.cpp:112-141)
  op.spts_to_fpts.assign((size_t)F * S, 0.);
  for (int fpt = 0; fpt < F; fpt++)
    for (int spt = 0; spt < S; spt++) {
      double v;
      if (nDims == 2) {
        int is = spt % P1, js = spt / P1;
        v = lagrange(s1, lf[fpt].x, is) * lagrange(s1, lf[fpt].y, js);
      } else {
        int ks = spt / (P1 * P1), js = (spt - P1 * P1 * ks) / P1, is = spt - P1 * js - P1 * P1 * ks;
        v = lagrange(s1, lf[fpt].x, is) * lagrange(s1, lf[fpt].y, js) * lagrange(s1, lf[fpt].z, ks);
      }
      op.spts_to_fpts[(size_t)fpt * S + spt] = v;
    }

  // opp_grad_spts (setupGradPts, src/operators.cpp:596-658); basis and
  // evaluation points are both the 1-D solution points.
  op.grad_spts.assign((size_t)nDims * S * S, 0.);
  for (int dim = 0; dim < nDims; dim++)
    for (int p = 0; p < S; p++)
      for (int b = 0; b < S; b++) {
        double v;
        if (nDims == 2) {
          int ip = p % P1, jp = p / P1, ib = b % P1, jb = b / P1;
          if (dim == 0) v = dlagrange(s1, s1[ip], ib) * lagrange(s1, s1[jp], jb);
          else v = dlagrange(s1, s1[jp], jb) * lagrange(s1, s1[ip], ib);
        } else {
          int kp = p / (P1 * P1), jp = (p - P1 * P1 * kp) / P1, ip = p - P1 * jp - P1 * P1 * kp;
          int kb = b / (P1 * P1), jb = (b - P1 * P1 * kb) / P1, ib = b - P1 * jb - P1 * P1 * kb;
          if (dim == 0) v = dlagrange(s1, s1[ip], ib) * lagrange(s1, s1[jp], jb) * lagrange(s1, s1[kp], kb);
          else if (dim == 1) v = dlagrange(s1, s1[jp], jb) * lagrange(s1, s1[ip], ib) * lagrange(s1, s1[kp], kb);
          else v = dlagrange(s1, s1[kp], kb) * lagrange(s1, s1[ip], ib) * lagrange(s1, s1[jp], jb);
        }
        op.grad_spts[((size_t)dim * S + p) * S + b] = v;
      }

  // opp_extrapolateFn (src/operators.cpp:715-797)
  op.extrap_fn.assign((size_t)nDims * F * S, 0.);
  for (int spt = 0; spt < S; spt++)
    for (int fpt = 0; fpt < F; fpt++) {
      double t[3];
      face_tnorm(nDims, fpt / Ff, t);
      for (int dim = 0; dim < nDims; dim++)
        op.extrap_fn[((size_t)dim * F + fpt) * S + spt] = op.spts_to_fpts[(size_t)fpt * S + spt] * t[dim];
    }

  // opp_correction (src/operators.cpp:800-824)
  double eta = (vcjh == VCJH_DG) ? 0. : vcjh_eta(vcjh, order);
  op.correction.assign((size_t)S * F, 0.);
  for (int spt = 0; spt < S; spt++)
    for (int fpt = 0; fpt < F; fpt++) op.correction[(size_t)spt * F + fpt] = div_vcjh(nDims, fpt, ls[spt], s1, order, eta);

  // opp_correctU (src/operators.cpp:915-1001)
  op.correct_u.assign((size_t)nDims * S * F, 0.);
  for (int fpt = 0; fpt < F; fpt++) {
    double t[3];
    face_tnorm(nDims, fpt / Ff, t);
    for (int spt = 0; spt < S; spt++)
      for (int dim = 0; dim < nDims; dim++)
        op.correct_u[((size_t)dim * S + spt) * F + fpt] = op.correction[(size_t)spt * F + fpt] * t[dim];
  }

  op.w_spts = weightsNd(order, nDims);
  op.w_face = weightsNd(order, nDims - 1);
}

// P-multigrid transfer operators of one order (oper::setupPMG,
// src/operators.cpp:1553-1634): the degree-`order` interpolant through the
// solution points, evaluated at the solution points of order+1 (prolongation,
// [(order+2)^D][S]) and of order-1 (restriction, [order^D][S]; empty at order 0).
// Products run x, y, z as in the reference; point p of a tensor grid with n
// points per direction has indices (p % n, p / n % n, p / n^2).
void pmg_operators(int nDims, int order, int ptsType, std::vector<double>& prolong, std::vector<double>& restrict_) {
  const int P1 = order + 1;
  const int S = nDims == 2 ? P1 * P1 : P1 * P1 * P1;
  const std::vector<double> s1 = pts1d(ptsType, order);
  auto build = [&](int n1, std::vector<double>& op) {
    const std::vector<double> t1 = pts1d(ptsType, n1 - 1);
    const int T = nDims == 2 ? n1 * n1 : n1 * n1 * n1;
    op.assign((size_t)T * S, 0.);
    for (int spt = 0; spt < S; spt++) {
      const int is = spt % P1, js = (spt / P1) % P1, ks = spt / (P1 * P1);
      for (int t = 0; t < T; t++) {
        const int it = t % n1, jt = (t / n1) % n1, kt = t / (n1 * n1);
        double v = lagrange(s1, t1[it], is) * lagrange(s1, t1[jt], js);
        if (nDims == 3) v = v * lagrange(s1, t1[kt], ks);
        op[(size_t)t * S + spt] = v;
      }
    }
  };
  build(order + 2, prolong);
  if (order > 0) build(order, restrict_);
  else restrict_.clear();
}

}  // namespace fr
