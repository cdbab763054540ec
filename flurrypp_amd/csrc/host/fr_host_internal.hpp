// Internal declarations shared by the host set-up sources.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

namespace fr {

struct Error : std::runtime_error {
  explicit Error(const std::string& s) : std::runtime_error(s) {}
};

enum { PTS_LEGENDRE = 0, PTS_LOBATTO = 1, PTS_EQUIDISTANT = 2 };
enum { VCJH_DG = 0, VCJH_SD = 1, VCJH_HU = 2, VCJH_CPLUS = 3 };
enum {
  BC_NONE = -1, BC_PERIODIC = 0, BC_CHAR = 1, BC_SUP_IN = 2, BC_SUP_OUT = 3, BC_SUB_IN = 4,
  BC_SUB_OUT = 5, BC_SLIP_WALL = 8, BC_ISOTHERMAL_NOSLIP = 9, BC_ADIABATIC_NOSLIP = 10,
  BC_OVERSET = 11, BC_SYMMETRY = 12, BC_ISOTHERMAL_NOSLIP_MOVING = 13
};

struct Pt {
  double x, y, z;
};

struct Operators {
  int nDims = 0, order = 0, nSpts = 0, nFpts = 0;
  std::vector<double> spts_to_fpts, grad_spts, extrap_fn, correction, correct_u, w_spts, w_face;
};

std::vector<double> pts1d(int ptsType, int order);
std::vector<double> weights1d(int order);
std::vector<double> weightsNd(int order, int nDims);
double lagrange(const std::vector<double>& x, double y, int mode);
double dlagrange(const std::vector<double>& x, double y, int mode);
double legendre(double r, int mode);
double dlegendre(double r, int mode);
double vcjh_eta(int vcjh, int order);
double dvcjh1d(double r, int mode, int order, double eta);
void sizes(int nDims, int order, int& nSpts, int& nFpts, int& nFptsFace);
void face_tnorm(int nDims, int iFace, double* t);
void ref_points(int nDims, int order, int ptsType, std::vector<Pt>& spts, std::vector<Pt>& fpts);
void operators(int nDims, int order, int ptsType, int vcjh, Operators& op);
void pmg_operators(int nDims, int order, int ptsType, std::vector<double>& prolong, std::vector<double>& restrict_);

}  // namespace fr
