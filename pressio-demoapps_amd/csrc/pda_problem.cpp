// pda_problem.cpp -- problem objects behind the C ABI: parameters, initial
// conditions, the fixed Jacobian pattern and the device-side index maps the
// kernels consume, and the evaluation entry points.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <string>

#include "pda_internal.hpp"

using namespace pda;

namespace {

struct ParamSpec {
  std::vector<std::string> phys_names;
  std::vector<double> phys_defaults;
  std::vector<std::string> ic_names;
  std::vector<double> ic_defaults;
};

// reference parameter names and defaults:
// impl/euler_2d_parametrization_helpers.hpp:57-97,
// impl/swe_2d_parametrization_helpers.hpp:58-104,
// impl/advection_diffusion_2d_parametrization_helpers.hpp:57-80,
// advection_diffusion_reaction2d.hpp:138-147
ParamSpec spec_for(int family, int ic_flag) {
  ParamSpec s;
  switch (family) {
    case PDA_EULER1D:
    case PDA_EULER3D:
      s.phys_names = {"gamma"};
      s.phys_defaults = {1.4};
      if (family == PDA_EULER3D) {
        // Sedov blast set-up; the reference hard-wires radius = 3 min(dx,dy,dz)
        // and background pressure 2.5e-5 (impl/euler_3d_initial_condition.hpp:97,132)
        s.ic_names = {"sedovRadiusCells", "sedovAmbientPressure", "sedovEnergy"};
        s.ic_defaults = {3., 2.5e-5, 0.851072};
      }
      break;
    case PDA_EULER2D:
      s.phys_names = {"gamma"};
      s.phys_defaults = {1.4};
      s.ic_names = {"normalShockMach", "crossShockDensity", "crossShockInletXVel",
                    "crossShockBottomYVel", "riemannTopRightPressure",
                    "riemannTopRightXVel", "riemannTopRightYVel",
                    "riemannTopRightDensity", "riemannBotLeftPressure"};
      s.ic_defaults = {9., 0.1, 10., 1., ic_flag == 1 ? 0.4 : 1.5, 0., 0., 1.5, 0.029};
      break;
    case PDA_SWE2D:
      s.phys_names = {"gravity", "coriolis"};
      s.phys_defaults = {9.8, -3.};
      s.ic_names = {"pulseMagnitude", "pulseX", "pulseY", "pulseMagnitude1", "pulseX1",
                    "pulseY1", "pulseMagnitude2", "pulseX2", "pulseY2"};
      s.ic_defaults = {1. / 8, 1., 1., 1. / 10, -2., -2., 1. / 8, 2., 2.};
      break;
    case PDA_BURGERS2D:
      s.phys_names = {"diffusion"};
      s.phys_defaults = {0.00001};
      s.ic_names = {"pulseMagnitude", "pulseSpread", "pulseX", "pulseY"};
      s.ic_defaults = {0.5, 0.15, 0.0, -0.2};
      break;
    case PDA_ADR2D:
      s.phys_names = {"ux", "uy", "diffusion", "sigma"};
      s.phys_defaults = {0.5 * std::cos(M_PI / 3), 0.5 * std::sin(M_PI / 3), 0.001, 1.0};
      break;
  }
  return s;
}

int family_m(int f) {
  switch (f) {
    case PDA_EULER1D: return 3;
    case PDA_EULER2D: return 4;
    case PDA_EULER3D: return 5;
    case PDA_SWE2D: return 3;
    case PDA_BURGERS2D: return 2;
    case PDA_ADR2D: return 1;
  }
  return -1;
}
int family_dim(int f) {
  return f == PDA_EULER1D ? 1 : f == PDA_EULER3D ? 3 : (f >= 1 && f <= 6) ? 2 : -1;
}

// which (family, problem) pairs the GPU engine implements, and the ghost rule
int ghost_rule_of(int family, int prob, bool *supported) {
  *supported = true;
  switch (family) {
    case PDA_EULER1D:
      if (prob == 0) return GHOST_NONE;             // PeriodicSmooth
      if (prob >= 1 && prob <= 3) return GHOST_OPPOSITE; // Sod, Lax, ShuOsher
      break;
    case PDA_EULER2D:
      if (prob == 0) return GHOST_NONE;             // PeriodicSmooth
      if (prob == 2 || prob == 4 || prob == 8) return GHOST_MIRROR; // SedovFull, Riemann, Neumann
      break;
    case PDA_EULER3D:
      if (prob == 0) return GHOST_NONE;
      if (prob == 1) return GHOST_OPPOSITE;         // SedovSymmetry
      break;
    case PDA_SWE2D:
      if (prob == 0) return GHOST_MIRROR;           // SlipWall
      break;
    case PDA_BURGERS2D:
      if (prob == 0) return GHOST_NONE;             // BurgersPeriodic
      if (prob == 1) return GHOST_MIRROR;           // BurgersOutflow (0 at the minus sides, gsign)
      break;
    case PDA_ADR2D:
      if (prob == 0) return GHOST_OPPOSITE;         // ProblemA, homogeneous Dirichlet
      break;
    default:
      break;
  }
  *supported = false;
  return GHOST_NONE;
}

double energy(double gm1inv, const double *prim, int nvel) {
  // euler_compute_energy.hpp:55-124
  double v2 = 0.;
  for (int d = 0; d < nvel; ++d) v2 += prim[1 + d] * prim[1 + d];
  return prim[nvel + 1] * gm1inv + 0.5 * prim[0] * v2;
}

void set_cons(double *u, const double *prim, int nvel, double gm1inv) {
  u[0] = prim[0];
  for (int d = 0; d < nvel; ++d) u[1 + d] = prim[0] * prim[1 + d];
  u[nvel + 1] = energy(gm1inv, prim, nvel);
}

// initial conditions, impl/*_initial_condition.hpp
int initial_condition_host(const pda_problem_s &p, std::vector<double> &U) {
  const pda_mesh_s &m = *p.mesh;
  const int64_t n = m.n_stencil;
  U.assign(n * p.m, 0.);
  const auto &x = m.x, &y = m.y, &z = m.z;
  switch (p.family) {
    case PDA_EULER1D: { // impl/euler_1d_initial_condition.hpp
      const double g = p.phys[0], gm1inv = 1. / (g - 1.);
      for (int64_t i = 0; i < n; ++i) {
        double prim[3] = {1, 0, 1};
        if (p.prob == 0) {
          prim[0] = 1. + 0.2 * std::sin(M_PI * x[i]); prim[1] = 1.; prim[2] = 1.;
        } else if (p.prob == 1) {
          if (x[i] <= 0.) { prim[0] = 1.; prim[1] = 0.; prim[2] = 1.; }
          else { prim[0] = 0.125; prim[1] = 0.; prim[2] = 0.1; }
        } else if (p.prob == 2) {
          if (x[i] <= 0.) { prim[0] = 0.445; prim[1] = 0.698; prim[2] = 3.528; }
          else { prim[0] = 0.5; prim[1] = 0.; prim[2] = 0.571; }
        } else {
          if (x[i] <= -4.) { prim[0] = 27. / 7.; prim[1] = 2.629369; prim[2] = 31. / 3.; }
          else { prim[0] = 1. + (1. / 5.) * std::sin(5. * x[i]); prim[1] = 0.; prim[2] = 1.; }
        }
        set_cons(&U[i * 3], prim, 1, gm1inv);
      }
      return PDA_OK;
    }
    case PDA_EULER2D: { // impl/euler_2d_initial_condition.hpp
      const double g = p.phys[0], gm1 = g - 1., gm1inv = 1. / gm1;
      if (p.prob == 0) {
        for (int64_t i = 0; i < n; ++i) {
          const double prim[4] = {1. + (1. / 5.) * std::sin(M_PI * (x[i] + y[i])), 1., 1., 1.};
          set_cons(&U[i * 4], prim, 2, gm1inv);
        }
      } else if (p.prob == 2) {
        const double sRad = 2. * std::min(m.d[0], m.d[1]);
        for (int64_t i = 0; i < n; ++i) {
          const double myR = std::sqrt(x[i] * x[i] + y[i] * y[i]);
          double prim[4] = {1., 0., 0., 5.e-5};
          if (myR <= sRad) prim[3] = gm1 / (M_PI * sRad * sRad);
          set_cons(&U[i * 4], prim, 2, gm1inv);
        }
      } else if (p.prob == 4 && p.ic_flag == 1) {
        const double trp = p.icp[4];
        for (int64_t i = 0; i < n; ++i) {
          double prim[4] = {0, 0, 0, 0};
          if (x[i] >= 0.5 && y[i] >= 0.5) { prim[0] = 0.5313; prim[1] = 0.; prim[2] = 0.; prim[3] = trp; }
          else if (x[i] < 0.5 && y[i] >= 0.5) { prim[0] = 1.; prim[1] = 0.7276; prim[2] = 0.; prim[3] = 1.; }
          else if (x[i] < 0.5 && y[i] < 0.5) { prim[0] = 0.8; prim[1] = 0.; prim[2] = 0.; prim[3] = 1.; }
          else if (x[i] > 0.5 && y[i] < 0.5) { prim[0] = 1.; prim[1] = 0.; prim[2] = 0.7276; prim[3] = 1.; }
          set_cons(&U[i * 4], prim, 2, gm1inv);
        }
      } else if (p.prob == 4 && p.ic_flag == 2) {
        // impl/euler_2d_initial_condition.hpp:308-384, compatibility relations
        const double trP = p.icp[4], trU = p.icp[5], trV = p.icp[6], trR = p.icp[7], blP = p.icp[8];
        if (trP <= blP) {
          set_error("INVALID: riemannTopRightPressure <= riemannBotLeftPressure");
          return PDA_E_INVALID_ARG;
        }
        const double eps = (g - 1.) / (g + 1.);
        const double fac13 = trP / blP;
        const double e13 = eps * (fac13 + 1.);
        const double fac43 = (1. / (2. * (1. + 2. * eps))) *
                             (e13 + std::sqrt(e13 * e13 + 4. * (1. + 2. * eps) * fac13));
        const double tlP = fac43 * blP, brP = tlP;
        const double tlV = trV, brU = trU;
        const double tlR = trR * (tlP / trP + eps) / (1 + eps * tlP / trR);
        const double brR = tlR;
        const double psi21 = (tlP - trP) * (tlR - trR) / (tlR * trR);
        const double tlU = std::sqrt(psi21) + trU;
        const double psi41 = (brP - trP) * (brR - trR) / (brR * trR);
        const double brV = std::sqrt(psi41) + trV;
        const double blU = tlU, blV = brV;
        const double blR = tlR * (blP - tlP) / ((blP - tlP) - psi41 * tlR);
        for (int64_t i = 0; i < n; ++i) {
          double prim[4] = {0, 0, 0, 0};
          if (x[i] >= 0.8 && y[i] >= 0.8) { prim[0] = trR; prim[1] = trU; prim[2] = trV; prim[3] = trP; }
          else if (x[i] < 0.8 && y[i] >= 0.8) { prim[0] = tlR; prim[1] = tlU; prim[2] = tlV; prim[3] = tlP; }
          else if (x[i] < 0.8 && y[i] < 0.8) { prim[0] = blR; prim[1] = blU; prim[2] = blV; prim[3] = blP; }
          else if (x[i] > 0.8 && y[i] < 0.8) { prim[0] = brR; prim[1] = brU; prim[2] = brV; prim[3] = brP; }
          set_cons(&U[i * 4], prim, 2, gm1inv);
        }
      }
      return PDA_OK;
    }
    case PDA_EULER3D: { // impl/euler_3d_initial_condition.hpp:54-141
      const double g = p.phys[0], gm1 = g - 1., gm1inv = 1. / gm1;
      if (p.prob == 0) {
        for (int64_t i = 0; i < n; ++i) {
          const double prim[5] = {1.0 + 0.2 * std::sin(M_PI * (x[i] + y[i] + z[i])), 1., 1., 1., 1.};
          set_cons(&U[i * 5], prim, 3, gm1inv);
        }
      } else {
        const double sRad = p.icp[0] * std::min(m.d[0], std::min(m.d[1], m.d[2]));
        for (int64_t i = 0; i < n; ++i) {
          const double myR = std::sqrt(x[i] * x[i] + y[i] * y[i] + z[i] * z[i]);
          double prim[5] = {1., 0., 0., 0., p.icp[1]};
          if (myR <= sRad) prim[4] = (3. * gm1 * p.icp[2]) / (4. * M_PI * sRad * sRad * sRad);
          set_cons(&U[i * 5], prim, 3, gm1inv);
        }
      }
      return PDA_OK;
    }
    case PDA_SWE2D: { // impl/swe_2d_initial_condition.hpp:55-110
      for (int64_t i = 0; i < n; ++i) {
        if (p.ic_flag == 2) {
          const double dx1 = x[i] - p.icp[4], dy1 = y[i] - p.icp[5];
          const double r1 = std::sqrt(dx1 * dx1 + dy1 * dy1);
          const double dx2 = x[i] - p.icp[7], dy2 = y[i] - p.icp[8];
          const double r2 = std::sqrt(dx2 * dx2 + dy2 * dy2);
          U[i * 3] = 1. + p.icp[3] * std::exp(-(r1 * r1)) + p.icp[6] * std::exp(-(r2 * r2));
        } else {
          const double dx1 = x[i] - p.icp[1], dy1 = y[i] - p.icp[2];
          const double r = std::sqrt(dx1 * dx1 + dy1 * dy1);
          U[i * 3] = 1. + p.icp[0] * std::exp(-(r * r));
        }
      }
      return PDA_OK;
    }
    case PDA_BURGERS2D: { // impl/advection_diffusion_2d_initial_condition.hpp:54-78
      const double mag = p.icp[0], spread = p.icp[1], x0 = p.icp[2], y0 = p.icp[3];
      for (int64_t i = 0; i < n; ++i) {
        const double dx = x[i] - x0, dy = y[i] - y0;
        const double v = mag * std::exp(-(dx * dx + dy * dy) / spread);
        U[i * 2] = v;
        U[i * 2 + 1] = v;
      }
      return PDA_OK;
    }
    case PDA_ADR2D: // zero state, impl/advection_diffusion_reaction_2d_initial_condition.hpp:55-60
      return PDA_OK;
  }
  return PDA_E_UNSUPPORTED;
}

// Jacobian BC factors of near-boundary rows (first-order ghost dependence):
// Euler 1D impl/euler_1d_prob_class.hpp:608-610, Euler 2D
// impl/euler_2d_prob_class.hpp:1115-1140, Euler 3D
// impl/euler_3d_prob_class.hpp:1017-1044, SWE impl/swe_2d_prob_class.hpp:839-854,
// ADR impl/advection_diffusion_reaction_2d_prob_class.hpp:640-641 and :809-812,
// Burgers outflow impl/advection_diffusion_2d_prob_class.hpp:1131-1156
void bc_factors(const pda_problem_s &p, int64_t r, int a, double *f) {
  for (int k = 0; k < 5; ++k) f[k] = 1.;
  if (p.family == PDA_ADR2D) f[0] = -1.;
  if (p.family == PDA_SWE2D) f[a == 0 ? 1 : 2] = -1.;
  if (p.family == PDA_EULER3D && p.prob == 1 && mesh_has_bd(*p.mesh, r, a, 0)) f[1 + a] = -1.;
  // Burgers outflow: homogeneous Dirichlet at the left and back boundaries,
  // homogeneous Neumann at the right and front ones
  if (p.family == PDA_BURGERS2D && p.prob == 1 && mesh_has_bd(*p.mesh, r, a, 0)) f[0] = f[1] = 0.;
}

int build(pda_problem_s &p) {
  const pda_mesh_s &M = *p.mesh;
  const int m = p.m, dim = p.dim, w = p.w;
  const int64_t nr = M.n_sample;
  const int nc = M.ncols();
  // row classes, impl/mesh_ccu.hpp:385-447
  std::vector<int32_t> ghost_row(nr, -1), near_rows;
  for (int64_t r = 0; r < nr; ++r)
    if (mesh_near_bd(M, r)) {
      ghost_row[r] = int32_t(near_rows.size());
      near_rows.push_back(int32_t(r));
    }
  p.n_near = int64_t(near_rows.size());
  p.n_inner = nr - p.n_near;
  // pattern: self + scheme-width neighbours for inner rows, self + non-missing
  // first neighbours for near rows; sorted unique cells (setFromTriplets)
  const int npos = 1 + 2 * dim * w;
  std::vector<int64_t> rowbase(nr);
  std::vector<int32_t> nblk(nr);
  std::vector<uint8_t> slot(size_t(nr) * npos, 0xff);
  std::vector<int32_t> cells;
  int maxb = 0;
  bool dup = false;
  int64_t nnz = 0;
  std::vector<std::vector<int32_t>> bcols(nr);
  for (int64_t r = 0; r < nr; ++r) {
    const bool near = ghost_row[r] >= 0;
    const int nnb = near ? 2 * dim : 2 * dim * w;
    cells.clear();
    for (int c = 0; c <= nnb; ++c) {
      const int32_t v = M.g(r, c);
      if (v != -1) cells.push_back(v);
    }
    std::sort(cells.begin(), cells.end());
    cells.erase(std::unique(cells.begin(), cells.end()), cells.end());
    int present = 0;
    for (int c = 0; c <= nnb; ++c) {
      const int32_t v = M.g(r, c);
      if (v == -1) continue;
      ++present;
      slot[size_t(r) * npos + c] = uint8_t(std::lower_bound(cells.begin(), cells.end(), v) - cells.begin());
    }
    if (present != int(cells.size())) dup = true;
    nblk[r] = int32_t(cells.size());
    maxb = std::max(maxb, int(cells.size()));
    rowbase[r] = nnz;
    nnz += int64_t(cells.size()) * m * m;
    bcols[r] = cells;
  }
  p.nnz = nnz;
  p.maxb = maxb;
  std::vector<int32_t> bcol(size_t(nr) * maxb, -1);
  for (int64_t r = 0; r < nr; ++r)
    std::copy(bcols[r].begin(), bcols[r].end(), bcol.begin() + size_t(r) * maxb);
  // BC factor table
  std::vector<double> bcf(size_t(std::max<int64_t>(p.n_near, 1)) * dim * m, 1.);
  for (int64_t g = 0; g < p.n_near; ++g)
    for (int a = 0; a < dim; ++a) {
      double f[5];
      bc_factors(p, near_rows[g], a, f);
      for (int k = 0; k < m; ++k) bcf[(g * dim + a) * m + k] = f[k];
    }
  // structured box: inner cells whose scheme stencil does not wrap
  p.box = Box();
  p.h_rem_rows.clear();
  const bool box2d = dim == 2 && (p.family == PDA_EULER2D || p.family == PDA_SWE2D || p.family == PDA_BURGERS2D);
  const bool box3d = dim == 3 && p.family == PDA_EULER3D;
  if (M.structured && (box2d || box3d)) {
    const int wm = (M.stencil - 1) / 2;
    bool ok = true;
    for (int a = 0; a < 3; ++a) {
      p.box.n[a] = M.n[a];
      if (a >= dim) { p.box.lo[a] = 0; p.box.hi[a] = 1; continue; }
      // rows within the mesh stencil of a non-periodic boundary are
      // near-boundary rows, but where the scheme's own stencil stays in the
      // grid (no ghost read, no BC factor) their Euler and shallow-water
      // values equal the inner-row values; the Burgers viscous term is an
      // inner-row term, so its box stays at the mesh stencil
      const bool per = (M.periodic_mask >> a) & 1;
      const int64_t lo = (per || p.family != PDA_BURGERS2D) ? w : wm;
      p.box.lo[a] = lo;
      p.box.hi[a] = M.n[a] - lo;
      if (p.box.hi[a] <= lo) ok = false;
    }
    // natural order: row r = (k * ny + j) * nx + i
    auto outside = [&](const Box &b, int64_t r) {
      const int64_t i = r % M.n[0], j = (r / M.n[0]) % M.n[1], k = r / (M.n[0] * M.n[1]);
      return i < b.lo[0] || i >= b.hi[0] || j < b.lo[1] || j >= b.hi[1] || k < b.lo[2] || k >= b.hi[2];
    };
    if (ok) {
      p.box.valid = true;
      for (int64_t r = 0; r < nr; ++r)
        if (outside(p.box, r)) p.h_rem_rows.push_back(int32_t(r));
      // velocity box
      p.vbox = p.box;
      for (int a = 0; a < dim; ++a) {
        const bool per = (M.periodic_mask >> a) & 1;
        const bool big = M.n[a] >= 2 * w + 1;
        if (big && per) p.vbox.mode[a] = 1;
        else if (big && !per && p.view.ghost_rule == GHOST_MIRROR && p.family != PDA_BURGERS2D)
          p.vbox.mode[a] = 2;
        if (p.vbox.mode[a]) {
          p.vbox.lo[a] = 0;
          p.vbox.hi[a] = M.n[a];
        }
      }
      p.h_vrem_rows.clear();
      for (int64_t r = 0; r < nr; ++r)
        if (outside(p.vbox, r)) p.h_vrem_rows.push_back(int32_t(r));
    }
  }
  p.h_ghost_row = std::move(ghost_row);
  p.h_near_rows = std::move(near_rows);
  p.h_bcf = std::move(bcf);
  p.h_rowbase = std::move(rowbase);
  p.h_nblk = std::move(nblk);
  p.h_slot = std::move(slot);
  p.h_bcol = std::move(bcol);
  View &v = p.view;
  v.ncols = nc;
  v.nrows = nr;
  v.n_near = p.n_near;
  v.npos = npos;
  v.dup = dup ? 1 : 0;
  for (int a = 0; a < 3; ++a) v.hinv[a] = M.dinv[a];
  for (size_t i = 0; i < p.phys.size() && i < 8; ++i) v.p[i] = p.phys[i];
  // ghost signs: reflective walls flip the wall-normal momentum
  for (int a = 0; a < 3; ++a)
    for (int s = 0; s < 2; ++s)
      for (int k = 0; k < 5; ++k) v.gsign[a][s][k] = 1.;
  if (p.family == PDA_SWE2D)
    for (int s = 0; s < 2; ++s) {
      v.gsign[0][s][1] = -1.;
      v.gsign[1][s][2] = -1.;
    }
  if (p.family == PDA_EULER3D && p.prob == 1)
    for (int a = 0; a < 3; ++a) v.gsign[a][0][1 + a] = -1.;
  // homogeneous Dirichlet: every ghost is the negated source cell
  // (impl/advection_diffusion_reaction_2d_ghost_filler_problemA.hpp:97-133)
  if (p.family == PDA_ADR2D)
    for (int a = 0; a < 3; ++a)
      for (int s = 0; s < 2; ++s) v.gsign[a][s][0] = -1.;
  // Burgers outflow: every ghost of the left and back sides is 0, those of
  // the right and front sides mirror the cells inside
  // (impl/advection_diffusion_2d_ghost_filler_outflow.hpp:105-231)
  if (p.family == PDA_BURGERS2D && p.prob == 1)
    for (int a = 0; a < 2; ++a)
      for (int k = 0; k < 2; ++k) v.gsign[a][0][k] = 0.;
  return PDA_OK;
}

int ensure_scalar_pattern_host(pda_problem_s &p) {
  if (!p.rowptr.empty()) return PDA_OK;
  const pda_mesh_s &M = *p.mesh;
  const int m = p.m;
  const int64_t nr = M.n_sample;
  const std::vector<int32_t> &nblk = p.h_nblk, &bcol = p.h_bcol;
  p.rowptr.resize(nr * m + 1);
  p.colidx.resize(p.nnz);
  int64_t e = 0;
  p.rowptr[0] = 0;
  for (int64_t r = 0; r < nr; ++r)
    for (int k = 0; k < m; ++k) {
      for (int s = 0; s < nblk[r]; ++s)
        for (int j = 0; j < m; ++j) p.colidx[e++] = bcol[size_t(r) * p.maxb + s] * m + j;
      p.rowptr[r * m + k + 1] = e;
    }
  return PDA_OK;
}

int ensure_scalar_pattern(pda_problem_s &p) {
  int rc = ensure_scalar_pattern_host(p);
  if (rc) return rc;
  rc = ensure_device(p);
  if (rc) return rc;
  if (!p.d_rowptr.ptr) {
    PDA_HIP_CHECK(p.d_rowptr.upload(p.rowptr));
    PDA_HIP_CHECK(p.d_colidx.upload(p.colidx));
  }
  return PDA_OK;
}

}  // namespace

namespace pda {
// scalar CSR pattern on the host and the device (made on first use)
int ensure_pattern_device(pda_problem_s &p) { return ensure_scalar_pattern(p); }

// device copies of the index maps, made on the first evaluation
int ensure_device(pda_problem_s &p) {
  if (p.on_device) return PDA_OK;
  const pda_mesh_s &M = *p.mesh;
  PDA_HIP_CHECK(p.d_graph.upload(M.graph));
  PDA_HIP_CHECK(p.d_ghost_row.upload(p.h_ghost_row));
  PDA_HIP_CHECK(p.d_near_rows.upload(p.h_near_rows));
  PDA_HIP_CHECK(p.d_bcf.upload(p.h_bcf));
  PDA_HIP_CHECK(p.d_rowbase.upload(p.h_rowbase));
  PDA_HIP_CHECK(p.d_nblk.upload(p.h_nblk));
  PDA_HIP_CHECK(p.d_slot.upload(p.h_slot));
  PDA_HIP_CHECK(p.d_bcol.upload(p.h_bcol));
  PDA_HIP_CHECK(p.d_rem_rows.upload(p.h_rem_rows));
  PDA_HIP_CHECK(p.d_vrem_rows.upload(p.h_vrem_rows));
  const size_t gn = size_t(2 * p.dim) * size_t(std::max<int64_t>(p.n_near, 1)) * p.w * p.m;
  // unused ghost slots hold numeric_limits<double>::min(),
  // impl/euler_2d_prob_class.hpp:1266-1270
  PDA_HIP_CHECK(p.d_ghost.upload(std::vector<double>(gn, 2.2250738585072014e-308)));
  View &v = p.view;
  v.graph = p.d_graph.as<int32_t>();
  v.ghost_row = p.d_ghost_row.as<int32_t>();
  v.near_rows = p.d_near_rows.as<int32_t>();
  v.ghost = p.d_ghost.as<double>();
  v.bcf = p.d_bcf.as<double>();
  v.rowbase = p.d_rowbase.as<int64_t>();
  v.nblk = p.d_nblk.as<int32_t>();
  v.slot = p.d_slot.as<uint8_t>();
  p.on_device = true;
  return PDA_OK;
}
}  // namespace pda

extern "C" {

int pda_problem_create(pda_mesh mesh, int family, int prob, int scheme, int ic_flag,
                       const char *const *keys, const double *vals, int nparams,
                       pda_problem *out) {
  if (!mesh || !out || scheme < 0 || scheme > 2 || (nparams > 0 && (!keys || !vals))) {
    set_error("pda_problem_create: invalid argument");
    return PDA_E_INVALID_ARG;
  }
  const int m = family_m(family), dim = family_dim(family);
  if (m < 0) {
    set_error("pda_problem_create: unknown family");
    return PDA_E_INVALID_ARG;
  }
  if (dim != mesh->dim) {
    set_error("pda_problem_create: mesh dimensionality does not match the problem");
    return PDA_E_INVALID_ARG;
  }
  const int w = scheme + 1;
  // schemes_info.hpp:96-104 stencilSizeCompatibleWithInviscidFluxReconstruction
  if (mesh->stencil < 2 * w + 1) {
    set_error("pda_problem_create: mesh stencil too small for the reconstruction");
    return PDA_E_INVALID_ARG;
  }
  bool supported = false;
  const int rule = ghost_rule_of(family, prob, &supported);
  if (!supported) {
    set_error("pda_problem_create: problem enum not implemented by the GPU engine");
    return PDA_E_UNSUPPORTED;
  }
  if (family == PDA_EULER2D && prob == 4 && ic_flag <= 0) ic_flag = 1;
  if (family == PDA_EULER2D && prob == 4 && ic_flag != 1 && ic_flag != 2) {
    set_error("Euler2d: invalid icFlag for the given problem enum");
    return PDA_E_INVALID_ARG;
  }
  if (ic_flag <= 0) ic_flag = 1;
  ParamSpec spec = spec_for(family, ic_flag);
  auto *p = new pda_problem_s();
  p->family = family;
  p->prob = prob;
  p->scheme = scheme;
  p->ic_flag = ic_flag;
  p->m = m;
  p->dim = dim;
  p->w = w;
  p->phys = spec.phys_defaults;
  p->icp = spec.ic_defaults;
  for (int i = 0; i < nparams; ++i) {
    const std::string k = keys[i] ? keys[i] : "";
    auto it = std::find(spec.phys_names.begin(), spec.phys_names.end(), k);
    if (it != spec.phys_names.end()) {
      p->phys[it - spec.phys_names.begin()] = vals[i];
      continue;
    }
    auto jt = std::find(spec.ic_names.begin(), spec.ic_names.end(), k);
    if (jt != spec.ic_names.end()) {
      p->icp[jt - spec.ic_names.begin()] = vals[i];
      continue;
    }
    delete p;
    set_error("pda_problem_create: invalid parameter name '" + k + "'");
    return PDA_E_INVALID_ARG;
  }
  // Euler 1D and 3D hard-wire gamma = 1.4 (impl/euler_1d_prob_class.hpp:106,
  // impl/euler_3d_prob_class.hpp:1061)
  if (family == PDA_EULER1D || family == PDA_EULER3D) p->phys[0] = 1.4;
  p->mesh = mesh;
  mesh_retain(mesh);
  p->view.ghost_rule = rule;
  const int rc = build(*p);
  if (rc) {
    pda_problem_destroy(p);
    return rc;
  }
  *out = p;
  return PDA_OK;
}

void pda_problem_destroy(pda_problem p) {
  if (!p) return;
  pda_mesh_s *m = p->mesh;
  delete p;
  mesh_release(m);
}

int pda_problem_sizes(pda_problem p, int *ndpc, int64_t *ns, int64_t *nst, int64_t *nnz) {
  if (!p) return PDA_E_INVALID_ARG;
  if (ndpc) *ndpc = p->m;
  if (ns) *ns = p->mesh->n_sample * p->m;
  if (nst) *nst = p->mesh->n_stencil * p->m;
  if (nnz) *nnz = p->nnz;
  return PDA_OK;
}

int pda_problem_parameter(pda_problem p, const char *name, double *value) {
  if (!p || !name || !value) return PDA_E_INVALID_ARG;
  ParamSpec spec = spec_for(p->family, p->ic_flag);
  const std::string k(name);
  for (size_t i = 0; i < spec.phys_names.size(); ++i)
    if (spec.phys_names[i] == k) { *value = p->phys[i]; return PDA_OK; }
  for (size_t i = 0; i < spec.ic_names.size(); ++i)
    if (spec.ic_names[i] == k) { *value = p->icp[i]; return PDA_OK; }
  set_error("unknown parameter '" + k + "'");
  return PDA_E_INVALID_ARG;
}

int pda_initial_condition(pda_problem p, double *d_U, pda_stream stream) {
  if (!p || !d_U) return PDA_E_INVALID_ARG;
  std::vector<double> U;
  const int rc = initial_condition_host(*p, U);
  if (rc) return rc;
  PDA_HIP_CHECK(hipMemcpyAsync(d_U, U.data(), sizeof(double) * U.size(), hipMemcpyHostToDevice,
                               hipStream_t(stream)));
  PDA_HIP_CHECK(hipStreamSynchronize(hipStream_t(stream)));
  return PDA_OK;
}

int pda_jacobian_pattern(pda_problem p, int64_t *n_rows, int64_t *nnz, const int64_t **d_rowptr,
                         const int32_t **d_colidx) {
  if (!p) return PDA_E_INVALID_ARG;
  const int rc = ensure_scalar_pattern(*p);
  if (rc) return rc;
  if (n_rows) *n_rows = p->mesh->n_sample * p->m;
  if (nnz) *nnz = p->nnz;
  if (d_rowptr) *d_rowptr = p->d_rowptr.as<int64_t>();
  if (d_colidx) *d_colidx = p->d_colidx.as<int32_t>();
  return PDA_OK;
}

int pda_initial_condition_host(pda_problem p, double *U) {
  if (!p || !U) return PDA_E_INVALID_ARG;
  std::vector<double> v;
  const int rc = initial_condition_host(*p, v);
  if (rc) return rc;
  std::memcpy(U, v.data(), sizeof(double) * v.size());
  return PDA_OK;
}

int pda_jacobian_pattern_host(pda_problem p, int64_t *rowptr, int32_t *colidx) {
  if (!p) return PDA_E_INVALID_ARG;
  const int rc = ensure_scalar_pattern_host(*p);
  if (rc) return rc;
  if (rowptr) std::memcpy(rowptr, p->rowptr.data(), sizeof(int64_t) * p->rowptr.size());
  if (colidx) std::memcpy(colidx, p->colidx.data(), sizeof(int32_t) * p->colidx.size());
  return PDA_OK;
}

int pda_velocity(pda_problem p, const double *d_U, double t, double *d_V, pda_stream stream) {
  if (!p || !d_U || !d_V) {
    set_error("pda_velocity: invalid argument");
    return PDA_E_INVALID_ARG;
  }
  const int rc = ensure_device(*p);
  if (rc) return rc;
  return launch_velocity(*p, d_U, t, d_V, hipStream_t(stream));
}

int pda_problem_autotune(pda_problem p, const double *d_U, pda_stream stream, int *rows_per_strip,
                         double *us) {
  if (!p || !d_U) {
    set_error("pda_problem_autotune: invalid argument");
    return PDA_E_INVALID_ARG;
  }
  int rc = ensure_device(*p);
  if (rc) return rc;
  hipStream_t s = hipStream_t(stream);
  const bool march = p->use_box && p->vbox.valid && p->vbox.cells() > 0 && has_box_velocity(*p);
  DevBuf V;
  PDA_HIP_CHECK(V.alloc(sizeof(double) * size_t(std::max<int64_t>(p->view.nrows * p->m, 1))));
  struct Events {
    hipEvent_t e[2] = {nullptr, nullptr};
    ~Events() {
      for (hipEvent_t x : e)
        if (x) (void)hipEventDestroy(x);
    }
  } ev;
  PDA_HIP_CHECK(hipEventCreate(&ev.e[0]));
  PDA_HIP_CHECK(hipEventCreate(&ev.e[1]));
  auto time_it = [&](double *t_us) -> int {
    int r = launch_velocity(*p, d_U, 0., V.as<double>(), s); // warm-up
    if (r) return r;
    PDA_HIP_CHECK(hipEventRecord(ev.e[0], s));
    for (int k = 0; k < 10; ++k) {
      r = launch_velocity(*p, d_U, 0., V.as<double>(), s);
      if (r) return r;
    }
    PDA_HIP_CHECK(hipEventRecord(ev.e[1], s));
    PDA_HIP_CHECK(hipEventSynchronize(ev.e[1]));
    float ms = 0.f;
    PDA_HIP_CHECK(hipEventElapsedTime(&ms, ev.e[0], ev.e[1]));
    *t_us = 1e3 * double(ms) / 10.;
    return PDA_OK;
  };
  int best = 0;
  double best_us = 1e300;
  if (!march) {
    p->vel_rps = 0;
    rc = time_it(&best_us);
    if (rc) return rc;
  } else {
    for (int cand : {1, 2, 3, 4, 5, 6, 8, 12, 16, 24, 32}) {
      p->vel_rps = cand;
      double t = 0.;
      rc = time_it(&t);
      if (rc) {
        p->vel_rps = 0;
        return rc;
      }
      if (t < best_us) {
        best_us = t;
        best = cand;
      }
    }
    p->vel_rps = best;
  }
  if (rows_per_strip) *rows_per_strip = best;
  if (us) *us = best_us;
  return PDA_OK;
}

int pda_problem_set_velocity_strip_height(pda_problem p, int rows_per_strip) {
  if (!p || rows_per_strip < 0 || rows_per_strip > 64) {
    set_error("pda_problem_set_velocity_strip_height: rows_per_strip must be 0..64");
    return PDA_E_INVALID_ARG;
  }
  p->vel_rps = rows_per_strip;
  return PDA_OK;
}

int pda_problem_set_jacobian_strip_height(pda_problem p, int rows_per_strip) {
  if (!p || rows_per_strip < 0 || rows_per_strip > 64) {
    set_error("pda_problem_set_jacobian_strip_height: rows_per_strip must be 0..64");
    return PDA_E_INVALID_ARG;
  }
  p->jac_rps = rows_per_strip;
  return PDA_OK;
}

int pda_jacobian(pda_problem p, const double *d_U, double t, double *d_J, double *d_V,
                 pda_stream stream) {
  if (!p || !d_U || !d_J) {
    set_error("pda_jacobian: invalid argument");
    return PDA_E_INVALID_ARG;
  }
  const int rc = ensure_device(*p);
  if (rc) return rc;
  return launch_jacobian(*p, d_U, t, d_J, d_V, hipStream_t(stream));
}

int pda_apply_jacobian(pda_problem p, const double *d_U, double t, const double *d_B,
                       int64_t ncols, int layout, double *d_AJ, pda_stream stream) {
  if (!p || !d_U || !d_B || !d_AJ || ncols < 0 || (layout != 0 && layout != 1)) {
    set_error("pda_apply_jacobian: invalid argument");
    return PDA_E_INVALID_ARG;
  }
  int rc = ensure_scalar_pattern(*p);
  if (rc) return rc;
  if (p->d_scratch_v.bytes < sizeof(double) * size_t(p->nnz))
    PDA_HIP_CHECK(p->d_scratch_v.alloc(sizeof(double) * size_t(p->nnz)));
  rc = launch_jacobian(*p, d_U, t, p->d_scratch_v.as<double>(), nullptr, hipStream_t(stream));
  if (rc) return rc;
  return launch_apply(*p, p->d_scratch_v.as<double>(), d_B, ncols, layout, d_AJ,
                      hipStream_t(stream));
}

int pda_jacobian_blocks(pda_problem p, int *maxb, const int32_t **d_nblk, const int32_t **d_bcol) {
  if (!p) return PDA_E_INVALID_ARG;
  const int rc = ensure_device(*p);
  if (rc) return rc;
  if (maxb) *maxb = p->maxb;
  if (d_nblk) *d_nblk = p->d_nblk.as<int32_t>();
  if (d_bcol) *d_bcol = p->d_bcol.as<int32_t>();
  return PDA_OK;
}

}  // extern "C"
