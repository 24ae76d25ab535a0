// pda_mesh.cpp -- host mesh layer: reference text format I/O, the
// natural-order structured generator and the sample-mesh builder.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <sys/stat.h>

#include "pda_internal.hpp"

namespace pda {

static thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }

bool mesh_has_bd(const pda_mesh_s &m, int64_t r, int a, int side) {
  const int layers = (m.stencil - 1) / 2;
  for (int L = 0; L < layers; ++L)
    if (m.g(r, col_of(m.dim, a, side, L)) == -1) return true;
  return false;
}

bool mesh_near_bd(const pda_mesh_s &m, int64_t r) {
  for (int a = 0; a < m.dim; ++a)
    for (int s = 0; s < 2; ++s)
      if (mesh_has_bd(m, r, a, s)) return true;
  return false;
}

void mesh_retain(pda_mesh_s *m) { m->refs.fetch_add(1); }
void mesh_release(pda_mesh_s *m) {
  if (m && m->refs.fetch_sub(1) == 1) delete m;
}

// value after the "%.14f" print / std::stod read that every coordinate and
// cell width goes through (meshing_scripts/create_full_mesh.py:152-213,
// impl/mesh_read_info.hpp:84-100)
static double roundtrip14(double v) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "%.14f", v);
  return std::strtod(buf, nullptr);
}

// neighbour index along one axis, as the _buildGraphStencil{3,5,7} methods
// of meshing_scripts/_impl/natural_order_mesh_2d.py:73-268 assign it
static int64_t nb_index(int64_t i, int64_t N, int side, int L, bool periodic) {
  const int64_t o = L + 1;
  if (side == 0) {
    if (i - o >= 0) return i - o;
    return periodic ? N + i - o : -1;
  }
  if (i + o < N) return i + o;
  return periodic ? i + o - N : -1;
}

static int build_structured(pda_mesh_s &m) {
  const int dim = m.dim, w = (m.stencil - 1) / 2;
  const int64_t Nx = m.n[0], Ny = m.n[1], Nz = m.n[2];
  const int64_t N = Nx * Ny * Nz;
  if (N > INT32_MAX) {
    set_error("mesh too large for int32 cell ids");
    return PDA_E_UNSUPPORTED;
  }
  m.n_sample = m.n_stencil = N;
  const int nc = m.ncols();
  m.graph.assign(N * nc, -1);
  m.x.assign(N, 0.);
  m.y.assign(N, 0.);
  m.z.assign(N, 0.);
  double draw[3];
  for (int a = 0; a < 3; ++a) {
    draw[a] = a < dim ? (m.bounds[2 * a + 1] - m.bounds[2 * a]) / double(m.n[a]) : 0.;
    m.d[a] = a < dim ? roundtrip14(draw[a]) : 0.;
    m.dinv[a] = a < dim ? 1. / m.d[a] : 0.;
  }
  const double ox = m.bounds[0] + 0.5 * draw[0];
  const double oy = dim > 1 ? m.bounds[2] + 0.5 * draw[1] : 0.;
  const double oz = dim > 2 ? m.bounds[4] + 0.5 * draw[2] : 0.;
  for (int64_t gk = 0; gk < Nz; ++gk)
    for (int64_t gj = 0; gj < Ny; ++gj)
      for (int64_t gi = 0; gi < Nx; ++gi) {
        const int64_t id = (gk * Ny + gj) * Nx + gi;
        int32_t *row = &m.graph[id * nc];
        row[0] = int32_t(id);
        const int64_t ijk[3] = {gi, gj, gk};
        for (int a = 0; a < dim; ++a)
          for (int side = 0; side < 2; ++side)
            for (int L = 0; L < w; ++L) {
              const int64_t q = nb_index(ijk[a], m.n[a], side, L, (m.periodic_mask >> a) & 1);
              int64_t gid = -1;
              if (q >= 0) {
                int64_t c[3] = {gi, gj, gk};
                c[a] = q;
                gid = (c[2] * Ny + c[1]) * Nx + c[0];
              }
              row[col_of(dim, a, side, L)] = int32_t(gid);
            }
        m.x[id] = roundtrip14(ox + double(gi) * draw[0]);
        if (dim > 1) m.y[id] = roundtrip14(oy + double(gj) * draw[1]);
        if (dim > 2) m.z[id] = roundtrip14(oz + double(gk) * draw[2]);
      }
  m.structured = true;
  return PDA_OK;
}

}  // namespace pda

using namespace pda;

extern "C" {

const char *pda_last_error(void) { return pda::g_last_error.c_str(); }
const char *pda_version(void) { return "pda-mi355x 0.1 (gfx950)"; }

int pda_mesh_create_structured(int dim, const int64_t *n, const double *bounds,
                               int stencil, int periodic_mask, pda_mesh *out) {
  if (!out || !n || !bounds || dim < 1 || dim > 3 ||
      (stencil != 3 && stencil != 5 && stencil != 7)) {
    set_error("pda_mesh_create_structured: invalid argument");
    return PDA_E_INVALID_ARG;
  }
  auto *m = new pda_mesh_s();
  m->dim = dim;
  m->stencil = stencil;
  m->periodic_mask = periodic_mask;
  for (int a = 0; a < 3; ++a) m->n[a] = a < dim ? n[a] : 1;
  for (int a = 0; a < dim; ++a) {
    if (m->n[a] < 1) {
      delete m;
      set_error("pda_mesh_create_structured: non-positive cell count");
      return PDA_E_INVALID_ARG;
    }
    m->bounds[2 * a] = bounds[2 * a];
    m->bounds[2 * a + 1] = bounds[2 * a + 1];
  }
  m->has_bounds = true;
  for (int a = 0; a < dim; ++a)
    if (((periodic_mask >> a) & 1) && m->n[a] < (stencil - 1) / 2) {
      delete m;
      set_error("pda_mesh_create_structured: periodic axis shorter than the stencil half-width");
      return PDA_E_INVALID_ARG;
    }
  const int rc = build_structured(*m);
  if (rc) {
    delete m;
    return rc;
  }
  *out = m;
  return PDA_OK;
}

int pda_mesh_load(const char *dir, pda_mesh *out) {
  if (!dir || !out) {
    set_error("pda_mesh_load: invalid argument");
    return PDA_E_INVALID_ARG;
  }
  const std::string base(dir);
  std::ifstream info(base + "/info.dat");
  if (!info) {
    set_error("file not found " + base + "/info.dat");
    return PDA_E_IO;
  }
  auto *m = new pda_mesh_s();
  std::string line;
  int64_t nxyz[3] = {0, 0, 0};
  while (std::getline(info, line)) {
    std::istringstream ss(line);
    std::string key, val;
    ss >> key >> val;
    if (key == "dim") m->dim = std::stoi(val);
    else if (key == "dx") m->d[0] = std::stod(val);
    else if (key == "dy") m->d[1] = std::stod(val);
    else if (key == "dz") m->d[2] = std::stod(val);
    else if (key == "sampleMeshSize") m->n_sample = std::stol(val);
    else if (key == "stencilMeshSize") m->n_stencil = std::stol(val);
    else if (key == "stencilSize") m->stencil = std::stoi(val);
    else if (key == "xMin") { m->bounds[0] = std::stod(val); m->has_bounds = true; }
    else if (key == "xMax") m->bounds[1] = std::stod(val);
    else if (key == "yMin") m->bounds[2] = std::stod(val);
    else if (key == "yMax") m->bounds[3] = std::stod(val);
    else if (key == "zMin") m->bounds[4] = std::stod(val);
    else if (key == "zMax") m->bounds[5] = std::stod(val);
    else if (key == "nx") nxyz[0] = std::stol(val);
    else if (key == "ny") nxyz[1] = std::stol(val);
    else if (key == "nz") nxyz[2] = std::stol(val);
  }
  if (m->dim < 1 || m->dim > 3 || m->n_sample <= 0 || m->n_stencil <= 0 ||
      (m->stencil != 3 && m->stencil != 5 && m->stencil != 7)) {
    delete m;
    set_error("invalid info.dat in " + base);
    return PDA_E_IO;
  }
  for (int a = 0; a < m->dim; ++a) m->dinv[a] = 1. / m->d[a];

  std::ifstream coords(base + "/coordinates.dat");
  if (!coords) {
    delete m;
    set_error("file not found " + base + "/coordinates.dat");
    return PDA_E_IO;
  }
  m->x.assign(m->n_stencil, 0.);
  m->y.assign(m->n_stencil, 0.);
  m->z.assign(m->n_stencil, 0.);
  while (std::getline(coords, line)) {
    std::istringstream ss(line);
    std::string c;
    if (!(ss >> c)) continue;
    const long gid = std::stol(c);
    if (gid < 0 || gid >= m->n_stencil) {
      delete m;
      set_error("coordinates.dat: id out of range");
      return PDA_E_IO;
    }
    ss >> c; m->x[gid] = std::stod(c);
    ss >> c; m->y[gid] = std::stod(c);
    if (m->dim == 3) { ss >> c; m->z[gid] = std::stod(c); }
  }

  std::ifstream conn(base + "/connectivity.dat");
  if (!conn) {
    delete m;
    set_error("file not found " + base + "/connectivity.dat");
    return PDA_E_IO;
  }
  const int nc = m->ncols();
  m->graph.assign(m->n_sample * nc, -1);
  int64_t row = 0;
  while (std::getline(conn, line) && row < m->n_sample) {
    std::istringstream ss(line);
    std::string c;
    if (!(ss >> c)) continue;
    m->graph[row * nc] = std::stoi(c);
    for (int i = 1; i < nc; ++i) {
      if (!(ss >> c)) {
        delete m;
        set_error("connectivity.dat: short row");
        return PDA_E_IO;
      }
      m->graph[row * nc + i] = std::stoi(c);
    }
    ++row;
  }
  if (row != m->n_sample) {
    delete m;
    set_error("connectivity.dat: row count mismatch");
    return PDA_E_IO;
  }
  for (int32_t v : m->graph)
    if (v < -1 || v >= m->n_stencil) {
      delete m;
      set_error("connectivity.dat: cell id out of range");
      return PDA_E_IO;
    }
  // recognise a natural-order full mesh written by the structured generator
  if (nxyz[0] > 0 && m->n_sample == m->n_stencil) {
    int64_t N = 1;
    for (int a = 0; a < m->dim; ++a) N *= nxyz[a];
    if (N == m->n_sample) {
      pda_mesh_s t;
      t.dim = m->dim;
      t.stencil = m->stencil;
      for (int a = 0; a < 3; ++a) t.n[a] = a < m->dim ? nxyz[a] : 1;
      for (int a = 0; a < m->dim; ++a)
        if (m->g(0, col_of(m->dim, a, 0, 0)) != -1) t.periodic_mask |= 1 << a;
      for (int i = 0; i < 6; ++i) t.bounds[i] = m->bounds[i];
      if (build_structured(t) == PDA_OK && t.graph == m->graph) {
        m->structured = true;
        m->periodic_mask = t.periodic_mask;
        for (int a = 0; a < 3; ++a) m->n[a] = t.n[a];
      }
    }
  }
  *out = m;
  return PDA_OK;
}

int pda_mesh_create_sample(pda_mesh full, const int32_t *sample_gids, int64_t n,
                           pda_mesh *out) {
  if (!full || !out || (n > 0 && !sample_gids)) {
    set_error("pda_mesh_create_sample: invalid argument");
    return PDA_E_INVALID_ARG;
  }
  if (full->n_sample != full->n_stencil) {
    set_error("pda_mesh_create_sample: the input must be a full mesh");
    return PDA_E_INVALID_ARG;
  }
  // create_sample_mesh.py:39-40: sorted sample ids (OrderedDict keys: unique)
  std::vector<int32_t> smp(sample_gids, sample_gids + n);
  for (int32_t v : smp)
    if (v < 0 || v >= full->n_sample) {
      set_error("pda_mesh_create_sample: sample id out of range");
      return PDA_E_INVALID_ARG;
    }
  std::sort(smp.begin(), smp.end());
  smp.erase(std::unique(smp.begin(), smp.end()), smp.end());
  // :46-56 stencil = sorted unique(sample U non -1 neighbours)
  const int nc = full->ncols();
  std::vector<int32_t> st;
  st.reserve(smp.size() * nc);
  for (int32_t gid : smp)
    for (int c = 0; c < nc; ++c) {
      const int32_t v = full->g(gid, c);
      if (v != -1) st.push_back(v);
    }
  std::sort(st.begin(), st.end());
  st.erase(std::unique(st.begin(), st.end()), st.end());
  // :75-80 fm_to_sm map = rank in the sorted stencil list
  std::vector<int32_t> fm2sm(full->n_stencil, -1);
  for (size_t i = 0; i < st.size(); ++i) fm2sm[st[i]] = int32_t(i);
  auto *m = new pda_mesh_s();
  m->dim = full->dim;
  m->stencil = full->stencil;
  for (int a = 0; a < 3; ++a) { m->d[a] = full->d[a]; m->dinv[a] = full->dinv[a]; }
  for (int i = 0; i < 6; ++i) m->bounds[i] = full->bounds[i];
  m->has_bounds = full->has_bounds;
  m->n_sample = int64_t(smp.size());
  m->n_stencil = int64_t(st.size());
  m->graph.resize(m->n_sample * nc);
  // :101-110 graph entries renumbered to stencil-local ids, rows in sample order
  for (int64_t r = 0; r < m->n_sample; ++r)
    for (int c = 0; c < nc; ++c) {
      const int32_t v = full->g(smp[r], c);
      m->graph[r * nc + c] = v == -1 ? -1 : fm2sm[v];
    }
  m->x.resize(st.size());
  m->y.resize(st.size());
  m->z.resize(st.size());
  for (size_t i = 0; i < st.size(); ++i) {
    m->x[i] = full->x[st[i]];
    m->y[i] = full->y[st[i]];
    m->z[i] = full->z[st[i]];
  }
  m->sample_gids = smp;
  m->stencil_gids = st;
  *out = m;
  return PDA_OK;
}

int pda_mesh_sample_gids(pda_mesh m, int32_t *sample_out, int32_t *stencil_out) {
  if (!m) return PDA_E_INVALID_ARG;
  if (sample_out) std::copy(m->sample_gids.begin(), m->sample_gids.end(), sample_out);
  if (stencil_out) std::copy(m->stencil_gids.begin(), m->stencil_gids.end(), stencil_out);
  return PDA_OK;
}

int pda_mesh_write(pda_mesh m, const char *dir) {
  if (!m || !dir) return PDA_E_INVALID_ARG;
  const std::string base(dir);
  mkdir(dir, 0755);
  FILE *f = std::fopen((base + "/info.dat").c_str(), "w");
  if (!f) {
    set_error("cannot write " + base + "/info.dat");
    return PDA_E_IO;
  }
  const char *axn[3] = {"x", "y", "z"};
  std::fprintf(f, "dim %1d\n", m->dim);
  for (int a = 0; a < m->dim; ++a) {
    std::fprintf(f, "%sMin %.14f\n", axn[a], m->bounds[2 * a]);
    std::fprintf(f, "%sMax %.14f\n", axn[a], m->bounds[2 * a + 1]);
  }
  const char *dn[3] = {"dx", "dy", "dz"};
  for (int a = 0; a < m->dim; ++a) std::fprintf(f, "%s %.14f\n", dn[a], m->d[a]);
  std::fprintf(f, "sampleMeshSize %8ld\n", long(m->n_sample));
  std::fprintf(f, "stencilMeshSize %8ld\n", long(m->n_stencil));
  std::fprintf(f, "stencilSize %2d\n", m->stencil);
  if (m->structured) {
    const char *nn[3] = {"nx", "ny", "nz"};
    for (int a = 0; a < m->dim; ++a) std::fprintf(f, "%s %8ld\n", nn[a], long(m->n[a]));
  }
  std::fclose(f);
  f = std::fopen((base + "/connectivity.dat").c_str(), "w");
  if (!f) return PDA_E_IO;
  const int nc = m->ncols();
  for (int64_t r = 0; r < m->n_sample; ++r) {
    for (int c = 0; c < nc; ++c) std::fprintf(f, "%8d ", m->graph[r * nc + c]);
    std::fprintf(f, "\n");
  }
  std::fclose(f);
  f = std::fopen((base + "/coordinates.dat").c_str(), "w");
  if (!f) return PDA_E_IO;
  for (int64_t i = 0; i < m->n_stencil; ++i) {
    std::fprintf(f, "%8ld %.14f %.14f ", long(i), m->x[i], m->y[i]);
    if (m->dim == 3) std::fprintf(f, "%.14f ", m->z[i]);
    std::fprintf(f, "\n");
  }
  std::fclose(f);
  if (!m->sample_gids.empty()) {
    f = std::fopen((base + "/sample_mesh_gids.dat").c_str(), "w");
    for (int32_t v : m->sample_gids) std::fprintf(f, "%8d\n", v);
    std::fclose(f);
    f = std::fopen((base + "/stencil_mesh_gids.dat").c_str(), "w");
    for (int32_t v : m->stencil_gids) std::fprintf(f, "%8d\n", v);
    std::fclose(f);
  }
  return PDA_OK;
}

int pda_mesh_info(pda_mesh m, int *dim, int *stencil, int64_t *n_sample,
                  int64_t *n_stencil, double *d) {
  if (!m) return PDA_E_INVALID_ARG;
  if (dim) *dim = m->dim;
  if (stencil) *stencil = m->stencil;
  if (n_sample) *n_sample = m->n_sample;
  if (n_stencil) *n_stencil = m->n_stencil;
  if (d) for (int a = 0; a < 3; ++a) d[a] = m->d[a];
  return PDA_OK;
}

const int32_t *pda_mesh_graph(pda_mesh m) { return m ? m->graph.data() : nullptr; }

const double *pda_mesh_coords(pda_mesh m, int axis) {
  if (!m) return nullptr;
  return axis == 0 ? m->x.data() : axis == 1 ? m->y.data() : m->z.data();
}

int pda_mesh_row_classes(pda_mesh m, int64_t *n_inner, int64_t *n_near) {
  if (!m) return PDA_E_INVALID_ARG;
  int64_t ni = 0, nn = 0;
  for (int64_t r = 0; r < m->n_sample; ++r) (mesh_near_bd(*m, r) ? nn : ni)++;
  if (n_inner) *n_inner = ni;
  if (n_near) *n_near = nn;
  return PDA_OK;
}

int pda_mesh_boundary_rows(pda_mesh m, int which, int32_t *rows, int64_t *count) {
  if (!m || !count || which < 0 || which > 2) {
    set_error("pda_mesh_boundary_rows: invalid argument");
    return PDA_E_INVALID_ARG;
  }
  if (which == 2 && m->dim != 2) {
    // impl/mesh_ccu.hpp:296-299 throws for meshes other than 2D
    set_error("pda_mesh_boundary_rows: cells strictly next to the boundary are defined for 2D meshes only");
    return PDA_E_UNSUPPORTED;
  }
  int64_t n = 0;
  for (int64_t r = 0; r < m->n_sample; ++r) {
    const bool near = mesh_near_bd(*m, r);
    bool take = which == 0 ? near : !near;
    if (which == 2) {
      // a first-order neighbour is missing: near rows whose face itself is
      // on the boundary, impl/mesh_ccu.hpp:276-305,440-445
      take = false;
      for (int a = 0; a < 2 && !take; ++a)
        for (int side = 0; side < 2 && !take; ++side) take = m->g(r, col_of(2, a, side, 0)) == -1;
    }
    if (take) {
      if (rows) rows[n] = int32_t(r);
      ++n;
    }
  }
  *count = n;
  return PDA_OK;
}

void pda_mesh_destroy(pda_mesh m) { mesh_release(m); }

}  // extern "C"
