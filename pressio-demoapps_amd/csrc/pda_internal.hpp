// pda_internal.hpp -- host-side types shared by the mesh, problem and kernel
// translation units of libpda.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <string>
#include <vector>

#include "pda_c_api.h"

namespace pda {

void set_error(const std::string &msg);

#define PDA_HIP_CHECK(expr)                                                    \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess) {                                                    \
      ::pda::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));     \
      return PDA_E_HIP;                                                        \
    }                                                                          \
  } while (0)

// Owned device allocation (freed with the owning object).
struct DevBuf {
  void *ptr = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { reset(); }
  void reset() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
  hipError_t alloc(size_t n) {
    reset();
    bytes = n;
    if (n == 0) return hipSuccess;
    return hipMalloc(&ptr, n);
  }
  template <class T>
  hipError_t upload(const std::vector<T> &v) {
    hipError_t e = alloc(v.size() * sizeof(T));
    if (e != hipSuccess || v.empty()) return e;
    return hipMemcpy(ptr, v.data(), bytes, hipMemcpyHostToDevice);
  }
  template <class T>
  T *as() const { return static_cast<T *>(ptr); }
};

}  // namespace pda

// Cell-centred uniform mesh, the data of impl/mesh_ccu.hpp:60-473.
struct pda_mesh_s {
  std::atomic<int> refs{1};
  int dim = 0;
  int stencil = 0;            // mesh stencil size 3, 5 or 7
  int64_t n_sample = 0;       // graph rows
  int64_t n_stencil = 0;      // state cells
  double d[3] = {0, 0, 0};    // dx dy dz as stored in info.dat
  double dinv[3] = {0, 0, 0}; // 1/d, impl/mesh_read_info.hpp:84-100
  double bounds[6] = {0, 0, 0, 0, 0, 0};
  bool has_bounds = false;
  std::vector<int32_t> graph; // n_sample x ncols()
  std::vector<double> x, y, z;
  // natural-order structured description, when known
  bool structured = false;
  int64_t n[3] = {1, 1, 1};
  int periodic_mask = 0;
  // sample-mesh provenance (ids w.r.t. the full mesh), empty for full meshes
  std::vector<int32_t> sample_gids, stencil_gids;

  int ncols() const { return 1 + (stencil - 1) * dim; }
  int32_t g(int64_t r, int c) const { return graph[r * ncols() + c]; }
};

namespace pda {
// graph column of the layer-L neighbour on side (0 minus, 1 plus) of axis a;
// layouts of meshing_scripts/_impl/{one_dim_mesh,natural_order_mesh_2d,
// natural_order_mesh_3d}.py and impl/functor_fill_stencil.hpp
inline int col_of(int dim, int a, int side, int L) {
  if (dim == 1) return 1 + 2 * L + side;
  if (dim == 2) {
    const int base = 1 + 4 * L;
    return a == 0 ? base + (side ? 2 : 0) : base + (side ? 1 : 3);
  }
  const int base = 1 + 6 * L;
  if (a == 0) return base + (side ? 2 : 0);
  if (a == 1) return base + (side ? 1 : 3);
  return base + (side ? 5 : 4);
}
// "boundary within the mesh stencil width" (impl/mesh_ccu.hpp:162-296)
bool mesh_has_bd(const pda_mesh_s &m, int64_t r, int a, int side);
bool mesh_near_bd(const pda_mesh_s &m, int64_t r);
void mesh_retain(pda_mesh_s *m);
void mesh_release(pda_mesh_s *m);
}  // namespace pda

namespace pda {

// Plain-data view of a problem handed to every kernel by value.
struct View {
  const int32_t *graph = nullptr; // nrows x ncols
  int ncols = 0;
  int64_t nrows = 0;
  const int32_t *ghost_row = nullptr; // per graph row: near-boundary index or -1
  const int32_t *near_rows = nullptr; // graph rows of near-boundary cells
  int64_t n_near = 0;
  double *ghost = nullptr;            // [axis][side][n_near][w*m]
  const double *bcf = nullptr;        // [n_near][dim][m] Jacobian BC factors
  const int64_t *rowbase = nullptr;   // per graph row: first CSR value of dof 0
  const int32_t *nblk = nullptr;      // per graph row: number of column blocks
  const uint8_t *slot = nullptr;      // per graph row: block slot of graph column c
  int npos = 0;                       // columns covered by the slot table
  int ghost_rule = 0;                 // 0 none, 1 opposite-cell, 2 mirror
  int dup = 0;                        // some row maps two columns to one slot
  double hinv[3] = {0, 0, 0};
  double p[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  double gsign[3][2][5];              // ghost value sign per axis, side, dof
  // rows the generic kernels visit: all rows when rows == nullptr
  const int32_t *rows = nullptr;
  int64_t nvisit = 0;
};

// Index box of a natural-order structured mesh whose cells are inner rows
// with their whole scheme stencil inside the grid (no periodic wrap): the
// structured kernels own these rows, the generic kernels the rest.
struct Box {
  int64_t lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0}; // half-open, per axis
  int64_t n[3] = {1, 1, 1};                     // grid extent
  // how a stencil index outside [0, n) along the axis is read:
  // 0 never happens (the box keeps stencils inside), 1 periodic wrap,
  // 2 mirror image about the boundary times the side's ghost sign
  int mode[3] = {0, 0, 0};
  bool valid = false;
  int64_t cells() const {
    return valid ? (hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]) : 0;
  }
};

enum GhostRule { GHOST_NONE = 0, GHOST_OPPOSITE = 1, GHOST_MIRROR = 2 };

// one instance of a batched evaluation: its problem view and velocity box,
// the offset of its state slice, its first workgroup and strip height
struct BatchInst {
  View P;
  Box B;
  int64_t off = 0;
  int64_t blk0 = 0;
  int rps = 1;
};

// An explicit stage fused into a velocity evaluation: with v the velocity
// of entry i, kout[i] = v (when kout is set) and
//   out[i] = c[0] x[0][i] + ... + c[nx-1] x[nx-1][i] + cv v,
// accumulated in that order (the order of the separate stage combination).
// out must not be the vector the velocity reads.
struct RkStage {
  double *kout = nullptr;
  double *out = nullptr;
  const double *x[4] = {nullptr, nullptr, nullptr, nullptr};
  double c[4] = {0., 0., 0., 0.};
  int nx = 0;
  double cv = 0.;
};

}  // namespace pda

struct pda_problem_s {
  pda_mesh_s *mesh = nullptr;
  int family = 0, prob = 0, scheme = 0, ic_flag = 1;
  int m = 0, dim = 0, w = 0;
  std::vector<double> phys;          // physical parameters (family order)
  std::vector<double> icp;           // initial-condition parameters
  int64_t n_inner = 0, n_near = 0;
  int maxb = 0;                      // max column blocks in a row
  int64_t nnz = 0;
  std::vector<int64_t> rowptr;       // host copy of the scalar CSR
  std::vector<int32_t> colidx;
  // host tables (built at creation, uploaded on first evaluation)
  std::vector<int32_t> h_ghost_row, h_near_rows, h_nblk, h_bcol;
  std::vector<int64_t> h_rowbase;
  std::vector<uint8_t> h_slot;
  std::vector<double> h_bcf;
  bool on_device = false;
  // structured fast path
  pda::Box box;
  std::vector<int32_t> h_rem_rows;   // rows outside the box
  pda::DevBuf d_rem_rows;
  // velocity box: with periodic or mirror-ghost axes it spans the whole
  // grid (ghost values of those rules do not depend on the row reading
  // them), else the same interior box
  pda::Box vbox;
  std::vector<int32_t> h_vrem_rows;  // rows outside the velocity box
  pda::DevBuf d_vrem_rows;
  int use_box = 1;                   // 0 forces the generic kernels everywhere
  int vel_rps = 0;                   // velocity march strip height (0: from the grid size)
  int jac_rps = 0;                   // first-order 2D/3D Euler Jacobian strip height (0: from the grid size)
  pda::View view;
  pda::DevBuf d_graph, d_ghost_row, d_near_rows, d_ghost, d_bcf;
  pda::DevBuf d_rowptr, d_colidx, d_rowbase, d_nblk, d_slot, d_bcol;
  pda::DevBuf d_scratch_v;           // Jacobian values for applyJacobian()
};

namespace pda {
int ensure_device(pda_problem_s &p);
// the scalar CSR pattern (d_rowptr, d_colidx) on the device
int ensure_pattern_device(pda_problem_s &p);
int launch_velocity(const pda_problem_s &p, const double *U, double t, double *V, hipStream_t s);
// velocity of every row with an explicit stage fused into its store, for a
// problem whose velocity is the generic row kernel alone (no structured
// box); PDA_E_UNSUPPORTED otherwise
bool velocity_stage_fits(const pda_problem_s &p);
int launch_velocity_stage(const pda_problem_s &p, const double *U, const RkStage &S, hipStream_t s);
// explicit steps of a 1D problem, up to steps1d_per_launch per launch
// (Uin != Uout)
constexpr int STEPS1D_MAX = 8;
bool steps1d_fits(const pda_problem_s &p, int scheme, bool *periodic);
int steps1d_per_launch(const pda_problem_s &p, int scheme);
int launch_steps1d(const pda_problem_s &p, int scheme, const double *Uin, double *Uout, int nsteps, double dt,
                   bool periodic, hipStream_t s);
int launch_jacobian(const pda_problem_s &p, const double *U, double t, double *J,
                    double *V, hipStream_t s);
int launch_apply(const pda_problem_s &p, const double *J, const double *B, int64_t ncols,
                 int layout, double *AJ, hipStream_t s);
// structured-box kernels; return PDA_E_UNSUPPORTED when the problem has no
// box kernel (the caller then covers all rows with the generic kernels)
int launch_velocity_box(const pda_problem_s &p, const double *U, double *V, hipStream_t s);
// the 2D box march forming an explicit stage in its store
int launch_velocity_box_stage(const pda_problem_s &p, const double *U, const RkStage &S, hipStream_t s);
// rest_done: the launch also covered the rows outside the box
int launch_jacobian_box(const pda_problem_s &p, const double *U, double *J, double *V,
                        hipStream_t s, bool *rest_done);
int64_t velocity_march_blocks(const Box &B, int dim, int w, int rows_per_strip);
int64_t velocity_strip_height(const Box &B, int dim, int w);
int64_t jacobian_strip_height(const Box &B);
// first-order 3D Euler Jacobian over the cells of p.box of a structured
// full mesh; the rows outside the box are left to the caller
int launch_jacobian_fo_euler3d_box(const pda_problem_s &p, const double *U, double *J, hipStream_t s);
int launch_velocity_batch(int family, int w, const BatchInst *inst, const int32_t *blk, int64_t nblocks,
                          const double *U, double *V, hipStream_t s);
int launch_velocity_batch_stage(int family, int w, const BatchInst *inst, const int32_t *blk, int64_t nblocks,
                                const double *U, const RkStage &S, hipStream_t s);
// view of the generic kernels: every row (rows == nullptr) or a row list
View generic_view(const pda_problem_s &p, const pda::DevBuf *rows, int64_t count);
bool has_box_kernels(const pda_problem_s &p);
// the velocity march also covers 3D Euler
bool has_box_velocity(const pda_problem_s &p);
}  // namespace pda
