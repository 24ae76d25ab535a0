// pda_kernels.hip -- gfx950 kernels of the finite-volume hot path:
// ghost fill, velocity (residual), velocity+Jacobian assembly into the fixed
// CSR pattern, and the Jacobian action.
//
// One thread owns one graph row (one sample-mesh cell) and produces all of
// its V entries and all of its Jacobian block-row, so every output value is
// written exactly once with plain stores (no atomics, no zero-fill pass);
// the per-entry accumulation order follows the reference:
//   - V: x, then y, then z axis, then sources
//     (impl/mixin_directional_flux_balance.hpp:68-84);
//   - interior Jacobian: chain rule flux-Jacobian x reconstruction gradient,
//     left face then right face per axis
//     (impl/mixin_directional_flux_balance_jacobian.hpp:142-284);
//   - near-boundary rows: first-order blocks with BC factors folded into the
//     diagonal block (impl/mixin_directional_flux_balance_jacobian.hpp:311-371).
#include "pda_internal.hpp"
#include "pda_physics.hpp"

namespace pda {
namespace {

// ------------------------------------------------------------------ physics
// Each family: M dofs per cell, DIM axes, face flux and its Jacobians for
// axis A, and the per-cell source terms applied after the axis sweep.
struct PhEuler1 {
  static constexpr int M = 3, DIM = 1;
  template <int A>
  __device__ static void flux(const View &P, const double *l, const double *r, double *F) {
    phys::euler_flux<1, A>(l, r, P.p[0], F);
  }
  template <int A>
  __device__ static void fjac(const View &P, const double *l, const double *r, double *JL, double *JR) {
    phys::euler_fjac<1, A>(l, r, P.p[0], JL, JR);
  }
};
struct PhEuler2 {
  static constexpr int M = 4, DIM = 2;
  template <int A>
  __device__ static void flux(const View &P, const double *l, const double *r, double *F) {
    phys::euler_flux<2, A>(l, r, P.p[0], F);
  }
  template <int A>
  __device__ static void fjac(const View &P, const double *l, const double *r, double *JL, double *JR) {
    phys::euler_fjac<2, A>(l, r, P.p[0], JL, JR);
  }
};
struct PhEuler3 {
  static constexpr int M = 5, DIM = 3;
  template <int A>
  __device__ static void flux(const View &P, const double *l, const double *r, double *F) {
    phys::euler_flux<3, A>(l, r, P.p[0], F);
  }
  template <int A>
  __device__ static void fjac(const View &P, const double *l, const double *r, double *JL, double *JR) {
    phys::euler_fjac<3, A>(l, r, P.p[0], JL, JR);
  }
};
struct PhSwe {
  static constexpr int M = 3, DIM = 2;
  template <int A>
  __device__ static void flux(const View &P, const double *l, const double *r, double *F) {
    phys::swe_flux<A>(l, r, P.p[0], F);
  }
  template <int A>
  __device__ static void fjac(const View &P, const double *l, const double *r, double *JL, double *JR) {
    phys::swe_fjac<A>(l, r, P.p[0], JL, JR);
  }
};
struct PhBurgers {
  static constexpr int M = 2, DIM = 2;
  template <int A>
  __device__ static void flux(const View &, const double *l, const double *r, double *F) {
    phys::burgers_flux<A>(l, r, F);
  }
  template <int A>
  __device__ static void fjac(const View &, const double *l, const double *r, double *JL, double *JR) {
    phys::burgers_fjac<A>(l, r, JL, JR);
  }
};

template <class Ph> struct Sources {
  // V sources after the axis sweep; the Jacobian part updates the diagonal
  // block held in registers (self) and returns the diagonal addend of
  // first-neighbour blocks through offdiag()
  __device__ static void apply(const View &, int64_t, bool, const double *, int32_t,
                               double *, double *) {}
  __device__ static double offdiag(const View &, bool, int, int, int) { return 0.; }
};

// Coriolis forcing, impl/swe_2d_prob_class.hpp:984-1012
template <> struct Sources<PhSwe> {
  __device__ static void apply(const View &P, int64_t, bool, const double *U, int32_t self,
                               double *v, double *J) {
    const double f = P.p[1];
    const double h = U[3 * int64_t(self)], hu = U[3 * int64_t(self) + 1], hv = U[3 * int64_t(self) + 2];
    v[1] -= f * hv / h;
    v[2] += f * hu / h;
    if (J) {
      J[1 * 3 + 0] += f * hv / (h * h);
      J[1 * 3 + 2] += -f / h;
      J[2 * 3 + 1] += f / h;
      J[2 * 3 + 0] += -f * hu / (h * h);
    }
  }
  __device__ static double offdiag(const View &, bool, int, int, int) { return 0.; }
};

// Burgers viscous term on inner cells,
// impl/advection_diffusion_2d_prob_class.hpp:1167-1204
template <> struct Sources<PhBurgers> {
  __device__ static void apply(const View &P, int64_t r, bool near, const double *U, int32_t self,
                               double *v, double *J) {
    if (near) return;
    const int32_t *gr = P.graph + r * P.ncols;
    const double dX = P.p[0] * (P.hinv[0] * P.hinv[0]);
    const double dY = P.p[0] * (P.hinv[1] * P.hinv[1]);
    const int64_t ui = 2 * int64_t(self), uL = 2 * int64_t(gr[1]), uF = 2 * int64_t(gr[2]);
    const int64_t uR = 2 * int64_t(gr[3]), uB = 2 * int64_t(gr[4]);
    for (int k = 0; k < 2; ++k) {
      v[k] += dX * (U[uR + k] - 2. * U[ui + k] + U[uL + k]);
      v[k] += dY * (U[uF + k] - 2. * U[ui + k] + U[uB + k]);
    }
    if (J) {
      J[0] += -2. * dX - 2. * dY;
      J[3] += -2. * dX - 2. * dY;
    }
  }
  // the diffusion adds to the diagonal of the four first-neighbour blocks
  __device__ static double offdiag(const View &P, bool near, int a, int k, int j) {
    if (near || k != j) return 0.;
    return P.p[0] * (P.hinv[a] * P.hinv[a]);
  }
};

// ------------------------------------------------------------------ helpers
template <int DIM>
__device__ __forceinline__ int colof(int a, int side, int L) {
  if (DIM == 1) return 1 + 2 * L + side;
  if (DIM == 2) return a == 0 ? 1 + 4 * L + (side ? 2 : 0) : 1 + 4 * L + (side ? 1 : 3);
  const int base = 1 + 6 * L;
  return a == 0 ? base + (side ? 2 : 0) : a == 1 ? base + (side ? 1 : 3) : base + (side ? 5 : 4);
}

__device__ __forceinline__ const double *ghost_ptr(const View &P, int a, int side, int64_t g,
                                                   int wm) {
  return P.ghost + ((int64_t(a) * 2 + side) * P.n_near + g) * wm;
}

// st[p][k], p = 0..2W, centre at W.  Missing neighbours (graph -1) of a
// near-boundary row are read from its ghost layers, as
// impl/functor_fill_stencil.hpp does.  GW is the ghost-layer width the ghost
// arrays were filled with (the scheme's), W the width gathered.
template <class Ph, int W, int A>
__device__ __forceinline__ void gather(const View &P, const int32_t *gr, int64_t g, int GW,
                                       const double *__restrict__ U, double st[][Ph::M]) {
  constexpr int M = Ph::M;
  const int64_t self = gr[0];
#pragma unroll
  for (int k = 0; k < M; ++k) st[W][k] = U[self * M + k];
#pragma unroll
  for (int side = 0; side < 2; ++side)
#pragma unroll
    for (int L = 0; L < W; ++L) {
      const int32_t nb = gr[colof<Ph::DIM>(A, side, L)];
      const int p = side ? W + 1 + L : W - 1 - L;
      if (nb >= 0) {
#pragma unroll
        for (int k = 0; k < M; ++k) st[p][k] = U[int64_t(nb) * M + k];
      } else {
        const double *gp = ghost_ptr(P, A, side, g, GW * M) + L * M;
#pragma unroll
        for (int k = 0; k < M; ++k) st[p][k] = gp[k];
      }
    }
}

// face states of the cell's left (minus) and right (plus) faces
template <int M, int W>
struct Faces {
  double lm[M], lp[M], rm[M], rp[M];
};
template <int M, int W>
struct FaceGrads {
  double glm[M][2 * W], glp[M][2 * W], grm[M][2 * W], grp[M][2 * W];
};

template <int M, int W>
__device__ __forceinline__ void reconstruct(const double st[][M], Faces<M, W> &f) {
#pragma unroll
  for (int k = 0; k < M; ++k) {
    double q[2 * W + 1];
#pragma unroll
    for (int p = 0; p <= 2 * W; ++p) q[p] = st[p][k];
    if constexpr (W == 1) {
      f.lm[k] = q[0]; f.lp[k] = q[1]; f.rm[k] = q[1]; f.rp[k] = q[2];
    } else if constexpr (W == 2) {
      phys::weno3(q, f.lm[k], f.lp[k]);
      phys::weno3(q + 1, f.rm[k], f.rp[k]);
    } else {
      phys::weno5(q, f.lm[k], f.lp[k]);
      phys::weno5(q + 1, f.rm[k], f.rp[k]);
    }
  }
}

template <int M, int W>
__device__ __forceinline__ void reconstruct_grad(const double st[][M], Faces<M, W> &f,
                                                 FaceGrads<M, W> &G) {
#pragma unroll
  for (int k = 0; k < M; ++k) {
    double q[2 * W + 1];
#pragma unroll
    for (int p = 0; p <= 2 * W; ++p) q[p] = st[p][k];
    if constexpr (W == 2) {
      phys::weno3_grad(q, f.lm[k], f.lp[k], G.glm[k], G.glp[k]);
      phys::weno3_grad(q + 1, f.rm[k], f.rp[k], G.grm[k], G.grp[k]);
    } else {
      phys::weno5_grad(q, f.lm[k], f.lp[k], G.glm[k], G.glp[k]);
      phys::weno5_grad(q + 1, f.rm[k], f.rp[k], G.grm[k], G.grp[k]);
    }
  }
}

// store (or, when a row maps two stencil columns onto one block, add) one
// m x m block at a slot of the row
template <int M>
__device__ __forceinline__ void put_block(const View &P, double *__restrict__ J, int64_t base,
                                          int rs, int s, const double (&b)[M][M]) {
#pragma unroll
  for (int k = 0; k < M; ++k)
#pragma unroll
    for (int j = 0; j < M; ++j) {
      double *dst = J + base + int64_t(k) * rs + s * M + j;
      if (P.dup) *dst += b[k][j];
      else *dst = b[k][j];
    }
}

// ---------------------------------------------------------------- ghosts
// Ghost layers of every near-boundary row.  Layer 0 copies the cell itself;
// for layer L>0 the source cell follows the family's rule:
//   GHOST_OPPOSITE: the opposite layer-(L-1) neighbour
//     (impl/euler_1d_ghost_filler.hpp:104-143, impl/euler_3d_ghost_filler_sedov.hpp:184-236,
//      impl/advection_diffusion_reaction_2d_ghost_filler_problemA.hpp:106-133);
//   GHOST_MIRROR: the mirror image about the boundary
//     (impl/euler_2d_ghost_filler_neumann.hpp:163-279,
//      impl/swe_2d_ghost_filler_inviscid_wall.hpp:160-270);
// each value is multiplied by the side's sign pattern (reflective walls,
// Dirichlet negation).  A slot is written only when its neighbour is -1,
// the same condition under which the stencil reads it.
template <int DIM, int M>
__global__ void k_ghosts(View P, int W, const double *__restrict__ U) {
  const int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (g >= P.n_near) return;
  const int64_t r = P.near_rows[g];
  const int32_t *gr = P.graph + r * P.ncols;
  const int32_t self = gr[0];
  for (int a = 0; a < DIM; ++a)
    for (int side = 0; side < 2; ++side) {
      double *dst = P.ghost + ((int64_t(a) * 2 + side) * P.n_near + g) * (W * M);
      const double *sg = P.gsign[a][side];
      for (int L = 0; L < W; ++L) {
        const int32_t nL = gr[colof<DIM>(a, side, L)];
        if (nL != -1) continue;
        int32_t src = self;
        if (L > 0) {
          if (P.ghost_rule == GHOST_OPPOSITE) {
            src = gr[colof<DIM>(a, 1 - side, L - 1)];
          } else {
            const int32_t n0 = gr[colof<DIM>(a, side, 0)];
            const int32_t o0 = gr[colof<DIM>(a, 1 - side, 0)];
            if (L == 1) {
              src = n0 == -1 ? o0 : n0;
            } else {
              const int32_t n1 = gr[colof<DIM>(a, side, 1)];
              const int32_t o1 = gr[colof<DIM>(a, 1 - side, 1)];
              if (n1 != -1 && n0 != -1) src = n1;
              if (n1 == -1 && n0 != -1) src = self;
              if (n1 == -1 && n0 == -1) src = o1;
            }
          }
        }
        if (src < 0) src = self; // degenerate meshes narrower than the stencil
        for (int k = 0; k < M; ++k) dst[L * M + k] = sg[k] * U[int64_t(src) * M + k];
      }
    }
}

// ---------------------------------------------------------------- velocity
template <class Ph, int W, int A>
__device__ __forceinline__ void velocity_axis(const View &P, const int32_t *gr, int64_t g,
                                              const double *__restrict__ U, double *acc) {
  constexpr int M = Ph::M;
  double st[2 * W + 1][M];
  gather<Ph, W, A>(P, gr, g, W, U, st);
  Faces<M, W> f;
  reconstruct<M, W>(st, f);
  double FL[M], FR[M];
  Ph::template flux<A>(P, f.lm, f.lp, FL);
  Ph::template flux<A>(P, f.rm, f.rp, FR);
  const double h = P.hinv[A];
#pragma unroll
  for (int k = 0; k < M; ++k) acc[k] += h * (FL[k] - FR[k]);
}

template <class Ph, int W>
__global__ void __launch_bounds__(256) k_velocity(View P, const double *__restrict__ U,
                                                   double *__restrict__ V) {
  constexpr int M = Ph::M;
  const int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= P.nrows) return;
  const int32_t *gr = P.graph + r * P.ncols;
  const int64_t g = P.ghost_row[r];
  double acc[M];
#pragma unroll
  for (int k = 0; k < M; ++k) acc[k] = 0.;
  velocity_axis<Ph, W, 0>(P, gr, g, U, acc);
  if constexpr (Ph::DIM > 1) velocity_axis<Ph, W, 1>(P, gr, g, U, acc);
  if constexpr (Ph::DIM > 2) velocity_axis<Ph, W, 2>(P, gr, g, U, acc);
  Sources<Ph>::apply(P, r, g >= 0, U, gr[0], acc, nullptr);
#pragma unroll
  for (int k = 0; k < M; ++k) V[r * M + k] = acc[k];
}

// ---------------------------------------------------------------- jacobian
// interior row, one axis
template <class Ph, int W, int A>
__device__ __forceinline__ void jac_axis_inner(const View &P, const int32_t *gr,
                                               const uint8_t *sl, const double *__restrict__ U,
                                               double *acc, double (&self)[Ph::M][Ph::M],
                                               double *__restrict__ J, int64_t base, int rs) {
  constexpr int M = Ph::M;
  double st[2 * W + 1][M];
  gather<Ph, W, A>(P, gr, -1, W, U, st);
  Faces<M, W> f;
  double FL[M], FR[M];
  double JLn[M * M], JLp[M * M], JRn[M * M], JRp[M * M];
  const double h = P.hinv[A];
  if constexpr (W == 1) {
    reconstruct<M, W>(st, f);
    Ph::template flux<A>(P, f.lm, f.lp, FL);
    Ph::template flux<A>(P, f.rm, f.rp, FR);
    Ph::template fjac<A>(P, f.lm, f.lp, JLn, JLp);
    Ph::template fjac<A>(P, f.rm, f.rp, JRn, JRp);
#pragma unroll
    for (int k = 0; k < M; ++k) acc[k] += h * (FL[k] - FR[k]);
    double bl[M][M], br[M][M];
#pragma unroll
    for (int k = 0; k < M; ++k)
#pragma unroll
      for (int j = 0; j < M; ++j) {
        bl[k][j] = 0. + JLn[k * M + j] * h + Sources<Ph>::offdiag(P, false, A, k, j);
        self[k][j] += (JLp[k * M + j] - JRn[k * M + j]) * h;
        br[k][j] = 0. + (-JRp[k * M + j]) * h + Sources<Ph>::offdiag(P, false, A, k, j);
      }
    put_block<M>(P, J, base, rs, sl[colof<Ph::DIM>(A, 0, 0)], bl);
    put_block<M>(P, J, base, rs, sl[colof<Ph::DIM>(A, 1, 0)], br);
  } else {
    FaceGrads<M, W> G;
    reconstruct_grad<M, W>(st, f, G);
    Ph::template flux<A>(P, f.lm, f.lp, FL);
    Ph::template flux<A>(P, f.rm, f.rp, FR);
    Ph::template fjac<A>(P, f.lm, f.lp, JLn, JLp);
    Ph::template fjac<A>(P, f.rm, f.rp, JRn, JRp);
#pragma unroll
    for (int k = 0; k < M; ++k) acc[k] += h * (FL[k] - FR[k]);
    // stencil point p (0..2W) collects the left face's gradient index p and
    // the right face's index p-1
#pragma unroll
    for (int p = 0; p <= 2 * W; ++p) {
      double b[M][M];
#pragma unroll
      for (int k = 0; k < M; ++k)
#pragma unroll
        for (int j = 0; j < M; ++j) {
          double v = (p == W) ? self[k][j] : 0.;
          if (p < 2 * W) {
            v += JLn[k * M + j] * G.glm[j][p] * h;
            v += JLp[k * M + j] * G.glp[j][p] * h;
          }
          if (p > 0) {
            v -= JRn[k * M + j] * G.grm[j][p - 1] * h;
            v -= JRp[k * M + j] * G.grp[j][p - 1] * h;
          }
          if (p == W - 1 || p == W + 1) v += Sources<Ph>::offdiag(P, false, A, k, j);
          b[k][j] = v;
        }
      if (p == W) {
#pragma unroll
        for (int k = 0; k < M; ++k)
#pragma unroll
          for (int j = 0; j < M; ++j) self[k][j] = b[k][j];
      } else {
        const int side = p > W ? 1 : 0;
        const int L = p > W ? p - W - 1 : W - 1 - p;
        put_block<M>(P, J, base, rs, sl[colof<Ph::DIM>(A, side, L)], b);
      }
    }
  }
}

// near-boundary row, one axis: velocity from the scheme's stencil (ghosts
// for missing cells), Jacobian first order with the BC factors
template <class Ph, int W, int A>
__device__ __forceinline__ void jac_axis_near(const View &P, const int32_t *gr, int64_t g,
                                              const uint8_t *sl, const double *__restrict__ U,
                                              double *acc, double (&self)[Ph::M][Ph::M],
                                              double *__restrict__ J, int64_t base, int rs) {
  constexpr int M = Ph::M;
  const double h = P.hinv[A];
  double FL[M], FR[M];
  double st1[3][M];
  gather<Ph, 1, A>(P, gr, g, W, U, st1);
  Faces<M, 1> f1;
  reconstruct<M, 1>(st1, f1);
  if constexpr (W == 1) {
    Ph::template flux<A>(P, f1.lm, f1.lp, FL);
    Ph::template flux<A>(P, f1.rm, f1.rp, FR);
  } else {
    double st[2 * W + 1][M];
    gather<Ph, W, A>(P, gr, g, W, U, st);
    Faces<M, W> f;
    reconstruct<M, W>(st, f);
    Ph::template flux<A>(P, f.lm, f.lp, FL);
    Ph::template flux<A>(P, f.rm, f.rp, FR);
  }
#pragma unroll
  for (int k = 0; k < M; ++k) acc[k] += h * (FL[k] - FR[k]);
  double JLn[M * M], JLp[M * M], JRn[M * M], JRp[M * M];
  Ph::template fjac<A>(P, f1.lm, f1.lp, JLn, JLp);
  Ph::template fjac<A>(P, f1.rm, f1.rp, JRn, JRp);
  const int32_t l0 = gr[colof<Ph::DIM>(A, 0, 0)], r0 = gr[colof<Ph::DIM>(A, 1, 0)];
  const double *fac = P.bcf + (g * Ph::DIM + A) * M;
#pragma unroll
  for (int k = 0; k < M; ++k)
#pragma unroll
    for (int j = 0; j < M; ++j) self[k][j] += (JLp[k * M + j] - JRn[k * M + j]) * h;
  if (l0 != -1) {
    double b[M][M];
#pragma unroll
    for (int k = 0; k < M; ++k)
#pragma unroll
      for (int j = 0; j < M; ++j) b[k][j] = 0. + JLn[k * M + j] * h;
    put_block<M>(P, J, base, rs, sl[colof<Ph::DIM>(A, 0, 0)], b);
  }
  if (r0 != -1) {
    double b[M][M];
#pragma unroll
    for (int k = 0; k < M; ++k)
#pragma unroll
      for (int j = 0; j < M; ++j) b[k][j] = 0. + (-JRp[k * M + j]) * h;
    put_block<M>(P, J, base, rs, sl[colof<Ph::DIM>(A, 1, 0)], b);
  }
  if (l0 == -1) {
#pragma unroll
    for (int k = 0; k < M; ++k)
#pragma unroll
      for (int j = 0; j < M; ++j) self[k][j] += (fac[j] * JLn[k * M + j]) * h;
  }
  if (r0 == -1) {
#pragma unroll
    for (int k = 0; k < M; ++k)
#pragma unroll
      for (int j = 0; j < M; ++j) self[k][j] += (fac[j] * -JRp[k * M + j]) * h;
  }
}

template <class Ph, int W>
__global__ void __launch_bounds__(256) k_jacobian(View P, const double *__restrict__ U,
                                                   double *__restrict__ V, double *__restrict__ J) {
  constexpr int M = Ph::M;
  const int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= P.nrows) return;
  const int32_t *gr = P.graph + r * P.ncols;
  const int64_t g = P.ghost_row[r];
  const uint8_t *sl = P.slot + r * P.npos;
  const int64_t base = P.rowbase[r];
  const int rs = P.nblk[r] * M;
  double acc[M];
  double self[M][M];
#pragma unroll
  for (int k = 0; k < M; ++k) {
    acc[k] = 0.;
#pragma unroll
    for (int j = 0; j < M; ++j) self[k][j] = 0.;
  }
  if (g < 0) {
    jac_axis_inner<Ph, W, 0>(P, gr, sl, U, acc, self, J, base, rs);
    if constexpr (Ph::DIM > 1) jac_axis_inner<Ph, W, 1>(P, gr, sl, U, acc, self, J, base, rs);
    if constexpr (Ph::DIM > 2) jac_axis_inner<Ph, W, 2>(P, gr, sl, U, acc, self, J, base, rs);
  } else {
    jac_axis_near<Ph, W, 0>(P, gr, g, sl, U, acc, self, J, base, rs);
    if constexpr (Ph::DIM > 1) jac_axis_near<Ph, W, 1>(P, gr, g, sl, U, acc, self, J, base, rs);
    if constexpr (Ph::DIM > 2) jac_axis_near<Ph, W, 2>(P, gr, g, sl, U, acc, self, J, base, rs);
  }
  Sources<Ph>::apply(P, r, g >= 0, U, gr[0], acc, &self[0][0]);
  put_block<M>(P, J, base, rs, sl[0], self);
  if (V) {
#pragma unroll
    for (int k = 0; k < M; ++k) V[r * M + k] = acc[k];
  }
}

// ------------------------------------------------- Jacobian action (SpMM)
// AJ = J * B.  One thread per (scalar row, column of B).
__global__ void k_apply(const int64_t *__restrict__ rowptr, const int32_t *__restrict__ col,
                        const double *__restrict__ J, const double *__restrict__ B,
                        int64_t nrows, int64_t ncolsB, int64_t ldB_rows, int layout,
                        double *__restrict__ AJ) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= nrows * ncolsB) return;
  const int64_t row = layout == 0 ? t % nrows : t / ncolsB;
  const int64_t c = layout == 0 ? t / nrows : t % ncolsB;
  double s = 0.;
  for (int64_t e = rowptr[row]; e < rowptr[row + 1]; ++e) {
    const int64_t bi = col[e];
    s += J[e] * (layout == 0 ? B[c * ldB_rows + bi] : B[bi * ncolsB + c]);
  }
  if (layout == 0) AJ[c * nrows + row] = s;
  else AJ[row * ncolsB + c] = s;
}

}  // namespace

int launch_ghosts(const pda_problem_s &p, const double *U, hipStream_t s) {
  if (p.view.ghost_rule == GHOST_NONE || p.n_near == 0) return PDA_OK;
  const int bs = 128;
  const unsigned nb = unsigned((p.n_near + bs - 1) / bs);
  switch (p.family) {
    case PDA_EULER1D: hipLaunchKernelGGL((k_ghosts<1, 3>), dim3(nb), dim3(bs), 0, s, p.view, p.w, U); break;
    case PDA_EULER2D: hipLaunchKernelGGL((k_ghosts<2, 4>), dim3(nb), dim3(bs), 0, s, p.view, p.w, U); break;
    case PDA_EULER3D: hipLaunchKernelGGL((k_ghosts<3, 5>), dim3(nb), dim3(bs), 0, s, p.view, p.w, U); break;
    case PDA_SWE2D: hipLaunchKernelGGL((k_ghosts<2, 3>), dim3(nb), dim3(bs), 0, s, p.view, p.w, U); break;
    case PDA_BURGERS2D: hipLaunchKernelGGL((k_ghosts<2, 2>), dim3(nb), dim3(bs), 0, s, p.view, p.w, U); break;
    default: set_error("ghost fill: unsupported family"); return PDA_E_UNSUPPORTED;
  }
  PDA_HIP_CHECK(hipGetLastError());
  return PDA_OK;
}

template <class Ph>
static int vel_dispatch(const pda_problem_s &p, const double *U, double *V, hipStream_t s) {
  const int bs = 256;
  const unsigned nb = unsigned((p.view.nrows + bs - 1) / bs);
  if (p.w == 1) hipLaunchKernelGGL((k_velocity<Ph, 1>), dim3(nb), dim3(bs), 0, s, p.view, U, V);
  else if (p.w == 2) hipLaunchKernelGGL((k_velocity<Ph, 2>), dim3(nb), dim3(bs), 0, s, p.view, U, V);
  else hipLaunchKernelGGL((k_velocity<Ph, 3>), dim3(nb), dim3(bs), 0, s, p.view, U, V);
  PDA_HIP_CHECK(hipGetLastError());
  return PDA_OK;
}

template <class Ph>
static int jac_dispatch(const pda_problem_s &p, const double *U, double *V, double *J,
                        hipStream_t s) {
  const int bs = 256;
  const unsigned nb = unsigned((p.view.nrows + bs - 1) / bs);
  if (p.w == 1) hipLaunchKernelGGL((k_jacobian<Ph, 1>), dim3(nb), dim3(bs), 0, s, p.view, U, V, J);
  else if (p.w == 2) hipLaunchKernelGGL((k_jacobian<Ph, 2>), dim3(nb), dim3(bs), 0, s, p.view, U, V, J);
  else hipLaunchKernelGGL((k_jacobian<Ph, 3>), dim3(nb), dim3(bs), 0, s, p.view, U, V, J);
  PDA_HIP_CHECK(hipGetLastError());
  return PDA_OK;
}

int launch_velocity(const pda_problem_s &p, const double *U, double, double *V, hipStream_t s) {
  if (p.view.nrows == 0) return PDA_OK;
  int rc = launch_ghosts(p, U, s);
  if (rc) return rc;
  switch (p.family) {
    case PDA_EULER1D: return vel_dispatch<PhEuler1>(p, U, V, s);
    case PDA_EULER2D: return vel_dispatch<PhEuler2>(p, U, V, s);
    case PDA_EULER3D: return vel_dispatch<PhEuler3>(p, U, V, s);
    case PDA_SWE2D: return vel_dispatch<PhSwe>(p, U, V, s);
    case PDA_BURGERS2D: return vel_dispatch<PhBurgers>(p, U, V, s);
    default: set_error("velocity: unsupported family"); return PDA_E_UNSUPPORTED;
  }
}

int launch_jacobian(const pda_problem_s &p, const double *U, double, double *J, double *V,
                    hipStream_t s) {
  if (p.view.nrows == 0) return PDA_OK;
  int rc = launch_ghosts(p, U, s);
  if (rc) return rc;
  if (p.view.dup) PDA_HIP_CHECK(hipMemsetAsync(J, 0, sizeof(double) * p.nnz, s));
  switch (p.family) {
    case PDA_EULER1D: return jac_dispatch<PhEuler1>(p, U, V, J, s);
    case PDA_EULER2D: return jac_dispatch<PhEuler2>(p, U, V, J, s);
    case PDA_EULER3D: return jac_dispatch<PhEuler3>(p, U, V, J, s);
    case PDA_SWE2D: return jac_dispatch<PhSwe>(p, U, V, J, s);
    case PDA_BURGERS2D: return jac_dispatch<PhBurgers>(p, U, V, J, s);
    default: set_error("jacobian: unsupported family"); return PDA_E_UNSUPPORTED;
  }
}

int launch_apply(const pda_problem_s &p, const double *J, const double *B, int64_t ncols,
                 int layout, double *AJ, hipStream_t s) {
  const int64_t nrows = p.view.nrows * p.m;
  const int64_t ldb = p.mesh->n_stencil * p.m;
  const int64_t n = nrows * ncols;
  if (n == 0) return PDA_OK;
  const int bs = 256;
  hipLaunchKernelGGL(k_apply, dim3(unsigned((n + bs - 1) / bs)), dim3(bs), 0, s,
                     p.d_rowptr.as<int64_t>(), p.d_colidx.as<int32_t>(), J, B, nrows, ncols, ldb,
                     layout, AJ);
  PDA_HIP_CHECK(hipGetLastError());
  return PDA_OK;
}

}  // namespace pda
