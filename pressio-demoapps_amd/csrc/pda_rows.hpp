// pda_rows.hpp -- per-row device code of the generic kernels: one thread
// owns one graph row (one sample-mesh cell) and produces all of its V
// entries and its whole Jacobian block-row, so every output value is written
// exactly once with plain stores (no atomics, no zero-fill pass).  The
// per-entry accumulation order follows the reference:
//   - V: x, then y, then z axis, then sources
//     (impl/mixin_directional_flux_balance.hpp:68-84);
//   - interior Jacobian: chain rule flux-Jacobian x reconstruction gradient,
//     left face then right face per axis
//     (impl/mixin_directional_flux_balance_jacobian.hpp:142-284);
//   - near-boundary rows: first-order blocks with BC factors folded into the
//     diagonal block (impl/mixin_directional_flux_balance_jacobian.hpp:311-371).
#pragma once
#include "pda_families.hpp"

namespace pda {
namespace rows {
using namespace fam;

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
__device__ __forceinline__ void fill_row_ghosts(const View &P, int W, int64_t g,
                                                const int32_t *gr, const double *__restrict__ U) {
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

// The row kernels fill the ghost slots of their own near-boundary row before
// reading them (the slots of a row are read by that row only), so no
// separate ghost pass runs ahead of an evaluation.

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

// velocity of graph row r into acc
template <class Ph, int W>
__device__ __forceinline__ void velocity_acc(const View &P, int64_t r, const double *__restrict__ U,
                                             double (&acc)[Ph::M]) {
  constexpr int M = Ph::M;
  const int32_t *gr = P.graph + r * P.ncols;
  const int64_t g = P.ghost_row[r];
  if (g >= 0) fill_row_ghosts<Ph::DIM, M>(P, W, g, gr, U);
#pragma unroll
  for (int k = 0; k < M; ++k) acc[k] = 0.;
  velocity_axis<Ph, W, 0>(P, gr, g, U, acc);
  if (g >= 0) AxisTerms<Ph>::template velocity<0>(P, gr, U, acc);
  if constexpr (Ph::DIM > 1) {
    // one axis at a time: its gathered stencil and temporaries are dead
    // before the next axis loads (the 3D WENO5 row otherwise holds all
    // three axes' stencils and runs at one wave per SIMD)
    __builtin_amdgcn_sched_barrier(0);
    velocity_axis<Ph, W, 1>(P, gr, g, U, acc);
    if (g >= 0) AxisTerms<Ph>::template velocity<1>(P, gr, U, acc);
  }
  if constexpr (Ph::DIM > 2) {
    __builtin_amdgcn_sched_barrier(0);
    velocity_axis<Ph, W, 2>(P, gr, g, U, acc);
    if (g >= 0) AxisTerms<Ph>::template velocity<2>(P, gr, U, acc);
  }
  Sources<Ph>::apply(P, r, g >= 0, U, gr[0], acc, nullptr);
}

// velocity of visited row t (row P.rows[t], or t when the list is absent)
template <class Ph, int W>
__device__ __forceinline__ void velocity_row(const View &P, int64_t t, const double *__restrict__ U,
                                             double *__restrict__ V) {
  constexpr int M = Ph::M;
  if (t >= P.nvisit) return;
  const int64_t r = P.rows ? P.rows[t] : t;
  double acc[M];
  velocity_acc<Ph, W>(P, r, U, acc);
#pragma unroll
  for (int k = 0; k < M; ++k) V[r * M + k] = acc[k];
}

// ---------------------------------------------------------------- jacobian
// interior row, one axis
template <class Ph, int W, int A, bool WV>
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
    if constexpr (WV) {
      Ph::template flux<A>(P, f.lm, f.lp, FL);
      Ph::template flux<A>(P, f.rm, f.rp, FR);
#pragma unroll
      for (int k = 0; k < M; ++k) acc[k] += h * (FL[k] - FR[k]);
    }
    Ph::template fjac<A>(P, f.lm, f.lp, JLn, JLp);
    Ph::template fjac<A>(P, f.rm, f.rp, JRn, JRp);
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
#pragma clang fp contract(off)
    FaceGrads<M, W> G;
    reconstruct_grad<M, W>(st, f, G);
    if constexpr (WV) {
      Ph::template flux<A>(P, f.lm, f.lp, FL);
      Ph::template flux<A>(P, f.rm, f.rp, FR);
#pragma unroll
      for (int k = 0; k < M; ++k) acc[k] += h * (FL[k] - FR[k]);
    }
    Ph::template fjac<A>(P, f.lm, f.lp, JLn, JLp);
    Ph::template fjac<A>(P, f.rm, f.rp, JRn, JRp);
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
template <class Ph, int W, int A, bool WV>
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
  if constexpr (WV) {
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
  }
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
      for (int j = 0; j < M; ++j) {
        b[k][j] = 0. + JLn[k * M + j] * h;
        if constexpr (AxisTerms<Ph>::near_offdiag) b[k][j] += Sources<Ph>::offdiag(P, true, A, k, j);
      }
    put_block<M>(P, J, base, rs, sl[colof<Ph::DIM>(A, 0, 0)], b);
  }
  if (r0 != -1) {
    double b[M][M];
#pragma unroll
    for (int k = 0; k < M; ++k)
#pragma unroll
      for (int j = 0; j < M; ++j) {
        b[k][j] = 0. + (-JRp[k * M + j]) * h;
        if constexpr (AxisTerms<Ph>::near_offdiag) b[k][j] += Sources<Ph>::offdiag(P, true, A, k, j);
      }
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

// Jacobian block-row (and velocity when WV, rightHandSideAndJacobian) of
// visited row t
template <class Ph, int W, bool WV>
__device__ __forceinline__ void jacobian_row(const View &P, int64_t t, const double *__restrict__ U,
                                             double *__restrict__ V, double *__restrict__ J) {
  constexpr int M = Ph::M;
  if (t >= P.nvisit) return;
  const int64_t r = P.rows ? P.rows[t] : t;
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
    jac_axis_inner<Ph, W, 0, WV>(P, gr, sl, U, acc, self, J, base, rs);
    if constexpr (Ph::DIM > 1) jac_axis_inner<Ph, W, 1, WV>(P, gr, sl, U, acc, self, J, base, rs);
    if constexpr (Ph::DIM > 2) jac_axis_inner<Ph, W, 2, WV>(P, gr, sl, U, acc, self, J, base, rs);
  } else {
    fill_row_ghosts<Ph::DIM, M>(P, W, g, gr, U);
    jac_axis_near<Ph, W, 0, WV>(P, gr, g, sl, U, acc, self, J, base, rs);
    if constexpr (WV) AxisTerms<Ph>::template velocity<0>(P, gr, U, acc);
    if constexpr (Ph::DIM > 1) {
      jac_axis_near<Ph, W, 1, WV>(P, gr, g, sl, U, acc, self, J, base, rs);
      if constexpr (WV) AxisTerms<Ph>::template velocity<1>(P, gr, U, acc);
    }
    if constexpr (Ph::DIM > 2) {
      jac_axis_near<Ph, W, 2, WV>(P, gr, g, sl, U, acc, self, J, base, rs);
      if constexpr (WV) AxisTerms<Ph>::template velocity<2>(P, gr, U, acc);
    }
  }
  Sources<Ph>::apply(P, r, g >= 0, U, gr[0], acc, &self[0][0]);
  put_block<M>(P, J, base, rs, sl[0], self);
  if constexpr (WV) {
#pragma unroll
    for (int k = 0; k < M; ++k) V[r * M + k] = acc[k];
  }
}

}  // namespace rows
}  // namespace pda
