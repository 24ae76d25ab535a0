// pda_families.hpp -- per-family policies used by every kernel: dofs per
// cell, dimension, face flux and flux Jacobians along axis A, and the source
// terms applied after the axis sweep; plus the stencil helpers.
#pragma once
#include "pda_internal.hpp"
#include "pda_physics.hpp"

namespace pda {
namespace fam {

// ------------------------------------------------------------------ physics
// Each family: M dofs per cell, DIM axes, face flux and its Jacobians for
// axis A, and the per-cell source terms applied after the axis sweep.
// sides()/entry() give the face Jacobian blocks in the family's compact form
// (Side): entry(S, i, j) is element (i, j) of that side's block, with i, j
// compile-time after unrolling.

// a side's block held as its M x M values
template <int M> struct MatSide {
  double J[M * M];
};

template <int D> struct PhEuler {
  static constexpr int M = D + 2, DIM = D;
  using Side = phys::EulerSide<D>;
  template <int A>
  __device__ static void flux(const View &P, const double *l, const double *r, double *F) {
    phys::euler_flux<D, A>(l, r, P.p[0], F);
  }
  template <int A>
  __device__ static void fjac(const View &P, const double *l, const double *r, double *JL, double *JR) {
    phys::euler_fjac<D, A>(l, r, P.p[0], JL, JR);
  }
  template <int A>
  __device__ static void sides(const View &P, const double *l, const double *r, Side &L, Side &R) {
    phys::euler_sides<D, A>(l, r, P.p[0], L, R);
  }
  template <int A>
  __device__ static double entry(const View &P, const Side &S, int i, int j) {
    return phys::euler_side_entry<D, A>(S, P.p[0], i, j);
  }
};
using PhEuler1 = PhEuler<1>;
using PhEuler2 = PhEuler<2>;
using PhEuler3 = PhEuler<3>;

// families whose face blocks are small enough to hold as values
template <class Derived, int M_> struct MatSided {
  using Side = MatSide<M_>;
  template <int A>
  __device__ static void sides(const View &P, const double *l, const double *r, Side &L, Side &R) {
    Derived::template fjac<A>(P, l, r, L.J, R.J);
  }
  template <int A>
  __device__ static double entry(const View &, const Side &S, int i, int j) {
    return S.J[i * M_ + j];
  }
};

struct PhSwe : MatSided<PhSwe, 3> {
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
struct PhBurgers : MatSided<PhBurgers, 2> {
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

// advection-diffusion-reaction problem A: one scalar advected with the
// constant velocity (p[0], p[1]); the face flux takes the minus-side state
// only (impl/advection_diffusion_reaction_2d_flux_mixin.hpp:95-105, 210-222)
struct PhAdr : MatSided<PhAdr, 1> {
  static constexpr int M = 1, DIM = 2;
  template <int A>
  __device__ static void flux(const View &P, const double *l, const double *, double *F) {
    F[0] = l[0] * P.p[A];
  }
  template <int A>
  __device__ static void fjac(const View &P, const double *, const double *, double *JL, double *JR) {
    JL[0] = P.p[A];
    JR[0] = 0.;
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
// impl/advection_diffusion_2d_prob_class.hpp:1167-1204; on near-boundary
// cells (the outflow problem) its velocity is added after each axis by
// AxisTerms<PhBurgers>, and its Jacobian here: the diagonal, then for each
// first neighbour (left, front, right, back) minus its weight on the
// diagonal where it is missing (a present one takes offdiag() in its
// block), impl/advection_diffusion_2d_prob_class.hpp:752-790
template <> struct Sources<PhBurgers> {
  __device__ static void apply(const View &P, int64_t r, bool near, const double *U, int32_t self,
                               double *v, double *J) {
    const int32_t *gr = P.graph + r * P.ncols;
    const double dX = P.p[0] * (P.hinv[0] * P.hinv[0]);
    const double dY = P.p[0] * (P.hinv[1] * P.hinv[1]);
    if (near) {
      if (J) {
#pragma clang fp contract(off)
        const double w[4] = {dX, dY, dX, dY};
        J[0] += -2. * dX - 2. * dY;
        J[3] += -2. * dX - 2. * dY;
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (gr[1 + q] == -1) {
            J[0] += -w[q];
            J[3] += -w[q];
          }
      }
      return;
    }
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
  // the diffusion adds to the diagonal of the first-neighbour blocks
  // (all four on inner rows, the present ones on near-boundary rows)
  __device__ static double offdiag(const View &P, bool, int a, int k, int j) {
    if (k != j) return 0.;
    return P.p[0] * (P.hinv[a] * P.hinv[a]);
  }
};

// ADR diffusion (D = p[2]) on inner rows, then the source f = 1 and the
// reaction -sigma*u (sigma = p[3]),
// impl/advection_diffusion_reaction_2d_prob_class.hpp:1040-1083 (velocity),
// :490-512 (inner Jacobian) and :840-872 (near-boundary Jacobian: a missing
// neighbour is the ghost -u, so its diffusion weight moves to the diagonal
// with a minus sign); the near-boundary diffusion itself is added per axis
// by AxisTerms<PhAdr>
template <> struct Sources<PhAdr> {
  __device__ static void apply(const View &P, int64_t r, bool near, const double *U, int32_t self,
                               double *v, double *J) {
#pragma clang fp contract(off)
    const int32_t *gr = P.graph + r * P.ncols;
    const double dX = P.p[2] * (P.hinv[0] * P.hinv[0]);
    const double dY = P.p[2] * (P.hinv[1] * P.hinv[1]);
    const double sigma = P.p[3];
    const double u = U[self];
    if (!near) {
      v[0] += dX * (U[gr[3]] - 2. * u + U[gr[1]]);
      v[0] += dY * (U[gr[2]] - 2. * u + U[gr[4]]);
    }
    v[0] += 1.;
    v[0] -= sigma * u;
    if (J) {
      if (!near) {
        J[0] += -2. * dX - 2. * dY;
        J[0] -= sigma;
      } else {
        double sv = -2. * dX - 2. * dY - sigma;
        if (gr[1] == -1) sv += -dX;
        if (gr[2] == -1) sv += -dY;
        if (gr[3] == -1) sv += -dX;
        if (gr[4] == -1) sv += -dY;
        J[0] += sv;
      }
    }
  }
  // the diffusion weight of every present first neighbour, inner and near
  __device__ static double offdiag(const View &P, bool, int a, int, int) {
    return P.p[2] * (P.hinv[a] * P.hinv[a]);
  }
};

// Terms a family adds to a near-boundary row right after each axis sweep
// (before the next axis), and whether the near-boundary Jacobian's
// first-neighbour blocks take Sources<Ph>::offdiag.
template <class Ph> struct AxisTerms {
  static constexpr bool near_offdiag = false;
  template <int A>
  __device__ static void velocity(const View &, const int32_t *, const double *, double *) {}
};

// ADR near-boundary diffusion along axis A from the cell's first
// neighbours, a missing one read as its ghost -u,
// impl/advection_diffusion_reaction_2d_prob_class.hpp:946-976
template <> struct AxisTerms<PhAdr> {
  static constexpr bool near_offdiag = true;
  template <int A>
  __device__ static void velocity(const View &P, const int32_t *gr, const double *U, double *acc) {
#pragma clang fp contract(off)
    const int32_t l = gr[A == 0 ? 1 : 4], r = gr[A == 0 ? 3 : 2];
    const double u = U[gr[0]];
    const double ul = l >= 0 ? U[l] : -u, ur = r >= 0 ? U[r] : -u;
    acc[0] += (P.p[2] * (P.hinv[A] * P.hinv[A])) * (ur - 2. * u + ul);
  }
};

// Burgers near-boundary diffusion along axis A from the cell's first
// neighbours, a missing one read as its ghost (0 at the outflow problem's
// left and back sides, the cell itself at the right and front ones: the
// side's sign times the cell), impl/advection_diffusion_2d_prob_class.hpp:
// 721-729 and :892-894
template <> struct AxisTerms<PhBurgers> {
  static constexpr bool near_offdiag = true;
  template <int A>
  __device__ static void velocity(const View &P, const int32_t *gr, const double *U, double *acc) {
#pragma clang fp contract(off)
    const int32_t l = gr[A == 0 ? 1 : 4], r = gr[A == 0 ? 3 : 2];
    const int64_t s = 2 * int64_t(gr[0]);
    const double d = P.p[0] * (P.hinv[A] * P.hinv[A]);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const double u = U[s + k];
      const double ul = l >= 0 ? U[2 * int64_t(l) + k] : P.gsign[A][0][k] * u;
      const double ur = r >= 0 ? U[2 * int64_t(r) + k] : P.gsign[A][1][k] * u;
      acc[k] += d * (ur - 2. * u + ul);
    }
  }
};

// ------------------------------------------------------------------ helpers
// LDS hand-off inside one wavefront (the workgroup is a single wave): LDS
// instructions of a wave execute in order, so only the compiler must be kept
// from reordering; unlike __syncthreads() this does not wait for the wave's
// outstanding global stores.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Streaming (non-temporal) stores for Jacobian values: each value is written
// once and not read again by the kernel, so the stores carry the nt hint and
// do not displace the states and tables the kernel re-reads from L2.
__device__ __forceinline__ void store_stream(double *p, double v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void store_stream(double2 *p, double2 v) {
  typedef double d2v __attribute__((ext_vector_type(2)));
  __builtin_nontemporal_store(d2v{v.x, v.y}, reinterpret_cast<d2v *>(p));
}

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
      else store_stream(dst, b[k][j]);
    }
}


}  // namespace fam
}  // namespace pda
