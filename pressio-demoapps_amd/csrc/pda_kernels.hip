// pda_kernels.hip -- generic kernels of the finite-volume hot path (any
// mesh: full, sample/stencil, 1D/2D/3D): velocity (residual), Jacobian
// assembly into the fixed CSR pattern, and the Jacobian action, plus the
// launch logic that pairs them with the structured box kernels.  The per-row
// device code is in pda_rows.hpp.
#include "pda_rows.hpp"

namespace pda {
namespace {
using namespace fam;

template <class Ph, int W>
__global__ void __launch_bounds__(64) k_velocity(View P, const double *__restrict__ U,
                                                   double *__restrict__ V) {
  rows::velocity_row<Ph, W>(P, int64_t(blockIdx.x) * blockDim.x + threadIdx.x, U, V);
}

// velocity with an explicit stage fused into the store (RkStage)
template <class Ph, int W>
__global__ void __launch_bounds__(64) k_velocity_stage(View P, const double *__restrict__ U, RkStage S) {
  constexpr int M = Ph::M;
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= P.nvisit) return;
  const int64_t r = P.rows ? P.rows[t] : t;
  double acc[M];
  rows::velocity_acc<Ph, W>(P, r, U, acc);
#pragma unroll
  for (int k = 0; k < M; ++k) {
    const int64_t e = r * M + k;
    if (S.kout) S.kout[e] = acc[k];
    double o = S.c[0] * S.x[0][e];
    for (int q = 1; q < S.nx; ++q) o += S.c[q] * S.x[q][e];
    o += S.cv * acc[k];
    S.out[e] = o;
  }
}

// ----------------------------------------------- 1D steps in one launch
// A small 1D problem is bound by launch latency: an SSPRK3 step is three
// velocity launches of a few waves each.  k_steps1d advances several steps
// in one launch.  Workgroup b owns cells [own b, own b + own) and holds them
// in LDS with H cells on each side, H = W x the number of stages it runs:
// each stage's values are valid on a region W cells narrower per side than
// the stage before, so after the last stage the owned cells are exact.  The
// region is 128 rows (256 lanes, one wave per SIMD of the workgroup's CU:
// a stage is one dependent chain of face arithmetic per lane, so a second
// wave per SIMD would double it).  Two lanes per row evaluate its left and
// right faces (the generic row code's stencil and ghost rule, the
// one-division WENO weights of the 2D march, the same flux) and exchange
// the fluxes; the row's velocity and its stage combination are formed as
// k_velocity_stage forms them, term by term.  Ghosts: the
// opposite rule (a row reads position g - d < 0 as cell g + d - 1, g + d
// >= n as g - d + 1, times the side's sign), or periodic wrap.  The state
// is read from Uin and written to Uout (workgroups read each other's cells).
constexpr int S1_ROWS = 128;

struct Steps1dCoef {
  double c[4][4];  // per stage: coefficients of the stage's terms
  double cv[4];    // per stage: coefficient of the velocity
};

template <class Ph, int W, int SCHEME>
__global__ void __launch_bounds__(2 * S1_ROWS) k_steps1d(View P, int64_t n, int periodic, int own, int H,
                                                        int nsteps, const double *__restrict__ Uin,
                                                        double *__restrict__ Uout, Steps1dCoef C) {
  constexpr int M = Ph::M;
  extern __shared__ double lds[];
  const int R = own + 2 * H;
  const int64_t a0 = int64_t(blockIdx.x) * own, base = a0 - H;
  // buffers of R x M doubles: 0 the step's state, 1 and 2 stage states,
  // 3..5 the RK4 stage velocities
  auto buf = [&](int b) { return lds + int64_t(b) * R * M; };
  for (int e = threadIdx.x; e < R * M; e += blockDim.x) {
    const int i = e / M, k = e - i * M;
    int64_t g = base + i;
    double v = 0.;
    if (periodic) {
      g %= n;
      if (g < 0) g += n;
      v = Uin[g * M + k];
    } else if (g >= 0 && g < n) {
      v = Uin[g * M + k];
    }
    buf(0)[e] = v;
  }
  __syncthreads();
  const int i = int(threadIdx.x) >> 1, f = int(threadIdx.x) & 1;
  const int64_t g = base + i;
  const bool row = i < R && (periodic || (g >= 0 && g < n));
  const bool store = row && f == 0;
  const double h = P.hinv[0];
  // velocity of row i from state X (every lane takes part in the exchange)
  auto vel = [&](const double *X, double (&acc)[M]) {
    double q[M][2 * W];
#pragma unroll
    for (int p = 0; p < 2 * W; ++p) {
      int64_t src = g + f - W + p;
      int side = -1;
      if (!periodic) {
        if (src < 0) {
          src = 2 * g - src - 1;
          side = 0;
        } else if (src >= n) {
          src = 2 * g - src + 1;
          side = 1;
        }
      }
      int64_t li = src - base;
      li = li < 0 ? 0 : (li >= R ? R - 1 : li);
#pragma unroll
      for (int k = 0; k < M; ++k) {
        const double v = X[li * M + k];
        q[k][p] = side < 0 ? v : P.gsign[0][side][k] * v;
      }
    }
    double um[M], up[M], F[M];
#pragma unroll
    for (int k = 0; k < M; ++k) {
      if constexpr (W == 1) {
        um[k] = q[k][0];
        up[k] = q[k][1];
      } else if constexpr (W == 2) {
        phys::weno3_1div(q[k], um[k], up[k]);
      } else {
        phys::weno5_1div(q[k], um[k], up[k]);
      }
    }
    Ph::template flux<0>(P, um, up, F);
#pragma unroll
    for (int k = 0; k < M; ++k) {
      const double o = __shfl_xor(F[k], 1);
      const double FL = f ? o : F[k], FR = f ? F[k] : o;
      acc[k] = 0.;
      acc[k] += h * (FL - FR);
    }
  };
  // out = c0 x0 + ... + cv v, the terms in k_velocity_stage's order
  auto comb = [&](double *out, int st, int nx, const double *x0, const double *x1, const double *x2,
                  const double *x3, const double (&acc)[M]) {
    if (!store) return;
#pragma unroll
    for (int k = 0; k < M; ++k) {
      const int e = i * M + k;
      double o = C.c[st][0] * x0[e];
      if (nx > 1) o += C.c[st][1] * x1[e];
      if (nx > 2) o += C.c[st][2] * x2[e];
      if (nx > 3) o += C.c[st][3] * x3[e];
      o += C.cv[st] * acc[k];
      out[e] = o;
    }
  };
  auto keep = [&](double *kb, const double (&acc)[M]) {
    if (!store) return;
#pragma unroll
    for (int k = 0; k < M; ++k) kb[i * M + k] = acc[k];
  };
  for (int s = 0; s < nsteps; ++s) {
    double acc[M];
    if constexpr (SCHEME == PDA_SSPRK3) {
      vel(buf(0), acc);  // u1 = u + dt f(u)
      comb(buf(1), 0, 1, buf(0), nullptr, nullptr, nullptr, acc);
      __syncthreads();
      vel(buf(1), acc);  // u2 = 1/4 u1 + 3/4 u + 1/4 dt f(u1)
      comb(buf(2), 1, 2, buf(1), buf(0), nullptr, nullptr, acc);
      __syncthreads();
      vel(buf(2), acc);  // u = 1/3 u + 2/3 u2 + 2/3 dt f(u2)
      comb(buf(0), 2, 2, buf(0), buf(2), nullptr, nullptr, acc);
      __syncthreads();
    } else {
      vel(buf(0), acc);  // k1; y1 = u + dt/2 k1
      keep(buf(3), acc);
      comb(buf(1), 0, 1, buf(0), nullptr, nullptr, nullptr, acc);
      __syncthreads();
      vel(buf(1), acc);  // k2; y2 = u + dt/2 k2
      keep(buf(4), acc);
      comb(buf(2), 1, 1, buf(0), nullptr, nullptr, nullptr, acc);
      __syncthreads();
      vel(buf(2), acc);  // k3; y3 = u + dt k3
      keep(buf(5), acc);
      comb(buf(1), 2, 1, buf(0), nullptr, nullptr, nullptr, acc);
      __syncthreads();
      vel(buf(1), acc);  // u = u + dt/6 k1 + dt/3 k2 + dt/3 k3 + dt/6 k4
      comb(buf(0), 3, 4, buf(0), buf(3), buf(4), buf(5), acc);
      __syncthreads();
    }
  }
  if (store && i >= H && i < H + own && g < n) {
#pragma unroll
    for (int k = 0; k < M; ++k) Uout[g * M + k] = buf(0)[i * M + k];
  }
}

template <class Ph, int W, bool WV>
__global__ void __launch_bounds__(64) k_jacobian(View P, const double *__restrict__ U,
                                                   double *__restrict__ V, double *__restrict__ J) {
  rows::jacobian_row<Ph, W, WV>(P, int64_t(blockIdx.x) * blockDim.x + threadIdx.x, U, V, J);
}

// ------------------------------- first-order Euler Jacobian, staged rows
// Rows of the first-order Euler Jacobian (Euler has no source term) with
// four lanes per row and 16 rows per wavefront.  Lane a < D evaluates the two
// faces of axis a.  The Euler flux along axis a is the flux along x with
// momentum components 1 and 1+a exchanged, so every lane runs the x-face
// code on exchanged states and writes entry (k, j) of its blocks to
// (p(k), p(j)), p the 1 <-> 1+a exchange: one instruction stream for all
// axes.  A row's blocks are assembled in LDS in CSR order (scalar row k,
// block slot, column), the diagonal block adding the axes in order as the
// generic row code does, and the wavefront then copies each staged row to
// its contiguous CSR values with coalesced streaming stores (one thread per
// row would scatter every store over 64 rows); 16 rows consecutive in the
// CSR leave as one range of 16-byte stores.  Used where no structured box
// kernel applies (sample meshes, 2D meshes given as a graph) and for the
// boundary rows of a structured 3D full mesh.  Near-boundary rows: a
// missing neighbour along axis a reads the ghost state sign_k U_k of the row's own cell (the first
// ghost layer of every ghost rule), writes no block, and adds fac_j JL (left
// side) or -fac_j JR (right side) of that face, times h, to the axis's
// diagonal part, fac the row's BC factors (the terms of
// impl/mixin_directional_flux_balance_jacobian.hpp:312-369, with the factors
// of impl/euler_3d_prob_class.hpp:626-651).
constexpr int SJ_LANES = 4, SJ_ROWS = 64 / SJ_LANES;

__device__ __forceinline__ int exch(int k, int a) { return k == 1 ? 1 + a : (k == 1 + a ? 1 : k); }

// ghost sign of native dof exch(k, a) on side `side` of axis a
template <int D>
__device__ __forceinline__ double ghost_sign(const View &P, int a, int side, int k) {
  const double s0 = P.gsign[0][side][exch(k, 0)];
  const double s1 = D > 1 ? P.gsign[1][side][exch(k, 1)] : 0.;
  const double s2 = D > 2 ? P.gsign[2][side][exch(k, 2)] : 0.;
  return a == 0 ? s0 : (a == 1 ? s1 : s2);
}

template <int D>
__global__ void __launch_bounds__(64) k_jac_fo_euler_staged(View P, const double *__restrict__ U,
                                                              double *__restrict__ J) {
  using Ph = PhEuler<D>;
  constexpr int M = D + 2, NB = 1 + 2 * D, RS = NB * M, PER = M * RS;
  __shared__ double stage[SJ_ROWS * PER];
  __shared__ int64_t row_of[SJ_ROWS];
  const int lane = threadIdx.x, g = lane / SJ_LANES, a = lane % SJ_LANES;
  const int64_t t = int64_t(blockIdx.x) * SJ_ROWS + g;
  int64_t r = -1;
  if (t < P.nvisit) {
    const int64_t rr = P.rows ? P.rows[t] : t;
    r = rr;
  }
  if (a == 0) row_of[g] = r;
  const bool act = r >= 0 && a < D;
  double *st = stage + g * PER;
  typename Ph::Side Ln, Lp, Rn, Rp;
  int sL = 0, sR = 0, s0 = 0;
  bool gl = false, gr_ = false; // no neighbour on the left / right
  double fac[M];                // BC factors of native dof exch(j, a)
  double h = 0.;
  if (act) {
    int64_t cells[3];
    {
      const int32_t *gr = P.graph + r * P.ncols;
      const uint8_t *sl = P.slot + r * P.npos;
      const int cl = colof<D>(a, 0, 0), cr = colof<D>(a, 1, 0);
      const int32_t il = gr[cl], ir = gr[cr];
      gl = il < 0;
      gr_ = ir < 0;
      cells[0] = gl ? gr[0] : il;
      cells[1] = gr[0];
      cells[2] = gr_ ? gr[0] : ir;
      sL = gl ? 0 : sl[cl];
      sR = gr_ ? 0 : sl[cr];
      s0 = sl[0];
    }
    double q[3][M];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int k = 0; k < M; ++k) q[c][k] = U[cells[c] * M + exch(k, a)];
#pragma unroll
    for (int k = 0; k < M; ++k) {
      if (gl) q[0][k] = ghost_sign<D>(P, a, 0, k) * q[0][k];
      if (gr_) q[2][k] = ghost_sign<D>(P, a, 1, k) * q[2][k];
      fac[k] = 0.;
    }
    const int64_t gi = P.ghost_row[r];
    if ((gl || gr_) && gi >= 0) {
      const double *f = P.bcf + (gi * D + a) * M;
#pragma unroll
      for (int k = 0; k < M; ++k) fac[k] = f[exch(k, a)];
    }
    Ph::template sides<0>(P, q[0], q[1], Ln, Lp);
    Ph::template sides<0>(P, q[1], q[2], Rn, Rp);
    h = P.hinv[a];
  }
  // a fence between the face math and the stage stores keeps the entry
  // evaluations from being hoisted into it (2D: 148 -> 126 registers)
  wave_sync();
  if (act) {
#pragma unroll
    for (int k = 0; k < M; ++k)
#pragma unroll
      for (int j = 0; j < M; ++j) {
        const int o = exch(k, a) * RS + exch(j, a);
        if (!gl) st[o + sL * M] = 0. + Ph::template entry<0>(P, Ln, k, j) * h;
        if (!gr_) st[o + sR * M] = 0. + (-Ph::template entry<0>(P, Rp, k, j)) * h;
      }
  }
  // every lane forms its axis's part of the diagonal block at once; the
  // axes then add their parts in order, one axis at a time, in LDS
  double dg[M][M];
  wave_sync();
  if (act) {
#pragma unroll
    for (int k = 0; k < M; ++k)
#pragma unroll
      for (int j = 0; j < M; ++j) {
        dg[k][j] = (Ph::template entry<0>(P, Lp, k, j) - Ph::template entry<0>(P, Rn, k, j)) * h;
        if (gl) dg[k][j] += (fac[j] * Ph::template entry<0>(P, Ln, k, j)) * h;
        if (gr_) dg[k][j] += (fac[j] * -Ph::template entry<0>(P, Rp, k, j)) * h;
      }
  }
#pragma unroll
  for (int ax = 0; ax < D; ++ax) {
    wave_sync();
    if (act && a == ax) {
#pragma unroll
      for (int k = 0; k < M; ++k)
#pragma unroll
        for (int j = 0; j < M; ++j) {
          const int o = exch(k, a) * RS + s0 * M + exch(j, a);
          st[o] = (ax == 0 ? 0. : st[o]) + dg[k][j];
        }
    }
  }
  wave_sync();
  // rows consecutive in the CSR (a run along x): the stage is their values
  // in order, copied as one range with 16-byte stores
  const int64_t r0 = row_of[0], rl = row_of[SJ_ROWS - 1];
  bool run = r0 >= 0;
#pragma unroll
  for (int gg = 1; gg < SJ_ROWS; ++gg) run = run && row_of[gg] == r0 + gg;
  int64_t rb0 = 0;
  if (run) {
    rb0 = P.rowbase[r0];
    run = P.rowbase[rl] == rb0 + int64_t(SJ_ROWS - 1) * PER;
    run = run && P.nblk[rl] == NB;
  }
  if (run) {
    constexpr int TOT = SJ_ROWS * PER;
    const int head = int(rb0 & 1); // a leading value that is not 16-byte aligned
    double *dst = J + rb0;
    if (head && lane == 0) store_stream(dst, stage[0]);
    const int npairs = (TOT - head) / 2;
    double2 *d2 = reinterpret_cast<double2 *>(dst + head);
    for (int q = lane; q < npairs; q += 64)
      store_stream(d2 + q, make_double2(stage[head + 2 * q], stage[head + 2 * q + 1]));
    if (((TOT - head) & 1) && lane == 0) store_stream(dst + TOT - 1, stage[TOT - 1]);
    return;
  }
  for (int gg = 0; gg < SJ_ROWS; ++gg) {
    const int64_t rg = row_of[gg];
    if (rg < 0) continue;
    double *dst = J + P.rowbase[rg];
    const double *src = stage + gg * PER;
    // a near-boundary row holds nblk of the NB block slots per scalar row
    const int rs = P.nblk[rg] * M;
    for (int e = lane; e < M * rs; e += 64) {
      int k = 0;
#pragma unroll
      for (int q = 1; q < M; ++q) k += e >= q * rs;
      store_stream(dst + e, src[k * RS + (e - k * rs)]);
    }
  }
}

// ------------------------------------ WENO3/WENO5 Jacobian, staged rows
// Inner rows of the WENO Jacobian (and, with WV, the velocity) with 2*D
// lanes per row: lane (a, f) evaluates face f (0 left, 1 right) of axis a,
// its reconstruction gradients for every dof and the face flux Jacobians.
// For these families the flux along axis a is the flux along x of the state
// with two dofs exchanged (AxisPerm), exchanged back, so every lane runs the
// x-face code on permuted states.  Face f reads the 2W cells at offsets
// o = f - W + i (i < 2W) and contributes to their blocks
//   block(o)[k][j] += s_f h (Jn[k][j] gm[j][i] + Jp[k][j] gp[j][i]),
// s_0 = +1, s_1 = -1 (impl/mixin_directional_flux_balance_jacobian.hpp:142-284
// sums the same terms, left face then right face per axis).  The blocks are
// assembled in an LDS stage in CSR order without races: the left face
// stores the negative offsets and the right face the positive ones, then
// each adds its part of the other side; the diagonal block and the velocity
// take the lanes' parts one lane at a time, and the row's first lane applies
// the family's sources.  Rows then leave as contiguous coalesced streaming
// stores.  Near-boundary rows and rows with merged columns stay with the
// generic kernel.  The per-lane live set is one face's gradients and
// Jacobians (2 M (2W) + 2 M^2 doubles), where one thread per row holds all
// faces of all axes and spills.
template <class Ph> struct AxisPerm {
  // Euler, SWE: (rho u) <-> (rho u_a)
  __host__ __device__ static constexpr int p(int k, int a) { return k == 1 ? 1 + a : (k == 1 + a ? 1 : k); }
};
template <> struct AxisPerm<PhBurgers> {
  __host__ __device__ static constexpr int p(int k, int a) { return k == 0 ? a : (k == a ? 0 : k); } // u <-> u_a
};

// value of the axis-a frame array x at the index of dof k of the
// physical frame: x[p_a(k)], selected among the D compile-time candidates
// (a is the lane's axis), so every index stays a compile-time constant
template <class Ph, class X>
__device__ __forceinline__ double perm_pick(const X &x, int a, int k) {
  double v = x(AxisPerm<Ph>::p(k, 0));
  if constexpr (Ph::DIM > 1) v = a == 1 ? x(AxisPerm<Ph>::p(k, 1)) : v;
  if constexpr (Ph::DIM > 2) v = a == 2 ? x(AxisPerm<Ph>::p(k, 2)) : v;
  return v;
}

// store (or add) the axis-frame block b at dst (row stride RS) in the
// physical frame: dst[K][J] = b[p_a(K)][p_a(J)] (p_a is an involution)
template <class Ph, int RS, bool ADD>
__device__ __forceinline__ void put_perm(double *dst, int a, const double (&b)[Ph::M][Ph::M]) {
  constexpr int M = Ph::M;
#pragma unroll
  for (int K = 0; K < M; ++K)
#pragma unroll
    for (int J = 0; J < M; ++J) {
      double v = b[AxisPerm<Ph>::p(K, 0)][AxisPerm<Ph>::p(J, 0)];
      if constexpr (Ph::DIM > 1) v = a == 1 ? b[AxisPerm<Ph>::p(K, 1)][AxisPerm<Ph>::p(J, 1)] : v;
      if constexpr (Ph::DIM > 2) v = a == 2 ? b[AxisPerm<Ph>::p(K, 2)][AxisPerm<Ph>::p(J, 2)] : v;
      if constexpr (ADD) dst[K * RS + J] += v;
      else dst[K * RS + J] = v;
    }
}

// graph column of the cell at signed offset o along axis a (o != 0)
template <int D> __device__ __forceinline__ int col_at(int a, int o) {
  return o < 0 ? colof<D>(a, 0, -o - 1) : colof<D>(a, 1, o - 1);
}

// one face's contribution to the block of its stencil point i, in the axis
// frame: s_f h (Jn diag(gm[:, i]) + Jp diag(gp[:, i]))
template <int M, int G>
__device__ __forceinline__ void face_block(double sh, const double (&Jn)[M * M], const double (&Jp)[M * M],
                                           const double (&gm)[M][G], const double (&gp)[M][G], int i0, int i1,
                                           bool sel, double (&b)[M][M]) {
#pragma clang fp contract(off)
#pragma unroll
  for (int j = 0; j < M; ++j) {
    const double gmv = sel ? gm[j][i1] : gm[j][i0], gpv = sel ? gp[j][i1] : gp[j][i0];
#pragma unroll
    for (int k = 0; k < M; ++k) b[k][j] = sh * (Jn[k * M + j] * gmv + Jp[k * M + j] * gpv);
  }
}

// rows per wavefront of the staged WENO kernel: 64 / (2 D), fewer for the
// 2D Euler WENO5 stage (16 rows would take 26.6 KB of LDS, 5 waves per CU;
// 12 rows fit 8, matching its 2 waves per SIMD)
template <class Ph, int W> constexpr int weno_staged_rows() {
  return (Ph::DIM == 2 && Ph::M == 4 && W == 3) ? 12 : 64 / (2 * Ph::DIM);
}

template <class Ph, int W, bool WV>
__global__ void __launch_bounds__(64) k_jac_weno_staged(View P, const double *__restrict__ U,
                                                          double *__restrict__ V, double *__restrict__ J) {
  constexpr int D = Ph::DIM, M = Ph::M, LPR = 2 * D, RPW = weno_staged_rows<Ph, W>();
  constexpr int NB = 1 + 2 * W * D, RS = NB * M, PER = M * RS, G = 2 * W;
  __shared__ double stage[RPW * PER];
  __shared__ double vst[WV ? RPW * M : 1];
  __shared__ int64_t row_of[RPW];
  const int lane = threadIdx.x, g = lane / LPR, role = lane % LPR;
  const int a = role >> 1, f = role & 1;
  const bool inwave = g < RPW; // 3D: 60 of the 64 lanes
  const int64_t t = int64_t(blockIdx.x) * RPW + g;
  int64_t r = -1;
  if (inwave && t < P.nvisit) {
    const int64_t rr = P.rows ? P.rows[t] : t;
    if (P.ghost_row[rr] < 0) r = rr;
  }
  if (inwave && role == 0) row_of[g] = r;
  const bool act = r >= 0;
  double *st = stage + (inwave ? g : 0) * PER;
  const int32_t *gr = P.graph + (act ? r : 0) * P.ncols;
  const uint8_t *sl = P.slot + (act ? r : 0) * P.npos;
  const double h = a == 0 ? P.hinv[0] : (a == 1 ? P.hinv[1] : P.hinv[2]);
  const double sh = f ? -h : h;
  double gm[M][G], gp[M][G], Jn[M * M], Jp[M * M];
  double fm[M], fp[M];
  if (act) {
    {
      double q[M][G];
#pragma unroll
      for (int i = 0; i < G; ++i) {
        const int o = f - W + i;
        const int64_t c = gr[o == 0 ? 0 : col_at<D>(a, o)];
        double u[M];
#pragma unroll
        for (int k = 0; k < M; ++k) u[k] = U[c * M + k];
#pragma unroll
        for (int k = 0; k < M; ++k) q[k][i] = perm_pick<Ph>([&](int x) { return u[x]; }, a, k);
      }
#pragma unroll
      for (int k = 0; k < M; ++k) {
        if constexpr (W == 2) phys::weno3_grad(q[k], fm[k], fp[k], gm[k], gp[k]);
        else phys::weno5_grad(q[k], fm[k], fp[k], gm[k], gp[k]);
      }
    }
    Ph::template fjac<0>(P, fm, fp, Jn, Jp);
    if constexpr (WV) {
      double F[M];
      Ph::template flux<0>(P, fm, fp, F);
      // the face's velocity part, held in fm until its turn below
#pragma unroll
      for (int k = 0; k < M; ++k) fm[k] = sh * F[k];
    }
  }
  // a fence between the face math and the block products: without it the
  // products are hoisted into the gradient computation and the 2D Euler
  // WENO5 kernel needs 512 registers and spills (with it: 236, no spills)
  wave_sync();
  if (act) {
    // phase 1: left face -> offsets -W..-1, right face -> offsets 1..W
#pragma unroll
    for (int m = 0; m < W; ++m) {
      const int o = f ? 1 + m : m - W;
      double b[M][M];
      face_block<M, G>(sh, Jn, Jp, gm, gp, m, W + m, f != 0, b);
      if (o == -1 || o == 1) {
#pragma unroll
        for (int k = 0; k < M; ++k)
#pragma unroll
          for (int j = 0; j < M; ++j) b[k][j] += Sources<Ph>::offdiag(P, false, a, k, j);
      }
      put_perm<Ph, RS, false>(st + sl[col_at<D>(a, o)] * M, a, b);
    }
  }
  wave_sync();
  if (act) {
    // phase 2: left face adds to offsets 1..W-1, right face to -W+1..-1
#pragma unroll
    for (int m = 0; m + 1 < W; ++m) {
      const int o = f ? m + 1 - W : m + 1;
      double b[M][M];
      face_block<M, G>(sh, Jn, Jp, gm, gp, W + 1 + m, m, f != 0, b);
      put_perm<Ph, RS, true>(st + sl[col_at<D>(a, o)] * M, a, b);
    }
  }
  // diagonal block and velocity: the row's lanes one at a time, in axis
  // order, left face before right face
  const int s0 = act ? sl[0] : 0;
#pragma unroll
  for (int ph = 0; ph < LPR; ++ph) {
    wave_sync();
    if (act && role == ph) {
      double b[M][M];
      face_block<M, G>(sh, Jn, Jp, gm, gp, W, W - 1, f != 0, b);
      if (ph == 0) put_perm<Ph, RS, false>(st + s0 * M, a, b);
      else put_perm<Ph, RS, true>(st + s0 * M, a, b);
      if constexpr (WV) {
#pragma unroll
        for (int k = 0; k < M; ++k) {
          const double v = perm_pick<Ph>([&](int x) { return fm[x]; }, a, k);
          vst[g * M + k] = (ph == 0 ? 0. : vst[g * M + k]) + v;
        }
      }
    }
  }
  wave_sync();
  if (act && role == 0) {
    double self[M][M], acc[M];
#pragma unroll
    for (int k = 0; k < M; ++k) {
      acc[k] = WV ? vst[g * M + k] : 0.;
#pragma unroll
      for (int j = 0; j < M; ++j) self[k][j] = st[k * RS + s0 * M + j];
    }
    Sources<Ph>::apply(P, r, false, U, gr[0], acc, &self[0][0]);
#pragma unroll
    for (int k = 0; k < M; ++k) {
#pragma unroll
      for (int j = 0; j < M; ++j) st[k * RS + s0 * M + j] = self[k][j];
      if constexpr (WV) V[r * M + k] = acc[k];
    }
  }
  wave_sync();
  for (int gg = 0; gg < RPW; ++gg) {
    const int64_t rg = row_of[gg];
    if (rg < 0) continue;
    double *dst = J + P.rowbase[rg];
    const double *src = stage + gg * PER;
    for (int e = lane; e < PER; e += 64) store_stream(dst + e, src[e]);
  }
}

template <class Ph, int W>
int launch_jac_weno_staged(const View &v, const double *U, double *V, double *J, hipStream_t s) {
  constexpr int RPW = weno_staged_rows<Ph, W>();
  const unsigned nb = unsigned((v.nvisit + RPW - 1) / RPW);
  if (V) hipLaunchKernelGGL((k_jac_weno_staged<Ph, W, true>), dim3(nb), dim3(64), 0, s, v, U, V, J);
  else hipLaunchKernelGGL((k_jac_weno_staged<Ph, W, false>), dim3(nb), dim3(64), 0, s, v, U, V, J);
  PDA_HIP_CHECK(hipGetLastError());
  return PDA_OK;
}

template <class Ph>
int jac_weno_staged(const View &v, int w, const double *U, double *V, double *J, hipStream_t s) {
  return w == 2 ? launch_jac_weno_staged<Ph, 2>(v, U, V, J, s) : launch_jac_weno_staged<Ph, 3>(v, U, V, J, s);
}

// ------------------------- WENO Jacobian of a 2D box, each face once
// The staged kernel above evaluates every face twice, once for each of its
// two cells.  For the cells of a structured 2D box (natural order, stencils
// inside the grid, no near-boundary row) this kernel gives each lane one
// face of a TX x TY tile of cells instead: its reconstruction gradients,
// flux Jacobians and 2W blocks are formed once and added to both cells' rows
// staged in LDS, the cell on the face's plus side taking +h (the face is its
// left face), the minus side -h (its right face), as
// impl/mixin_directional_flux_balance_jacobian.hpp:142-284 sums them.  A
// tile of TX x TY cells has (TX+1) TY x-faces and TX (TY+1) y-faces (those
// on the tile's edges are also evaluated by the neighbouring tile, for its
// own cells).  Per row the stage receives, as in the staged kernel: phase 1
// stores (left face: offsets -W..-1, right face: 1..W, with the family's
// first-neighbour sources at -+1), phase 2 adds (left face: 1..W-1, right
// face: -W+1..-1), then the diagonal block and the velocity from the four
// faces one at a time (x left, x right, y left, y right), then the row's
// sources; no two lanes write one address in the same phase.  In the box's
// natural order a tile row's TX rows are consecutive in the CSR, and
// where no stencil of the tile wraps around a periodic edge the sorted CSR
// blocks of a row are the y neighbours below (offsets -W..-1), the x
// neighbours and the cell (-W..W), the y neighbours above (1..W); the
// tiles at a periodic edge take the stencil cells from the mesh graph and
// the block slots from the rows' slot tables, so a fully periodic grid is
// covered whole.
// LPF lanes share a face: with two, each forms the reconstruction gradients
// of half the dofs and the matching half of the block columns, the face
// states are exchanged between the pair and both form the flux Jacobians.
template <class Ph> struct WenoTile {
  static constexpr int TX = 8, TY = 3, LPF = 1; // 27 x-faces + 32 y-faces = 59 lanes
};
template <> struct WenoTile<PhEuler2> {
  static constexpr int TX = 4, TY = 3, LPF = 2; // (15 + 16) faces x 2 lanes = 62 lanes
};

// this lane's block columns j0 + v of an axis-frame face Jacobian, rows in
// the physical frame: out[K][v] = Jx[p_a(K)][j0 + v] (j0 = 0, or NV when
// sel), chosen once per face so the block products below need no selects
template <class Ph, int NV>
__device__ __forceinline__ void lane_cols(const double (&Jx)[Ph::M * Ph::M], int a, bool sel,
                                          double (&out)[Ph::M][NV]) {
  constexpr int M = Ph::M;
#pragma unroll
  for (int K = 0; K < M; ++K)
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      const int r0 = AxisPerm<Ph>::p(K, 0), r1 = AxisPerm<Ph>::p(K, 1);
      const double x0 = sel ? Jx[r0 * M + NV + v] : Jx[r0 * M + v];
      const double x1 = sel ? Jx[r1 * M + NV + v] : Jx[r1 * M + v];
      out[K][v] = a == 1 ? x1 : x0;
    }
}

// the face's block at stencil point i, physical rows, the lane's columns:
// b[K][v] = sh (Jn[K][v] gm[v][i] + Jp[K][v] gp[v][i])
template <int M, int NV, int G>
__device__ __forceinline__ void face_cols_p(double sh, const double (&Jn)[M][NV], const double (&Jp)[M][NV],
                                            const double (&gm)[NV][G], const double (&gp)[NV][G], int i,
                                            double (&b)[M][NV]) {
#pragma clang fp contract(off)
#pragma unroll
  for (int v = 0; v < NV; ++v)
#pragma unroll
    for (int k = 0; k < M; ++k) b[k][v] = sh * (Jn[k][v] * gm[v][i] + Jp[k][v] * gp[v][i]);
}

// store (or add) physical-row block columns at dst (row stride RS): column
// j0 + v of the axis frame is column p_a(j0 + v) of the physical frame
template <class Ph, int RS, bool ADD, int NV>
__device__ __forceinline__ void put_cols(double *dst, int a, int j0, const double (&b)[Ph::M][NV]) {
  constexpr int M = Ph::M;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const int j = j0 + v;
    const int J = a == 0 ? AxisPerm<Ph>::p(j, 0) : AxisPerm<Ph>::p(j, 1);
#pragma unroll
    for (int K = 0; K < M; ++K) {
      if constexpr (ADD) dst[K * RS + J] += b[K][v];
      else dst[K * RS + J] = b[K][v];
    }
  }
}

template <int W> __device__ __forceinline__ int tile_slot(int a, int o) {
  return a == 0 ? 2 * W + o : (o < 0 ? W + o : 3 * W + o);
}

// LDS of a tile: stage[NR * PERP], vst[NR * M].  A row's PER = M (1 + 4W) M
// staged values are held at a stride PERP = PER rounded up to an odd number
// of doubles: with PER a multiple of 16 doubles (128 bytes, all LDS banks)
// the face lanes' writes to one block entry of different rows would all
// fall in one bank.
template <class Ph, int W> struct WenoTileLds {
  static constexpr int NR = WenoTile<Ph>::TX * WenoTile<Ph>::TY, M = Ph::M;
  static constexpr int PER = M * (1 + 4 * W) * M, PERP = PER | 1;
  static constexpr int DOUBLES = NR * PERP + NR * M;
};

template <class Ph, int W, bool WV>
__device__ __forceinline__ void jac_weno_tiled_body(const View &P, const Box &B, int64_t bid,
                                                    const double *__restrict__ U, double *__restrict__ V,
                                                    double *__restrict__ J, double *stage) {
  static_assert(Ph::DIM == 2, "2D box");
  constexpr int M = Ph::M, TX = WenoTile<Ph>::TX, TY = WenoTile<Ph>::TY, NR = TX * TY;
  constexpr int LPF = WenoTile<Ph>::LPF, NV = M / LPF;
  constexpr int NB = 1 + 4 * W, RS = NB * M, PER = M * RS, G = 2 * W;
  constexpr int PERP = WenoTileLds<Ph, W>::PERP;
  constexpr int NXF = (TX + 1) * TY, NYF = TX * (TY + 1);
  static_assert(LPF * (NXF + NYF) <= 64 && NV * LPF == M, "LPF lanes per face");
  static_assert(WenoTileLds<Ph, W>::PER == PER, "stage layout");
  double *vst = stage + NR * PERP;
  const int lane = threadIdx.x, fl = lane / LPF, hv = lane % LPF; // face, half
  const int j0 = hv * NV;                                           // the lane's first dof
  const int64_t nx = B.n[0];
  const int64_t bw = B.hi[0] - B.lo[0];
  const int64_t ntx = (bw + TX - 1) / TX;
  const int64_t x0 = B.lo[0] + (bid % ntx) * TX, y0 = B.lo[1] + (bid / ntx) * TY;
  const int nax = int(B.hi[0] - x0 < TX ? B.hi[0] - x0 : TX);
  const int nay = int(B.hi[1] - y0 < TY ? B.hi[1] - y0 : TY);
  // a tile some stencil of which wraps around a periodic edge
  const bool edge = x0 < W || x0 + nax + W > nx || y0 < W || y0 + nay + W > B.n[1];
  // this lane's face: axis a, tile rows on its minus (rL) and plus (rR)
  // sides (-1 outside the tile's active cells), its minus-side cell cm (a
  // tile cell, or the first tile cell's minus neighbour from the graph)
  int a = 0, rL = -1, rR = -1;
  int64_t cm = 0;
  if (fl < NXF) {
    const int cj = fl / (TX + 1), fi = fl % (TX + 1);
    if (cj < nay && fi <= nax) {
      rL = fi >= 1 ? cj * TX + fi - 1 : -1;
      rR = fi < nax ? cj * TX + fi : -1;
      const int64_t r0 = (y0 + cj) * nx + x0;
      cm = fi >= 1 || !edge ? r0 + fi - 1 : int64_t(P.graph[r0 * P.ncols + col_at<2>(0, -1)]);
    }
  } else if (fl < NXF + NYF) {
    a = 1;
    const int l = fl - NXF, fj = l / TX, ci = l % TX;
    if (ci < nax && fj <= nay) {
      rL = fj >= 1 ? (fj - 1) * TX + ci : -1;
      rR = fj < nay ? fj * TX + ci : -1;
      const int64_t r0 = y0 * nx + x0 + ci;
      cm = fj >= 1 || !edge ? r0 + (fj - 1) * nx : int64_t(P.graph[r0 * P.ncols + col_at<2>(1, -1)]);
    }
  }
  const bool act = rL >= 0 || rR >= 0;
  const double h = a == 0 ? P.hinv[0] : P.hinv[1];
  // the CSR slot tables of the two rows (row = cell on a full mesh)
  const int64_t gL = (y0 + (rL >= 0 ? rL : 0) / TX) * nx + x0 + (rL >= 0 ? rL : 0) % TX;
  const int64_t gR = (y0 + (rR >= 0 ? rR : 0) / TX) * nx + x0 + (rR >= 0 ? rR : 0) % TX;
  const uint8_t *slL = P.slot + gL * P.npos, *slR = P.slot + gR * P.npos;
  // block slot of offset o along a (o != 0) in the row's CSR
  auto slotL = [&](int o) { return edge ? int(slL[col_at<2>(a, o)]) : tile_slot<W>(a, o); };
  auto slotR = [&](int o) { return edge ? int(slR[col_at<2>(a, o)]) : tile_slot<W>(a, o); };
  // gm, gp: the lane's NV dofs (axis frame j0..j0+NV-1); fm, fp: all
  double gm[NV][G], gp[NV][G], Jn[M * M], Jp[M * M];
  double fm[M] = {}, fp[M] = {}, F[M];
  if (act) {
    double q[NV][G];
    const int32_t *grm = P.graph + cm * P.ncols;
    const int64_t stride = a == 0 ? 1 : nx;
#pragma unroll
    for (int i = 0; i < G; ++i) {
      const int o = i - W + 1;
      const int64_t c = o == 0 ? cm : (edge ? int64_t(grm[col_at<2>(a, o)]) : cm + o * stride);
      double u[M];
#pragma unroll
      for (int k = 0; k < M; ++k) u[k] = U[c * M + k];
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        const double x0v = perm_pick<Ph>([&](int x) { return u[x]; }, a, v);
        q[v][i] = LPF == 1 || hv == 0 ? x0v : perm_pick<Ph>([&](int x) { return u[x]; }, a, (NV + v) % M);
      }
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      double um, up;
      if constexpr (W == 2) phys::weno3_grad(q[v], um, up, gm[v], gp[v]);
      else phys::weno5_grad(q[v], um, up, gm[v], gp[v]);
#pragma unroll
      for (int k = 0; k < M; ++k)
        if (k == j0 + v) {
          fm[k] = um;
          fp[k] = up;
        }
    }
  }
  if constexpr (LPF == 2) {
    // the pair's other half of the face states (every lane takes part)
#pragma unroll
    for (int k = 0; k < M; ++k) {
      const double tm = __shfl_xor(fm[k], 1), tp = __shfl_xor(fp[k], 1);
      if (k / NV != hv) {
        fm[k] = tm;
        fp[k] = tp;
      }
    }
  }
  if (act) {
    Ph::template fjac<0>(P, fm, fp, Jn, Jp);
    if constexpr (WV) Ph::template flux<0>(P, fm, fp, F);
  }
  // the lane's columns of the two flux Jacobians, rows in the physical frame
  double JnP[M][NV], JpP[M][NV];
  lane_cols<Ph, NV>(Jn, a, hv != 0, JnP);
  lane_cols<Ph, NV>(Jp, a, hv != 0, JpP);
  // the face math and the block products apart (see k_jac_weno_staged)
  wave_sync();
  double *stL = stage + (rL >= 0 ? rL : 0) * PERP, *stR = stage + (rR >= 0 ? rR : 0) * PERP;
  // phase 1: as the plus cell's left face, offsets -W..-1 (stencil points
  // 0..W-1); as the minus cell's right face, offsets 1..W (points W..2W-1);
  // the first neighbours' source term enters entry (p_a(K), j) of the axis
  // frame, row K of the physical one
#pragma unroll
  for (int m = 0; m < W; ++m) {
    double b[M][NV];
    if (rR >= 0) {
      face_cols_p<M, NV, G>(h, JnP, JpP, gm, gp, m, b);
      if (m == W - 1) {
#pragma unroll
        for (int K = 0; K < M; ++K)
#pragma unroll
          for (int v = 0; v < NV; ++v)
            b[K][v] += Sources<Ph>::offdiag(P, false, a, a == 0 ? AxisPerm<Ph>::p(K, 0) : AxisPerm<Ph>::p(K, 1),
                                            j0 + v);
      }
      put_cols<Ph, RS, false, NV>(stR + slotR(m - W) * M, a, j0, b);
    }
    if (rL >= 0) {
      face_cols_p<M, NV, G>(-h, JnP, JpP, gm, gp, W + m, b);
      if (m == 0) {
#pragma unroll
        for (int K = 0; K < M; ++K)
#pragma unroll
          for (int v = 0; v < NV; ++v)
            b[K][v] += Sources<Ph>::offdiag(P, false, a, a == 0 ? AxisPerm<Ph>::p(K, 0) : AxisPerm<Ph>::p(K, 1),
                                            j0 + v);
      }
      put_cols<Ph, RS, false, NV>(stL + slotL(m + 1) * M, a, j0, b);
    }
  }
  wave_sync();
  // phase 2: plus cell's offsets 1..W-1 (points W+1..2W-1), minus cell's
  // offsets -W+1..-1 (points 0..W-2)
#pragma unroll
  for (int m = 0; m + 1 < W; ++m) {
    double b[M][NV];
    if (rR >= 0) {
      face_cols_p<M, NV, G>(h, JnP, JpP, gm, gp, W + 1 + m, b);
      put_cols<Ph, RS, true, NV>(stR + slotR(1 + m) * M, a, j0, b);
    }
    if (rL >= 0) {
      face_cols_p<M, NV, G>(-h, JnP, JpP, gm, gp, m, b);
      put_cols<Ph, RS, true, NV>(stL + slotL(m - W + 1) * M, a, j0, b);
    }
  }
  // diagonal blocks and velocity: x faces as left faces (store), x faces as
  // right faces, y faces as left faces, y faces as right faces (add)
#pragma unroll
  for (int ph = 0; ph < 4; ++ph) {
    wave_sync();
    const int pa = ph >> 1;
    const bool plus = (ph & 1) == 0; // the face is the left face of its plus cell
    const int rr = plus ? rR : rL;
    if (act && a == pa && rr >= 0) {
      double b[M][NV];
      const double sh = plus ? h : -h;
      const int i = plus ? W : W - 1;
      face_cols_p<M, NV, G>(sh, JnP, JpP, gm, gp, i, b);
      double *dst = stage + rr * PERP + (edge ? int(plus ? slR[0] : slL[0]) : 2 * W) * M;
      if (ph == 0) put_cols<Ph, RS, false, NV>(dst, a, j0, b);
      else put_cols<Ph, RS, true, NV>(dst, a, j0, b);
      if (WV && hv == 0) {
#pragma unroll
        for (int k = 0; k < M; ++k) {
          const double v = sh * perm_pick<Ph>([&](int x) { return F[x]; }, a, k);
          vst[rr * M + k] = (ph == 0 ? 0. : vst[rr * M + k]) + v;
        }
      }
    }
  }
  wave_sync();
  // the rows' sources, one lane per row
  if (lane < NR && lane % TX < nax && lane / TX < nay) {
    const int64_t r = (y0 + lane / TX) * nx + x0 + lane % TX;
    double *st = stage + lane * PERP;
    const int s0 = edge ? int(P.slot[r * P.npos]) : 2 * W;
    double self[M][M], acc[M];
#pragma unroll
    for (int k = 0; k < M; ++k) {
      acc[k] = WV ? vst[lane * M + k] : 0.;
#pragma unroll
      for (int j = 0; j < M; ++j) self[k][j] = st[k * RS + s0 * M + j];
    }
    Sources<Ph>::apply(P, r, false, U, int32_t(r), acc, &self[0][0]);
#pragma unroll
    for (int k = 0; k < M; ++k) {
#pragma unroll
      for (int j = 0; j < M; ++j) st[k * RS + s0 * M + j] = self[k][j];
      if constexpr (WV) V[r * M + k] = acc[k];
    }
  }
  wave_sync();
  // each tile row: nax consecutive CSR rows, one contiguous range
  for (int cj = 0; cj < nay; ++cj) {
    const int64_t r0 = (y0 + cj) * nx + x0;
    double *dst = J + P.rowbase[r0];
    const double *src = stage + cj * TX * PERP;
    for (int c = 0; c < nax; ++c)
#pragma unroll
      for (int e = lane; e < PER; e += 64) store_stream(dst + c * PER + e, src[c * PERP + e]);
  }
}

template <class Ph, int W, bool WV>
__global__ void __launch_bounds__(64) k_jac_weno_tiled(View P, Box B, const double *__restrict__ U,
                                                         double *__restrict__ V, double *__restrict__ J) {
  __shared__ double smem[WenoTileLds<Ph, W>::DOUBLES];
  jac_weno_tiled_body<Ph, W, WV>(P, B, int64_t(blockIdx.x), U, V, J, smem);
}

template <class Ph, int W>
int launch_jac_weno_tiled(const View &v, const Box &B, const double *U, double *V, double *J, hipStream_t s) {
  constexpr int TX = WenoTile<Ph>::TX, TY = WenoTile<Ph>::TY;
  const int64_t ntx = (B.hi[0] - B.lo[0] + TX - 1) / TX, nty = (B.hi[1] - B.lo[1] + TY - 1) / TY;
  const unsigned nb = unsigned(ntx * nty);
  if (V) hipLaunchKernelGGL((k_jac_weno_tiled<Ph, W, true>), dim3(nb), dim3(64), 0, s, v, B, U, V, J);
  else hipLaunchKernelGGL((k_jac_weno_tiled<Ph, W, false>), dim3(nb), dim3(64), 0, s, v, B, U, V, J);
  PDA_HIP_CHECK(hipGetLastError());
  return PDA_OK;
}

template <class Ph>
int jac_weno_tiled(const View &v, const Box &B, int w, const double *U, double *V, double *J, hipStream_t s) {
  return w == 2 ? launch_jac_weno_tiled<Ph, 2>(v, B, U, V, J, s) : launch_jac_weno_tiled<Ph, 3>(v, B, U, V, J, s);
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

template <class Ph>
static int vel_dispatch(const View &v, int w, const double *U, double *V, hipStream_t s) {
  const int bs = 64; // one-wave groups fit next to the structured kernels
  if (v.nvisit == 0) return PDA_OK;
  const unsigned nb = unsigned((v.nvisit + bs - 1) / bs);
  if (w == 1) hipLaunchKernelGGL((k_velocity<Ph, 1>), dim3(nb), dim3(bs), 0, s, v, U, V);
  else if (w == 2) hipLaunchKernelGGL((k_velocity<Ph, 2>), dim3(nb), dim3(bs), 0, s, v, U, V);
  else hipLaunchKernelGGL((k_velocity<Ph, 3>), dim3(nb), dim3(bs), 0, s, v, U, V);
  PDA_HIP_CHECK(hipGetLastError());
  return PDA_OK;
}

template <class Ph>
static int vel_stage_dispatch(const View &v, int w, const double *U, const RkStage &S, hipStream_t s) {
  const int bs = 64;
  if (v.nvisit == 0) return PDA_OK;
  const unsigned nb = unsigned((v.nvisit + bs - 1) / bs);
  if (w == 1) hipLaunchKernelGGL((k_velocity_stage<Ph, 1>), dim3(nb), dim3(bs), 0, s, v, U, S);
  else if (w == 2) hipLaunchKernelGGL((k_velocity_stage<Ph, 2>), dim3(nb), dim3(bs), 0, s, v, U, S);
  else hipLaunchKernelGGL((k_velocity_stage<Ph, 3>), dim3(nb), dim3(bs), 0, s, v, U, S);
  PDA_HIP_CHECK(hipGetLastError());
  return PDA_OK;
}

template <class Ph>
static int jac_dispatch(const View &v, int w, const double *U, double *V, double *J,
                        hipStream_t s) {
  const int bs = 64; // one-wave groups fit next to the structured kernels
  if (v.nvisit == 0) return PDA_OK;
  const unsigned nb = unsigned((v.nvisit + bs - 1) / bs);
  if (V) {
    if (w == 1) hipLaunchKernelGGL((k_jacobian<Ph, 1, true>), dim3(nb), dim3(bs), 0, s, v, U, V, J);
    else if (w == 2) hipLaunchKernelGGL((k_jacobian<Ph, 2, true>), dim3(nb), dim3(bs), 0, s, v, U, V, J);
    else hipLaunchKernelGGL((k_jacobian<Ph, 3, true>), dim3(nb), dim3(bs), 0, s, v, U, V, J);
  } else {
    if (w == 1) hipLaunchKernelGGL((k_jacobian<Ph, 1, false>), dim3(nb), dim3(bs), 0, s, v, U, V, J);
    else if (w == 2) hipLaunchKernelGGL((k_jacobian<Ph, 2, false>), dim3(nb), dim3(bs), 0, s, v, U, V, J);
    else hipLaunchKernelGGL((k_jacobian<Ph, 3, false>), dim3(nb), dim3(bs), 0, s, v, U, V, J);
  }
  PDA_HIP_CHECK(hipGetLastError());
  return PDA_OK;
}

// zero fill as a kernel (a memset node would not be safe inside captured graphs)
__global__ void k_fill(double *x, int64_t n, double v) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    x[i] = v;
}

// the rows the generic kernels cover: all of them, or those outside the box
View generic_view(const pda_problem_s &p, const DevBuf *rows, int64_t count) {
  View v = p.view;
  if (rows) {
    v.rows = rows->as<int32_t>();
    v.nvisit = count;
  } else {
    v.rows = nullptr;
    v.nvisit = v.nrows;
  }
  return v;
}

static int vel_generic(const View &v, const pda_problem_s &p, const double *U, double *V,
                       hipStream_t s) {
  switch (p.family) {
    case PDA_EULER1D: return vel_dispatch<PhEuler1>(v, p.w, U, V, s);
    case PDA_EULER2D: return vel_dispatch<PhEuler2>(v, p.w, U, V, s);
    case PDA_EULER3D: return vel_dispatch<PhEuler3>(v, p.w, U, V, s);
    case PDA_SWE2D: return vel_dispatch<PhSwe>(v, p.w, U, V, s);
    case PDA_BURGERS2D: return vel_dispatch<PhBurgers>(v, p.w, U, V, s);
    case PDA_ADR2D: return vel_dispatch<PhAdr>(v, p.w, U, V, s);
    default: set_error("velocity: unsupported family"); return PDA_E_UNSUPPORTED;
  }
}

static int jac_generic(const View &v, const pda_problem_s &p, const double *U, double *V,
                       double *J, hipStream_t s) {
  switch (p.family) {
    case PDA_EULER1D: return jac_dispatch<PhEuler1>(v, p.w, U, V, J, s);
    case PDA_EULER2D: return jac_dispatch<PhEuler2>(v, p.w, U, V, J, s);
    case PDA_EULER3D: return jac_dispatch<PhEuler3>(v, p.w, U, V, J, s);
    case PDA_SWE2D: return jac_dispatch<PhSwe>(v, p.w, U, V, J, s);
    case PDA_BURGERS2D: return jac_dispatch<PhBurgers>(v, p.w, U, V, J, s);
    case PDA_ADR2D: return jac_dispatch<PhAdr>(v, p.w, U, V, J, s);
    default: set_error("jacobian: unsupported family"); return PDA_E_UNSUPPORTED;
  }
}

static bool box_velocity(const pda_problem_s &p) {
  return p.use_box && p.vbox.valid && p.vbox.cells() > 0 && has_box_velocity(p);
}
static bool box_jacobian(const pda_problem_s &p) {
  return p.use_box && p.box.valid && p.box.cells() > 0 && has_box_kernels(p) && !p.view.dup &&
         p.w == 1;
}

// With a structured box the box kernel and the generic kernel over the rows
// outside the box run back to back on the caller's stream (disjoint rows);
// a box launcher may take the outside rows into its own launch.
int launch_velocity(const pda_problem_s &p, const double *U, double, double *V, hipStream_t s) {
  if (p.view.nrows == 0) return PDA_OK;
  if (!box_velocity(p)) return vel_generic(generic_view(p, nullptr, 0), p, U, V, s);
  int rc = p.h_vrem_rows.empty()
               ? PDA_OK
               : vel_generic(generic_view(p, &p.d_vrem_rows, int64_t(p.h_vrem_rows.size())), p, U, V, s);
  if (rc) return rc;
  return launch_velocity_box(p, U, V, s);
}

// the problems whose velocity launches can form an explicit stage in their
// stores: the generic row kernel, and the 2D box march (with the generic
// rows outside its box)
bool velocity_stage_fits(const pda_problem_s &p) {
  return p.view.nrows > 0 && (!box_velocity(p) || p.dim == 2);
}

int launch_velocity_stage(const pda_problem_s &p, const double *U, const RkStage &S, hipStream_t s) {
  if (!velocity_stage_fits(p)) return PDA_E_UNSUPPORTED;
  if (!S.out || S.nx < 1 || S.nx > 4) {
    set_error("velocity stage: invalid stage");
    return PDA_E_INVALID_ARG;
  }
  if (box_velocity(p)) {
    if (!p.h_vrem_rows.empty()) {
      const View v = generic_view(p, &p.d_vrem_rows, int64_t(p.h_vrem_rows.size()));
      int rc = PDA_E_UNSUPPORTED;
      switch (p.family) {
        case PDA_EULER2D: rc = vel_stage_dispatch<PhEuler2>(v, p.w, U, S, s); break;
        case PDA_SWE2D: rc = vel_stage_dispatch<PhSwe>(v, p.w, U, S, s); break;
        case PDA_BURGERS2D: rc = vel_stage_dispatch<PhBurgers>(v, p.w, U, S, s); break;
      }
      if (rc) return rc;
    }
    return launch_velocity_box_stage(p, U, S, s);
  }
  const View v = generic_view(p, nullptr, 0);
  switch (p.family) {
    case PDA_EULER1D: return vel_stage_dispatch<PhEuler1>(v, p.w, U, S, s);
    case PDA_EULER2D: return vel_stage_dispatch<PhEuler2>(v, p.w, U, S, s);
    case PDA_EULER3D: return vel_stage_dispatch<PhEuler3>(v, p.w, U, S, s);
    case PDA_SWE2D: return vel_stage_dispatch<PhSwe>(v, p.w, U, S, s);
    case PDA_BURGERS2D: return vel_stage_dispatch<PhBurgers>(v, p.w, U, S, s);
    case PDA_ADR2D: return vel_stage_dispatch<PhAdr>(v, p.w, U, S, s);
    default: set_error("velocity: unsupported family"); return PDA_E_UNSUPPORTED;
  }
}

// the fused 1D stepping path: 1D Euler on a natural-order full mesh (row r
// is cell r, its neighbours r -+ 1..w, wrapped when periodic), SSPRK3 or
// RK4, with the opposite ghost rule or a periodic grid
bool steps1d_fits(const pda_problem_s &p, int scheme, bool *periodic_out) {
  const pda_mesh_s &M = *p.mesh;
  if (p.family != PDA_EULER1D || M.dim != 1 || (scheme != PDA_SSPRK3 && scheme != PDA_RK4)) return false;
  if (M.n_sample != M.n_stencil || p.view.nrows != M.n_sample) return false;
  const int64_t n = M.n_sample;
  if (n < 2 * p.w + 1 || M.stencil < 2 * p.w + 1) return false;
  bool periodic = true, open = true;
  for (int64_t r = 0; r < n; ++r) {
    if (M.g(r, 0) != r) return false;
    for (int L = 0; L < p.w; ++L) {
      const int32_t lft = M.g(r, col_of(1, 0, 0, L)), rgt = M.g(r, col_of(1, 0, 1, L));
      const int64_t wl = ((r - 1 - L) % n + n) % n, wr = (r + 1 + L) % n;
      const int64_t ol = r - 1 - L >= 0 ? r - 1 - L : -1, orr = r + 1 + L < n ? r + 1 + L : -1;
      periodic = periodic && lft == wl && rgt == wr;
      open = open && lft == ol && rgt == orr;
    }
  }
  *periodic_out = periodic;
  if (periodic) return true;
  return open && p.view.ghost_rule == GHOST_OPPOSITE;
}

// steps per launch: the most that leave at least 32 owned rows of the
// 128-row region (SSPRK3 WENO5: 5 steps, 38 owned rows)
int steps1d_per_launch(const pda_problem_s &p, int scheme) {
  const int per = (scheme == PDA_RK4 ? 4 : 3) * p.w;  // halo rows per side per step
  const int k = (S1_ROWS - 32) / (2 * per);
  return k < 1 ? 1 : (k > STEPS1D_MAX ? STEPS1D_MAX : k);
}

int launch_steps1d(const pda_problem_s &p, int scheme, const double *Uin, double *Uout, int nsteps, double dt,
                   bool periodic, hipStream_t s) {
  const int kmax = steps1d_per_launch(p, scheme);
  if (nsteps < 1 || nsteps > kmax) {
    set_error("steps1d: steps per launch out of range");
    return PDA_E_INVALID_ARG;
  }
  const int nst = scheme == PDA_RK4 ? 4 : 3, nbuf = scheme == PDA_RK4 ? 6 : 3;
  const int H = nst * p.w * kmax, own = S1_ROWS - 2 * H;
  const int64_t n = p.mesh->n_sample;
  constexpr int M = PhEuler1::M;
  const unsigned threads = 2 * S1_ROWS;
  const size_t lds = sizeof(double) * size_t(nbuf) * size_t(S1_ROWS) * M;
  const dim3 grid{unsigned((n + own - 1) / own), 1, 1};
  if (own < 1 || lds > 64 * 1024) {
    set_error("steps1d: region too large");
    return PDA_E_INVALID_ARG;
  }
  // the coefficients exactly as the stage-by-stage step forms them
  Steps1dCoef C{};
  if (scheme == PDA_SSPRK3) {
    const double third = 1. / 3., twoThird = 2. / 3., quarter = 1. / 4., threeQuarter = 3. / 4.;
    C.c[0][0] = 1.;
    C.cv[0] = dt;
    C.c[1][0] = quarter;
    C.c[1][1] = threeQuarter;
    C.cv[1] = quarter * dt;
    C.c[2][0] = third;
    C.c[2][1] = twoThird;
    C.cv[2] = twoThird * dt;
  } else {
    const double h2 = dt / 2., h6 = dt / 6., h3 = dt / 3.;
    C.c[0][0] = C.c[1][0] = C.c[2][0] = 1.;
    C.cv[0] = h2;
    C.cv[1] = h2;
    C.cv[2] = dt;
    C.c[3][0] = 1.;
    C.c[3][1] = h6;
    C.c[3][2] = h3;
    C.c[3][3] = h3;
    C.cv[3] = h6;
  }
  const View &v = p.view;
  const int per = periodic ? 1 : 0;
#define PDA_S1(WW, SCH) \
  hipLaunchKernelGGL((k_steps1d<PhEuler1, WW, SCH>), grid, dim3(threads), lds, s, v, n, per, own, H, nsteps, Uin, \
                     Uout, C)
  if (scheme == PDA_SSPRK3) {
    if (p.w == 1) PDA_S1(1, PDA_SSPRK3);
    else if (p.w == 2) PDA_S1(2, PDA_SSPRK3);
    else PDA_S1(3, PDA_SSPRK3);
  } else {
    if (p.w == 1) PDA_S1(1, PDA_RK4);
    else if (p.w == 2) PDA_S1(2, PDA_RK4);
    else PDA_S1(3, PDA_RK4);
  }
#undef PDA_S1
  PDA_HIP_CHECK(hipGetLastError());
  return PDA_OK;
}

int launch_jacobian(const pda_problem_s &p, const double *U, double, double *J, double *V,
                    hipStream_t s) {
  if (p.view.nrows == 0) return PDA_OK;
  if (!box_jacobian(p)) {
    if (p.view.dup) {
      hipLaunchKernelGGL(k_fill, dim3(1024), dim3(256), 0, s, J, p.nnz, 0.);
      PDA_HIP_CHECK(hipGetLastError());
    }
    if (!p.view.dup && p.w > 1 && p.family != PDA_ADR2D) {
      // WENO Jacobian: staged inner rows, then the near-boundary rows (first
      // order with BC factors) with the generic row code; for the 2D
      // two- and three-dof families, the box's rows with each face once
      // (when the box holds no near-boundary row) and the staged kernel on
      // the rows outside it
      if (p.dim == 2 && (p.family == PDA_BURGERS2D || p.family == PDA_SWE2D || p.family == PDA_EULER2D) &&
          p.box.valid && p.box.cells() > 0) {
        // tiles over the whole grid when it is periodic along both axes,
        // else over the box when it holds no near-boundary row (its
        // non-periodic sides at least the mesh stencil's half width in)
        const int wm = (p.mesh->stencil - 1) / 2;
        const int per = p.mesh->periodic_mask & 3;
        Box tb = p.box;
        bool ok = true;
        if (per == 3) {
          for (int ax = 0; ax < 2; ++ax) {
            tb.lo[ax] = 0;
            tb.hi[ax] = tb.n[ax];
          }
        } else if (per == 0) {
          for (int ax = 0; ax < 2; ++ax) ok = ok && p.box.lo[ax] >= wm;
        } else {
          ok = false;
        }
        if (ok) {
          const View vb = generic_view(p, nullptr, 0);
          int rc = p.family == PDA_BURGERS2D ? jac_weno_tiled<PhBurgers>(vb, tb, p.w, U, V, J, s)
                   : p.family == PDA_SWE2D   ? jac_weno_tiled<PhSwe>(vb, tb, p.w, U, V, J, s)
                                             : jac_weno_tiled<PhEuler2>(vb, tb, p.w, U, V, J, s);
          if (rc) return rc;
          if (per == 0 && int64_t(p.h_rem_rows.size()) > p.n_near) {
            // the rows outside the box that are not near-boundary rows (the
            // staged kernel skips those)
            const View vr = generic_view(p, &p.d_rem_rows, int64_t(p.h_rem_rows.size()));
            rc = p.family == PDA_BURGERS2D ? jac_weno_staged<PhBurgers>(vr, p.w, U, V, J, s)
                 : p.family == PDA_SWE2D   ? jac_weno_staged<PhSwe>(vr, p.w, U, V, J, s)
                                           : jac_weno_staged<PhEuler2>(vr, p.w, U, V, J, s);
            if (rc) return rc;
          }
          if (p.n_near == 0) return PDA_OK;
          return jac_generic(generic_view(p, &p.d_near_rows, p.n_near), p, U, V, J, s);
        }
      }
      const View v = generic_view(p, nullptr, 0);
      int rc;
      switch (p.family) {
        case PDA_EULER1D: rc = jac_weno_staged<PhEuler1>(v, p.w, U, V, J, s); break;
        case PDA_EULER2D: rc = jac_weno_staged<PhEuler2>(v, p.w, U, V, J, s); break;
        case PDA_EULER3D: rc = jac_weno_staged<PhEuler3>(v, p.w, U, V, J, s); break;
        case PDA_SWE2D: rc = jac_weno_staged<PhSwe>(v, p.w, U, V, J, s); break;
        case PDA_BURGERS2D: rc = jac_weno_staged<PhBurgers>(v, p.w, U, V, J, s); break;
        default: set_error("jacobian: unsupported family"); return PDA_E_UNSUPPORTED;
      }
      if (rc || p.n_near == 0) return rc;
      return jac_generic(generic_view(p, &p.d_near_rows, p.n_near), p, U, V, J, s);
    }
    if (!V && !p.view.dup && p.w == 1 && (p.family == PDA_EULER2D || p.family == PDA_EULER3D)) {
      // first-order Euler Jacobian alone
      const View v = generic_view(p, nullptr, 0);
      if (p.family == PDA_EULER3D && p.use_box && p.box.valid && p.box.cells() > 0) {
        // the box's x lines marched in z, then the rows outside the box
        // (boundary rows, with ghosts) staged four lanes per row
        const int rc = launch_jacobian_fo_euler3d_box(p, U, J, s);
        if (rc || p.h_rem_rows.empty()) return rc;
        const View vr = generic_view(p, &p.d_rem_rows, int64_t(p.h_rem_rows.size()));
        hipLaunchKernelGGL((k_jac_fo_euler_staged<3>), dim3(unsigned((vr.nvisit + SJ_ROWS - 1) / SJ_ROWS)),
                           dim3(64), 0, s, vr, U, J);
        PDA_HIP_CHECK(hipGetLastError());
        return PDA_OK;
      }
      // every row, near-boundary rows included, staged four lanes per row
      const unsigned nb = unsigned((v.nvisit + SJ_ROWS - 1) / SJ_ROWS);
      if (p.family == PDA_EULER2D)
        hipLaunchKernelGGL((k_jac_fo_euler_staged<2>), dim3(nb), dim3(64), 0, s, v, U, J);
      else hipLaunchKernelGGL((k_jac_fo_euler_staged<3>), dim3(nb), dim3(64), 0, s, v, U, J);
      PDA_HIP_CHECK(hipGetLastError());
      return PDA_OK;
    }
    return jac_generic(generic_view(p, nullptr, 0), p, U, V, J, s);
  }
  bool rest_done = false;
  int rc = launch_jacobian_box(p, U, J, V, s, &rest_done);
  if (rc || rest_done || p.h_rem_rows.empty()) return rc;
  return jac_generic(generic_view(p, &p.d_rem_rows, int64_t(p.h_rem_rows.size())), p, U, V, J, s);
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
