// pda_box_kernels.hip -- kernels for the cells of a natural-order structured
// 2D mesh whose whole scheme stencil lies inside the grid (the "box"); the
// generic row kernels cover the rows outside it.  They compute the same
// quantities as the per-row kernels (same face states, fluxes and flux
// Jacobians, same per-cell accumulation order; values agree to rounding)
// but are laid out for the hardware:
//   - velocity: one wavefront marches a row segment up a strip of rows,
//     evaluating each face once (x faces shared by lane shuffles, y faces
//     carried in registers); in 2D each cell is reconstructed once per axis
//     from rows held in registers;
//   - first-order Jacobian: the blocks are assembled in LDS in CSR order and
//     leave as one contiguous run per row segment (coalesced 16-byte stores);
//     for 2D Euler two lanes per cell evaluate each face once, for 3D Euler
//     three lanes per cell (one per axis) march the x lines of a strip of
//     planes.
// Cell ids follow the structured layout gid = j*nx + i (graph row == gid).
#include "pda_rows.hpp"

namespace pda {
namespace {
using namespace fam;

// Workgroups are dealt round-robin to the 8 XCDs (each with its own L2):
// workgroup b runs on XCD b % 8.  Renumbering b -> the b/8-th item of XCD
// b%8's contiguous share of [0, n) gives each XCD a contiguous range of
// strips, so state rows shared by neighbouring strips are read once per L2.
__device__ __forceinline__ int64_t xcd_contiguous(int64_t b, int64_t n) {
  const int64_t q = n / 8, r = n % 8, x = b % 8, k = b / 8;
  return x * q + (x < r ? x : r) + k;
}

// ------------------------------------------------ velocity 2D, row marching
// A wavefront owns a segment of up to 63 consecutive box cells of a grid row
// and marches up a strip of rows.  Lane l evaluates the x face on the left
// of cell l (lane 63 the face on the right of cell 62) and the y face above
// cell l; the x face on the right comes from lane l+1 by a lane shuffle and
// the y face below is the one evaluated for the previous row, carried in
// registers.  Every face is reconstructed once (plus one row of y faces per
// strip) and no workgroup barrier is needed.
constexpr int VSEG = 63;

// state of cell (i, j, k) with the index along axis A possibly outside
// [0, n): periodic wrap, or the mirror cell times the side's ghost sign (the
// GHOST_MIRROR rule, whose ghost values do not depend on the reading row)
template <int M, int A>
__device__ __forceinline__ void load_cell_mapped(const View &P, const Box &B, const double *__restrict__ U,
                                                 int64_t i, int64_t j, int64_t k, double *q) {
  int64_t c = A == 0 ? i : (A == 1 ? j : k);
  const int64_t n = B.n[A];
  int side = -1;
  if (c < 0) {
    if (B.mode[A] == 1) c += n;
    else { c = -1 - c; side = 0; }
  } else if (c >= n) {
    if (B.mode[A] == 1) c -= n;
    else { c = 2 * n - 1 - c; side = 1; }
  }
  const int64_t ii = A == 0 ? c : i, jj = A == 1 ? c : j, kk = A == 2 ? c : k;
  const int64_t cell = (kk * B.n[1] + jj) * B.n[0] + ii;
#pragma unroll
  for (int k = 0; k < M; ++k) {
    const double v = U[cell * M + k];
    q[k] = side < 0 ? v : P.gsign[A][side][k] * v;
  }
}

// flux of the face on the low side (along A) of cell (i, j, k)
template <class Ph, int W, int A>
__device__ __forceinline__ void face_flux_fast(const View &P, const Box &B, const double *__restrict__ U,
                                               int64_t i, int64_t j, int64_t k, double *F) {
  constexpr int M = Ph::M;
  double st[2 * W][M];
#pragma unroll
  for (int p = 0; p < 2 * W; ++p)
    load_cell_mapped<M, A>(P, B, U, A == 0 ? i + p - W : i, A == 1 ? j + p - W : j, A == 2 ? k + p - W : k,
                           st[p]);
  double qm[M], qp[M];
#pragma unroll
  for (int k = 0; k < M; ++k) {
    double v[2 * W];
#pragma unroll
    for (int p = 0; p < 2 * W; ++p) v[p] = st[p][k];
    if constexpr (W == 1) {
      qm[k] = v[0];
      qp[k] = v[1];
    } else if constexpr (W == 2) {
      phys::weno3_1div(v, qm[k], qp[k]);
    } else {
      phys::weno5_1div(v, qm[k], qp[k]);
    }
  }
  Ph::template flux<A>(P, qm, qp, F);
}

template <class Ph, int W>
__device__ __forceinline__ void vel_march_body(const View &P, const Box &B, int rows_per_strip, int64_t bid,
                                               const double *__restrict__ U, double *__restrict__ V) {
  constexpr int M = Ph::M;
  const int lane = threadIdx.x;
  const int64_t nx = B.n[0], ny = B.n[1];
  const int64_t bw = B.hi[0] - B.lo[0];
  const int64_t nseg = (bw + VSEG - 1) / VSEG;
  const int64_t nstrip = (B.hi[1] - B.lo[1] + rows_per_strip - 1) / rows_per_strip;
  // workgroup -> (plane, strip, segment); one plane in 2D
  const int64_t plane = bid / (nseg * nstrip), rem = bid - plane * nseg * nstrip;
  const int64_t seg = rem % nseg, strip = rem / nseg;
  const int64_t kz = B.lo[2] + plane;
  const int64_t iseg = B.lo[0] + seg * VSEG;
  const int ncell = int(bw - seg * VSEG < VSEG ? bw - seg * VSEG : VSEG);
  const int64_t j0 = B.lo[1] + strip * rows_per_strip;
  const int64_t j1 = j0 + rows_per_strip < B.hi[1] ? j0 + rows_per_strip : B.hi[1];
  // Lane l evaluates the x face on the left of column i = iseg + l; lane
  // ncell (and the idle lanes after it) the face on the right of the last
  // cell, i = iseg + ncell <= B.hi[0].  With a full-grid box that column is
  // n, outside the grid: x stencils there are mapped by the axis mode, and
  // everything else those lanes compute (y faces, sources; never stored)
  // uses the last column ic instead.
  const int64_t i = iseg + (lane <= ncell ? lane : ncell);
  const int64_t ic = i < nx ? i : nx - 1;
  const bool own = lane < ncell;
  const double hx = P.hinv[0], hy = P.hinv[1], hz = P.hinv[2];
  // one call site per face kind, so that every face value is the same
  // whatever the strip height (the first pass only evaluates the face below
  // the strip)
  double fyb[M];
  for (int64_t j = j0 - 1; j < j1; ++j) {
    double fy[M];
    face_flux_fast<Ph, W, 1>(P, B, U, ic, j + 1, kz, fy);
    if (j < j0) {
#pragma unroll
      for (int k = 0; k < M; ++k) fyb[k] = fy[k];
      continue;
    }
    const int64_t c = (kz * ny + j) * nx + ic;
    double fx[M], acc[M];
    face_flux_fast<Ph, W, 0>(P, B, U, i, j, kz, fx);
#pragma unroll
    for (int k = 0; k < M; ++k) {
      const double fr = __shfl_down(fx[k], 1, 64);
      acc[k] = 0. + hx * (fx[k] - fr);
      acc[k] += hy * (fyb[k] - fy[k]);
      fyb[k] = fy[k];
    }
    if constexpr (Ph::DIM == 3) {
      // z faces below and above the cell (no carry across planes)
      double fzb[M], fza[M];
      face_flux_fast<Ph, W, 2>(P, B, U, ic, j, kz, fzb);
      face_flux_fast<Ph, W, 2>(P, B, U, ic, j, kz + 1, fza);
#pragma unroll
      for (int k = 0; k < M; ++k) acc[k] += hz * (fzb[k] - fza[k]);
    }
    Sources<Ph>::apply(P, c, false, U, int32_t(c), acc, nullptr);
    if (own) {
#pragma unroll
      for (int k = 0; k < M; ++k) V[c * M + k] = acc[k];
    }
  }
}

// ------------------------------------------ velocity 2D, cell-centred march
// A wavefront holds 64 consecutive columns of a grid row, the 64 - 2W box
// cells it owns and W columns on each side, and marches up a strip of rows
// with the 2W - 1 rows a y reconstruction reads held in registers: one row
// is loaded per step.  Each lane reconstructs its own cell once per axis,
// the cell's low-side and high-side values sharing the WENO windows
// (weno5_cell_1div).  The x face on the low side of lane l's cell takes the
// high-side value of lane l - 1 and the low-side value of lane l, the x
// stencil being the neighbouring lanes' row values (DPP wavefront shifts,
// no LDS); the face on the high side is lane l + 1's.  The y face above row r takes the high-side
// value of row r, carried from the previous step, and the low-side value of
// row r + 1.  Every face value equals the face-by-face march's bit for bit.
template <int W> __host__ __device__ constexpr int vcc_width() { return 64 - 2 * W; }

// state of cell (i, j) of a 2D grid, each index possibly outside the grid
// and mapped along its axis as load_cell_mapped maps it
template <int M>
__device__ __forceinline__ void load_cell_mapped2(const View &P, const Box &B, const double *__restrict__ U,
                                                  int64_t i, int64_t j, double *q) {
  double sg[M];
#pragma unroll
  for (int k = 0; k < M; ++k) sg[k] = 1.;
  int64_t c[2] = {i, j};
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    const int64_t n = B.n[a];
    int side = -1;
    if (c[a] < 0) {
      if (B.mode[a] == 1) c[a] += n;
      else { c[a] = -1 - c[a]; side = 0; }
    } else if (c[a] >= n) {
      if (B.mode[a] == 1) c[a] -= n;
      else { c[a] = 2 * n - 1 - c[a]; side = 1; }
    }
    if (side >= 0) {
#pragma unroll
      for (int k = 0; k < M; ++k) sg[k] *= P.gsign[a][side][k];
    }
  }
  const int64_t cell = c[1] * B.n[0] + c[0];
#pragma unroll
  for (int k = 0; k < M; ++k) q[k] = sg[k] * U[cell * M + k];
}

// value of lane l - 1 (lane_prev) or l + 1 (lane_next) of the wavefront,
// by DPP wavefront shifts of the two halves; lane 0 (lane 63) gets 0
__device__ __forceinline__ double lane_dpp(double v, int ctrl_is_next) {
  const long long b = __double_as_longlong(v);
  int lo, hi;
  if (ctrl_is_next) {
    lo = __builtin_amdgcn_update_dpp(0, int(b), 0x130, 0xf, 0xf, false);        // wave_shl:1
    hi = __builtin_amdgcn_update_dpp(0, int(b >> 32), 0x130, 0xf, 0xf, false);
  } else {
    lo = __builtin_amdgcn_update_dpp(0, int(b), 0x138, 0xf, 0xf, false);        // wave_shr:1
    hi = __builtin_amdgcn_update_dpp(0, int(b >> 32), 0x138, 0xf, 0xf, false);
  }
  return __longlong_as_double((long long)(unsigned(lo)) | ((long long)hi << 32));
}
__device__ __forceinline__ double lane_prev(double v) { return lane_dpp(v, 0); }
__device__ __forceinline__ double lane_next(double v) { return lane_dpp(v, 1); }

// both face values of the middle cell of s[0 .. 2W-2] (W = 1: the value)
template <int W>
__device__ __forceinline__ void recon_cell(const double *s, double &hi, double &lo) {
  if constexpr (W == 1) {
    hi = s[0];
    lo = s[0];
  } else if constexpr (W == 2) {
    phys::weno3_cell_1div(s[0], s[1], s[2], hi, lo);
  } else {
    phys::weno5_cell_1div(s[0], s[1], s[2], s[3], s[4], hi, lo);
  }
}

// STG: instead of V, each row writes the explicit stage of S (k_i = v into
// S.kout when set, and S.out = c_0 x_0 + ... + cv v in the order of
// k_velocity_stage and k_lincomb, so a fused step equals the launches it
// replaces bit for bit); the stage's vectors are indexed from soff (a
// batch instance's state offset)
template <class Ph, int W, bool STG>
__device__ __forceinline__ void vel_cc_body(const View &P, const Box &B, int rows_per_strip, int64_t bid,
                                            const double *__restrict__ U, double *__restrict__ V,
                                            const RkStage &S, int64_t soff) {
  static_assert(Ph::DIM == 2, "2D march");
  constexpr int M = Ph::M, NC = vcc_width<W>(), NR = 2 * W - 1;
  const int lane = threadIdx.x;
  const int64_t nx = B.n[0];
  const int64_t bw = B.hi[0] - B.lo[0];
  const int64_t nseg = (bw + NC - 1) / NC;
  const int64_t seg = bid % nseg, strip = bid / nseg;
  const int64_t iseg = B.lo[0] + seg * NC;
  const int ncell = int(bw - seg * NC < NC ? bw - seg * NC : NC);
  const int64_t j0 = B.lo[1] + strip * rows_per_strip;
  const int64_t j1 = j0 + rows_per_strip < B.hi[1] ? j0 + rows_per_strip : B.hi[1];
  // lane l holds column iseg - W + l; lanes past the last column that the
  // segment's cells read hold that column again
  const int lc = lane < ncell + 2 * W - 1 ? lane : ncell + 2 * W - 1;
  const int64_t i = iseg - W + lc;
  const bool own = lane >= W && lane < W + ncell;
  const double hx = P.hinv[0], hy = P.hinv[1];
  // win[p] = row r + 2 - W + p at step r (W = 1: row r + 1); nxt = the
  // row that enters the window at the next step, loaded one step ahead
  double win[NR][M], nxt[M];
#pragma unroll
  for (int p = 0; p < NR; ++p) load_cell_mapped2<M>(P, B, U, i, j0 - W + p, win[p]);
  load_cell_mapped2<M>(P, B, U, i, j0 - 1 + W, nxt);
  double qhi[M], fb[M];  // high-side y value of row r; flux of the face below row r
  // The loop is neither unrolled nor peeled: each face kind has one copy of
  // its code, so a face's value does not depend on the step that forms it
  // (the first two steps of a strip form the faces below it).
#pragma clang loop unroll(disable)
  for (int64_t r = j0 - 2; r < j1; ++r) {
    if (r > j0 - 2) {
#pragma unroll
      for (int p = 0; p + 1 < NR; ++p)
#pragma unroll
        for (int k = 0; k < M; ++k) win[p][k] = win[p + 1][k];
#pragma unroll
      for (int k = 0; k < M; ++k) win[NR - 1][k] = nxt[k];
      if (r + 1 < j1) load_cell_mapped2<M>(P, B, U, i, r + 1 + W, nxt);
    }
    // y face values of cell (i, r + 1)
    double yhi[M], ylo[M];
#pragma unroll
    for (int k = 0; k < M; ++k) {
      double s[NR];
#pragma unroll
      for (int p = 0; p < NR; ++p) s[p] = win[p][k];
      recon_cell<W>(s, yhi[k], ylo[k]);
    }
    if (r >= j0 - 1) {
      double fa[M];  // the face above row r
      Ph::template flux<1>(P, qhi, ylo, fa);
      if (r >= j0) {
        // x face values of cell (i, r), the stencil from the lanes beside
        double xhi[M], xlo[M];
#pragma unroll
        for (int k = 0; k < M; ++k) {
          double s[NR];
          s[W - 1] = W >= 2 ? win[W >= 2 ? W - 2 : 0][k] : qhi[k];
#pragma unroll
          for (int o = 1; o < W; ++o) {
            s[W - 1 - o] = lane_prev(s[W - o]);
            s[W - 1 + o] = lane_next(s[W - 2 + o]);
          }
          recon_cell<W>(s, xhi[k], xlo[k]);
        }
        double ql[M], fx[M];
#pragma unroll
        for (int k = 0; k < M; ++k) ql[k] = lane_prev(xhi[k]);
        Ph::template flux<0>(P, ql, xlo, fx);
        double acc[M];
#pragma unroll
        for (int k = 0; k < M; ++k) {
          const double fr = lane_next(fx[k]);
          acc[k] = 0. + hx * (fx[k] - fr);
          acc[k] += hy * (fb[k] - fa[k]);
        }
        if (own) {
          const int64_t c = r * nx + i;
          Sources<Ph>::apply(P, c, false, U, int32_t(c), acc, nullptr);
          if constexpr (STG) {
#pragma unroll
            for (int k = 0; k < M; ++k) {
              const int64_t e = soff + c * M + k;
              if (S.kout) S.kout[e] = acc[k];
              double o = S.c[0] * S.x[0][e];
              for (int q = 1; q < S.nx; ++q) o += S.c[q] * S.x[q][e];
              o += S.cv * acc[k];
              S.out[e] = o;
            }
          } else {
#pragma unroll
            for (int k = 0; k < M; ++k) V[c * M + k] = acc[k];
          }
        }
      }
#pragma unroll
      for (int k = 0; k < M; ++k) fb[k] = fa[k];
    }
#pragma unroll
    for (int k = 0; k < M; ++k) qhi[k] = yhi[k];
  }
}

// waves per SIMD the 2D march is compiled for: 3 (up to 168 registers)
// where the rows of a WENO window of three or more dofs do not fit 128
// registers without spilling
template <class Ph, int W> constexpr int vcc_waves() { return W >= 2 && Ph::M >= 3 ? 3 : 4; }

template <class Ph, int W, bool STG>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(vcc_waves<Ph, W>()))) k_vel_cc(View P, Box B, int rows_per_strip,
                                                  const double *__restrict__ U, double *__restrict__ V,
                                                  RkStage S) {
  vel_cc_body<Ph, W, STG>(P, B, rows_per_strip, xcd_contiguous(int64_t(blockIdx.x), int64_t(gridDim.x)), U, V, S, 0);
}

template <class Ph, int W>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(Ph::DIM == 3 ? 2 : 4))) k_vel_march(View P, Box B, int rows_per_strip,
                                                     const double *__restrict__ U,
                                                     double *__restrict__ V) {
  vel_march_body<Ph, W>(P, B, rows_per_strip, xcd_contiguous(int64_t(blockIdx.x), int64_t(gridDim.x)), U, V);
}

// many independent instances (same family and scheme, each with its own
// grid, parameters and state slice) in one launch: blk_inst maps a
// workgroup to its instance, whose workgroups are numbered from blk0
template <class Ph, int W, bool STG>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(vcc_waves<Ph, W>()))) k_vel_march2d_batch(const BatchInst *__restrict__ inst,
                                                           const int32_t *__restrict__ blk_inst,
                                                           const double *__restrict__ U,
                                                           double *__restrict__ V, RkStage S) {
  const BatchInst &I = inst[blk_inst[blockIdx.x]];
  // the stage's vectors are batch states: the instance's slices from I.off
  vel_cc_body<Ph, W, STG>(I.P, I.B, I.rps, int64_t(blockIdx.x) - I.blk0, U + I.off, STG ? nullptr : V + I.off, S,
                          I.off);
}

// ------------------------------------------------- first-order Jacobian 2D
// One wavefront per segment of 64 consecutive box cells of a grid row.  The
// segment's Jacobian rows are contiguous in the CSR value array (all its
// cells have the five blocks [back, left, self, right, front] in that
// column order), so after the blocks are formed the wavefront writes the
// segment as one contiguous run, 32 cells per LDS round.

template <int M> struct Chunk { using T = double2; };
template <> struct Chunk<1> { using T = double; };
template <> struct Chunk<3> { using T = double; };
template <> struct Chunk<5> { using T = double; };

template <class Ph>
__global__ void __launch_bounds__(64) k_jac_fo_box2d(View P, Box B, const double *__restrict__ U,
                                                      double *__restrict__ V,
                                                      double *__restrict__ J) {
  constexpr int M = Ph::M, NB = 5, PER = NB * M * M;
  using CT = typename Chunk<M>::T;
  constexpr int CW = sizeof(CT) / sizeof(double);
  constexpr int NCH = PER / CW;
  __shared__ CT stage[32][NCH + 1];
  const int lane = threadIdx.x;
  const int64_t nx = B.n[0];
  const int64_t bw = B.hi[0] - B.lo[0];
  const int64_t nseg = (bw + 63) / 64;
  const int64_t j = B.lo[1] + int64_t(blockIdx.x) / nseg;
  const int64_t iseg = B.lo[0] + (int64_t(blockIdx.x) % nseg) * 64;
  const int ncell = int(bw - (iseg - B.lo[0]) < 64 ? bw - (iseg - B.lo[0]) : 64);
  const bool act = lane < ncell;
  double blk[NB][M][M];
  if (act) {
    const int64_t c = j * nx + iseg + lane;
    double qS[M], qL[M], qR[M], qB[M], qF[M];
#pragma unroll
    for (int k = 0; k < M; ++k) {
      qS[k] = U[c * M + k];
      qL[k] = U[(c - 1) * M + k];
      qR[k] = U[(c + 1) * M + k];
      qB[k] = U[(c - nx) * M + k];
      qF[k] = U[(c + nx) * M + k];
    }
    const double hx = P.hinv[0], hy = P.hinv[1];
    double Ja[M * M], Jb[M * M], keep[M * M];
    // x: left face (l0 | self) then right face (self | r0)
    Ph::template fjac<0>(P, qL, qS, Ja, Jb);
#pragma unroll
    for (int e = 0; e < M * M; ++e) {
      blk[1][e / M][e % M] = 0. + Ja[e] * hx + Sources<Ph>::offdiag(P, false, 0, e / M, e % M);
      keep[e] = Jb[e];
    }
    Ph::template fjac<0>(P, qS, qR, Ja, Jb);
#pragma unroll
    for (int e = 0; e < M * M; ++e) {
      blk[2][e / M][e % M] = 0. + (keep[e] - Ja[e]) * hx;
      blk[3][e / M][e % M] = 0. + (-Jb[e]) * hx + Sources<Ph>::offdiag(P, false, 0, e / M, e % M);
    }
    // y: back face (b0 | self) then front face (self | f0)
    Ph::template fjac<1>(P, qB, qS, Ja, Jb);
#pragma unroll
    for (int e = 0; e < M * M; ++e) {
      blk[0][e / M][e % M] = 0. + Ja[e] * hy + Sources<Ph>::offdiag(P, false, 1, e / M, e % M);
      keep[e] = Jb[e];
    }
    Ph::template fjac<1>(P, qS, qF, Ja, Jb);
#pragma unroll
    for (int e = 0; e < M * M; ++e) {
      blk[2][e / M][e % M] += (keep[e] - Ja[e]) * hy;
      blk[4][e / M][e % M] = 0. + (-Jb[e]) * hy + Sources<Ph>::offdiag(P, false, 1, e / M, e % M);
    }
    double acc[M];
#pragma unroll
    for (int k = 0; k < M; ++k) acc[k] = 0.;
    if (V) {
      double FL[M], FR[M];
      Ph::template flux<0>(P, qL, qS, FL);
      Ph::template flux<0>(P, qS, qR, FR);
#pragma unroll
      for (int k = 0; k < M; ++k) acc[k] = 0. + hx * (FL[k] - FR[k]);
      Ph::template flux<1>(P, qB, qS, FL);
      Ph::template flux<1>(P, qS, qF, FR);
#pragma unroll
      for (int k = 0; k < M; ++k) acc[k] += hy * (FL[k] - FR[k]);
    }
    Sources<Ph>::apply(P, c, false, U, int32_t(c), acc, &blk[2][0][0]);
    if (V) {
#pragma unroll
      for (int k = 0; k < M; ++k) V[c * M + k] = acc[k];
    }
  }
  const int64_t base0 = P.rowbase[j * nx + iseg];
#pragma unroll
  for (int g = 0; g < 2; ++g) {
    if (act && (lane >> 5) == g) {
      // CSR order inside the cell: row k, then block slot s, then column dof
#pragma unroll
      for (int ch = 0; ch < NCH; ++ch) {
        double v[CW];
#pragma unroll
        for (int w = 0; w < CW; ++w) {
          const int e = ch * CW + w;          // 0..PER-1 in (k, s, jj) order
          const int k = e / (NB * M), s = (e / M) % NB, jj = e % M;
          v[w] = blk[s][k][jj];
        }
        CT x;
        if constexpr (CW == 2) x = make_double2(v[0], v[1]);
        else x = v[0];
        stage[lane & 31][ch] = x;
      }
    }
    wave_sync();
    const int ng = ncell - 32 * g < 0 ? 0 : (ncell - 32 * g > 32 ? 32 : ncell - 32 * g);
    CT *dst = reinterpret_cast<CT *>(J + base0 + int64_t(32 * g) * PER);
    for (int q = lane; q < ng * NCH; q += 64) store_stream(dst + q, stage[q / NCH][q % NCH]);
    wave_sync();
  }
}

// ------------------------- first-order Euler 2D Jacobian, two lanes per cell
// A wavefront owns a segment of up to ESEG consecutive box cells of a grid
// row and marches up a strip of rows.  Lanes 0..31 ("x lanes") evaluate the
// x faces of the segment (lane l: the face on the left of cell l, lane ncell
// the face on the right of the last cell), lanes 32..62 ("y lanes") the y
// face above cell l-32; the y lane carries that face to the next row, where
// it is the face below.  The 2D Euler flux along y is the flux along x with
// the two momentum components exchanged, so y lanes swap (rho u, rho v) of
// both states, run the same x-face code as the x lanes (no divergence) and
// write entry (a, b) of their blocks to row p(a), column p(b) with p the
// 1 <-> 2 exchange.  The row's blocks are assembled in LDS in CSR order
// (per cell: scalar row k, then block [back, left, self, right, front],
// then column) and copied to the contiguous CSR values of the segment with
// coalesced 16-byte stores.  Euler has no source terms.
constexpr int ESEG = 31;
constexpr int ESTR = 81; // LDS doubles per cell: 80 values + 1 pad against bank conflicts

template <int M>
__device__ __forceinline__ void load_cell(const double *__restrict__ U, int64_t c, double *q) {
#pragma unroll
  for (int k = 0; k < M; ++k) q[k] = U[c * M + k];
}

__device__ __forceinline__ void swap_mom(bool sw, double *q) {
  const double a = q[1], b = q[2];
  q[1] = sw ? b : a;
  q[2] = sw ? a : b;
}

template <class S>
__device__ __forceinline__ S pick(bool c, const S &a, const S &b) {
  S r;
  const double *pa = reinterpret_cast<const double *>(&a);
  const double *pb = reinterpret_cast<const double *>(&b);
  double *pr = reinterpret_cast<double *>(&r);
#pragma unroll
  for (int q = 0; q < int(sizeof(S) / sizeof(double)); ++q) pr[q] = c ? pa[q] : pb[q];
  return r;
}

template <class S>
__device__ __forceinline__ S shfl_next(const S &a) {
  S r;
  const double *pa = reinterpret_cast<const double *>(&a);
  double *pr = reinterpret_cast<double *>(&r);
#pragma unroll
  for (int q = 0; q < int(sizeof(S) / sizeof(double)); ++q) pr[q] = __shfl_down(pa[q], 1, 64);
  return r;
}

__device__ __forceinline__ void jac_fo_euler2d_seg(const View &P, const Box &B, int rows_per_strip,
                                                  int64_t bid, const double *__restrict__ U,
                                                  double *__restrict__ J) {
  using Ph = PhEuler2;
  using Side = Ph::Side;
  constexpr int M = 4, PER = 80;
  __shared__ double stage[ESEG * ESTR];
  const int lane = threadIdx.x;
  const bool yl = lane >= 32;
  const int l = lane & 31;
  const int64_t nx = B.n[0];
  const int64_t bw = B.hi[0] - B.lo[0];
  const int64_t nseg = (bw + ESEG - 1) / ESEG;
  const int64_t seg = bid % nseg, strip = bid / nseg;
  const int64_t iseg = B.lo[0] + seg * ESEG;
  const int ncell = int(bw - seg * ESEG < ESEG ? bw - seg * ESEG : ESEG);
  const int64_t j0 = B.lo[1] + strip * rows_per_strip;
  const int64_t j1 = j0 + rows_per_strip < B.hi[1] ? j0 + rows_per_strip : B.hi[1];
  // x lane l: face between cells l-1 and l (valid l <= ncell); y lane l: cell l (valid l < ncell)
  const int ci = yl ? (l < ncell ? l : ncell - 1) : (l <= ncell ? l : ncell);
  const int64_t da = yl ? 0 : -1;       // first state: cell on the left (x) or the cell itself (y)
  const int64_t db = yl ? nx : 1;       // second state relative to the first
  const double h = yl ? P.hinv[1] : P.hinv[0];
  // block 1: x lane -> left block of cell l (slot 1, +h JL); y lane -> front block (slot 4, -h JR)
  // block 2: x lane -> right block of cell l-1 (slot 3, -h JR); y lane -> back block (slot 0, +h JL below)
  const bool v1 = yl ? l < ncell : l < ncell;
  const bool v2 = yl ? l < ncell : (l >= 1 && l <= ncell);
  const int c1 = l, c2 = yl ? l : l - 1;
  const int s1 = yl ? 4 : 1, s2 = yl ? 0 : 3;
  const double g1 = yl ? -h : h, g2 = yl ? h : -h;
  // LDS offset of entry (a, b) of a block, with the y lanes' exchange of 1 and 2
  auto pr = [&](int a) { return yl ? (a == 1 ? 2 : a == 2 ? 1 : a) : a; };
  Side bl0, bl1; // y lanes: the face below the current row (left, right state)
  {
    double qa[M], qb[M];
    const int64_t c = (j0 - 1) * nx + iseg + ci;
    load_cell<M>(U, c, qa);
    load_cell<M>(U, c + nx, qb);
    swap_mom(yl, qa);
    swap_mom(yl, qb);
    Ph::sides<0>(P, qa, qb, bl0, bl1);
  }
  double qa[M], qb[M];
  {
    const int64_t c = j0 * nx + iseg + ci + da;
    load_cell<M>(U, c, qa);
    load_cell<M>(U, c + db, qb);
  }
  // Consume the first row's states here (an empty asm statement that only
  // reads them, so the wait for these loads is placed before the loop): the
  // loop head then only has to wait for the states prefetched in the
  // previous iteration, which were issued before that row's stores, and
  // the stores stay in flight while the next row is computed.
#pragma unroll
  for (int k = 0; k < M; ++k) asm volatile("" : : "v"(qa[k]), "v"(qb[k]));
  for (int64_t j = j0; j < j1; ++j) {
    swap_mom(yl, qa);
    swap_mom(yl, qb);
    Side S0, S1;
    Ph::sides<0>(P, qa, qb, S0, S1);
    {
      // the row's states are consumed: prefetch the next row's now (the last
      // row reloads itself, so the loads and their wait counts are static),
      // and read the row's first CSR offset, so that both arrive while the
      // blocks are formed
      const int64_t jn = j + 1 < j1 ? j + 1 : j;
      const int64_t c = jn * nx + iseg + ci + da;
      load_cell<M>(U, c, qa);
      load_cell<M>(U, c + db, qb);
    }
    const int cell0 = __builtin_amdgcn_readfirstlane(int(j * nx + iseg));
    const int64_t rb0 = P.rowbase[cell0];
    const Side nS0 = shfl_next(S0); // x lanes: left-state side of the next x face
    {
      const Side X1 = pick(yl, S1, S0);
      double *d1 = stage + c1 * ESTR + s1 * M;
      if (v1) {
#pragma unroll
        for (int a = 0; a < M; ++a)
#pragma unroll
          for (int b = 0; b < M; ++b) d1[pr(a) * 5 * M + pr(b)] = g1 * Ph::entry<0>(P, X1, a, b);
      }
      const Side X2 = pick(yl, bl0, S1);
      double *d2 = stage + c2 * ESTR + s2 * M;
      if (v2) {
#pragma unroll
        for (int a = 0; a < M; ++a)
#pragma unroll
          for (int b = 0; b < M; ++b) d2[pr(a) * 5 * M + pr(b)] = g2 * Ph::entry<0>(P, X2, a, b);
      }
    }
    // diagonal block: x part (JR of the left face - JL of the right face) * hx
    // written by the x lanes, then the y part (JR below - JL above) * hy
    // added by the y lanes
    {
      const Side Y1 = pick(yl, bl1, S1), Y2 = pick(yl, S0, nS0);
      double part[M * M];
#pragma unroll
      for (int a = 0; a < M; ++a)
#pragma unroll
        for (int b = 0; b < M; ++b)
          part[a * M + b] = (Ph::entry<0>(P, Y1, a, b) - Ph::entry<0>(P, Y2, a, b)) * h;
      double *ds = stage + l * ESTR + 2 * M;
      if (!yl && l < ncell) {
#pragma unroll
        for (int a = 0; a < M; ++a)
#pragma unroll
          for (int b = 0; b < M; ++b) ds[a * 5 * M + b] = part[a * M + b];
      }
      wave_sync();
      if (yl && l < ncell) {
#pragma unroll
        for (int a = 0; a < M; ++a)
#pragma unroll
          for (int b = 0; b < M; ++b) ds[pr(a) * 5 * M + pr(b)] += part[a * M + b];
      }
    }
    bl0 = S0;
    bl1 = S1;
    wave_sync();
    // copy the row's blocks out: 16-byte chunks, 40 per cell.  A fixed,
    // unrolled count of stores per lane (lanes past the end store the
    // segment's last chunk again, with the same value) lets the compiler
    // count them, so the next row waits only for its prefetched states and
    // not for these stores to reach memory.
    double2 *dst = reinterpret_cast<double2 *>(J + rb0);
    const int total = ncell * (PER / 2);
    // (an opaque copy of the lane index: the chunk addresses are recomputed
    // per row instead of being held in registers across the whole march)
    int lq = lane;
    asm volatile("" : "+v"(lq));
#pragma unroll
    for (int it = 0; it < (ESEG * (PER / 2) + 63) / 64; ++it) {
      const int q0 = lq + 64 * it;
      const int q = q0 < total ? q0 : total - 1;
      const int cc = q / (PER / 2), o = q - cc * (PER / 2);
      const double *src = stage + cc * ESTR + 2 * o;
      store_stream(dst + q, make_double2(src[0], src[1]));
    }
    wave_sync();
  }
}

// One launch for the whole first-order Euler 2D Jacobian: the first
// nframe workgroups run the generic row code over the rows outside the box
// (boundary ring, ghosts and BC factors), the rest the box segments.  Lower
// workgroup ids are dispatched first, so the latency-bound ring rows start
// ahead of the box and finish in its shadow; both paths need about the same
// registers, so the fusion costs the box no occupancy.
__global__ void __launch_bounds__(64) k_jac_fo_euler2d(View P, Box B, int rows_per_strip,
                                                        int nframe, const double *__restrict__ U,
                                                        double *__restrict__ J) {
  if (int(blockIdx.x) < nframe) {
    rows::jacobian_row<PhEuler2, 1, false>(P, int64_t(blockIdx.x) * 64 + threadIdx.x, U, nullptr, J);
    return;
  }
  // (the XCD-contiguous renumbering of the velocity kernel measured 9 %
  // slower here: 171.6 vs 157.9 us per call at 1024 x 1024)
  jac_fo_euler2d_seg(P, B, rows_per_strip, int64_t(blockIdx.x) - nframe, U, J);
}

// ------------------------- first-order Euler 3D Jacobian, marching x lines
// A wavefront owns a segment of up to E consecutive box cells of an x line
// (j, k) and marches in z over a strip of planes; consecutive workgroups take
// consecutive lines j of the same strip, so the waves in flight write
// neighbouring ranges of the CSR values.  Lane a E + l (a = 0, 1, 2) holds
// cell l of the segment and evaluates its two faces along axis a (lanes past
// 3E, and lanes past the segment's last cell, repeat the last cell and write
// nothing).  The 3D Euler flux along axis a is the flux along x with the
// momentum components 1 and 1+a exchanged, so every lane runs the x-face
// code on exchanged states (one instruction stream for the three axes) and
// writes entry (r, c) of its blocks to row p(r), column p(c), p that
// exchange.  For a cell and an axis with faces lo (below) and up (above):
//   block -a = h JL(lo),  block +a = -h JR(up),  diagonal += h (JR(lo) - JL(up)).
// The wave's two face evaluations per lane are the staged kernel's; the x
// faces inside the segment are evaluated twice (by the cells on both sides,
// from the same states, so with the same values), which costs no issue
// slots: the y and z lanes evaluate two faces in the same instructions.  The
// diagonal adds the axes in LDS in the order x, y, z, as the generic row
// code does.  The line's blocks are assembled in LDS in CSR order (per
// cell: scalar row r, then the blocks [-z, -y, -x, self, +x, +y, +z], then
// column) and leave as one contiguous range of 16-byte streaming stores
// that the next line does not wait for; the next line's states are loaded
// while the current line's blocks are formed.  The box keeps every face's
// two cells inside the grid (no ghost, no BC factor), and its cells' CSR
// rows hold the seven blocks in that order, consecutive cells consecutive.
// (Running the rows outside the box in the same launch, as the 2D kernel
// does, measured slower here: 6.24 vs 5.09 ms per jacobian() at 256^3 with
// the generic row code first, and 5.16 vs 4.93-4.95 ms with the staged rows
// interleaved among the box's workgroups, one in seven: their workgroups
// hold this kernel's registers and LDS.)
template <int E>
__global__ void __launch_bounds__(64) k_jac_fo_euler3d(View P, Box B, int rows_per_strip,
                                                        const double *__restrict__ U,
                                                        double *__restrict__ J) {
  using Ph = PhEuler3;
  using Side = Ph::Side;
  constexpr int M = 5, RS = 35, PER = M * RS;
  static_assert(3 * E <= 64, "one lane per cell and axis");
  __shared__ double stage[E * PER];
  const int lane = threadIdx.x;
  const int ax = lane < E ? 0 : (lane < 2 * E ? 1 : 2);
  const int l = lane - ax * E;
  const int64_t nx = B.n[0], nxy = B.n[0] * B.n[1];
  const int64_t bw = B.hi[0] - B.lo[0], bh = B.hi[1] - B.lo[1];
  const int64_t nseg = (bw + E - 1) / E;
  const int64_t bid = blockIdx.x;
  const int64_t seg = bid % nseg, jj = B.lo[1] + (bid / nseg) % bh;
  const int64_t strip = bid / (nseg * bh);
  const int64_t iseg = B.lo[0] + seg * E;
  const int ncell = int(bw - seg * E < E ? bw - seg * E : E);
  const int64_t k0 = B.lo[2] + strip * rows_per_strip;
  const int64_t k1 = k0 + rows_per_strip < B.hi[2] ? k0 + rows_per_strip : B.hi[2];
  // the lane's cell and its neighbours along its axis: (cell-s, cell, cell+s)
  const bool own = l < ncell;
  const int ci = own ? l : ncell - 1;
  const int64_t st = ax == 0 ? 1 : (ax == 1 ? nx : nxy);
  const double h = P.hinv[ax];
  // row r of an entry goes to row p(r) of the stage, column c to p(c)
  auto pr = [&](int r) { return r == 1 ? 1 + ax : (r == 1 + ax ? 1 : r); };
  double *d1 = stage + (own ? l : 0) * PER + (2 - ax) * M;  // block -a
  double *d2 = stage + (own ? l : 0) * PER + (4 + ax) * M;  // block +a
  double *dd = stage + (own ? l : 0) * PER + 3 * M;         // diagonal
  double qa[M], qb[M], qc[M];
  {
    const int64_t c = k0 * nxy + jj * nx + iseg + ci;
    load_cell<M>(U, c - st, qa);
    load_cell<M>(U, c, qb);
    load_cell<M>(U, c + st, qc);
  }
  // the first line's states are consumed here, so the loop head waits only
  // for the states prefetched in the previous iteration
#pragma unroll
  for (int k = 0; k < M; ++k) asm volatile("" : : "v"(qa[k]), "v"(qb[k]), "v"(qc[k]));
#pragma clang loop unroll(disable)
  for (int64_t kk = k0; kk < k1; ++kk) {
    double ea[M], eb[M], ec[M];
#pragma unroll
    for (int k = 0; k < M; ++k) {
      ea[k] = qa[pr(k)];
      eb[k] = qb[pr(k)];
      ec[k] = qc[pr(k)];
    }
    Side A0, A1, B0, B1;  // sides of the faces below (A) and above (B) the cell
    Ph::sides<0>(P, ea, eb, A0, A1);
    Ph::sides<0>(P, eb, ec, B0, B1);
    {
      // prefetch the next line (the last line reloads itself, so the loads
      // and their wait counts are static) and read the line's first CSR
      // offset, both arriving while the blocks are formed
      const int64_t kn = kk + 1 < k1 ? kk + 1 : kk;
      const int64_t c = kn * nxy + jj * nx + iseg + ci;
      load_cell<M>(U, c - st, qa);
      load_cell<M>(U, c, qb);
      load_cell<M>(U, c + st, qc);
    }
    const int cell0 = __builtin_amdgcn_readfirstlane(int(kk * nxy + jj * nx + iseg));
    const int64_t rb0 = P.rowbase[cell0];
    if (own) {
#pragma unroll
      for (int r = 0; r < M; ++r)
#pragma unroll
        for (int c = 0; c < M; ++c) {
          d1[pr(r) * RS + pr(c)] = h * Ph::entry<0>(P, A0, r, c);
          d2[pr(r) * RS + pr(c)] = -h * Ph::entry<0>(P, B1, r, c);
        }
    }
    double part[M * M];
#pragma unroll
    for (int r = 0; r < M; ++r)
#pragma unroll
      for (int c = 0; c < M; ++c) part[r * M + c] = (Ph::entry<0>(P, A1, r, c) - Ph::entry<0>(P, B0, r, c)) * h;
    if (own && ax == 0) {
#pragma unroll
      for (int r = 0; r < M; ++r)
#pragma unroll
        for (int c = 0; c < M; ++c) dd[r * RS + c] = part[r * M + c];
    }
#pragma unroll
    for (int phase = 1; phase < 3; ++phase) {
      wave_sync();
      if (own && ax == phase) {
#pragma unroll
        for (int r = 0; r < M; ++r)
#pragma unroll
          for (int c = 0; c < M; ++c) dd[pr(r) * RS + pr(c)] += part[r * M + c];
      }
    }
    wave_sync();
    // copy the line's values out: 16-byte chunks after a leading value that
    // is not 16-byte aligned; a fixed, unrolled count per lane (lanes past
    // the end store the last chunk again, with the same value) lets the
    // compiler count the stores, so the next line does not wait for them
    const int head = int(rb0 & 1);
    const int total = ncell * PER;
    const int npairs = (total - head) >> 1;
    double2 *dst = reinterpret_cast<double2 *>(J + rb0 + head);
    int lq = lane;
    asm volatile("" : "+v"(lq));
#pragma unroll
    for (int it = 0; it < (E * PER / 2 + 63) / 64; ++it) {
      const int q0 = lq + 64 * it;
      const int q = q0 < npairs ? q0 : npairs - 1;
      const double *src = stage + head + 2 * q;
      store_stream(dst + q, make_double2(src[0], src[1]));
    }
    if (lane == 0 && head) store_stream(J + rb0, stage[0]);
    if (lane == 1 && ((total - head) & 1)) store_stream(J + rb0 + total - 1, stage[total - 1]);
    wave_sync();
  }
}

}  // namespace

bool has_box_kernels(const pda_problem_s &p) {
  return p.dim == 2 &&
         (p.family == PDA_EULER2D || p.family == PDA_SWE2D || p.family == PDA_BURGERS2D);
}
bool has_box_velocity(const pda_problem_s &p) {
  return has_box_kernels(p) || (p.dim == 3 && p.family == PDA_EULER3D);
}

template <class Ph>
static int vel_box(const pda_problem_s &p, const double *U, double *V, hipStream_t s) {
  const Box &B = p.vbox;
  int64_t rps = velocity_strip_height(B, Ph::DIM, p.w);
  if (p.vel_rps > 0) rps = p.vel_rps; // measured by pda_problem_autotune
  const dim3 grid{unsigned(velocity_march_blocks(B, Ph::DIM, p.w, int(rps))), 1, 1};
  if constexpr (Ph::DIM == 2) {
    const RkStage S{};
    if (p.w == 1) hipLaunchKernelGGL((k_vel_cc<Ph, 1, false>), grid, dim3(64), 0, s, p.view, B, int(rps), U, V, S);
    else if (p.w == 2) hipLaunchKernelGGL((k_vel_cc<Ph, 2, false>), grid, dim3(64), 0, s, p.view, B, int(rps), U, V, S);
    else hipLaunchKernelGGL((k_vel_cc<Ph, 3, false>), grid, dim3(64), 0, s, p.view, B, int(rps), U, V, S);
  } else {
    if (p.w == 1) hipLaunchKernelGGL((k_vel_march<Ph, 1>), grid, dim3(64), 0, s, p.view, B, int(rps), U, V);
    else if (p.w == 2) hipLaunchKernelGGL((k_vel_march<Ph, 2>), grid, dim3(64), 0, s, p.view, B, int(rps), U, V);
    else hipLaunchKernelGGL((k_vel_march<Ph, 3>), grid, dim3(64), 0, s, p.view, B, int(rps), U, V);
  }
  PDA_HIP_CHECK(hipGetLastError());
  return PDA_OK;
}

// the 2D box march with an explicit stage formed in its store
template <class Ph>
static int vel_box_stage(const pda_problem_s &p, const double *U, const RkStage &S, hipStream_t s) {
  const Box &B = p.vbox;
  int64_t rps = velocity_strip_height(B, Ph::DIM, p.w);
  if (p.vel_rps > 0) rps = p.vel_rps;
  const dim3 grid{unsigned(velocity_march_blocks(B, Ph::DIM, p.w, int(rps))), 1, 1};
  if (p.w == 1) hipLaunchKernelGGL((k_vel_cc<Ph, 1, true>), grid, dim3(64), 0, s, p.view, B, int(rps), U, nullptr, S);
  else if (p.w == 2) hipLaunchKernelGGL((k_vel_cc<Ph, 2, true>), grid, dim3(64), 0, s, p.view, B, int(rps), U, nullptr, S);
  else hipLaunchKernelGGL((k_vel_cc<Ph, 3, true>), grid, dim3(64), 0, s, p.view, B, int(rps), U, nullptr, S);
  PDA_HIP_CHECK(hipGetLastError());
  return PDA_OK;
}

int launch_velocity_box_stage(const pda_problem_s &p, const double *U, const RkStage &S, hipStream_t s) {
  if (!p.vbox.valid || p.vbox.cells() == 0) return PDA_OK;
  switch (p.family) {
    case PDA_EULER2D: return vel_box_stage<PhEuler2>(p, U, S, s);
    case PDA_SWE2D: return vel_box_stage<PhSwe>(p, U, S, s);
    case PDA_BURGERS2D: return vel_box_stage<PhBurgers>(p, U, S, s);
  }
  return PDA_E_UNSUPPORTED;
}

int launch_velocity_box(const pda_problem_s &p, const double *U, double *V, hipStream_t s) {
  if (!p.vbox.valid || p.vbox.cells() == 0) return PDA_OK;
  switch (p.family) {
    case PDA_EULER2D: return vel_box<PhEuler2>(p, U, V, s);
    case PDA_SWE2D: return vel_box<PhSwe>(p, U, V, s);
    case PDA_BURGERS2D: return vel_box<PhBurgers>(p, U, V, s);
    case PDA_EULER3D: return vel_box<PhEuler3>(p, U, V, s);
  }
  return PDA_E_UNSUPPORTED;
}

template <class Ph>
static int jac_box(const pda_problem_s &p, const double *U, double *J, double *V, hipStream_t s,
                   bool *rest_done) {
  const Box &B = p.box;
  if constexpr (std::is_same<Ph, PhEuler2>::value) {
    if (V == nullptr) {
      // strips of rows_per_strip rows: ~16k one-wave groups for the 1024^2 grid
      const int64_t nseg = (B.hi[0] - B.lo[0] + ESEG - 1) / ESEG;
      const int64_t bh = B.hi[1] - B.lo[1];
      int64_t rps = jacobian_strip_height(B);
      if (p.jac_rps > 0) rps = p.jac_rps; // pda_problem_set_jacobian_strip_height
      const int64_t nstrip = (bh + rps - 1) / rps;
      const View v = generic_view(p, &p.d_rem_rows, int64_t(p.h_rem_rows.size()));
      const int64_t nframe = (v.nvisit + 63) / 64;
      hipLaunchKernelGGL(k_jac_fo_euler2d, dim3(unsigned(nframe + nseg * nstrip)), dim3(64), 0, s, v,
                         B, int(rps), int(nframe), U, J);
      PDA_HIP_CHECK(hipGetLastError());
      *rest_done = true;
      return PDA_OK;
    }
  }
  const int64_t nseg = (B.hi[0] - B.lo[0] + 63) / 64;
  const int64_t nblocks = nseg * (B.hi[1] - B.lo[1]);
  hipLaunchKernelGGL((k_jac_fo_box2d<Ph>), dim3(unsigned(nblocks)), dim3(64), 0, s, p.view, B, U, V, J);
  PDA_HIP_CHECK(hipGetLastError());
  return PDA_OK;
}

int launch_jacobian_box(const pda_problem_s &p, const double *U, double *J, double *V,
                        hipStream_t s, bool *rest_done) {
  *rest_done = false;
  if (p.w != 1) return PDA_E_UNSUPPORTED;
  if (!p.box.valid || p.box.cells() == 0) return PDA_OK;
  switch (p.family) {
    case PDA_EULER2D: return jac_box<PhEuler2>(p, U, J, V, s, rest_done);
    case PDA_SWE2D: return jac_box<PhSwe>(p, U, J, V, s, rest_done);
    case PDA_BURGERS2D: return jac_box<PhBurgers>(p, U, J, V, s, rest_done);
  }
  return PDA_E_UNSUPPORTED;
}

// cells of a row that one velocity wavefront owns
static int64_t velocity_segment(int dim, int w) {
  return dim == 2 ? (w == 1 ? vcc_width<1>() : (w == 2 ? vcc_width<2>() : vcc_width<3>())) : VSEG;
}

// default strip height of the single-problem velocity march.  Each strip
// evaluates extra rows of y values, so taller strips do less work but
// give fewer waves; 2D grids aim at ~3 waves per SIMD (1024^2: 6 rows,
// 3,078 waves on 1,024 SIMDs), 3D grids at ~20 strips per plane (the
// measured best: 12 rows at 256^3, 5 at 128^3).
int64_t velocity_strip_height(const Box &B, int dim, int w) {
  const int64_t seg = velocity_segment(dim, w);
  const int64_t nseg = (B.hi[0] - B.lo[0] + seg - 1) / seg;
  const int64_t bh = B.hi[1] - B.lo[1];
  int64_t rps = dim == 3 ? bh / 20 : (bh * nseg + 3071) / 3072;
  return rps < 1 ? 1 : (rps > 64 ? 64 : rps);
}

// default strip height of the first-order 2D Euler Jacobian march: about
// 16k one-wave groups (2 rows per strip at 1024^2)
int64_t jacobian_strip_height(const Box &B) {
  const int64_t nseg = (B.hi[0] - B.lo[0] + ESEG - 1) / ESEG;
  const int64_t rps = ((B.hi[1] - B.lo[1]) * nseg) / 16384;
  return rps < 1 ? 1 : (rps > 64 ? 64 : rps);
}

// cells per x segment of the first-order 3D Euler Jacobian march: 3 x 21
// lanes fill the wavefront (29.4 KB of LDS, 5 waves per CU); at 256^3 and
// 6-plane strips 5.02 ms per jacobian() against 5.09 with 17 cells and 5.13
// with 16 (at its best height, 4), measured with one lane per x face
constexpr int J3SEG = 21;

// default strip height (planes per workgroup) of that march, the fastest
// at 256^3 (5.02 ms per jacobian() at 6, 5.11 at 5, 5.28 at 8, 5.46 at 2)
constexpr int64_t J3_STRIP = 6;

int launch_jacobian_fo_euler3d_box(const pda_problem_s &p, const double *U, double *J, hipStream_t s) {
  const Box &B = p.box;
  for (int a = 0; a < 3; ++a)
    if (B.mode[a] != 0 || B.lo[a] < 1 || B.hi[a] > B.n[a] - 1) {
      set_error("first-order 3D Jacobian march: box stencil leaves the grid");
      return PDA_E_INVALID_ARG;
    }
  int64_t rps = J3_STRIP;
  if (p.jac_rps > 0) rps = p.jac_rps; // pda_problem_set_jacobian_strip_height
  const int64_t nseg = (B.hi[0] - B.lo[0] + J3SEG - 1) / J3SEG;
  const int64_t nstrip = (B.hi[2] - B.lo[2] + rps - 1) / rps;
  const int64_t nb = nseg * (B.hi[1] - B.lo[1]) * nstrip;
  hipLaunchKernelGGL((k_jac_fo_euler3d<J3SEG>), dim3(unsigned(nb)), dim3(64), 0, s, generic_view(p, nullptr, 0),
                     B, int(rps), U, J);
  PDA_HIP_CHECK(hipGetLastError());
  return PDA_OK;
}

// number of one-wave groups the velocity march uses for a box
int64_t velocity_march_blocks(const Box &B, int dim, int w, int rows_per_strip) {
  const int64_t seg = velocity_segment(dim, w);
  const int64_t nseg = (B.hi[0] - B.lo[0] + seg - 1) / seg;
  const int64_t bh = B.hi[1] - B.lo[1];
  return nseg * ((bh + rows_per_strip - 1) / rows_per_strip) * (B.hi[2] - B.lo[2]);
}

template <class Ph, bool STG>
static void vel_batch(int w, const BatchInst *inst, const int32_t *blk, int64_t nblocks, const double *U,
                      double *V, const RkStage &S, hipStream_t s) {
  const dim3 grid{unsigned(nblocks), 1, 1};
  if (w == 1) hipLaunchKernelGGL((k_vel_march2d_batch<Ph, 1, STG>), grid, dim3(64), 0, s, inst, blk, U, V, S);
  else if (w == 2) hipLaunchKernelGGL((k_vel_march2d_batch<Ph, 2, STG>), grid, dim3(64), 0, s, inst, blk, U, V, S);
  else hipLaunchKernelGGL((k_vel_march2d_batch<Ph, 3, STG>), grid, dim3(64), 0, s, inst, blk, U, V, S);
}

template <bool STG>
static int vel_batch_family(int family, int w, const BatchInst *inst, const int32_t *blk, int64_t nblocks,
                            const double *U, double *V, const RkStage &S, hipStream_t s) {
  if (nblocks <= 0) return PDA_OK;
  switch (family) {
    case PDA_EULER2D: vel_batch<PhEuler2, STG>(w, inst, blk, nblocks, U, V, S, s); break;
    case PDA_SWE2D: vel_batch<PhSwe, STG>(w, inst, blk, nblocks, U, V, S, s); break;
    case PDA_BURGERS2D: vel_batch<PhBurgers, STG>(w, inst, blk, nblocks, U, V, S, s); break;
    default: set_error("batch velocity: unsupported family"); return PDA_E_UNSUPPORTED;
  }
  PDA_HIP_CHECK(hipGetLastError());
  return PDA_OK;
}

int launch_velocity_batch(int family, int w, const BatchInst *inst, const int32_t *blk, int64_t nblocks,
                          const double *U, double *V, hipStream_t s) {
  return vel_batch_family<false>(family, w, inst, blk, nblocks, U, V, RkStage{}, s);
}

int launch_velocity_batch_stage(int family, int w, const BatchInst *inst, const int32_t *blk, int64_t nblocks,
                                const double *U, const RkStage &S, hipStream_t s) {
  if (!S.out || S.nx < 1 || S.nx > 4) {
    set_error("batch velocity stage: invalid stage");
    return PDA_E_INVALID_ARG;
  }
  return vel_batch_family<true>(family, w, inst, blk, nblocks, U, nullptr, S, s);
}

}  // namespace pda
