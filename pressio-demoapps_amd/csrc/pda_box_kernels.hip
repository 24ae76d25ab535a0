// pda_box_kernels.hip -- kernels for the cells of a natural-order structured
// mesh whose whole scheme stencil lies inside the grid (the "box").  They
// give the same values as the per-row kernels (same face states, same
// fluxes, same accumulation order) but:
//   - velocity: every face is reconstructed once per tile and shared by the
//     two cells it separates through LDS, instead of once per cell side;
//   - first-order Jacobian: the five m x m blocks of a cell are formed in
//     registers and written through an LDS stage so that HBM stores are
//     contiguous across the wavefront.
// Cell ids follow the structured layout gid = j*nx + i (graph row == gid).
#include "pda_families.hpp"

namespace pda {
namespace {
using namespace fam;

// ------------------------------------------------------------ velocity 2D
constexpr int VTX = 64, VTY = 4; // tile: 64 cells in x, 4 rows in y (256 threads)

template <class Ph, int W, int A>
__device__ __forceinline__ void face_flux_at(const View &P, const double *__restrict__ U,
                                             int64_t c_right, int64_t stride, double *F) {
  // face between cells c_right - stride and c_right; stencil c_right+(p-W)*stride, p=0..2W-1
  constexpr int M = Ph::M;
  double q[2 * W][M];
#pragma unroll
  for (int p = 0; p < 2 * W; ++p)
#pragma unroll
    for (int k = 0; k < M; ++k) q[p][k] = U[(c_right + int64_t(p - W) * stride) * M + k];
  double qm[M], qp[M];
#pragma unroll
  for (int k = 0; k < M; ++k) {
    double v[2 * W];
#pragma unroll
    for (int p = 0; p < 2 * W; ++p) v[p] = q[p][k];
    if constexpr (W == 1) {
      qm[k] = v[0];
      qp[k] = v[1];
    } else if constexpr (W == 2) {
      phys::weno3(v, qm[k], qp[k]);
    } else {
      phys::weno5(v, qm[k], qp[k]);
    }
  }
  Ph::template flux<A>(P, qm, qp, F);
}

template <class Ph, int W>
__global__ void __launch_bounds__(256) k_vel_box2d(View P, Box B, const double *__restrict__ U,
                                                    double *__restrict__ V) {
  constexpr int M = Ph::M;
  __shared__ double fx[VTY][VTX + 1][M];
  __shared__ double fy[VTY + 1][VTX][M];
  const int64_t nx = B.n[0];
  const int64_t i0 = B.lo[0] + int64_t(blockIdx.x) * VTX;
  const int64_t j0 = B.lo[1] + int64_t(blockIdx.y) * VTY;
  const int t = threadIdx.x;
  // x faces: (r, q) is the face on the left of cell (i0+q, j0+r); q = 0..VTX
  for (int f = t; f < VTY * (VTX + 1); f += 256) {
    const int r = f / (VTX + 1), q = f % (VTX + 1);
    const int64_t i = i0 + q, j = j0 + r;
    if (i > B.hi[0] || j >= B.hi[1]) continue;
    face_flux_at<Ph, W, 0>(P, U, j * nx + i, 1, fx[r][q]);
  }
  // y faces: (r, q) is the face below cell (i0+q, j0+r); r = 0..VTY
  for (int f = t; f < (VTY + 1) * VTX; f += 256) {
    const int r = f / VTX, q = f % VTX;
    const int64_t i = i0 + q, j = j0 + r;
    if (i >= B.hi[0] || j > B.hi[1]) continue;
    face_flux_at<Ph, W, 1>(P, U, j * nx + i, nx, fy[r][q]);
  }
  __syncthreads();
  const int q = t % VTX, r = t / VTX;
  const int64_t i = i0 + q, j = j0 + r;
  if (i >= B.hi[0] || j >= B.hi[1]) return;
  const int64_t c = j * nx + i;
  double acc[M];
#pragma unroll
  for (int k = 0; k < M; ++k) acc[k] = 0. + P.hinv[0] * (fx[r][q][k] - fx[r][q + 1][k]);
#pragma unroll
  for (int k = 0; k < M; ++k) acc[k] += P.hinv[1] * (fy[r][q][k] - fy[r + 1][q][k]);
  Sources<Ph>::apply(P, c, false, U, int32_t(c), acc, nullptr);
#pragma unroll
  for (int k = 0; k < M; ++k) V[c * M + k] = acc[k];
}

// ------------------------------------------------- first-order Jacobian 2D
// One wavefront per segment of 64 consecutive box cells of a grid row.  The
// segment's Jacobian rows are contiguous in the CSR value array (all its
// cells have the five blocks [back, left, self, right, front] in that
// column order), so after the blocks are formed the wavefront writes the
// segment as one contiguous run, 32 cells per LDS round.
// LDS hand-off inside one wavefront (the workgroup is a single wave): LDS
// instructions of a wave execute in order, so only the compiler must be kept
// from reordering; unlike __syncthreads() this does not wait for the wave's
// outstanding global stores.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

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
    for (int q = lane; q < ng * NCH; q += 64) dst[q] = stage[q / NCH][q % NCH];
    wave_sync();
  }
}

}  // namespace

bool has_box_kernels(const pda_problem_s &p) {
  return p.dim == 2 &&
         (p.family == PDA_EULER2D || p.family == PDA_SWE2D || p.family == PDA_BURGERS2D);
}

template <class Ph>
static int vel_box(const pda_problem_s &p, const double *U, double *V, hipStream_t s) {
  const Box &B = p.box;
  const dim3 grid(unsigned((B.hi[0] - B.lo[0] + VTX - 1) / VTX),
                  unsigned((B.hi[1] - B.lo[1] + VTY - 1) / VTY));
  if (p.w == 1) hipLaunchKernelGGL((k_vel_box2d<Ph, 1>), grid, dim3(256), 0, s, p.view, B, U, V);
  else if (p.w == 2) hipLaunchKernelGGL((k_vel_box2d<Ph, 2>), grid, dim3(256), 0, s, p.view, B, U, V);
  else hipLaunchKernelGGL((k_vel_box2d<Ph, 3>), grid, dim3(256), 0, s, p.view, B, U, V);
  PDA_HIP_CHECK(hipGetLastError());
  return PDA_OK;
}

int launch_velocity_box(const pda_problem_s &p, const double *U, double *V, hipStream_t s) {
  if (!p.box.valid || p.box.cells() == 0) return PDA_OK;
  switch (p.family) {
    case PDA_EULER2D: return vel_box<PhEuler2>(p, U, V, s);
    case PDA_SWE2D: return vel_box<PhSwe>(p, U, V, s);
    case PDA_BURGERS2D: return vel_box<PhBurgers>(p, U, V, s);
  }
  return PDA_E_UNSUPPORTED;
}

template <class Ph>
static int jac_box(const pda_problem_s &p, const double *U, double *J, double *V, hipStream_t s) {
  const Box &B = p.box;
  const int64_t nseg = (B.hi[0] - B.lo[0] + 63) / 64;
  const int64_t nblocks = nseg * (B.hi[1] - B.lo[1]);
  hipLaunchKernelGGL((k_jac_fo_box2d<Ph>), dim3(unsigned(nblocks)), dim3(64), 0, s, p.view, B, U, V, J);
  PDA_HIP_CHECK(hipGetLastError());
  return PDA_OK;
}

int launch_jacobian_box(const pda_problem_s &p, const double *U, double *J, double *V,
                        hipStream_t s) {
  if (p.w != 1) return PDA_E_UNSUPPORTED;
  if (!p.box.valid || p.box.cells() == 0) return PDA_OK;
  switch (p.family) {
    case PDA_EULER2D: return jac_box<PhEuler2>(p, U, J, V, s);
    case PDA_SWE2D: return jac_box<PhSwe>(p, U, J, V, s);
    case PDA_BURGERS2D: return jac_box<PhBurgers>(p, U, J, V, s);
  }
  return PDA_E_UNSUPPORTED;
}

}  // namespace pda
