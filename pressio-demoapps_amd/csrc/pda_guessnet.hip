// pda_guessnet.hip -- the learned initial guess for implicit Newton steps on
// periodic 2D two-component fields (the Burgers problem): a 4-layer 3x3
// convolutional network, 32 channels, BF16 weights and activations with
// FP32 accumulation on the MFMA units (v_mfma_f32_16x16x32_bf16), trained
// on the device by the kernels below (forward, backward-data,
// backward-weight, Adam).
//
//   The guess is an FP64 predictor of y_{n+1} of a fractional extrapolation
//   order t in [1, 4] that the network picks per step.  The predictors are
//   the BDF2 relation y = (4/3) y_n - (1/3) y_{n-1} + (2/3) dt f(y) with
//   f(y_{n+1}) extrapolated in time with order k from f_n, ..., f_{n-k}:
//     P_k = (4/3) y_n - (1/3) y_{n-1} + (2/3) dt E_k,
//     E_1 = 2 f_n - f_{n-1},  E_2 = 3 f_n - 3 f_{n-1} + f_{n-2},
//     E_3 = 4 f_n - 6 f_{n-1} + 4 f_{n-2} - f_{n-3},
//     E_4 = 5 f_n - 10 f_{n-1} + 10 f_{n-2} - 5 f_{n-3} + f_{n-4};
//   P(t) interpolates linearly between P_k and P_{k+1} (k = floor(t)).
//   The network's NOUT = 7 output channels, averaged over the grid, are
//   its estimates m_j of what the step costs from P(T_j), T_j = 1 + j/2
//   (j = 0..6), relative to the mean over the orders; t = sum_j p_j T_j
//   with p = softmax(-m / kTau), a soft arg-min (kTau = 1/4 iteration).
//   Training minimises sum_j (m_j - (c_j - mean c))^2 / 7 of a sample, c_j
//   being what the step cost from P(T_j) (the package measures the Newton +
//   Krylov iterations when it makes the sample).  The order that pays
//   changes with the flow: on
//   smooth fields the highest one (the guess error is O(dt^{k+2})), near
//   steep fronts a low one (higher-order extrapolation in time of a front
//   passing a cell amplifies grid-scale error, which costs Newton
//   iterations even where it lowers ||guess - y_{n+1}||_2).
//   Every field the guess adds is FP64 and the network contributes one
//   number, t, so its BF16 rounding cannot put pixel-scale noise into the
//   guess (the Newton and Krylov work grows with the high-frequency content
//   of the guess's error, which pixel noise is).
//   features (8 ch): dt D^k f / s_k for k = 1, 2, 3 (2 channels each, D^k
//                    the k-th backward time difference of f), nu_hat,
//                    |dt D^4 f| / s_4 (f = the problem's velocity, FP64)
//   L1 8->32 SiLU, L2 32->32 SiLU, L3 32->32 SiLU, L4 32->7 (linear)
//   (periodic padding)
// Every convolution is an implicit GEMM: D[pixel][cout] += A[pixel][k] B[k][cout]
// over K = taps x input channels, 32 K per MFMA.  An MFMA pairs A's and B's K
// elements by register position, so any consistent assignment of K to
// (lane group, register element) is valid; the kernels use K = 8*g + j for
// lane group g = lane/16 and element j.  The output layout used is the
// 16x16 MFMA's: lane l holds D[4*(l/16) + i][l%16], i = 0..3; A row = l%16,
// B column = l%16.
#include <cmath>

#include "pda_internal.hpp"

namespace pda {
namespace nn {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int C = 32;      // hidden channels
constexpr int CIN0 = 8;    // feature channels (7 used)
constexpr int NOUT = 7;    // output channels: logits of the orders 1, 1.5, ..., 4
constexpr int NCOMP = 2;   // field components per pixel
constexpr int NL = 4;      // layers
constexpr int MT = 4;      // 16-pixel M tiles per wavefront in the forward kernels

// layer geometry: input channels, output channels
__host__ __device__ constexpr int cin_of(int l) { return l == 0 ? CIN0 : C; }
__host__ __device__ constexpr int cout_of(int l) { return l == NL - 1 ? NOUT : C; }
// K steps of 32 of a forward layer: 9 taps x 32 channels, or 9 taps x 8
// channels packed four taps per step
__host__ __device__ constexpr int ksteps(int cin) { return cin == C ? 9 : 3; }

__device__ __forceinline__ float silu(float z) { return z / (1.f + __expf(-z)); }
__device__ __forceinline__ float dsilu(float z) {
  const float s = 1.f / (1.f + __expf(-z));
  return s * (1.f + z * (1.f - s));
}

__device__ __forceinline__ int64_t shifted(int64_t p, int tap, int H, int W) {
  const int y = int(p / W), x = int(p - int64_t(y) * W);
  int yy = y + tap / 3 - 1, xx = x + tap % 3 - 1;
  yy = yy < 0 ? yy + H : (yy >= H ? yy - H : yy);
  xx = xx < 0 ? xx + W : (xx >= W ? xx - W : xx);
  return int64_t(yy) * W + xx;
}

// A operand: 8 bf16 of the pixel at row y, column x for K step s, lane
// group g (the tap's neighbour, periodic)
template <int CIN>
__device__ __forceinline__ bf16x8 load_a(const __bf16 *__restrict__ X, int y, int x, int s, int g, int H, int W) {
  const int tap = CIN == C ? s : 4 * s + g;
  if (CIN != C && tap >= 9) {
    bf16x8 z;
#pragma unroll
    for (int j = 0; j < 8; ++j) z[j] = __bf16(0.f);
    return z;
  }
  int yy = y + tap / 3 - 1, xx = x + tap % 3 - 1;
  yy = yy < 0 ? yy + H : (yy >= H ? yy - H : yy);
  xx = xx < 0 ? xx + W : (xx >= W ? xx - W : xx);
  const int64_t q = int64_t(yy) * W + xx;
  return *reinterpret_cast<const bf16x8 *>(X + q * CIN + (CIN == C ? 8 * g : 0));
}

// Forward convolution (and, with MODE_BWD, the backward-data convolution).
//   Wp: packed weights [kstep][ncout_pad][32] bf16 (B operand rows)
//   MODE_HIDDEN: Z = acc + bias (FP32, kept for the backward; not stored
//                when Z is null, as in inference), Y = SiLU(Z) (bf16); with
//                csum, also the block's per-channel sums of Y (FP32,
//                csum[c][block])
//   MODE_OUT:    out[p][c] = acc + bias for c < NOUT (FP32)
//   MODE_BWD:    dZprev[p][c] = acc * SiLU'(Zprev[p][c]) (bf16), no bias
enum { MODE_HIDDEN = 0, MODE_OUT = 1, MODE_BWD = 2 };

template <int CIN, int NT, int MODE>
__global__ void __launch_bounds__(64) k_conv(const __bf16 *__restrict__ X, const __bf16 *__restrict__ Wp,
                                              const float *__restrict__ bias, int H, int W,
                                              float *__restrict__ Z, __bf16 *__restrict__ Y,
                                              float *__restrict__ out, const float *__restrict__ Zprev,
                                              float *__restrict__ csum) {
  const int lane = threadIdx.x, r = lane & 15, g = lane >> 4;
  const int64_t p0 = int64_t(blockIdx.x) * (16 * MT);
  f32x4 acc[MT][NT];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
  // the row and column of this lane's A-row pixel in each M tile
  int py[MT], px[MT];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt) {
    const int p = int(p0) + 16 * mt + r;
    py[mt] = p / W;
    px[mt] = p - py[mt] * W;
  }
  constexpr int KS = ksteps(CIN);
#pragma unroll
  for (int s = 0; s < KS; ++s) {
    bf16x8 b[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
      b[nt] = *reinterpret_cast<const bf16x8 *>(Wp + (int64_t(s) * (16 * NT) + 16 * nt + r) * 32 + 8 * g);
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
      const bf16x8 a = load_a<CIN>(X, py[mt], px[mt], s, g, H, W);
#pragma unroll
      for (int nt = 0; nt < NT; ++nt)
        acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[nt], acc[mt][nt], 0, 0, 0);
    }
  }
  float ysum[NT];
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) ysum[nt] = 0.f;
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int64_t p = p0 + 16 * mt + 4 * g + i;
        const int c = 16 * nt + r;
        if constexpr (MODE == MODE_HIDDEN) {
          const float z = acc[mt][nt][i] + bias[c];
          if (Z) Z[p * C + c] = z;
          const __bf16 y = __bf16(silu(z));
          Y[p * C + c] = y;
          ysum[nt] += float(y);
        } else if constexpr (MODE == MODE_OUT) {
          if (c < NOUT) out[p * NOUT + c] = acc[mt][nt][i] + bias[c];
        } else {
          Y[p * C + c] = __bf16(acc[mt][nt][i] * dsilu(Zprev[p * C + c]));
        }
      }
  if constexpr (MODE == MODE_HIDDEN) {
    if (csum) {
      // channel c = 16 nt + r is held by the four lane groups g: sum them
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) {
        float v = ysum[nt];
        v += __shfl_xor(v, 16);
        v += __shfl_xor(v, 32);
        if (g == 0) csum[int64_t(16 * nt + r) * gridDim.x + blockIdx.x] = v;
      }
    }
  }
}

// Backward-weight: partial dW[tap][cout][cin] over a chunk of 32*KCH pixels,
// as MFMAs with M = cout, N = cin, K = pixels.  dZ (bf16 [p][32], zero
// beyond the layer's outputs) and X (bf16 [p][CIN]) tiles of 32 pixels are
// staged in LDS so that each lane gathers 8 consecutive pixels of one
// channel.
constexpr int KCH = 16; // 32-pixel K steps per chunk

template <int CIN>
__global__ void __launch_bounds__(64) k_conv_wgrad(const __bf16 *__restrict__ dZ, const __bf16 *__restrict__ X,
                                                    int H, int W, int coutp, float *__restrict__ part) {
  constexpr int NTN = CIN == C ? 2 : 1; // 16-wide cin tiles
  __shared__ __bf16 sz[32][C + 2];
  __shared__ __bf16 sx[32][CIN + 2];
  const int lane = threadIdx.x, r = lane & 15, g = lane >> 4;
  const int tap = blockIdx.y, chunk = blockIdx.x;
  const int MTN = coutp / 16; // 16-wide cout tiles
  f32x4 acc[2][NTN];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < NTN; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int ks = 0; ks < KCH; ++ks) {
    const int64_t pb = (int64_t(chunk) * KCH + ks) * 32;
    // stage dZ[pb .. pb+31][0..31] and X[shift(pb + i, tap)][0..CIN-1]
    for (int q = lane; q < 32 * C; q += 64) sz[q / C][q % C] = dZ[(pb + q / C) * C + q % C];
    for (int q = lane; q < 32 * CIN; q += 64)
      sx[q / CIN][q % CIN] = X[shifted(pb + q / CIN, tap, H, W) * CIN + q % CIN];
    __syncthreads();
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
      if (mt >= MTN) break;
      bf16x8 a;
#pragma unroll
      for (int j = 0; j < 8; ++j) a[j] = sz[8 * g + j][16 * mt + r];
#pragma unroll
      for (int nt = 0; nt < NTN; ++nt) {
        bf16x8 b;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int cin = 16 * nt + r;
          b[j] = cin < CIN ? sx[8 * g + j][cin < CIN ? cin : 0] : __bf16(0.f);
        }
        acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[mt][nt], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  // part[chunk][tap][cout][cin]
  float *dst = part + (int64_t(chunk) * 9 + tap) * (C * C);
#pragma unroll
  for (int mt = 0; mt < 2; ++mt) {
    if (mt >= MTN) break;
#pragma unroll
    for (int nt = 0; nt < NTN; ++nt)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int cout = 16 * mt + 4 * g + i, cin = 16 * nt + r;
        if (cin < CIN) dst[cout * C + cin] = acc[mt][nt][i];
      }
  }
}

// sum of the partials over chunks -> dW[tap][cout][cin] (FP32); bias
// gradient db[cout] = sum_p dZ[p][cout]
__global__ void k_wgrad_reduce(const float *__restrict__ part, int nchunk, int cin, int cout,
                               float *__restrict__ dW) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x; // tap*cout*cin
  if (e >= 9 * cout * cin) return;
  const int tap = e / (cout * cin), co = (e / cin) % cout, ci = e % cin;
  float s = 0.f;
  for (int c = 0; c < nchunk; ++c) s += part[(int64_t(c) * 9 + tap) * (C * C) + co * C + ci];
  dW[e] = s;
}

__global__ void k_bias_grad(const __bf16 *__restrict__ dZ, int64_t npix, int cout, float *__restrict__ db) {
  __shared__ float sh[256];
  const int c = blockIdx.x;
  float s = 0.f;
  for (int64_t p = threadIdx.x; p < npix; p += 256) s += float(dZ[p * C + c]);
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[c] = sh[0];
}

// the velocities the predictors read: f[0] = f(y_n), ..., f[4] = f(y_{n-4})
struct FHist {
  const double *f[5];
};

// features: dt D^k f / s_k (k = 1, 2, 3; 2 channels each), nu_hat,
// |dt D^4 f| / s_4
__global__ void k_features(FHist F, double c1, double c2, double c3, double c4, double nu_hat, int64_t npix,
                           __bf16 *__restrict__ X0) {
  const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= npix) return;
  __bf16 *x = X0 + p * CIN0;
  double n4 = 0.;
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const int64_t i = 2 * p + c;
    const double f0 = F.f[0][i], f1 = F.f[1][i], f2 = F.f[2][i], f3 = F.f[3][i], f4 = F.f[4][i];
    const double d1 = f0 - f1, d2 = f0 - 2. * f1 + f2, d3 = f0 - 3. * f1 + 3. * f2 - f3;
    const double d4 = f0 - 4. * f1 + 6. * f2 - 4. * f3 + f4;
    x[c] = __bf16(float(d1 * c1));
    x[2 + c] = __bf16(float(d2 * c2));
    x[4 + c] = __bf16(float(d3 * c3));
    n4 += (d4 * c4) * (d4 * c4);
  }
  x[6] = __bf16(float(nu_hat));
  x[7] = __bf16(float(sqrt(n4)));
}

// sums over the blocks of the last hidden layer's per-block channel sums:
// one block per channel
__global__ void __launch_bounds__(256) k_chan_sums(const float *__restrict__ csum, int nblk,
                                                   double *__restrict__ sums) {
  __shared__ double sh[256];
  const int c = blockIdx.x;
  double v = 0.;
  for (int b = threadIdx.x; b < nblk; b += 256) v += double(csum[int64_t(c) * nblk + b]);
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (int(threadIdx.x) < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) sums[c] = sh[0];
}

// a sample's costs at the orders T_j = 1 + j/2, by value
struct Costs {
  double c[NOUT];
};

constexpr double kTau = 0.25; // temperature of the soft arg-min over the cost estimates

// policy layout: pol[0..NOUT-1] = the weights p, pol[NOUT] = t,
// pol[NOUT+1 .. NOUT+5] = the extrapolation weights e_j of f_{n-j} in P(t),
// pol[NOUT+6 ..] = the cost estimates m
constexpr int POL_T = NOUT, POL_E = NOUT + 1, POL_M = NOUT + 6, POL_LEN = 2 * NOUT + 6;

// binomial extrapolation weights of order k (1..4) for f_n, ..., f_{n-4}
__device__ __forceinline__ double extrap_weight(int k, int j) {
  // (-1)^j C(k+1, j+1), zero for j > k
  constexpr double E[4][5] = {{2., -1., 0., 0., 0.},
                              {3., -3., 1., 0., 0.},
                              {4., -6., 4., -1., 0.},
                              {5., -10., 10., -5., 1.}};
  return E[k - 1][j];
}

// The cost estimates m_j = grid mean of output channel j.  The output layer
// is linear and its padding periodic, so that mean is
//   m_j = bias_j + sum_i Weff[j][i] mean(Y[:, i]),  Weff[j][i] = sum_tap W[tap][j][i]
// (the BF16 weights the MFMAs use), from the channel sums of the last
// hidden layer; p = softmax(-m / kTau), t = sum_j p_j T_j, the weights of
// P(t) and m into pol; with costs c, also lg[0] = sum_j (m_j - (c_j -
// mean c))^2 / NOUT and lg[1 + j] = d lg[0] / d m_j.  One thread.
__global__ void k_policy(const double *__restrict__ sums, const float *__restrict__ weff,
                         const float *__restrict__ bias, double inv_npix, Costs costs, int with_costs,
                         double *__restrict__ pol, double *__restrict__ lg) {
  if (threadIdx.x != 0) return;
  double m[NOUT], mn = 1e300;
#pragma unroll
  for (int c = 0; c < NOUT; ++c) {
    double v = 0.;
    for (int i = 0; i < C; ++i) v += double(weff[c * C + i]) * sums[i];
    m[c] = double(bias[c]) + v * inv_npix;
    pol[POL_M + c] = m[c];
    mn = fmin(mn, m[c]);
  }
  double z = 0., t = 0., e[NOUT];
#pragma unroll
  for (int c = 0; c < NOUT; ++c) {
    e[c] = exp(-(m[c] - mn) / kTau);
    z += e[c];
  }
#pragma unroll
  for (int c = 0; c < NOUT; ++c) {
    e[c] /= z;
    pol[c] = e[c];
    t += e[c] * (1. + 0.5 * c);
  }
  t = fmin(fmax(t, 1.), 4.);
  const int k = t >= 4. ? 3 : int(t);
  const double w = t - k;
  pol[POL_T] = t;
#pragma unroll
  for (int j = 0; j < 5; ++j) pol[POL_E + j] = (1. - w) * extrap_weight(k, j) + w * extrap_weight(k + 1, j);
  if (with_costs) {
    double cbar = 0., l = 0.;
#pragma unroll
    for (int c = 0; c < NOUT; ++c) cbar += costs.c[c];
    cbar /= NOUT;
#pragma unroll
    for (int c = 0; c < NOUT; ++c) {
      const double r = m[c] - (costs.c[c] - cbar);
      l += r * r;
      lg[1 + c] = 2. * r / NOUT;
    }
    lg[0] = l / NOUT;
  }
}

// guess = (4/3) y_n - (1/3) y_{n-1} + (2/3) dt sum_j e_j f_{n-j}, e = the
// weights of P(t) in pol
__global__ void k_guess(const double *__restrict__ un, const double *__restrict__ unm1, FHist F, double dt,
                        const double *__restrict__ pol, int64_t n, double *__restrict__ guess) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double *e = pol + POL_E;
  const double E = e[0] * F.f[0][i] + e[1] * F.f[1][i] + e[2] * F.f[2][i] + e[3] * F.f[3][i] + e[4] * F.f[4][i];
  guess[i] = (4. / 3.) * un[i] - (1. / 3.) * unm1[i] + ((2. / 3.) * dt) * E;
}

// dZ of the output layer: d loss / d out[p][c] = (d loss / d m_c) / npix
// for every pixel (the logits are means over the pixels), zero beyond NOUT
__global__ void k_dz_uniform(const double *__restrict__ lg, int64_t npix, __bf16 *__restrict__ dZ) {
  const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= npix) return;
  const double inv = 1. / double(npix);
#pragma unroll
  for (int c = 0; c < NOUT; ++c) dZ[p * C + c] = __bf16(float(lg[1 + c] * inv));
  for (int c = NOUT; c < C; ++c) dZ[p * C + c] = __bf16(0.f);
}

// Adam on the FP32 master parameters
__global__ void k_adam(float *__restrict__ w, const float *__restrict__ gr, float *__restrict__ m1,
                       float *__restrict__ m2, int64_t n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = gr[i];
  m1[i] = b1 * m1[i] + (1.f - b1) * gi;
  m2[i] = b2 * m2[i] + (1.f - b2) * gi * gi;
  w[i] -= lr * (m1[i] * c1) / (sqrtf(m2[i] * c2) + eps);
}

// Pack the FP32 master parameters of every layer into the bf16 B-operand
// layouts: forward Wf[l][s][coutp][32] and backward-data Wb[l][tap'][cin][32]
// (= W[8 - tap'][cout][cin], K = cout); and the output layer's tap sums of
// its BF16 weights, weff[cout][cin] (FP32).
__global__ void k_pack(const float *__restrict__ w, const int64_t *__restrict__ woff, __bf16 *__restrict__ Wf,
                       const int64_t *__restrict__ foff, __bf16 *__restrict__ Wb,
                       const int64_t *__restrict__ boff, float *__restrict__ weff) {
  const int l = blockIdx.y;
  if (l == NL - 1 && blockIdx.x == 0)
    for (int e = threadIdx.x; e < NOUT * C; e += blockDim.x) {
      const float *Wl = w + woff[l];
      float v = 0.f;
      for (int tap = 0; tap < 9; ++tap) v += float(__bf16(Wl[tap * NOUT * C + e]));
      weff[e] = v;
    }
  const int cin = cin_of(l), cout = cout_of(l), coutp = cout == NOUT ? 16 : C;
  const int KS = ksteps(cin);
  const float *W = w + woff[l]; // [tap][cout][cin]
  const int nf = KS * coutp * 32;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nf; e += gridDim.x * blockDim.x) {
    const int s = e / (coutp * 32), co = (e / 32) % coutp, k = e % 32;
    float v = 0.f;
    if (co < cout) {
      if (cin == C) {
        v = W[(s * cout + co) * cin + k];
      } else {
        const int tap = 4 * s + k / 8, ci = k % 8;
        v = tap < 9 ? W[(tap * cout + co) * cin + ci] : 0.f;
      }
    }
    Wf[foff[l] + e] = __bf16(v);
  }
  if (l == 0) return; // no input gradient for the first layer
  const int nb = 9 * C * 32;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nb; e += gridDim.x * blockDim.x) {
    const int tp = e / (C * 32), ci = (e / 32) % C, k = e % 32;
    const float v = k < cout ? W[((8 - tp) * cout + k) * cin + ci] : 0.f;
    Wb[boff[l] + e] = __bf16(v);
  }
}

}  // namespace nn
}  // namespace pda

struct pda_guessnet_s {
  int H = 0, W = 0;
  int64_t npix = 0;
  double dt = 5e-3;                 // time step of the predictors
  double s[4] = {1e-3, 1e-4, 1e-5, 1e-6}; // spreads of dt D^k f, k = 1..4
  double nu_mean = 5.5e-3, nu_scale = 4.5e-3;
  int64_t nparam = 0;
  int64_t woff[pda::nn::NL + 1] = {0};
  int64_t foff[pda::nn::NL + 1] = {0}, boff[pda::nn::NL + 1] = {0};
  pda::DevBuf w, g, m1, m2;          // FP32 master params, grads, Adam moments
  pda::DevBuf Wf, Wb, d_woff, d_foff, d_boff;
  pda::DevBuf X0, Y[3], Z[3], dZa, dZb, out, part;
  pda::DevBuf csum, sums, weff;      // last hidden layer's channel sums; output tap sums
  pda::DevBuf pol;                   // the last forward's policy: p, t, weights of P(t)
  pda::DevBuf lg;                    // expected cost, its derivatives w.r.t. the NOUT logits
  int adam_t = 0;
};

namespace {
using namespace pda;
using namespace pda::nn;

int pack(pda_guessnet_s &net, hipStream_t st) {
  hipLaunchKernelGGL(k_pack, dim3(16, NL), dim3(256), 0, st, net.w.as<float>(), net.d_woff.as<int64_t>(),
                     net.Wf.as<__bf16>(), net.d_foff.as<int64_t>(), net.Wb.as<__bf16>(),
                     net.d_boff.as<int64_t>(), net.weff.as<float>());
  PDA_HIP_CHECK(hipGetLastError());
  return PDA_OK;
}

const float *bias_of(const pda_guessnet_s &net, int l) {
  return net.w.as<float>() + net.woff[l] + 9 * cin_of(l) * cout_of(l);
}

// the velocity history a forward reads
int history(const double *const *f, FHist *F, const char *who) {
  if (!f) {
    set_error(std::string(who) + ": null velocity history");
    return PDA_E_INVALID_ARG;
  }
  for (int k = 0; k < 5; ++k) {
    if (!f[k]) {
      set_error(std::string(who) + ": null velocity f_{n-" + std::to_string(k) + "}");
      return PDA_E_INVALID_ARG;
    }
    F->f[k] = f[k];
  }
  return PDA_OK;
}

int features(pda_guessnet_s &net, const FHist &F, double nu, hipStream_t st) {
  hipLaunchKernelGGL(k_features, dim3(unsigned((net.npix + 255) / 256)), dim3(256), 0, st, F,
                     net.dt / net.s[0], net.dt / net.s[1], net.dt / net.s[2], net.dt / net.s[3],
                     (nu - net.nu_mean) / net.nu_scale, net.npix, net.X0.as<__bf16>());
  PDA_HIP_CHECK(hipGetLastError());
  return PDA_OK;
}

// the order policy of the last forward into net.pol; with costs, the
// cost-estimate loss and its derivatives into net.lg
int policy(pda_guessnet_s &net, const Costs *costs, hipStream_t st) {
  const int nblk = int(net.npix / (16 * MT));
  hipLaunchKernelGGL(k_chan_sums, dim3(C), dim3(256), 0, st, net.csum.as<float>(), nblk, net.sums.as<double>());
  hipLaunchKernelGGL(k_policy, dim3(1), dim3(64), 0, st, net.sums.as<double>(), net.weff.as<float>(),
                     bias_of(net, NL - 1), 1. / double(net.npix), costs ? *costs : Costs{}, costs ? 1 : 0,
                     net.pol.as<double>(), net.lg.as<double>());
  PDA_HIP_CHECK(hipGetLastError());
  return PDA_OK;
}

// forward through the hidden layers from the features in X0, with the
// last one's channel sums (all the policy needs: the output layer is
// linear); the pre-activations Z only when a backward follows; with
// full_out, the output layer too, out = L4 (FP32 [p][NOUT])
int forward(pda_guessnet_s &net, bool keep_z, bool full_out, hipStream_t st) {
  const unsigned grid = unsigned(net.npix / (16 * MT));
  const __bf16 *Wf = net.Wf.as<__bf16>();
  hipLaunchKernelGGL((k_conv<CIN0, 2, MODE_HIDDEN>), dim3(grid), dim3(64), 0, st, net.X0.as<__bf16>(),
                     Wf + net.foff[0], bias_of(net, 0), net.H, net.W, keep_z ? net.Z[0].as<float>() : nullptr,
                     net.Y[0].as<__bf16>(), nullptr, nullptr, nullptr);
  for (int l = 1; l < NL - 1; ++l)
    hipLaunchKernelGGL((k_conv<C, 2, MODE_HIDDEN>), dim3(grid), dim3(64), 0, st, net.Y[l - 1].as<__bf16>(),
                       Wf + net.foff[l], bias_of(net, l), net.H, net.W, keep_z ? net.Z[l].as<float>() : nullptr,
                       net.Y[l].as<__bf16>(), nullptr, nullptr, l == NL - 2 ? net.csum.as<float>() : nullptr);
  if (full_out)
    hipLaunchKernelGGL((k_conv<C, 1, MODE_OUT>), dim3(grid), dim3(64), 0, st, net.Y[NL - 2].as<__bf16>(),
                       Wf + net.foff[NL - 1], bias_of(net, NL - 1), net.H, net.W, nullptr, nullptr,
                       net.out.as<float>(), nullptr, nullptr);
  PDA_HIP_CHECK(hipGetLastError());
  return PDA_OK;
}

// Gradients (into net.g) of the loss whose output-layer dZ is in dZa.
// Layer l: dW_l from (dZ_l, input_l); dZ_{l-1} = (dZ_l conv flipped
// transposed W_l) * SiLU'(Z_{l-1}), alternating between dZa and dZb.
int backward(pda_guessnet_s &net, hipStream_t st) {
  const unsigned grid = unsigned(net.npix / (16 * MT));
  const int nchunk = int(net.npix / (32 * KCH));
  float *gW = net.g.as<float>();
  const __bf16 *Wb = net.Wb.as<__bf16>();
  __bf16 *dz = net.dZa.as<__bf16>(), *dz_next = net.dZb.as<__bf16>();
  for (int l = NL - 1; l >= 0; --l) {
    const int cin = cin_of(l), cout = cout_of(l), coutp = cout == NOUT ? 16 : C;
    const __bf16 *Xin = l == 0 ? net.X0.as<__bf16>() : net.Y[l - 1].as<__bf16>();
    if (cin == C)
      hipLaunchKernelGGL((k_conv_wgrad<C>), dim3(nchunk, 9), dim3(64), 0, st, dz, Xin, net.H, net.W, coutp,
                         net.part.as<float>());
    else
      hipLaunchKernelGGL((k_conv_wgrad<CIN0>), dim3(nchunk, 9), dim3(64), 0, st, dz, Xin, net.H, net.W,
                         coutp, net.part.as<float>());
    const int ne = 9 * cout * cin;
    hipLaunchKernelGGL(k_wgrad_reduce, dim3((ne + 255) / 256), dim3(256), 0, st, net.part.as<float>(), nchunk,
                       cin, cout, gW + net.woff[l]);
    hipLaunchKernelGGL(k_bias_grad, dim3(cout), dim3(256), 0, st, dz, net.npix, cout,
                       gW + net.woff[l] + 9 * cin * cout);
    PDA_HIP_CHECK(hipGetLastError());
    if (l == 0) break;
    hipLaunchKernelGGL((k_conv<C, 2, MODE_BWD>), dim3(grid), dim3(64), 0, st, dz, Wb + net.boff[l], nullptr,
                       net.H, net.W, nullptr, dz_next, nullptr, net.Z[l - 1].as<float>(), nullptr);
    PDA_HIP_CHECK(hipGetLastError());
    std::swap(dz, dz_next);
  }
  return PDA_OK;
}

// loss of the current forward's cost estimates against a sample's costs;
// leaves the output dZ in dZa
int loss(pda_guessnet_s &net, const double *costs, float *lval, hipStream_t st) {
  Costs c;
  for (int j = 0; j < NOUT; ++j) c.c[j] = costs[j];
  int rc = policy(net, &c, st);
  if (rc) return rc;
  const int nb = int((net.npix + 255) / 256);
  hipLaunchKernelGGL(k_dz_uniform, dim3(unsigned(nb)), dim3(256), 0, st, net.lg.as<double>(), net.npix,
                     net.dZa.as<__bf16>());
  PDA_HIP_CHECK(hipGetLastError());
  if (lval) {
    double l = 0.;
    PDA_HIP_CHECK(hipMemcpyAsync(&l, net.lg.ptr, sizeof(double), hipMemcpyDeviceToHost, st));
    PDA_HIP_CHECK(hipStreamSynchronize(st));
    *lval = float(l);
  }
  return PDA_OK;
}

int check_costs(const double *costs, const char *who) {
  if (!costs) {
    set_error(std::string(who) + ": null costs");
    return PDA_E_INVALID_ARG;
  }
  for (int j = 0; j < NOUT; ++j)
    if (!std::isfinite(costs[j])) {
      set_error(std::string(who) + ": the costs must be finite");
      return PDA_E_INVALID_ARG;
    }
  return PDA_OK;
}

}  // namespace

extern "C" {

int pda_guessnet_create(int nx, int ny, pda_guessnet *out) {
  if (!out || nx < 4 || ny < 4 || (int64_t(nx) * ny) % (32 * KCH) != 0 || int64_t(nx) * ny > (int64_t(1) << 30)) {
    set_error("pda_guessnet_create: need nx, ny >= 4 and nx*ny a multiple of 512, at most 2^30");
    return PDA_E_INVALID_ARG;
  }
  auto *net = new pda_guessnet_s();
  net->W = nx;
  net->H = ny;
  net->npix = int64_t(nx) * ny;
  int64_t wo = 0, fo = 0, bo = 0;
  for (int l = 0; l < NL; ++l) {
    net->woff[l] = wo;
    net->foff[l] = fo;
    net->boff[l] = bo;
    const int cin = cin_of(l), cout = cout_of(l), coutp = cout == NOUT ? 16 : C;
    wo += 9 * cin * cout + cout;
    fo += int64_t(ksteps(cin)) * coutp * 32;
    bo += l == 0 ? 0 : 9 * C * 32;
  }
  net->woff[NL] = wo;
  net->foff[NL] = fo;
  net->boff[NL] = bo;
  net->nparam = wo;
  const int64_t np = net->npix;
  hipError_t e = net->w.alloc(sizeof(float) * wo);
  if (e == hipSuccess) e = net->g.alloc(sizeof(float) * wo);
  if (e == hipSuccess) e = net->m1.alloc(sizeof(float) * wo);
  if (e == hipSuccess) e = net->m2.alloc(sizeof(float) * wo);
  if (e == hipSuccess) e = net->Wf.alloc(sizeof(__bf16) * fo);
  if (e == hipSuccess) e = net->Wb.alloc(sizeof(__bf16) * std::max<int64_t>(bo, 1));
  if (e == hipSuccess) e = net->d_woff.upload(std::vector<int64_t>(net->woff, net->woff + NL + 1));
  if (e == hipSuccess) e = net->d_foff.upload(std::vector<int64_t>(net->foff, net->foff + NL + 1));
  if (e == hipSuccess) e = net->d_boff.upload(std::vector<int64_t>(net->boff, net->boff + NL + 1));
  if (e == hipSuccess) e = net->X0.alloc(sizeof(__bf16) * np * CIN0);
  for (int i = 0; i < 3 && e == hipSuccess; ++i) e = net->Y[i].alloc(sizeof(__bf16) * np * C);
  for (int i = 0; i < 3 && e == hipSuccess; ++i) e = net->Z[i].alloc(sizeof(float) * np * C);
  if (e == hipSuccess) e = net->dZa.alloc(sizeof(__bf16) * np * C);
  if (e == hipSuccess) e = net->dZb.alloc(sizeof(__bf16) * np * C);
  if (e == hipSuccess) e = net->out.alloc(sizeof(float) * np * NOUT);
  if (e == hipSuccess) e = net->part.alloc(sizeof(float) * (np / (32 * KCH)) * 9 * C * C);
  if (e == hipSuccess) e = net->csum.alloc(sizeof(float) * (np / (16 * MT)) * C);
  if (e == hipSuccess) e = net->sums.alloc(sizeof(double) * C);
  if (e == hipSuccess) e = net->weff.alloc(sizeof(float) * NOUT * C);
  if (e == hipSuccess) e = net->pol.alloc(sizeof(double) * POL_LEN);
  if (e == hipSuccess) e = net->lg.alloc(sizeof(double) * (1 + NOUT));
  if (e == hipSuccess) e = hipMemset(net->pol.ptr, 0, sizeof(double) * POL_LEN);
  if (e == hipSuccess) e = hipMemset(net->w.ptr, 0, sizeof(float) * wo);
  if (e == hipSuccess) e = hipMemset(net->m1.ptr, 0, sizeof(float) * wo);
  if (e == hipSuccess) e = hipMemset(net->m2.ptr, 0, sizeof(float) * wo);
  if (e != hipSuccess) {
    set_error(std::string("pda_guessnet_create: ") + hipGetErrorString(e));
    delete net;
    return PDA_E_HIP;
  }
  *out = net;
  return PDA_OK;
}

void pda_guessnet_destroy(pda_guessnet net) { delete net; }

int pda_guessnet_num_params(pda_guessnet net, int64_t *n) {
  if (!net || !n) return PDA_E_INVALID_ARG;
  *n = net->nparam;
  return PDA_OK;
}

int pda_guessnet_set_params(pda_guessnet net, const float *w, int64_t n) {
  if (!net || !w || n != net->nparam) {
    set_error("pda_guessnet_set_params: expected the network's parameter count");
    return PDA_E_INVALID_ARG;
  }
  PDA_HIP_CHECK(hipMemcpy(net->w.ptr, w, sizeof(float) * n, hipMemcpyHostToDevice));
  PDA_HIP_CHECK(hipMemset(net->m1.ptr, 0, sizeof(float) * n));
  PDA_HIP_CHECK(hipMemset(net->m2.ptr, 0, sizeof(float) * n));
  net->adam_t = 0;
  const int rc = pack(*net, nullptr);
  if (rc) return rc;
  PDA_HIP_CHECK(hipDeviceSynchronize());
  return PDA_OK;
}

int pda_guessnet_get_params(pda_guessnet net, float *w, int64_t n) {
  if (!net || !w || n != net->nparam) return PDA_E_INVALID_ARG;
  PDA_HIP_CHECK(hipMemcpy(w, net->w.ptr, sizeof(float) * n, hipMemcpyDeviceToHost));
  return PDA_OK;
}

int pda_guessnet_set_scales(pda_guessnet net, double dt, const double *spreads, double nu_mean, double nu_scale) {
  if (!net || !spreads || !(dt > 0.) || !(nu_scale > 0.)) {
    set_error("pda_guessnet_set_scales: dt, the spreads and nu_scale must be positive");
    return PDA_E_INVALID_ARG;
  }
  for (int k = 0; k < 4; ++k)
    if (!(spreads[k] > 0.)) {
      set_error("pda_guessnet_set_scales: dt, the spreads and nu_scale must be positive");
      return PDA_E_INVALID_ARG;
    }
  net->dt = dt;
  for (int k = 0; k < 4; ++k) net->s[k] = spreads[k];
  net->nu_mean = nu_mean;
  net->nu_scale = nu_scale;
  return PDA_OK;
}

int pda_guessnet_predict(pda_guessnet net, const double *d_un, const double *d_unm1, const double *const *d_f,
                         double nu, double *d_guess, pda_stream stream) {
  FHist F;
  if (!net || !d_un || !d_unm1 || !d_guess) {
    set_error("pda_guessnet_predict: null argument");
    return PDA_E_INVALID_ARG;
  }
  int rc = history(d_f, &F, "pda_guessnet_predict");
  if (rc) return rc;
  hipStream_t st = hipStream_t(stream);
  rc = features(*net, F, nu, st);
  if (!rc) rc = forward(*net, false, false, st);
  if (!rc) rc = policy(*net, nullptr, st);
  if (rc) return rc;
  const int64_t n = net->npix * NCOMP;
  hipLaunchKernelGGL(k_guess, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, d_un, d_unm1, F, net->dt,
                     net->pol.as<double>(), n, d_guess);
  PDA_HIP_CHECK(hipGetLastError());
  return PDA_OK;
}

int pda_guessnet_policy(pda_guessnet net, double *order, double *probs, double *estimates) {
  if (!net || !order) {
    set_error("pda_guessnet_policy: null argument");
    return PDA_E_INVALID_ARG;
  }
  double pol[POL_LEN];
  PDA_HIP_CHECK(hipMemcpy(pol, net->pol.ptr, sizeof(pol), hipMemcpyDeviceToHost));
  *order = pol[POL_T];
  for (int j = 0; j < NOUT; ++j) {
    if (probs) probs[j] = pol[j];
    if (estimates) estimates[j] = pol[POL_M + j];
  }
  return PDA_OK;
}

int pda_guessnet_gradients(pda_guessnet net, const double *const *d_f, double nu, const double *costs,
                           float *grad, float *lossv, float *outv, pda_stream stream) {
  FHist F;
  if (!net) {
    set_error("pda_guessnet_gradients: null network");
    return PDA_E_INVALID_ARG;
  }
  int rc = history(d_f, &F, "pda_guessnet_gradients");
  if (!rc) rc = check_costs(costs, "pda_guessnet_gradients");
  if (rc) return rc;
  hipStream_t st = hipStream_t(stream);
  rc = features(*net, F, nu, st);
  if (!rc) rc = forward(*net, true, outv != nullptr, st);
  if (!rc) rc = loss(*net, costs, lossv, st);
  if (!rc) rc = backward(*net, st);
  if (rc) return rc;
  if (grad)
    PDA_HIP_CHECK(hipMemcpyAsync(grad, net->g.ptr, sizeof(float) * net->nparam, hipMemcpyDeviceToHost, st));
  if (outv)
    PDA_HIP_CHECK(hipMemcpyAsync(outv, net->out.ptr, sizeof(float) * net->npix * NOUT, hipMemcpyDeviceToHost, st));
  PDA_HIP_CHECK(hipStreamSynchronize(st));
  return PDA_OK;
}

int pda_guessnet_train_step(pda_guessnet net, const double *const *d_f, double nu, const double *costs,
                            float lr, float *lossv, pda_stream stream) {
  FHist F;
  if (!net || !(lr > 0.f)) {
    set_error("pda_guessnet_train_step: invalid argument");
    return PDA_E_INVALID_ARG;
  }
  int rc = history(d_f, &F, "pda_guessnet_train_step");
  if (!rc) rc = check_costs(costs, "pda_guessnet_train_step");
  if (rc) return rc;
  hipStream_t st = hipStream_t(stream);
  rc = features(*net, F, nu, st);
  if (!rc) rc = forward(*net, true, false, st);
  if (!rc) rc = loss(*net, costs, lossv, st);
  if (!rc) rc = backward(*net, st);
  if (rc) return rc;
  const float b1 = 0.9f, b2 = 0.999f;
  net->adam_t += 1;
  const float c1 = 1.f / (1.f - std::pow(b1, float(net->adam_t)));
  const float c2 = 1.f / (1.f - std::pow(b2, float(net->adam_t)));
  hipLaunchKernelGGL(k_adam, dim3(unsigned((net->nparam + 255) / 256)), dim3(256), 0, st, net->w.as<float>(),
                     net->g.as<float>(), net->m1.as<float>(), net->m2.as<float>(), net->nparam, lr, b1, b2,
                     1e-8f, c1, c2);
  PDA_HIP_CHECK(hipGetLastError());
  return pack(*net, st);
}

}  // extern "C"
