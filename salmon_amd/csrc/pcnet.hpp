// Learned CG preconditioner: a three-layer periodic 3x3x3 convolution
// network (channels 2 -> 16 -> 16 -> 2, ReLU, no biases) applied to the
// phase-canonicalised residual, and the kernels that train it.  No reference
// counterpart (CG_ompk applies no preconditioner, ARTED/GS/CG.f90:85-92).
//
// Direction of a residual g (NL complex):
//   phi = arg(sum g^2)/2,  u = g e^{-i phi} sqrt(NL)/|g|  (re, im channels)
//   d   = (y_re + i y_im) e^{i phi} |g|/sqrt(NL),  y = net(u)
// so d(c g) = c d(g) for complex c (up to the sign of e^{i phi}).
//
// Training objective of a sample (g, x, band ib, theta = <x|H|x>, target
// eigenvalue lam) with the system's lower bands Z_j (j < ib): with
// q = d - sum_{z in {Z_j, x}} z <z|d> h and p = q/|q|_h,
//   c = Re<p|Hp> h, b = <g|p> h, ev = (theta + c - sqrt((theta-c)^2 + 4|b|^2))/2
//   loss = -mean(min((theta - ev)/(theta - lam), 1))
// the fraction of the band's eigenvalue error one 2x2 Rayleigh-Ritz step in
// span{x, p} removes (CG.f90:107-113).
//
// Weights W[t][ci][co], tap t = (dx+1)*9 + (dy+1)*3 + (dz+1) reading the
// point (ix+dx, iy+dy, iz+dz).  Activations are [sample][point][channel].
#pragma once
#include "common.hpp"

namespace salmon {
namespace pc {

constexpr int C0 = 2, C1 = 16, C2 = 16, C3 = 2;
constexpr int NW1 = 27 * C0 * C1, NW2 = 27 * C1 * C2, NW3 = 27 * C2 * C3;
constexpr int NW = NW1 + NW2 + NW3;
constexpr int kBlk = 256;
constexpr int kRedChunk = 2048;  // points per partial of the per-sample reductions

__device__ __forceinline__ int wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// grid point of tap t around p
__device__ __forceinline__ int nbr(int ix, int iy, int iz, int t, const Grid& g) {
  const int dx = t / 9 - 1, dy = (t / 3) % 3 - 1, dz = t % 3 - 1;
  return (wrap(ix + dx, g.NLx) * g.NLy + wrap(iy + dy, g.NLy)) * g.NLz + wrap(iz + dz, g.NLz);
}

// ---------------------------------------------------------------------------
// Per-vector reductions in two passes (partials per chunk, then a fixed-order
// sum per vector).  Vectors are rows of NL complex values, stride NL.
// ---------------------------------------------------------------------------

// stats partials of row s: (Re sum g^2, Im sum g^2, sum |g|^2).  Grid (nchunk, n).
__global__ void __launch_bounds__(kBlk) k_pc_stats(const cplx<double>* __restrict__ g, int NL, double* __restrict__ part) {
  __shared__ double smem[4 * 3];
  const int s = blockIdx.y, c = blockIdx.x;
  const cplx<double>* v = g + (size_t)s * NL;
  double a[3] = {0.0, 0.0, 0.0};
  for (int p = c * kRedChunk + threadIdx.x; p < min(NL, (c + 1) * kRedChunk); p += kBlk) {
    const cplx<double> e = v[p];
    a[0] += e.re * e.re - e.im * e.im;
    a[1] += 2.0 * e.re * e.im;
    a[2] += e.re * e.re + e.im * e.im;
  }
  block_sum256<3>(a, smem);
  if (threadIdx.x == 0)
    for (int q = 0; q < 3; ++q) part[((size_t)s * gridDim.x + c) * 3 + q] = a[q];
}

// stats(s) = (cos phi, sin phi, sqrt(NL)/|g|, |g|/sqrt(NL)).  One thread per row.
__global__ void k_pc_stats_fin(const double* __restrict__ part, int nchunk, int n, int NL, double* __restrict__ stats) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  double a[3] = {0.0, 0.0, 0.0};
#pragma unroll 8
  for (int c = 0; c < nchunk; ++c)
    for (int q = 0; q < 3; ++q) a[q] += part[((size_t)s * nchunk + c) * 3 + q];
  const double phi = 0.5 * atan2(a[1], a[0]);
  const double nrm = sqrt(a[2]);
  stats[4 * s] = cos(phi);
  stats[4 * s + 1] = sin(phi);
  stats[4 * s + 2] = nrm > 0.0 ? sqrt((double)NL) / nrm : 0.0;
  stats[4 * s + 3] = nrm / sqrt((double)NL);
}

// u = g e^{-i phi} sqrt(NL)/|g| as two FP32 channels.  Grid (ceil(NL/256), n).
__global__ void __launch_bounds__(kBlk) k_pc_input(const cplx<double>* __restrict__ g, const double* __restrict__ stats,
                                                   int NL, float* __restrict__ u) {
  const int s = blockIdx.y, p = blockIdx.x * kBlk + threadIdx.x;
  if (p >= NL) return;
  const double cr = stats[4 * s], ci = stats[4 * s + 1], sc = stats[4 * s + 2];
  const cplx<double> e = g[(size_t)s * NL + p];
  // e * (cr - i ci)
  u[((size_t)s * NL + p) * 2] = (float)((e.re * cr + e.im * ci) * sc);
  u[((size_t)s * NL + p) * 2 + 1] = (float)((e.im * cr - e.re * ci) * sc);
}

// d = (y0 + i y1) e^{i phi} * (scaled ? |g|/sqrt(NL) : 1).  Grid (ceil(NL/256), n).
__global__ void __launch_bounds__(kBlk) k_pc_output(const float* __restrict__ y, const double* __restrict__ stats, int NL,
                                                    int scaled, cplx<double>* __restrict__ d) {
  const int s = blockIdx.y, p = blockIdx.x * kBlk + threadIdx.x;
  if (p >= NL) return;
  const double cr = stats[4 * s], ci = stats[4 * s + 1], sc = scaled ? stats[4 * s + 3] : 1.0;
  const double y0 = y[((size_t)s * NL + p) * 2], y1 = y[((size_t)s * NL + p) * 2 + 1];
  d[(size_t)s * NL + p] = cplx<double>{(y0 * cr - y1 * ci) * sc, (y0 * ci + y1 * cr) * sc};
}

// ---------------------------------------------------------------------------
// Convolutions (FP32).  Forward: out[s][p][co] = act(sum_t sum_ci W[t][ci][co]
// in[s][nbr(p,t)][ci]).  Grid (ceil(NL/256), n).
// ---------------------------------------------------------------------------
template <int CI, int CO, bool RELU>
__global__ void __launch_bounds__(kBlk) k_conv_fwd(const float* __restrict__ in, const float* __restrict__ W, Grid g,
                                                   float* __restrict__ out) {
  __shared__ float w[27 * CI * CO];
  for (int i = threadIdx.x; i < 27 * CI * CO; i += kBlk) w[i] = W[i];
  __syncthreads();
  const int s = blockIdx.y, p = blockIdx.x * kBlk + threadIdx.x;
  if (p >= g.NL) return;
  const int iz = p % g.NLz, iy = (p / g.NLz) % g.NLy, ix = p / (g.NLz * g.NLy);
  const float* a = in + (size_t)s * g.NL * CI;
  float acc[CO];
#pragma unroll
  for (int co = 0; co < CO; ++co) acc[co] = 0.0f;
  for (int t = 0; t < 27; ++t) {
    const float* an = a + (size_t)nbr(ix, iy, iz, t, g) * CI;
#pragma unroll
    for (int ci = 0; ci < CI; ++ci) {
      const float v = an[ci];
#pragma unroll
      for (int co = 0; co < CO; ++co) acc[co] = fmaf(w[(t * CI + ci) * CO + co], v, acc[co]);
    }
  }
  float* o = out + ((size_t)s * g.NL + p) * CO;
#pragma unroll
  for (int co = 0; co < CO; ++co) o[co] = RELU ? fmaxf(acc[co], 0.0f) : acc[co];
}

// gin[s][p][ci] = [ain > 0] sum_t sum_co W[t][ci][co] gout[s][p - d_t][co]
// (p - d_t is tap 26 - t of p).  Grid (ceil(NL/256), n).
template <int CI, int CO>
__global__ void __launch_bounds__(kBlk) k_conv_bwd_data(const float* __restrict__ gout, const float* __restrict__ ain,
                                                        const float* __restrict__ W, Grid g, float* __restrict__ gin) {
  __shared__ float w[27 * CI * CO];
  for (int i = threadIdx.x; i < 27 * CI * CO; i += kBlk) w[i] = W[i];
  __syncthreads();
  const int s = blockIdx.y, p = blockIdx.x * kBlk + threadIdx.x;
  if (p >= g.NL) return;
  const int iz = p % g.NLz, iy = (p / g.NLz) % g.NLy, ix = p / (g.NLz * g.NLy);
  const float* go = gout + (size_t)s * g.NL * CO;
  float acc[CI];
#pragma unroll
  for (int ci = 0; ci < CI; ++ci) acc[ci] = 0.0f;
  for (int t = 0; t < 27; ++t) {
    const float* gn = go + (size_t)nbr(ix, iy, iz, 26 - t, g) * CO;
#pragma unroll
    for (int co = 0; co < CO; ++co) {
      const float v = gn[co];
#pragma unroll
      for (int ci = 0; ci < CI; ++ci) acc[ci] = fmaf(w[(t * CI + ci) * CO + co], v, acc[ci]);
    }
  }
  const float* ai = ain + ((size_t)s * g.NL + p) * CI;
  float* o = gin + ((size_t)s * g.NL + p) * CI;
#pragma unroll
  for (int ci = 0; ci < CI; ++ci) o[ci] = ai[ci] > 0.0f ? acc[ci] : 0.0f;
}

// Weight-gradient partials: part[(t*nblk + b)*CI*CO + ci*CO + co] =
// sum over the block's points of in[s][nbr(p,t)][ci] gout[s][p][co].
// Grid (nblk = n * ceil(NL/kWChunk), 27); thread i < CI*CO owns (ci, co).
constexpr int kWChunk = 1024, kWSub = 64;
template <int CI, int CO>
__global__ void __launch_bounds__(kBlk) k_conv_bwd_w(const float* __restrict__ ain, const float* __restrict__ gout, Grid g,
                                                     float* __restrict__ part) {
  static_assert(CI * CO <= kBlk, "one (ci, co) pair per thread");
  __shared__ float sa[kWSub][CI + 1], sg[kWSub][CO + 1];
  const int t = blockIdx.y;
  const int per = (g.NL + kWChunk - 1) / kWChunk;
  const int s = blockIdx.x / per, c = blockIdx.x % per;
  const int ci = threadIdx.x / CO, co = threadIdx.x % CO;
  const int p0 = c * kWChunk, p1 = min(g.NL, p0 + kWChunk);
  const float* a = ain + (size_t)s * g.NL * CI;
  const float* go = gout + (size_t)s * g.NL * CO;
  float acc = 0.0f;
  for (int q0 = p0; q0 < p1; q0 += kWSub) {
    for (int i = threadIdx.x; i < kWSub * (CI + CO); i += kBlk) {
      const int r = i / (CI + CO), k = i % (CI + CO), p = q0 + r;
      float v = 0.0f;
      if (p < p1) {
        if (k < CI) {
          const int iz = p % g.NLz, iy = (p / g.NLz) % g.NLy, ix = p / (g.NLz * g.NLy);
          v = a[(size_t)nbr(ix, iy, iz, t, g) * CI + k];
        } else {
          v = go[(size_t)p * CO + (k - CI)];
        }
      }
      if (k < CI) sa[r][k] = v; else sg[r][k - CI] = v;
    }
    __syncthreads();
    if (threadIdx.x < CI * CO) {
#pragma unroll 8
      for (int r = 0; r < kWSub; ++r) acc = fmaf(sa[r][ci], sg[r][co], acc);
    }
    __syncthreads();
  }
  if (threadIdx.x < CI * CO) part[((size_t)t * gridDim.x + blockIdx.x) * CI * CO + threadIdx.x] = acc;
}

// grad[t*CI*CO + i] = sum over blocks (in order) of part, in FP64.
__global__ void __launch_bounds__(kBlk) k_conv_bwd_w_sum(const float* __restrict__ part, int nblk, int cico,
                                                         double* __restrict__ grad) {
  const int i = blockIdx.x * kBlk + threadIdx.x;
  if (i >= 27 * cico) return;
  const int t = i / cico, k = i % cico;
  double s = 0.0;
#pragma unroll 8
  for (int b = 0; b < nblk; ++b) s += part[((size_t)t * nblk + b) * cico + k];
  grad[i] = s;
}

// ---------------------------------------------------------------------------
// Training: projections, normalisation, the Rayleigh-Ritz objective and its
// gradient with respect to d.
// ---------------------------------------------------------------------------

// Batch description (device arrays, n samples).
struct Batch {
  int n, NL, NB;
  double h;
  const cplx<double>* Z;   // (NB, NL) lower bands of the system
  const cplx<double>* x;   // (n, NL)
  const cplx<double>* gr;  // (n, NL) residuals
  const int* ib;           // (n)
  const double* theta;     // (n)
  const double* lam;       // (n)
};

// dots partials: part[((s*(NB+1) + j)*nchunk + c)*2] = sum conj(V_j) v_s over chunk c,
// V_j = Z_j (j < ib_s; zero otherwise), V_NB = x_s.  Grid (nchunk, NB+1, n).
__global__ void __launch_bounds__(kBlk) k_pc_dots(Batch B, const cplx<double>* __restrict__ v, double* __restrict__ part) {
  __shared__ double smem[4 * 2];
  const int s = blockIdx.z, j = blockIdx.y, c = blockIdx.x;
  const bool use = j == B.NB || j < B.ib[s];
  const cplx<double>* V = j == B.NB ? B.x + (size_t)s * B.NL : B.Z + (size_t)j * B.NL;
  const cplx<double>* w = v + (size_t)s * B.NL;
  double a[2] = {0.0, 0.0};
  if (use) {
    for (int p = c * kRedChunk + threadIdx.x; p < min(B.NL, (c + 1) * kRedChunk); p += kBlk) {
      const cplx<double> z = V[p], e = w[p];
      a[0] += z.re * e.re + z.im * e.im;
      a[1] += z.re * e.im - z.im * e.re;
    }
  }
  block_sum256<2>(a, smem);
  if (threadIdx.x == 0) {
    const size_t o = (((size_t)s * (B.NB + 1) + j) * gridDim.x + c) * 2;
    part[o] = a[0];
    part[o + 1] = a[1];
  }
}

// coef[s*(NB+1)+j] = h * sum over chunks (in order).  One thread per (s, j).
__global__ void k_pc_dots_fin(Batch B, const double* __restrict__ part, int nchunk, cplx<double>* __restrict__ coef) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B.n * (B.NB + 1)) return;
  double re = 0.0, im = 0.0;
#pragma unroll 8
  for (int c = 0; c < nchunk; ++c) {
    re += part[((size_t)i * nchunk + c) * 2];
    im += part[((size_t)i * nchunk + c) * 2 + 1];
  }
  coef[i] = cplx<double>{re * B.h, im * B.h};
}

// out = v - sum_j coef_j V_j.  Grid (ceil(NL/256), n).
__global__ void __launch_bounds__(kBlk) k_pc_sub(Batch B, const cplx<double>* __restrict__ v,
                                                 const cplx<double>* __restrict__ coef, cplx<double>* __restrict__ out) {
  const int s = blockIdx.y, p = blockIdx.x * kBlk + threadIdx.x;
  if (p >= B.NL) return;
  cplx<double> e = v[(size_t)s * B.NL + p];
  const cplx<double>* cf = coef + (size_t)s * (B.NB + 1);
  const int nb = B.ib[s];
  for (int j = 0; j < nb; ++j) {
    const cplx<double> z = B.Z[(size_t)j * B.NL + p], a = cf[j];
    e.re -= z.re * a.re - z.im * a.im;
    e.im -= z.re * a.im + z.im * a.re;
  }
  const cplx<double> z = B.x[(size_t)s * B.NL + p], a = cf[B.NB];
  e.re -= z.re * a.re - z.im * a.im;
  e.im -= z.re * a.im + z.im * a.re;
  out[(size_t)s * B.NL + p] = e;
}

// partials of (|q|^2) per sample.  Grid (nchunk, n).
__global__ void __launch_bounds__(kBlk) k_pc_norm_part(const cplx<double>* __restrict__ q, int NL, double* __restrict__ part) {
  __shared__ double smem[4];
  const int s = blockIdx.y, c = blockIdx.x;
  double a[1] = {0.0};
  for (int p = c * kRedChunk + threadIdx.x; p < min(NL, (c + 1) * kRedChunk); p += kBlk) {
    const cplx<double> e = q[(size_t)s * NL + p];
    a[0] += e.re * e.re + e.im * e.im;
  }
  block_sum256<1>(a, smem);
  if (threadIdx.x == 0) part[(size_t)s * gridDim.x + c] = a[0];
}

// nu[s] = sqrt(h sum part); p = q / nu.  Grid (ceil(NL/256), n); nu computed per block.
__global__ void __launch_bounds__(kBlk) k_pc_normalise(const cplx<double>* __restrict__ q, const double* __restrict__ part,
                                                       int nchunk, int NL, double h, cplx<double>* __restrict__ p,
                                                       double* __restrict__ nu) {
  const int s = blockIdx.y, i = blockIdx.x * kBlk + threadIdx.x;
  double a = 0.0;
#pragma unroll 8
  for (int c = 0; c < nchunk; ++c) a += part[(size_t)s * nchunk + c];
  const double n = sqrt(h * a);
  if (blockIdx.x == 0 && threadIdx.x == 0) nu[s] = n;
  if (i >= NL) return;
  const cplx<double> e = q[(size_t)s * NL + i];
  p[(size_t)s * NL + i] = cplx<double>{e.re / n, e.im / n};
}

// partials of (Re<p|Hp>, Re<g|p>, Im<g|p>).  Grid (nchunk, n).
__global__ void __launch_bounds__(kBlk) k_pc_bc_part(Batch B, const cplx<double>* __restrict__ p,
                                                     const cplx<double>* __restrict__ hp, double* __restrict__ part) {
  __shared__ double smem[4 * 3];
  const int s = blockIdx.y, c = blockIdx.x;
  double a[3] = {0.0, 0.0, 0.0};
  for (int i = c * kRedChunk + threadIdx.x; i < min(B.NL, (c + 1) * kRedChunk); i += kBlk) {
    const size_t o = (size_t)s * B.NL + i;
    const cplx<double> e = p[o], f = hp[o], r = B.gr[o];
    a[0] += e.re * f.re + e.im * f.im;
    a[1] += r.re * e.re + r.im * e.im;
    a[2] += r.re * e.im - r.im * e.re;
  }
  block_sum256<3>(a, smem);
  if (threadIdx.x == 0)
    for (int q = 0; q < 3; ++q) part[((size_t)s * gridDim.x + c) * 3 + q] = a[q];
}

// Per sample: loss contribution and the coefficients of
//   grad_p = alpha Hp + beta g,  r = Re<p|grad_p>:
// co[s*6 + 0..5] = (-min(gain,1), alpha, Re beta, Im beta, r, unused).
// The loss is averaged over the n samples (gradients scaled by 1/n).
__global__ void k_pc_objective(Batch B, const double* __restrict__ part, int nchunk, double* __restrict__ co) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B.n) return;
  double a[3] = {0.0, 0.0, 0.0};
#pragma unroll 8
  for (int c = 0; c < nchunk; ++c)
    for (int q = 0; q < 3; ++q) a[q] += part[((size_t)s * nchunk + c) * 3 + q];
  const double h = B.h, th = B.theta[s], lam = B.lam[s];
  const double c = a[0] * h, br = a[1] * h, bi = a[2] * h;
  const double R = sqrt((th - c) * (th - c) + 4.0 * (br * br + bi * bi));
  const double ev = 0.5 * (th + c - R);
  const double gain = (th - ev) / (th - lam);
  double* o = co + 6 * s;
  if (!(gain < 1.0)) {
    o[0] = -1.0;
    o[1] = o[2] = o[3] = o[4] = o[5] = 0.0;
    return;
  }
  const double dl = 1.0 / ((th - lam) * B.n);
  const double alpha = dl * 0.5 * (1.0 - (c - th) / R) * 2.0 * h;
  const double bs = dl * (-1.0 / R) * 2.0 * h;
  const double betr = bs * br, beti = bs * bi;
  // r = Re<p|alpha Hp + beta g> = alpha c/h + Re(beta conj(b))/h
  const double r = alpha * c / h + (betr * br + beti * bi) / h;
  o[0] = -gain;
  o[1] = alpha;
  o[2] = betr;
  o[3] = beti;
  o[4] = r;
  o[5] = 0.0;
}

// gq = (alpha Hp + beta g - p h r) / nu.  Grid (ceil(NL/256), n).
__global__ void __launch_bounds__(kBlk) k_pc_gradq(Batch B, const cplx<double>* __restrict__ p,
                                                   const cplx<double>* __restrict__ hp, const double* __restrict__ co,
                                                   const double* __restrict__ nu, cplx<double>* __restrict__ gq) {
  const int s = blockIdx.y, i = blockIdx.x * kBlk + threadIdx.x;
  if (i >= B.NL) return;
  const double* c = co + 6 * s;
  const double al = c[1], br = c[2], bi = c[3], r = c[4] * B.h, n = nu[s];
  const size_t o = (size_t)s * B.NL + i;
  const cplx<double> e = p[o], f = hp[o], gg = B.gr[o];
  gq[o] = cplx<double>{(al * f.re + (br * gg.re - bi * gg.im) - e.re * r) / n,
                       (al * f.im + (br * gg.im + bi * gg.re) - e.im * r) / n};
}

// gy = gd e^{-i phi} as FP32 channels.  Grid (ceil(NL/256), n).
__global__ void __launch_bounds__(kBlk) k_pc_gy(const cplx<double>* __restrict__ gd, const double* __restrict__ stats, int NL,
                                                float* __restrict__ gy) {
  const int s = blockIdx.y, p = blockIdx.x * kBlk + threadIdx.x;
  if (p >= NL) return;
  const double cr = stats[4 * s], ci = stats[4 * s + 1];
  const cplx<double> e = gd[(size_t)s * NL + p];
  gy[((size_t)s * NL + p) * 2] = (float)(e.re * cr + e.im * ci);
  gy[((size_t)s * NL + p) * 2 + 1] = (float)(e.im * cr - e.re * ci);
}

// Adam on FP32 master weights; grad in FP64.
__global__ void __launch_bounds__(kBlk) k_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                                               const double* __restrict__ grad, int n, float lr, float b1, float b2,
                                               float eps, float c1, float c2) {
  const int i = blockIdx.x * kBlk + threadIdx.x;
  if (i >= n) return;
  const float gi = (float)grad[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ---------------------------------------------------------------------------
// Inference with BF16 weights and activations on the matrix cores
// (v_mfma_f32_16x16x32_bf16, FP32 accumulation).  A convolution layer is the
// GEMM out[p][n] = sum_kk A[p][kk] B[kk][n] with kk = t*CI + ci (27*CI padded
// to a multiple of 32), A[p][kk] = a[nbr(p,t)][ci], B[kk][n] = W[t][ci][n].
// One wave computes 16 points x 16 outputs; per 32-wide K step lane l holds
// A[l%16][8*(l/16) .. +7] and B[8*(l/16) .. +7][l%16] and receives
// D[4*(l/16) .. +3][l%16].
// ---------------------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short f2bf(float f) {  // round to nearest even
  const unsigned int u = __float_as_uint(f);
  return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

template <int CI>
constexpr int kpad() { return (27 * CI + 31) / 32 * 32; }

// W (FP32 [t][ci][co]) -> Bt (bf16 [n][kk], n < 16, zero padded).  One thread per entry.
template <int CI, int CO>
__global__ void k_pc_pack_bf16(const float* __restrict__ W, unsigned short* __restrict__ Bt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  constexpr int K = kpad<CI>();
  if (i >= 16 * K) return;
  const int n = i / K, kk = i % K;
  float v = 0.0f;
  if (n < CO && kk < 27 * CI) v = W[(size_t)kk * CO + n];
  Bt[i] = f2bf(v);
}

// u (bf16 [s][p][2]) from the residual rows, as k_pc_input.
__global__ void __launch_bounds__(kBlk) k_pc_input_bf16(const cplx<double>* __restrict__ g, const double* __restrict__ stats,
                                                        int NL, unsigned short* __restrict__ u) {
  const int s = blockIdx.y, p = blockIdx.x * kBlk + threadIdx.x;
  if (p >= NL) return;
  const double cr = stats[4 * s], ci = stats[4 * s + 1], sc = stats[4 * s + 2];
  const cplx<double> e = g[(size_t)s * NL + p];
  u[((size_t)s * NL + p) * 2] = f2bf((float)((e.re * cr + e.im * ci) * sc));
  u[((size_t)s * NL + p) * 2 + 1] = f2bf((float)((e.im * cr - e.re * ci) * sc));
}

// One layer.  in: bf16 [s][p][CI]; out: bf16 [s][p][CO] (OUT_BF16) or FP32.
// Block = 4 waves = 64 points; grid (ceil(NL/64), n).
template <int CI, int CO, bool RELU, bool OUT_BF16>
__global__ void __launch_bounds__(kBlk) k_conv_mfma(const unsigned short* __restrict__ in,
                                                    const unsigned short* __restrict__ Bt, Grid g, void* __restrict__ out) {
  constexpr int K = kpad<CI>();
  __shared__ __attribute__((aligned(16))) unsigned short sb[16 * K];
  for (int i = threadIdx.x; i < 16 * K / 8; i += kBlk) ((uint4*)sb)[i] = ((const uint4*)Bt)[i];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, s = blockIdx.y;
  const int p0 = (blockIdx.x * 4 + wave) * 16;
  const int m = lane & 15, kq = lane >> 4;
  const int p = p0 + m;
  const bool ok = p < g.NL;
  const int pc = ok ? p : 0;  // points past the grid read point 0, weighted by zero
  const int iz = pc % g.NLz, iy = (pc / g.NLz) % g.NLy, ix = pc / (g.NLz * g.NLy);
  const unsigned short* a = in + (size_t)s * g.NL * CI;
  // every reduction block's neighbour values are gathered before the first
  // MFMA (taps past 27 and points past the grid load valid data and are
  // zeroed by a select), so the gathers are in flight together
  constexpr int NKB = K / 32;
  union Frag { bf16x8 v; unsigned short h[8]; uint4 q; };
  Frag A[NKB];
#pragma unroll
  for (int i = 0; i < NKB; ++i) {
    const int kb = 32 * i + 8 * kq;
    if constexpr (CI % 8 == 0) {
      const int t = kb / CI, c0 = kb % CI;
      const uint4 q = *(const uint4*)(a + (size_t)nbr(ix, iy, iz, t < 27 ? t : 26, g) * CI + c0);
      A[i].q = (ok && t < 27) ? q : make_uint4(0u, 0u, 0u, 0u);
    } else {
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int kk = kb + j, t = kk / CI, ci = kk % CI;
        const unsigned short x = a[(size_t)nbr(ix, iy, iz, t < 27 ? t : 26, g) * CI + ci];
        A[i].h[j] = (ok && t < 27) ? x : (unsigned short)0;
      }
    }
  }
  f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int i = 0; i < NKB; ++i) {
    Frag B;
    B.q = *(const uint4*)(sb + m * K + 32 * i + 8 * kq);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[i].v, B.v, acc, 0, 0, 0);
  }
  const int n = m;
  if (n < CO) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int q = p0 + 4 * kq + i;
      if (q < g.NL) {
        float v = acc[i];
        if (RELU) v = fmaxf(v, 0.0f);
        if constexpr (OUT_BF16) ((unsigned short*)out)[((size_t)s * g.NL + q) * CO + n] = f2bf(v);
        else ((float*)out)[((size_t)s * g.NL + q) * CO + n] = v;
      }
    }
  }
}

}  // namespace pc
}  // namespace salmon
