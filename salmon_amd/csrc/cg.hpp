// Ground-state kernels: the band-by-band CG of CG_ompk
// (ARTED/GS/CG.f90:34-139), the kinetic preconditioner applied to its
// residual, and the subspace matrix / rotation of diag_omp
// (ARTED/GS/diag.f90:95-130).  FP64 throughout.
//
// CG runs all k points of one band at once (grid.y = k).  Every scalar of the
// iteration (<x|H|x>, |g|^2, the 2x2 Rayleigh-Ritz coefficients, the exit
// flag) lives in device memory, so a band's whole iteration sequence is
// enqueued without a host round trip: reductions end in the last-arriving
// block of each k, which sums the per-block partials in block order (a fixed
// order, so results do not depend on scheduling) and computes the scalar.
// A k whose band met the exit test (|ev - <x|H|x>| < 1e-15, CG.f90:114)
// skips the remaining iteration kernels.
#pragma once
#include "common.hpp"

namespace salmon {
namespace cg {

#ifndef SALMON_CG_PPT
#define SALMON_CG_PPT 4
#endif
constexpr int kPPT = SALMON_CG_PPT;  // points per thread per chunk
constexpr int kChunk = 256 * kPPT;   // points per block
// the lower-band subtractions (k_cg_sub, k_gs_sub) read ib bands per point and
// run on finer chunks: more workgroups keep more of those loads in flight
#ifndef SALMON_CG_SUB_PPT
#define SALMON_CG_SUB_PPT 1
#endif
constexpr int kSubPPT = SALMON_CG_SUB_PPT;
constexpr int kSubChunk = 256 * kSubPPT;
constexpr int kJG = 8;               // lower bands per projection block

// per-k scalars (double slots)
enum : int {
  XHX = 0, XTX, GKGK, UK, SN, EVX, CXR, CXI, CP, ZSR, ZSI, EFIN, kNS = 16
};

struct State {
  double* sc;        // (NK*kNS)
  int* done;         // (NK) exit flag of the current band
  int* iters;        // (NK*NB) iterations used by each band
  double* esp_var;   // (NK*NB)
  double* part;      // (NK*maxblk*kPartQ) per-block partial sums
  unsigned* cnt;     // (NK) arrival counters of the last-block reductions
  cplx<double>* zs;  // (NK*NB) projections on the lower bands
  cplx<double>* gram;  // (NK*NB*NB) G(j,i) = Hxyz <B_j|B_i>, i < j, of the stored bands
  cplx<double>* minv;  // (NK*NB*NB) rows of (I + L)^-1, L(j,i) = G(j,i) for i < j
  double* minvr;       // (NK*NB*NB) rows of (I + Re L)^-1
  const double* occ; // (NK*NB)
  int NK, NB, NL, nchunk;
  int nsub;          // chunks of the subtraction kernels (kSubChunk points)
  int pstride;       // doubles of part per k: nchunk*ceil(NB/kJG)*kPartQ
  double Hxyz;
};
constexpr int kPartQ = 2 * kJG;

// Sums NQ per-block values over the blocks of row k = blockIdx.y in block
// order.  Returns true in the block that arrives last; there, tot[] holds the
// totals in thread 0.  nblk = number of blocks of the row.
template <int NQ>
__device__ bool last_block_sum(double (&v)[NQ], const State& S, int nblk, int bid, double (&tot)[NQ]) {
  __shared__ double smem[4 * NQ];
  __shared__ int is_last;
  block_sum256<NQ>(v, smem);
  const int k = blockIdx.y;
  double* part = S.part + (size_t)k * S.pstride;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) __hip_atomic_store(&part[bid * NQ + q], v[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned t = __hip_atomic_fetch_add(&S.cnt[k], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    is_last = (t == (unsigned)(nblk - 1));
  }
  __syncthreads();
  if (!is_last) return false;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) tot[q] = 0.0;
#pragma unroll 8
    for (int b = 0; b < nblk; ++b) {
#pragma unroll
      for (int q = 0; q < NQ; ++q) tot[q] += __hip_atomic_load(&part[b * NQ + q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __hip_atomic_store(&S.cnt[k], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  return true;
}

__device__ __forceinline__ cplx<double> cmulc(cplx<double> a, cplx<double> b) {  // a*b
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

// Source vector of the projection and subtraction kernels: either a stored
// vector V, or the residual H x - <x|H|x> x formed on the fly (CG.f90:85).
struct Src {
  const cplx<double>* v;   // stored vector (nullptr: residual)
  size_t vstride;          // distance between the vectors of consecutive k
  const cplx<double>* hx;  // residual operands, NL apart per k
  const cplx<double>* x;
  __device__ __forceinline__ cplx<double> at(int k, int NL, int p, double xhx) const {
    if (v) return v[(size_t)k * vstride + p];
    const size_t o = (size_t)k * NL + p;
    const cplx<double> h = hx[o], e = x[o];
    return {h.re - xhx * e.re, h.im - xhx * e.im};
  }
};

// Projections on the lower bands.  The reference orthogonalises with the
// sequential (modified) Gram-Schmidt loop (CG.f90:69-71, 86-89): the
// coefficient of band j is taken against the vector already stripped of bands
// 0..j-1.  Its coefficients c solve c_j = b_j - sum_{k<j} G_jk c_k with
// b_j = Hxyz <B_j|v> and the Gram matrix G_jk = Hxyz <B_j|B_k> of the lower
// bands, which the lower bands need not make diagonal (the start vector's
// overlap keeps only its real part, CG.f90:46,70, so a band may keep an
// imaginary overlap with the bands below it).  k_cg_proj forms b in one
// parallel pass and solves the triangle in the last block, as c = M b with
// M = (I + L)^-1 (L the strictly lower part of G): G does not change while a
// band iterates, so M gains one row per stored band and each solve is a
// parallel matrix-vector product instead of ib dependent steps.
//   PROJ_START  real coefficients (real(8) :: s, CG.f90:46,70): the
//               recurrence keeps c real, so c = (I + Re L)^-1 Re b
//   PROJ_VEC    complex coefficients (CG.f90:86-89)
//   PROJ_GRAM   stores row ib of G from the stored band ib, and row ib of
//               both inverses: M(ib, l) = -sum_{l<=q<ib} G(ib, q) M(q, l)
//               (unit diagonal implicit).
// Grid: (nchunk, ceil(ib/kJG), NK).
enum : int { PROJ_START = 0, PROJ_VEC = 1, PROJ_GRAM = 2, PROJ_PLAIN = 3 };
constexpr int kCgMaxNB = 1024;

__global__ void __launch_bounds__(256) k_cg_proj(State S, const cplx<double>* __restrict__ zu, int ib, Src src,
                                                 int mode, int check_done) {
  const int k = blockIdx.z;
  if (check_done && S.done[k]) return;
  const int c = blockIdx.x, g0 = blockIdx.y * kJG;
  const int nj = min(kJG, ib - g0);
  const double xhx = S.sc[k * kNS + XHX];
  const cplx<double>* B = zu + (size_t)k * S.NB * S.NL;
  double acc[2 * kJG];
#pragma unroll
  for (int q = 0; q < 2 * kJG; ++q) acc[q] = 0.0;
  // every load of the block is unconditional (points past the grid read
  // point 0, bands past ib the last lower band; their terms are zeroed by a
  // select), so no branch separates the loads from their uses and they are
  // in flight together; each band's sum takes its points in the same order
  cplx<double> e[kPPT];
  int pc[kPPT];
  bool okp[kPPT];
#pragma unroll
  for (int i = 0; i < kPPT; ++i) {
    const int p = c * kChunk + i * 256 + threadIdx.x;
    okp[i] = p < S.NL;
    pc[i] = okp[i] ? p : 0;
  }
  // (Src::at with its source choice outside the point loop)
  if (src.v) {
#pragma unroll
    for (int i = 0; i < kPPT; ++i) e[i] = src.v[(size_t)k * src.vstride + pc[i]];
  } else {
#pragma unroll
    for (int i = 0; i < kPPT; ++i) e[i] = src.at(k, S.NL, pc[i], xhx);
  }
#pragma unroll
  for (int j = 0; j < kJG; ++j) {
    const int jb = g0 + (j < nj ? j : nj - 1);
#pragma unroll
    for (int i = 0; i < kPPT; ++i) {
      const cplx<double> b0 = B[(size_t)jb * S.NL + pc[i]];
      const bool w = okp[i] && j < nj;
      const double br = w ? b0.re : 0.0, bi = w ? b0.im : 0.0;
      acc[2 * j] += br * e[i].re + bi * e[i].im;
      acc[2 * j + 1] += br * e[i].im - bi * e[i].re;
    }
  }
  __shared__ double smem[4 * 2 * kJG];
  block_sum256<2 * kJG>(acc, smem);
  // partials: [group][chunk][2*kJG]; the last block of row k sums over chunks
  __shared__ int is_last;
  double* part = S.part + (size_t)k * S.pstride;
  const int ngroup = gridDim.y, nblk = S.nchunk * ngroup;
  const int bid = blockIdx.y * S.nchunk + c;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < 2 * kJG; ++q)
      __hip_atomic_store(&part[(size_t)bid * kPartQ + q], acc[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned t = __hip_atomic_fetch_add(&S.cnt[k], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    is_last = (t == (unsigned)(nblk - 1));
  }
  __syncthreads();
  if (!is_last) return;
  __shared__ cplx<double> zsh[kCgMaxNB];
  // one thread per lower band sums its chunks in chunk order; the partials of
  // kTailBatch chunks are loaded (clamped chunk index) before any is added,
  // so one cache round trip serves the batch
  constexpr int kTailBatch = 32;
  for (int j = threadIdx.x; j < ib; j += 256) {
    const int grp = j / kJG, jj = j % kJG;
    double sr = 0.0, si = 0.0;
    for (int c0 = 0; c0 < S.nchunk; c0 += kTailBatch) {
      double vr[kTailBatch], vi[kTailBatch];
#pragma unroll
      for (int u = 0; u < kTailBatch; ++u) {
        const size_t b = (size_t)(grp * S.nchunk + min(c0 + u, S.nchunk - 1)) * kPartQ;
        vr[u] = __hip_atomic_load(&part[b + 2 * jj], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        vi[u] = __hip_atomic_load(&part[b + 2 * jj + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
#pragma unroll
      for (int u = 0; u < kTailBatch; ++u) {
        const bool ok = c0 + u < S.nchunk;
        sr = ok ? sr + vr[u] : sr;
        si = ok ? si + vi[u] : si;
      }
    }
    zsh[j] = cplx<double>{sr * S.Hxyz, si * S.Hxyz};
  }
  if (threadIdx.x == 0) __hip_atomic_store(&S.cnt[k], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  cplx<double>* gram = S.gram + (size_t)k * S.NB * S.NB;
  cplx<double>* M = S.minv + (size_t)k * S.NB * S.NB;
  double* MR = S.minvr + (size_t)k * S.NB * S.NB;
  if (mode == PROJ_GRAM) {
    // raw_j = Hxyz <B_j|B_ib>, so G(ib, j) = conj(raw_j)
    for (int j = threadIdx.x; j < ib; j += 256) gram[(size_t)ib * S.NB + j] = cplx<double>{zsh[j].re, -zsh[j].im};
    // the unit diagonal of M is implicit (band 0 has no projection launch)
    for (int l = threadIdx.x; l < ib; l += 256) {
      double sr = zsh[l].re, si = -zsh[l].im, rr = zsh[l].re;  // q = l: G(ib, l) M(l, l)
      for (int q = l + 1; q < ib; ++q) {
        const double gr = zsh[q].re, gi = -zsh[q].im;
        const cplx<double> m = M[(size_t)q * S.NB + l];
        sr += gr * m.re - gi * m.im;
        si += gr * m.im + gi * m.re;
        rr += gr * MR[(size_t)q * S.NB + l];
      }
      M[(size_t)ib * S.NB + l] = cplx<double>{-sr, -si};
      MR[(size_t)ib * S.NB + l] = -rr;
    }
    return;
  }
  if (mode == PROJ_PLAIN) {  // coefficients against the unmodified vector
    for (int j = threadIdx.x; j < ib; j += 256) S.zs[(size_t)k * S.NB + j] = zsh[j];
    return;
  }
  // c = M b, one thread per row in column order; the row's entries of
  // kTailBatch columns are loaded (clamped column index) before any is used
  for (int j = threadIdx.x; j < ib; j += 256) {
    double cr = zsh[j].re, ci = mode == PROJ_START ? 0.0 : zsh[j].im;
    if (mode == PROJ_START) {
      for (int l0 = 0; l0 < j; l0 += kTailBatch) {
        double mr[kTailBatch];
#pragma unroll
        for (int u = 0; u < kTailBatch; ++u) mr[u] = MR[(size_t)j * S.NB + min(l0 + u, j - 1)];
#pragma unroll
        for (int u = 0; u < kTailBatch; ++u) {
          const int l = l0 + u;
          cr = l < j ? cr + mr[u] * zsh[min(l, j - 1)].re : cr;
        }
      }
    } else {
      for (int l0 = 0; l0 < j; l0 += kTailBatch) {
        cplx<double> mm[kTailBatch];
#pragma unroll
        for (int u = 0; u < kTailBatch; ++u) mm[u] = M[(size_t)j * S.NB + min(l0 + u, j - 1)];
#pragma unroll
        for (int u = 0; u < kTailBatch; ++u) {
          const int l = l0 + u;
          const cplx<double> m = mm[u], z = zsh[min(l, j - 1)];
          const double tr = cr + (m.re * z.re - m.im * z.im), ti = ci + (m.re * z.im + m.im * z.re);
          cr = l < j ? tr : cr;
          ci = l < j ? ti : ci;
        }
      }
    }
    S.zs[(size_t)k * S.NB + j] = cplx<double>{cr, ci};
  }
}

// Roles of the subtraction kernel's reduction.
enum : int {
  SUB_START = 0,  // start vector: SN = 1/sqrt(|out|^2 Hxyz); resets the band's flags
  SUB_GRAD = 1,   // residual, no preconditioner: s = |g|^2 Hxyz (CG.f90:90)
  SUB_PCDIR = 2,  // preconditioned residual: s = Re<Kg|g> Hxyz
  SUB_NONE = 3,   // no reduction
  SUB_LEARN = 4   // learned direction: no conjugation (uk = 0); the network is
                  // not a fixed linear operator, so each iteration is a
                  // preconditioned steepest-descent step with the 2x2 Rayleigh-Ritz
};

// out = src - sum_{j<ib} zs(j) B_j  (CG.f90:69-71, 86-89), then the role's
// reduction.  For SUB_GRAD / SUB_PCDIR the finisher forms the conjugation
// ratio uk = s/gkgk (CG.f90:92-98; iteration 1 takes p = direction).
// Grid (nsub, NK).
__global__ void __launch_bounds__(256) k_cg_sub(State S, const cplx<double>* __restrict__ zu, int ib, Src src,
                                                cplx<double>* out, const cplx<double>* __restrict__ other,
                                                int role, int it, int ncg) {
  const int k = blockIdx.y;
  if (role != SUB_START && S.done[k]) return;
  const int c = blockIdx.x;
  const size_t off = (size_t)k * S.NL;
  const double xhx = S.sc[k * kNS + XHX];
  const cplx<double>* B = zu + (size_t)k * S.NB * S.NL;
  const cplx<double>* zs = S.zs + (size_t)k * S.NB;
  double v[1] = {0.0};
#pragma unroll
  for (int i = 0; i < kSubPPT; ++i) {
    const int p = c * kSubChunk + i * 256 + threadIdx.x;
    if (p < S.NL) {
      cplx<double> e = src.at(k, S.NL, p, xhx);
#pragma unroll 8
      for (int j = 0; j < ib; ++j) {
        const cplx<double> b = B[(size_t)j * S.NL + p], z = zs[j];
        e.re -= b.re * z.re - b.im * z.im;
        e.im -= b.re * z.im + b.im * z.re;
      }
      out[off + p] = e;
      if (role == SUB_PCDIR) {
        const cplx<double> o = other[off + p];
        v[0] += e.re * o.re + e.im * o.im;
      } else {
        v[0] += e.re * e.re + e.im * e.im;
      }
    }
  }
  if (role == SUB_NONE) return;
  if (role == SUB_LEARN) {
    if (c == 0 && threadIdx.x == 0) S.sc[k * kNS + UK] = 0.0;
    return;
  }
  double tot[1];
  if (!last_block_sum<1>(v, S, S.nsub, c, tot)) return;
  if (threadIdx.x != 0) return;
  double* sc = S.sc + k * kNS;
  if (role == SUB_START) {
    sc[SN] = 1.0 / sqrt(tot[0] * S.Hxyz);
    S.done[k] = 0;
    S.iters[(size_t)k * S.NB + ib] = ncg;
  } else {
    const double s = tot[0] * S.Hxyz;
    sc[UK] = it == 1 ? 0.0 : s / sc[GKGK];
    sc[GKGK] = s;
  }
}

// v *= SN (normalisation, CG.f90:72-73,104-105); also copies into dst when given.
__global__ void __launch_bounds__(256) k_cg_scale(State S, cplx<double>* __restrict__ v, cplx<double>* __restrict__ dst,
                                                  size_t dst_stride, int check_done) {
  const int k = blockIdx.y;
  if (check_done && S.done[k]) return;
  const double s = S.sc[k * kNS + SN];
  const size_t off = (size_t)k * S.NL;
#pragma unroll
  for (int i = 0; i < kPPT; ++i) {
    const int p = blockIdx.x * kChunk + i * 256 + threadIdx.x;
    if (p < S.NL) {
      cplx<double> e = v[off + p];
      e.re *= s;
      e.im *= s;
      v[off + p] = e;
      if (dst) dst[(size_t)k * dst_stride + p] = e;
    }
  }
}

// Reductions of the form sum conj(a) b over the chunk; NQ = 1 takes the real
// part, NQ = 2 the complex value.
template <int NQ>
__device__ __forceinline__ void dot_acc(double (&v)[NQ], cplx<double> a, cplx<double> b) {
  v[0] += a.re * b.re + a.im * b.im;
  if constexpr (NQ > 1) v[1] += a.re * b.im - a.im * b.re;
}

// xkHxk, xkTxk of the start vector (CG.f90:79-80).
__global__ void __launch_bounds__(256) k_cg_xhx(State S, const cplx<double>* __restrict__ x,
                                                const cplx<double>* __restrict__ hx, const cplx<double>* __restrict__ tx) {
  const int k = blockIdx.y, c = blockIdx.x;
  const size_t off = (size_t)k * S.NL;
  double v[2] = {0.0, 0.0};
#pragma unroll
  for (int i = 0; i < kPPT; ++i) {
    const int p = c * kChunk + i * 256 + threadIdx.x;
    if (p < S.NL) {
      const cplx<double> e = x[off + p], h = hx[off + p], t = tx[off + p];
      v[0] += e.re * h.re + e.im * h.im;
      v[1] += e.re * t.re + e.im * t.im;
    }
  }
  double tot[2];
  if (!last_block_sum<2>(v, S, S.nchunk, c, tot)) return;
  if (threadIdx.x == 0) {
    S.sc[k * kNS + XHX] = tot[0] * S.Hxyz;
    S.sc[k * kNS + XTX] = tot[1] * S.Hxyz;
  }
}

// p = dir + uk p (CG.f90:92-98), then zs = <x|p> Hxyz (CG.f90:103).
__global__ void __launch_bounds__(256) k_cg_pupd(State S, const cplx<double>* __restrict__ dir,
                                                 cplx<double>* __restrict__ P, const cplx<double>* __restrict__ x) {
  const int k = blockIdx.y, c = blockIdx.x;
  if (S.done[k]) return;
  const size_t off = (size_t)k * S.NL;
  const double uk = S.sc[k * kNS + UK];
  double v[2] = {0.0, 0.0};
#pragma unroll
  for (int i = 0; i < kPPT; ++i) {
    const int p = c * kChunk + i * 256 + threadIdx.x;
    if (p < S.NL) {
      cplx<double> d = dir[off + p];
      if (uk != 0.0) {
        const cplx<double> q = P[off + p];
        d.re += uk * q.re;
        d.im += uk * q.im;
      }
      P[off + p] = d;
      dot_acc<2>(v, x[off + p], d);
    }
  }
  double tot[2];
  if (!last_block_sum<2>(v, S, S.nchunk, c, tot)) return;
  if (threadIdx.x == 0) {
    S.sc[k * kNS + ZSR] = tot[0] * S.Hxyz;
    S.sc[k * kNS + ZSI] = tot[1] * S.Hxyz;
  }
}

// pko = p - x zs, SN = 1/sqrt(|pko|^2 Hxyz) (CG.f90:103-105).
__global__ void __launch_bounds__(256) k_cg_pko(State S, const cplx<double>* __restrict__ P,
                                                const cplx<double>* __restrict__ x, cplx<double>* __restrict__ pko) {
  const int k = blockIdx.y, c = blockIdx.x;
  if (S.done[k]) return;
  const size_t off = (size_t)k * S.NL;
  const cplx<double> z{S.sc[k * kNS + ZSR], S.sc[k * kNS + ZSI]};
  double v[1] = {0.0};
#pragma unroll
  for (int i = 0; i < kPPT; ++i) {
    const int p = c * kChunk + i * 256 + threadIdx.x;
    if (p < S.NL) {
      const cplx<double> e = x[off + p], q = P[off + p];
      cplx<double> o;
      o.re = q.re - (e.re * z.re - e.im * z.im);
      o.im = q.im - (e.re * z.im + e.im * z.re);
      pko[off + p] = o;
      v[0] += o.re * o.re + o.im * o.im;
    }
  }
  double tot[1];
  if (!last_block_sum<1>(v, S, S.nchunk, c, tot)) return;
  if (threadIdx.x == 0) S.sc[k * kNS + SN] = 1.0 / sqrt(tot[0] * S.Hxyz);
}

// 2x2 Rayleigh-Ritz of span{x, pko} (CG.f90:107-114).
__global__ void __launch_bounds__(256) k_cg_rr(State S, const cplx<double>* __restrict__ x,
                                               const cplx<double>* __restrict__ pko, const cplx<double>* __restrict__ hp,
                                               int ib, int it) {
  const int k = blockIdx.y, c = blockIdx.x;
  if (S.done[k]) return;
  const size_t off = (size_t)k * S.NL;
  double v[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int i = 0; i < kPPT; ++i) {
    const int p = c * kChunk + i * 256 + threadIdx.x;
    if (p < S.NL) {
      const cplx<double> e = x[off + p], q = pko[off + p], h = hp[off + p];
      v[0] += e.re * h.re + e.im * h.im;
      v[1] += e.re * h.im - e.im * h.re;
      v[2] += q.re * h.re + q.im * h.im;
    }
  }
  double tot[3];
  if (!last_block_sum<3>(v, S, S.nchunk, c, tot)) return;
  if (threadIdx.x != 0) return;
  double* sc = S.sc + k * kNS;
  const double xhx = sc[XHX];
  const double xhp_r = tot[0] * S.Hxyz, xhp_i = tot[1] * S.Hxyz, php = tot[2] * S.Hxyz;
  const double a2 = xhp_r * xhp_r + xhp_i * xhp_i;
  const double ev = 0.5 * ((xhx + php) - sqrt((xhx - php) * (xhx - php) + 4.0 * a2));
  double cxr = xhp_r / (ev - xhx), cxi = xhp_i / (ev - xhx);
  const double cp = 1.0 / sqrt(1.0 + (cxr * cxr + cxi * cxi));
  cxr *= cp;
  cxi *= cp;
  sc[EVX] = ev;
  sc[CXR] = cxr;
  sc[CXI] = cxi;
  sc[CP] = cp;
  if (fabs(ev - xhx) < 1e-15) {
    S.done[k] = 1;
    S.iters[(size_t)k * S.NB + ib] = it;
  }
}

// x = cx x + cp pko, hx, tx likewise (CG.f90:115-117); then <x|H|x>, <x|T|x>
// (CG.f90:118-119).
__global__ void __launch_bounds__(256) k_cg_rotate(State S, cplx<double>* __restrict__ x, cplx<double>* __restrict__ hx,
                                                   cplx<double>* __restrict__ tx, const cplx<double>* __restrict__ pko,
                                                   const cplx<double>* __restrict__ hp, const cplx<double>* __restrict__ tp) {
  const int k = blockIdx.y, c = blockIdx.x;
  if (S.done[k]) return;
  const size_t off = (size_t)k * S.NL;
  const double* sc = S.sc + k * kNS;
  const cplx<double> cx{sc[CXR], sc[CXI]};
  const double cp = sc[CP];
  double v[2] = {0.0, 0.0};
#pragma unroll
  for (int i = 0; i < kPPT; ++i) {
    const int p = c * kChunk + i * 256 + threadIdx.x;
    if (p < S.NL) {
      cplx<double> e = cmulc(cx, x[off + p]), h = cmulc(cx, hx[off + p]), t = cmulc(cx, tx[off + p]);
      const cplx<double> q = pko[off + p], hq = hp[off + p], tq = tp[off + p];
      e.re += cp * q.re;
      e.im += cp * q.im;
      h.re += cp * hq.re;
      h.im += cp * hq.im;
      t.re += cp * tq.re;
      t.im += cp * tq.im;
      x[off + p] = e;
      hx[off + p] = h;
      tx[off + p] = t;
      v[0] += e.re * h.re + e.im * h.im;
      v[1] += e.re * t.re + e.im * t.im;
    }
  }
  double tot[2];
  if (!last_block_sum<2>(v, S, S.nchunk, c, tot)) return;
  if (threadIdx.x == 0) {
    S.sc[k * kNS + XHX] = tot[0] * S.Hxyz;
    S.sc[k * kNS + XTX] = tot[1] * S.Hxyz;
  }
}

// Final normalisation factor of the band (CG.f90:122).
__global__ void __launch_bounds__(256) k_cg_norm(State S, const cplx<double>* __restrict__ x) {
  const int k = blockIdx.y, c = blockIdx.x;
  const size_t off = (size_t)k * S.NL;
  double v[1] = {0.0};
#pragma unroll
  for (int i = 0; i < kPPT; ++i) {
    const int p = c * kChunk + i * 256 + threadIdx.x;
    if (p < S.NL) {
      const cplx<double> e = x[off + p];
      v[0] += e.re * e.re + e.im * e.im;
    }
  }
  double tot[1];
  if (!last_block_sum<1>(v, S, S.nchunk, c, tot)) return;
  if (threadIdx.x == 0) S.sc[k * kNS + SN] = 1.0 / sqrt(tot[0] * S.Hxyz);
}

// e = <x|Hx> Hxyz of the stored band (CG.f90:125).
__global__ void __launch_bounds__(256) k_cg_efin(State S, const cplx<double>* __restrict__ x,
                                                 const cplx<double>* __restrict__ hp) {
  const int k = blockIdx.y, c = blockIdx.x;
  const size_t off = (size_t)k * S.NL;
  double v[1] = {0.0};
#pragma unroll
  for (int i = 0; i < kPPT; ++i) {
    const int p = c * kChunk + i * 256 + threadIdx.x;
    if (p < S.NL) dot_acc<1>(v, x[off + p], hp[off + p]);
  }
  double tot[1];
  if (!last_block_sum<1>(v, S, S.nchunk, c, tot)) return;
  if (threadIdx.x == 0) S.sc[k * kNS + EFIN] = tot[0] * S.Hxyz;
}

// esp_var(ib,k) = sqrt(|Hx - e x|^2 Hxyz) occ(ib,k) (CG.f90:126).
__global__ void __launch_bounds__(256) k_cg_resid(State S, const cplx<double>* __restrict__ x,
                                                  const cplx<double>* __restrict__ hp, int ib) {
  const int k = blockIdx.y, c = blockIdx.x;
  const size_t off = (size_t)k * S.NL;
  const double e = S.sc[k * kNS + EFIN];
  double v[1] = {0.0};
#pragma unroll
  for (int i = 0; i < kPPT; ++i) {
    const int p = c * kChunk + i * 256 + threadIdx.x;
    if (p < S.NL) {
      const cplx<double> a = x[off + p], h = hp[off + p];
      const double r = h.re - e * a.re, q = h.im - e * a.im;
      v[0] += r * r + q * q;
    }
  }
  double tot[1];
  if (!last_block_sum<1>(v, S, S.nchunk, c, tot)) return;
  if (threadIdx.x == 0) {
    const size_t o = (size_t)k * S.NB + ib;
    S.esp_var[o] = sqrt(tot[0] * S.Hxyz) * S.occ[o];
  }
}

// ---------------------------------------------------------------------------
// Kinetic preconditioner K g = F^-1 [ k(x) F g ], x = t(q)/<x|T|x>, with t(q)
// the finite-difference kinetic symbol -1/2 sum_axes (lap0 + 2 sum_m lap_m
// cos(m q h)) and k(x) = f/(f + 16 x^4), f = 27 + 18x + 12x^2 + 8x^3 (the
// Teter-Payne-Allan form).  The 3-D transform is three passes of 1-D DFTs
// along z, y and x (direct sums over a line held in LDS, twiddles from a
// table, line length <= kDftMax); the x pass applies k(x) between its forward
// and inverse transforms.  No reference counterpart: CG_ompk applies no
// preconditioner (CG.f90:85-92).
// ---------------------------------------------------------------------------
constexpr int kDftMax = 64;

struct DftAxis {
  int n;       // line length
  int es;      // element stride
  int nlines;  // lines of the axis
  int lpb;     // lines per block
};

// line l of axis a -> base offset (z: l*NLz; y: ix*NLy*NLz + iz with l = ix*NLz + iz;
// x: l with l = iy*NLz + iz)
__device__ __forceinline__ int dft_line_base(int axis, int l, const Grid& g) {
  if (axis == 2) return l * g.NLz;
  if (axis == 1) return (l / g.NLz) * g.NLy * g.NLz + (l % g.NLz);
  return l;
}

// One DFT pass.  dir = -1 forward (exp(-2 pi i c m/n)), +1 inverse (scaled by
// 1/n).  For axis x with symb != nullptr: forward, times k(x), inverse.
// tw = (cos, sin)(2 pi j/n), j = 0..n-1, per axis at tw + axis*kDftMax.
__global__ void __launch_bounds__(256) k_dft_pass(const State S, const cplx<double>* __restrict__ in,
                                                  cplx<double>* __restrict__ out, Grid g, int axis, int dir,
                                                  const double* __restrict__ tw, const double* __restrict__ symb,
                                                  int check_done) {
  const int k = blockIdx.y;
  if (check_done && S.done[k]) return;
  const int n = axis == 2 ? g.NLz : (axis == 1 ? g.NLy : g.NLx);
  const int es = axis == 2 ? 1 : (axis == 1 ? g.NLz : g.NLy * g.NLz);
  const int nlines = g.NL / n;
  const int lpb = 256 / n;
  const int l0 = blockIdx.x * lpb;
  __shared__ cplx<double> s[256 + 8];
  __shared__ double2 w[kDftMax];
  for (int j = threadIdx.x; j < n; j += 256) w[j] = make_double2(tw[2 * (axis * kDftMax + j)], tw[2 * (axis * kDftMax + j) + 1]);
  const size_t off = (size_t)k * g.NL;
  // element e of line l sits at s[li*n + e]; z lines load contiguously,
  // y and x lines with consecutive lines on consecutive threads
  const int t = threadIdx.x;
  int li, e;
  if (axis == 2) { li = t / n; e = t % n; } else { li = t % lpb; e = t / lpb; }
  const bool active = t < lpb * n && (l0 + li) < nlines;
  const int base = active ? dft_line_base(axis, l0 + li, g) : 0;
  if (active) s[li * n + e] = in[off + base + (size_t)e * es];
  __syncthreads();
  auto dft = [&](int sgn) -> cplx<double> {
    double re = 0.0, im = 0.0;
    const cplx<double>* a = s + li * n;
    int j = 0;
    for (int m = 0; m < n; ++m) {
      const double c = w[j].x, sn = sgn * w[j].y;
      re += c * a[m].re - sn * a[m].im;
      im += c * a[m].im + sn * a[m].re;
      j += e;
      if (j >= n) j -= n;
    }
    return cplx<double>{re, im};
  };
  cplx<double> r{0.0, 0.0};
  if (active) r = dft(dir);
  if (symb) {
    // x pass: forward done; multiply by k(x) and transform back
    __syncthreads();
    if (active) {
      const int l = l0 + li, iy = l / g.NLz, iz = l % g.NLz;
      const double tq = (symb[e] + symb[kDftMax + iy]) + symb[2 * kDftMax + iz];
      double txx = S.sc[k * kNS + XTX];
      if (!(txx > 1e-12)) txx = 1e-12;
      const double x = tq / txx;
      const double f = 27.0 + 18.0 * x + 12.0 * x * x + 8.0 * x * x * x;
      const double kf = f / (f + 16.0 * x * x * x * x);
      s[li * n + e] = cplx<double>{kf * r.re, kf * r.im};
    }
    __syncthreads();
    if (active) r = dft(+1);
  }
  if (active) {
    if (dir > 0 || symb) {
      const double inv = 1.0 / n;
      r.re *= inv;
      r.im *= inv;
    }
    out[off + base + (size_t)e * es] = r;
  }
}

// ---------------------------------------------------------------------------
// Subspace matrix and rotation of diag_omp (ARTED/GS/diag.f90:111-127).
// ---------------------------------------------------------------------------

// M(i,j) partials = sum_p conj(A_i(p)) B_j(p) over a slice of points; 16x16
// tiles of (i, j), one thread per pair, points staged through LDS.
// Grid (ceil(NB/16)^2, nslice, NK); out: part[(k*nslice + slice)*NB*NB + i*NB + j].
constexpr int kTile = 16, kSlicePts = 64;
__global__ void __launch_bounds__(256) k_sub_matrix(const cplx<double>* __restrict__ A, const cplx<double>* __restrict__ B,
                                                    int NB, int NL, int pts_per_slice, double* __restrict__ part) {
  const int ntile = (NB + kTile - 1) / kTile;
  const int ti = blockIdx.x / ntile, tj = blockIdx.x % ntile;
  const int k = blockIdx.z, slice = blockIdx.y, nslice = gridDim.y;
  const int i = ti * kTile + threadIdx.x / kTile, j = tj * kTile + threadIdx.x % kTile;
  __shared__ cplx<double> sa[kTile][kSlicePts + 1], sb[kTile][kSlicePts + 1];
  const cplx<double>* Ak = A + (size_t)k * NB * NL;
  const cplx<double>* Bk = B + (size_t)k * NB * NL;
  const int p0 = slice * pts_per_slice, p1 = min(NL, p0 + pts_per_slice);
  double re = 0.0, im = 0.0;
  for (int q0 = p0; q0 < p1; q0 += kSlicePts) {
    for (int idx = threadIdx.x; idx < kTile * kSlicePts; idx += 256) {
      const int r = idx / kSlicePts, pp = idx % kSlicePts, p = q0 + pp;
      const int ia = ti * kTile + r, jb = tj * kTile + r;
      sa[r][pp] = (ia < NB && p < p1) ? Ak[(size_t)ia * NL + p] : cplx<double>{0.0, 0.0};
      sb[r][pp] = (jb < NB && p < p1) ? Bk[(size_t)jb * NL + p] : cplx<double>{0.0, 0.0};
    }
    __syncthreads();
    const int r = threadIdx.x / kTile, cc = threadIdx.x % kTile;
#pragma unroll 8
    for (int pp = 0; pp < kSlicePts; ++pp) {
      const cplx<double> a = sa[r][pp], b = sb[cc][pp];
      re += a.re * b.re + a.im * b.im;
      im += a.re * b.im - a.im * b.re;
    }
    __syncthreads();
  }
  if (i < NB && j < NB) {
    double* o = part + ((size_t)k * nslice + slice) * NB * NB * 2;
    o[2 * ((size_t)i * NB + j)] = re;
    o[2 * ((size_t)i * NB + j) + 1] = im;
  }
}

// sums the slices in order: M(k, i, j) * scale
__global__ void __launch_bounds__(256) k_sub_matrix_sum(const double* __restrict__ part, int NB, int nslice,
                                                        double scale, double* __restrict__ M) {
  const int k = blockIdx.y;
  const int idx = blockIdx.x * 256 + threadIdx.x;
  if (idx >= NB * NB * 2) return;
  double s = 0.0;
#pragma unroll 8
  for (int sl = 0; sl < nslice; ++sl) s += part[((size_t)k * nslice + sl) * NB * NB * 2 + idx];
  M[(size_t)k * NB * NB * 2 + idx] = s * scale;
}

// C_i(p) = sum_j A_j(p) U(j, i) for 16 outputs i per thread (U column-major
// per k: U[(k*NB + i)*NB + j]).  Grid (ceil(NL/256), ceil(NB/16), NK).
__global__ void __launch_bounds__(256) k_rotate_bands(const cplx<double>* __restrict__ A, const cplx<double>* __restrict__ U,
                                                     cplx<double>* __restrict__ C, int NB, int NL) {
  const int k = blockIdx.z, i0 = blockIdx.y * kTile;
  const int p = blockIdx.x * 256 + threadIdx.x;
  extern __shared__ cplx<double> su[];  // NB x kTile
  for (int idx = threadIdx.x; idx < NB * kTile; idx += 256) {
    const int j = idx % NB, ii = idx / NB;
    su[idx] = (i0 + ii < NB) ? U[((size_t)k * NB + i0 + ii) * NB + j] : cplx<double>{0.0, 0.0};
  }
  __syncthreads();
  if (p >= NL) return;
  const cplx<double>* Ak = A + (size_t)k * NB * NL;
  double re[kTile], im[kTile];
#pragma unroll
  for (int ii = 0; ii < kTile; ++ii) { re[ii] = 0.0; im[ii] = 0.0; }
  for (int j = 0; j < NB; ++j) {
    const cplx<double> a = Ak[(size_t)j * NL + p];
#pragma unroll
    for (int ii = 0; ii < kTile; ++ii) {
      const cplx<double> u = su[ii * NB + j];
      re[ii] += a.re * u.re - a.im * u.im;
      im[ii] += a.re * u.im + a.im * u.re;
    }
  }
#pragma unroll
  for (int ii = 0; ii < kTile; ++ii)
    if (i0 + ii < NB) C[((size_t)k * NB + i0 + ii) * NL + p] = cplx<double>{re[ii], im[ii]};
}

// Gram-Schmidt step of band ib in place (Gram_Schmidt_ompk,
// ARTED/GS/Gram_Schmidt.f90:43-50): v -= sum_{j<ib} zs_j B_j; with_norm ends
// with SN = 1/sqrt(|v|^2 Hxyz).  Grid (nsub, NK).
__global__ void __launch_bounds__(256) k_gs_sub(State S, cplx<double>* __restrict__ zu, int ib, int with_norm) {
  const int k = blockIdx.y, c = blockIdx.x;
  cplx<double>* B = zu + (size_t)k * S.NB * S.NL;
  cplx<double>* v = B + (size_t)ib * S.NL;
  const cplx<double>* zs = S.zs + (size_t)k * S.NB;
  double a[1] = {0.0};
#pragma unroll
  for (int i = 0; i < kSubPPT; ++i) {
    const int p = c * kSubChunk + i * 256 + threadIdx.x;
    if (p < S.NL) {
      cplx<double> e = v[p];
#pragma unroll 8
      for (int j = 0; j < ib; ++j) {
        const cplx<double> b = B[(size_t)j * S.NL + p], z = zs[j];
        e.re -= b.re * z.re - b.im * z.im;
        e.im -= b.re * z.im + b.im * z.re;
      }
      v[p] = e;
      a[0] += e.re * e.re + e.im * e.im;
    }
  }
  if (!with_norm) return;
  double tot[1];
  if (!last_block_sum<1>(a, S, S.nsub, c, tot)) return;
  if (threadIdx.x == 0) S.sc[k * kNS + SN] = 1.0 / sqrt(tot[0] * S.Hxyz);
}

// band ib *= SN.  Grid (nchunk, NK).
__global__ void __launch_bounds__(256) k_gs_scale(State S, cplx<double>* __restrict__ zu, int ib) {
  const int k = blockIdx.y;
  const double f = S.sc[k * kNS + SN];
  cplx<double>* v = zu + ((size_t)k * S.NB + ib) * S.NL;
#pragma unroll
  for (int i = 0; i < kPPT; ++i) {
    const int p = blockIdx.x * kChunk + i * 256 + threadIdx.x;
    if (p < S.NL) {
      v[p].re *= f;
      v[p].im *= f;
    }
  }
}

// Training sample of the learned preconditioner: the orthogonalised residual
// and the iterate of k = 0, with the band and <x|H|x>.  valid = 0 when the
// band has already left its iterations.
__global__ void __launch_bounds__(256) k_pc_record(State S, const cplx<double>* __restrict__ g,
                                                   const cplx<double>* __restrict__ x, cplx<double>* __restrict__ sg,
                                                   cplx<double>* __restrict__ sx, double* __restrict__ theta,
                                                   int* __restrict__ sib, int* __restrict__ valid, int ib) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < S.NL) {
    sg[i] = g[i];
    sx[i] = x[i];
  }
  if (i == 0) {
    theta[0] = S.sc[XHX];
    sib[0] = ib;
    valid[0] = S.done[0] ? 0 : 1;
  }
}

}  // namespace cg
}  // namespace salmon
