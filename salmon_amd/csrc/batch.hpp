// Batched real-time step over independent systems of different grids, band
// counts, k points and atoms: the system descriptors, the KB phases and the
// one-launch-per-phase kernels.  In that kernel set each phase of the step
// (KB projection and stencil for each of the four Taylor terms, the Taylor
// fold, the observables, their per-system sums) is one launch covering every
// system: blocks are laid out system after system and a block finds its
// system in a per-phase block -> system table (one load).  The stencil's z
// blocking (points per thread along z) is chosen per grid shape by timing the
// candidates on the system's own data.  The persistent one-launch-per-step
// kernel is in batch_sweep.hpp.  Either way one step is captured in a HIP
// graph and replayed; A(t) comes from a device table indexed by a device step
// counter.
#pragma once
#include "streamed.hpp"

namespace salmon {
namespace batch {

constexpr int kPhases = 5;  // first-block offsets: projection, stencil, fold, observables, KB apply

template <typename T>
struct Sys {
  Grid g;
  KBTables kb;
  cplx<T>* zu;
  cplx<T>* h[4];
  cplx<T>* uv;  // (NKB, Nlma)
  cplx<T>* ph;  // (NK, NPI) KB phases at this step's midpoint k
  cplx<T>* phn; // (NK, NPI) KB phases at k + A(t+dt), for the observables
  cplx<T>* kbv; // (NKB, NT) KB term of every touched point of the current H application
  const T* vloc;
  const double* vloc64;
  const double* kac0;
  const double* occ;
  double* rows;  // (NKB, 8)
  int NK, NB, NKB, zb;
  int sw_zb, sw_tw;          // persistent sweep: column length along z, waves per team
  const int* kb_order;       // persistent sweep: atoms grouped by colour (no shared grid point within a colour)
  const int* kb_cstart;      // (ncolor+1) first entry of each colour in kb_order
  int ncolor;
  int pblk, hblk, kblk;      // projection, stencil and KB-apply blocks per orbital
  long first[kPhases + 1];   // first block of the system in each phase
};

struct Step {
  const double* ac;  // (nsteps+1, 3), or (nsys, nsteps+1, 3) with ac_stride = 3 (nsteps+1)
  const int* step;   // device step counter
  double dt;
  const unsigned short* owner[kPhases];  // system of every block, per phase
  long ac_stride;    // doubles between the A(t) tables of consecutive systems (0: one shared table)
  int obs_only;      // persistent sweep: the launch that closes a run (observables only)
};

__device__ __forceinline__ int find_sys(const Step& st, int phase, long b) { return st.owner[phase][b]; }

// A at the step's midpoint and at t+dt for system s (control_ms.f90 gives
// every macro-point its own Ac_m, Ac_new_m: dt_evolve.f90:85 for the
// midpoint, control_ms.f90:316-318 for the current)
__device__ __forceinline__ void step_fields(const Step& st, int s, double amid[3], double anew[3]) {
  const int n = *st.step;
  const double* ac = st.ac + (size_t)s * st.ac_stride;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    amid[d] = 0.5 * (ac[3 * n + d] + ac[3 * (n + 1) + d]);
    anew[d] = ac[3 * (n + 1) + d];
  }
}

// KB projection of Taylor term `term` (input zu for term 0, else h[term-1]).
template <typename T>
__global__ void __launch_bounds__(256) k_b_project(const Sys<T>* __restrict__ S, int nsys, int term, Step st) {
  const long b = blockIdx.x;
  const int s = find_sys(st, 0, b);
  const Sys<T>& y = S[s];
  const long lb = b - y.first[0];
  const int o = (int)(lb / y.pblk);
  const int ilma = (int)(lb % y.pblk) * 4 + (threadIdx.x / kWave);
  if (o >= y.NKB || ilma >= y.kb.Nlma) return;
  double am[3], an[3];
  step_fields(st, s, am, an);
  double kx, ky, kz;
  orbital_k(y.kac0, y.NK, o / y.NB, am[0], am[1], am[2], kx, ky, kz);
  const cplx<T>* in = (term == 0 ? y.zu : y.h[term - 1]) + (size_t)o * y.g.NL;
  kb_project_wave<T>(in, ilma, kx, ky, kz, y.g, y.kb, y.uv + (size_t)o * y.kb.Nlma + ilma,
                     y.ph + (size_t)(o / y.NB) * y.kb.NPI);
}

// The KB phases of every system's k points (one grid row per system) at k +
// A of the step n the device counter holds: which = 0, the midpoint
// (A(t_n) + A(t_{n+1}))/2, into ph, shared by the step's four projections and
// stencils; 1, A(t_{n+1}), into phn, for the observables after the step (one
// launch per phase); 2, A(t_n), into phn, for the observables of the
// orbitals entering step n (persistent sweep; at the end of a run n = nsteps,
// the last row of the A table).
template <typename T>
__global__ void __launch_bounds__(256) k_b_phase(const Sys<T>* __restrict__ S, Step st, int which) {
  const Sys<T>& y = S[blockIdx.y];
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= (long)y.NK * y.kb.NPI) return;
  const int ik = (int)(i / y.kb.NPI), jp = (int)(i - (long)ik * y.kb.NPI);
  const int n = *st.step;
  const double* ac = st.ac + (size_t)blockIdx.y * st.ac_stride;
  double a[3];
#pragma unroll
  for (int d = 0; d < 3; ++d)
    a[d] = which == 0 ? 0.5 * (ac[3 * n + d] + ac[3 * (n + 1) + d]) : ac[3 * (n + (which == 1 ? 1 : 0)) + d];
  double kx, ky, kz;
  orbital_k(y.kac0, y.NK, ik, a[0], a[1], a[2], kx, ky, kz);
  double s, c;
  sincos(kx * y.kb.pos[3 * jp] + ky * y.kb.pos[3 * jp + 1] + kz * y.kb.pos[3 * jp + 2], &s, &c);
  (which ? y.phn : y.ph)[i] = cplx<T>{(T)c, (T)s};
}

// KB term of every touched grid point of every orbital (pseudo_pt phase 2,
// common/hpsi.f90:195-206, gathered through the transposed table): one
// thread per (orbital, touched point) sums its table entries, so the stencil
// adds one value per touched point instead of looping over entries with most
// lanes of the wave idle.
template <typename T>
__global__ void __launch_bounds__(256) k_b_kbapply(const Sys<T>* __restrict__ S, Step st) {
  const long b = blockIdx.x;
  const int s = find_sys(st, 4, b);
  const Sys<T>& y = S[s];
  const long lb = b - y.first[4];
  const int o = (int)(lb / y.kblk);
  const int t = (int)(lb % y.kblk) * 256 + threadIdx.x;
  const KBTables& kb = y.kb;
  if (o >= y.NKB || t >= kb.NT) return;
  const cplx<T>* ph = y.ph + (size_t)(o / y.NB) * kb.NPI;
  const cplx<T>* U = y.uv + (size_t)o * kb.Nlma;
  T fr = 0, fi = 0;
  for (int r = kb.t_start[t]; r < kb.t_start[t + 1]; ++r) {
    const int il = kb.e_lma[r], jp = kb.e_jp[r];
    const int a = kb.lma_atom[il];
    const T cs = ph[jp].re, sn = ph[jp].im;
    const T u = (T)kb.uv[(size_t)il * kb.Nps + (jp - kb.atom_start[a])];
    // zproj * uVpsi = u (c - i s) (U.re + i U.im)
    const cplx<T> w = U[il];
    fr += u * (cs * w.re + sn * w.im);
    fi += u * (cs * w.im - sn * w.re);
  }
  y.kbv[(size_t)o * kb.NT + t] = cplx<T>{fr, fi};
}

// ZB consecutive z points of one (ix, iy) line: the z neighbours come from a
// register window of ZB+8 values; the y and x terms and the potential are
// added per point in the order of hpsi_point, then the touched point's KB term
// (k_b_kbapply).
// ZB consecutive values from src; complex64 runs of an even length as 16-byte
// loads (the callers' runs start at even indices of 16-byte aligned orbitals)
template <typename T, int ZB>
__device__ __forceinline__ void load_run(const cplx<T>* __restrict__ src, cplx<T> (&v)[ZB]) {
  if constexpr (sizeof(T) == 4 && ZB % 2 == 0) {
    const float4* s4 = reinterpret_cast<const float4*>(src);
#pragma unroll
    for (int h = 0; h < ZB / 2; ++h) {
      const float4 f = s4[h];
      v[2 * h] = cplx<T>{f.x, f.y};
      v[2 * h + 1] = cplx<T>{f.z, f.w};
    }
  } else {
#pragma unroll
    for (int q = 0; q < ZB; ++q) v[q] = src[q];
  }
}

template <typename T, int ZB>
__device__ __forceinline__ void store_run(cplx<T>* __restrict__ dst, const cplx<T> (&v)[ZB]) {
  if constexpr (sizeof(T) == 4 && ZB % 2 == 0) {
    float4* d4 = reinterpret_cast<float4*>(dst);
#pragma unroll
    for (int h = 0; h < ZB / 2; ++h) d4[h] = float4{v[2 * h].re, v[2 * h].im, v[2 * h + 1].re, v[2 * h + 1].im};
  } else {
#pragma unroll
    for (int q = 0; q < ZB; ++q) dst[q] = v[q];
  }
}

template <typename T, int ZB>
__device__ __forceinline__ void hpsi_zblock(const cplx<T>* __restrict__ E, int p0, double kx, double ky, double kz,
                                            const T* __restrict__ vloc, const Grid& g, const KBTables& kb,
                                            const cplx<T>* __restrict__ kbv, cplx<T>* __restrict__ out) {
  const int NLy = g.NLy, NLz = g.NLz;
  const int iz0 = p0 % NLz, iy = (p0 / NLz) % NLy, ix = p0 / (NLz * NLy);
  const int line = p0 - iz0;
  cplx<T> win[ZB + 8];
#pragma unroll
  for (int j = 0; j < ZB + 8; ++j) {
    int z = iz0 - 4 + j;
    z = z < 0 ? z + NLz : (z >= NLz ? z - NLz : z);
    win[j] = E[line + z];
  }
  const T A = (T)((kx * kx + ky * ky + kz * kz - (g.lap[0][0] + g.lap[1][0] + g.lap[2][0])) * 0.5);
  const double kk[3] = {kx, ky, kz};
  // per own point: z neighbours (window), then y, then x, each n = 1..4
  T vr[ZB], vi[ZB], wr[ZB], wi[ZB];
#pragma unroll
  for (int q = 0; q < ZB; ++q) {
    vr[q] = vi[q] = wr[q] = wi[q] = 0;
#pragma unroll
    for (int n = 1; n <= 4; ++n) {
      const cplx<T> ep = win[4 + q + n], em = win[4 + q - n];
      const T C = (T)g.lap[2][n], D = (T)(kz * g.nab[2][n]);
      vr[q] += C * (ep.re + em.re);
      vi[q] += C * (ep.im + em.im);
      wr[q] += D * (ep.re - em.re);
      wi[q] += D * (ep.im - em.im);
    }
  }
#pragma unroll
  for (int ax3 = 1; ax3 >= 0; --ax3) {
    const int n_ax = ax3 == 1 ? NLy : g.NLx;
    const int c_ax = ax3 == 1 ? iy : ix;
    const int stride = ax3 == 1 ? NLz : NLz * NLy;
    const int base = p0 - c_ax * stride;
#pragma unroll
    for (int n = 1; n <= 4; ++n) {
      int cp = c_ax + n;
      if (cp >= n_ax) cp -= n_ax;
      int cm = c_ax - n;
      if (cm < 0) cm += n_ax;
      cplx<T> ep[ZB], em[ZB];
      load_run<T, ZB>(E + base + cp * stride, ep);
      load_run<T, ZB>(E + base + cm * stride, em);
      const T C = (T)g.lap[ax3][n], D = (T)(kk[ax3] * g.nab[ax3][n]);
#pragma unroll
      for (int q = 0; q < ZB; ++q) {
        vr[q] += C * (ep[q].re + em[q].re);
        vi[q] += C * (ep[q].im + em[q].im);
        wr[q] += D * (ep[q].re - em[q].re);
        wi[q] += D * (ep[q].im - em[q].im);
      }
    }
  }
  cplx<T> f[ZB];
#pragma unroll
  for (int q = 0; q < ZB; ++q) {
    const int p = p0 + q;
    const cplx<T> e = win[4 + q];
    const T V = vloc[p];
    f[q].re = (V + A) * e.re - (T)0.5 * vr[q] + wi[q];
    f[q].im = (V + A) * e.im - (T)0.5 * vi[q] - wr[q];
    const int t = kb.pt2t[p];
    if (t >= 0) {
      const cplx<T> v = kbv[t];
      f[q].re += v.re;
      f[q].im += v.im;
    }
  }
  store_run<T, ZB>(out + p0, f);
}

// Stencil + local + KB of Taylor term `term`: h[term] = H (term == 0 ? zu : h[term-1]).
// One launch per z-blocking class (the systems tuned to ZB); block `base +
// blockIdx.x` of the stencil phase.
template <typename T, int ZB>
__global__ void __launch_bounds__(256) k_b_hpsi(const Sys<T>* __restrict__ S, int term, Step st, long base) {
  const long b = base + blockIdx.x;
  const int s = find_sys(st, 1, b);
  const Sys<T>& y = S[s];
  const long lb = b - y.first[1];
  const int o = (int)(lb / y.hblk);
  const int p0 = ((int)(lb % y.hblk) * 256 + threadIdx.x) * ZB;
  if (o >= y.NKB || p0 >= y.g.NL) return;
  double am[3], an[3];
  step_fields(st, s, am, an);
  double kx, ky, kz;
  orbital_k(y.kac0, y.NK, o / y.NB, am[0], am[1], am[2], kx, ky, kz);
  const cplx<T>* in = (term == 0 ? y.zu : y.h[term - 1]) + (size_t)o * y.g.NL;
  hpsi_zblock<T, ZB>(in, p0, kx, ky, kz, y.vloc, y.g, y.kb, y.kbv + (size_t)o * y.kb.NT, y.h[term] + (size_t)o * y.g.NL);
}

// zu += sum_n zfac(n) h_n, zfac(n) = (-i dt)^n/n! (ARTED/RT/hamiltonian.f90:46-49,99-122).
template <typename T>
__global__ void __launch_bounds__(256) k_b_fold(const Sys<T>* __restrict__ S, int nsys, Step st) {
  const long b = blockIdx.x;
  const int s = find_sys(st, 2, b);
  const Sys<T>& y = S[s];
  const size_t i = (size_t)(b - y.first[2]) * 256 + threadIdx.x;
  if (i >= (size_t)y.NKB * y.g.NL) return;
  const double dt = st.dt;
  const T z1i = (T)(-dt), z2r = (T)(-dt * dt / 2.0), z3i = (T)(dt * dt * dt / 6.0), z4r = (T)(dt * dt * dt * dt / 24.0);
  cplx<T> z = y.zu[i];
  const cplx<T> a = y.h[0][i], bb = y.h[1][i], c = y.h[2][i], d = y.h[3][i];
  z.re += -z1i * a.im;
  z.im += z1i * a.re;
  z.re += z2r * bb.re;
  z.im += z2r * bb.im;
  z.re += -z3i * c.im;
  z.im += z3i * c.re;
  z.re += z4r * d.re;
  z.im += z4r * d.im;
  y.zu[i] = z;
}

// Observables of every orbital at kAc0 + A(t+dt): one block per orbital.
template <typename T>
__global__ void __launch_bounds__(256) k_b_observe(const Sys<T>* __restrict__ S, int nsys, Step st) {
  const long b = blockIdx.x;
  const int s = find_sys(st, 3, b);
  const Sys<T>& y = S[s];
  const int o = (int)(b - y.first[3]);
  double am[3], an[3];
  step_fields(st, s, am, an);
  double kx, ky, kz;
  orbital_k(y.kac0, y.NK, o / y.NB, an[0], an[1], an[2], kx, ky, kz);
  observe_orbital<T>(y.zu + (size_t)o * y.g.NL, kx, ky, kz, y.occ[o], y.vloc64, y.g, y.kb, y.rows + (size_t)o * 8,
                     y.phn + (size_t)(o / y.NB) * y.kb.NPI);
}

// Per-system sums of the rows (one block per system) into res(n - lag, s, 0..7),
// n the device step counter (lag 1: the rows hold the previous step's
// observables; nothing before the first step).
template <typename T>
__global__ void __launch_bounds__(256) k_b_sum(const Sys<T>* __restrict__ S, int nsys, int* step, double* res,
                                               int lag) {
  __shared__ double smem[4 * 8];
  if (*step - lag < 0) return;
  const Sys<T>& y = S[blockIdx.x];
  double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int r = threadIdx.x; r < y.NKB; r += 256) {
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] += y.rows[(size_t)r * 8 + i];
  }
  block_sum256<8>(v, smem);
  if (threadIdx.x == 0) {
    const int n = *step - lag;
#pragma unroll
    for (int i = 0; i < 8; ++i) res[((size_t)n * nsys + blockIdx.x) * 8 + i] = v[i];
  }
}

__global__ void k_b_tick(int* step) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *step += 1;
}

}  // namespace batch
}  // namespace salmon
