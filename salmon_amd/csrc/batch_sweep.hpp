// Persistent batched step: one launch per RT step covers every system.
//
// A workgroup (256 threads for complex64, 512 for FP64) pulls work items (system, first orbital) from a
// device counter, costliest first, and runs the whole step of its orbitals:
// for each of the four Taylor terms the KB projection, one sweep over the
// grid for the stencil and the local potential, and the KB term added to the
// touched points; the Taylor fold is part of the fourth sweep, and a fifth
// sweep with its own projections evaluates the observables at k + A(t+dt).
//
// Sweep layout: the workgroup's waves form teams of TW waves, one orbital per
// team.  Thread i of a team owns the column of ZB consecutive z points at
// in-plane offset i*ZB of every x plane and walks x = 0 .. NLx-1 with the
// column's planes x-4 .. x+5 in registers (one global load per value and
// sweep, the x neighbours never leave the thread).  The y and z neighbours
// come from the current plane, which the team stores in one of two LDS plane
// buffers (x parity) before a barrier; a column takes col_stride<ZB> slots
// there, one more than ZB for even ZB, so that the lanes' 8- and 16-byte
// reads fall in distinct banks.  The per-point operands of plane x+1 (the
// potential, and for the fold the orbital and its first two H applications)
// are loaded during step x, before the window load of plane x+5.
#pragma once
#include "batch.hpp"

namespace salmon {
namespace batch {

// waves per workgroup: complex64 teams need at most 4 waves (two 256-thread
// workgroups share a compute unit, each with its own barriers); FP64 columns
// of at most 2 points need 8 for a 32^2 plane
template <typename T>
constexpr int sweep_waves() {
  return sizeof(T) == 4 ? 4 : 8;
}
constexpr int kSweepMaxWaves = 8;
constexpr int kColWin = 10;      // register window of a column: planes x-4 .. x+5
constexpr int kRedStride = 10;   // doubles per wave in the reduction area
constexpr int kSweepHead = 16 + kSweepMaxWaves * kRedStride * 8;  // bytes before the team areas

// Per-point operands of a sweep besides its input.
template <typename T>
struct SweepIO {
  const cplx<T>* in;   // orbital the stencil applies to
  cplx<T>* out;        // H in (term < 3), or the folded orbital (term 3)
  cplx<T>* zu;         // term 3: the orbital at t (read, then overwritten by out == zu)
  const cplx<T>* h0;   // term 3: H zu
  const cplx<T>* h1;   // term 3: H^2 zu
};

__device__ __forceinline__ int wrap_once(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

// LDS slots per column of ZB points
template <int ZB>
__host__ __device__ constexpr int col_stride() {
  return ZB % 2 == 0 ? ZB + 1 : ZB;
}

// LDS slots of one plane buffer: the columns, then one column that threads
// without a column of their own write to
template <int ZB>
__host__ __device__ constexpr int plane_slots(int P) {
  return (P / ZB + 1) * col_stride<ZB>();
}

typedef unsigned int u32x2_t __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));

// Buffer resource over nbytes at base (wave-uniform base): stores at offsets
// at or past nbytes are dropped by the hardware, which lets threads without
// work issue the same stores as the others.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t orbital_rsrc(const void* base, unsigned nbytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)nbytes, 0x00020000);
}
constexpr int kDropOffset = 0x7ffffff0;

// ZB consecutive values at byte offset off of the resource (16-byte stores for
// complex64 pairs and FP64 values, 8-byte stores otherwise)
template <typename T, int ZB>
__device__ __forceinline__ void store_run_rsrc(__amdgpu_buffer_rsrc_t rs, int off, const cplx<T> (&v)[ZB]) {
  if constexpr (sizeof(T) == 4 && ZB % 2 == 0) {
#pragma unroll
    for (int h = 0; h < ZB / 2; ++h) {
      const u32x4_t d = {__float_as_uint(v[2 * h].re), __float_as_uint(v[2 * h].im), __float_as_uint(v[2 * h + 1].re),
                         __float_as_uint(v[2 * h + 1].im)};
      __builtin_amdgcn_raw_buffer_store_b128(d, rs, off + 16 * h, 0, 0);
    }
  } else if constexpr (sizeof(T) == 4) {
#pragma unroll
    for (int q = 0; q < ZB; ++q) {
      const u32x2_t d = {__float_as_uint(v[q].re), __float_as_uint(v[q].im)};
      __builtin_amdgcn_raw_buffer_store_b64(d, rs, off + 8 * q, 0, 0);
    }
  } else {
#pragma unroll
    for (int q = 0; q < ZB; ++q) {
      const unsigned long long r = (unsigned long long)__double_as_longlong((double)v[q].re);
      const unsigned long long i = (unsigned long long)__double_as_longlong((double)v[q].im);
      const u32x4_t d = {(unsigned)r, (unsigned)(r >> 32), (unsigned)i, (unsigned)(i >> 32)};
      __builtin_amdgcn_raw_buffer_store_b128(d, rs, off + 16 * q, 0, 0);
    }
  }
}

// One sweep of a team over its orbital.  MODE 0: out = (T + V) in (stencil
// and local potential, no KB term); MODE 1: zu += z1 h0 + z2 h1 + z3 in +
// z4 (T + V) in; MODE 2: observables of in, the per-thread sums
// acc = {|psi|^2, V|psi|^2, Re conj(psi) L, Im conj(psi) G_x, _y, _z}
// (stencil sums L and G as in observe_orbital, streamed.hpp; V the potential
// the propagation uses); MODE 3: MODE 0 and MODE 2 in one sweep.  Per point
// the neighbour sums L = sum_a sum_n lap_n (psi+ + psi-) and, per axis,
// G_a = sum_n nab_n (psi+ - psi-); the H applications form the k term as
// sum_a k_a G_a in all four Taylor terms (the same arithmetic in each keeps
// the complex64 Taylor step norm-consistent).
template <typename T, int ZB, int MODE>
__device__ __forceinline__ void col_sweep(const Grid& g, bool act, int tl, cplx<T>* __restrict__ pl,
                                          const SweepIO<T>& io, const T* __restrict__ vl, double kx, double ky,
                                          double kz, double dt, double (&acc)[6]) {
  const int NLx = g.NLx, NLy = g.NLy, NLz = g.NLz, P = NLy * NLz;
  const bool col = act && tl * ZB < P;
  const int w0 = col ? tl * ZB : 0;
  const int iy = w0 / NLz, iz0 = w0 - iy * NLz;
  const int line = iy * NLz;
  constexpr int CS = col_stride<ZB>();
  const int PS = plane_slots<ZB>(P);
  // LDS slot of in-plane offset w: column w / ZB, point w % ZB
  auto slot = [](int w) { return (w / ZB) * CS + w % ZB; };
  // LDS slots of the z window outside the column (z = iz0-4 .. iz0-1 and
  // iz0+ZB .. iz0+ZB+3) and of the y neighbour runs (iy+n, then iy-n)
  int zo[8], yo[8];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    zo[j] = slot(line + wrap_once(iz0 - 4 + j, NLz));
    zo[4 + j] = slot(line + wrap_once(iz0 + ZB + j, NLz));
  }
#pragma unroll
  for (int n = 1; n <= 4; ++n) {
    yo[n - 1] = slot(wrap_once(iy + n, NLy) * NLz + iz0);
    yo[3 + n] = slot(wrap_once(iy - n, NLy) * NLz + iz0);
  }
  const int own = col ? (w0 / ZB) * CS : (P / ZB) * CS;  // threads without a column: the spare column
  // the output: stores of threads without a column go past its end
  const __amdgpu_buffer_rsrc_t ors =
      orbital_rsrc(MODE != 2 ? (const void*)io.out : (const void*)io.in, (unsigned)(g.NL * sizeof(cplx<T>)));
  // coefficients per axis (x, y, z) and n: C = lap, D = nab
  const T kt[3] = {(T)kx, (T)ky, (T)kz};
  T C[3][5], D[3][5];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int n = 1; n <= 4; ++n) {
      C[a][n] = (T)g.lap[a][n];
      D[a][n] = (T)g.nab[a][n];
    }
  const T A = (T)((kx * kx + ky * ky + kz * kz - (g.lap[0][0] + g.lap[1][0] + g.lap[2][0])) * 0.5);
  const T z1i = (T)(-dt), z2r = (T)(-dt * dt / 2.0), z3i = (T)(dt * dt * dt / 6.0),
          z4r = (T)(dt * dt * dt * dt / 24.0);
  const cplx<T>* __restrict__ in = io.in;

  cplx<T> win[kColWin][ZB];
  auto ld = [&](int X, cplx<T>(&d)[ZB]) {
    const int Xs = wrap_once(X, NLx);
    load_run<T, ZB>(in + (size_t)Xs * P + w0, d);
  };
  // per-point operands, double-buffered by plane parity
  T V[2][ZB];
  cplx<T> zz[2][ZB], a0[2][ZB], a1[2][ZB];
  auto ld_ops = [&](int X, int b) {
    const size_t pb = (size_t)X * P + w0;
#pragma unroll
    for (int q = 0; q < ZB; ++q) V[b][q] = vl[pb + q];
    if constexpr (MODE == 1) {
      load_run<T, ZB>(io.zu + pb, zz[b]);
      load_run<T, ZB>(io.h0 + pb, a0[b]);
      load_run<T, ZB>(io.h1 + pb, a1[b]);
    }
  };
  // every thread loads (threads without a column from a valid point of the
  // orbital) so that no branch separates the loads from their uses
  ld_ops(0, 0);
#pragma unroll
  for (int X = -4; X <= 4; ++X) ld(X, win[(X + kColWin) % kColWin]);
  for (int x0 = 0; x0 < NLx; x0 += kColWin) {
#pragma unroll
    for (int j = 0; j < kColWin; ++j) {
      const int x = x0 + j;
      if (x >= NLx) break;
      const size_t pb = (size_t)x * P + w0;
      const int soff = col ? (int)(pb * sizeof(cplx<T>)) : kDropOffset;
      const int cb = j & 1;  // operand buffer of plane x (x0 is even)
      ld_ops(x + 1 < NLx ? x + 1 : 0, cb ^ 1);
      ld(x + 5, win[(j + 5) % kColWin]);
      cplx<T>* pln = pl + cb * PS;
#pragma unroll
      for (int q = 0; q < ZB; ++q) pln[own + q] = win[j][q];
      lds_barrier();
      // z window: 4 values either side from LDS, the column itself from registers
      cplx<T> zw[ZB + 8];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        zw[t] = pln[zo[t]];
        zw[ZB + 4 + t] = pln[zo[4 + t]];
      }
#pragma unroll
      for (int q = 0; q < ZB; ++q) zw[4 + q] = win[j][q];
      // per point: L (vr, vi), the k term sum_a k_a G_a (wr, wi), the
      // observables' sums from L and each G_a as its axis completes
      T vr[ZB], vi[ZB], wr[ZB], wi[ZB];
      double er[ZB], ei[ZB];
#pragma unroll
      for (int q = 0; q < ZB; ++q) {
        // threads without a column add zeros to the observables
        er[q] = col ? (double)win[j][q].re : 0.0;
        ei[q] = col ? (double)win[j][q].im : 0.0;
        T gr = 0, gi = 0;
        vr[q] = vi[q] = 0;
#pragma unroll
        for (int n = 1; n <= 4; ++n) {
          const cplx<T> ep = zw[4 + q + n], em = zw[4 + q - n];
          vr[q] += C[2][n] * (ep.re + em.re);
          vi[q] += C[2][n] * (ep.im + em.im);
          gr += D[2][n] * (ep.re - em.re);
          gi += D[2][n] * (ep.im - em.im);
        }
        wr[q] = kt[2] * gr;
        wi[q] = kt[2] * gi;
        if constexpr (MODE >= 2) acc[5] += er[q] * (double)gi - ei[q] * (double)gr;
      }
      {
        T gr[ZB], gi[ZB];
#pragma unroll
        for (int q = 0; q < ZB; ++q) gr[q] = gi[q] = 0;
#pragma unroll
        for (int n = 1; n <= 4; ++n) {
#pragma unroll
          for (int q = 0; q < ZB; ++q) {
            const cplx<T> ep = pln[yo[n - 1] + q], em = pln[yo[3 + n] + q];
            vr[q] += C[1][n] * (ep.re + em.re);
            vi[q] += C[1][n] * (ep.im + em.im);
            gr[q] += D[1][n] * (ep.re - em.re);
            gi[q] += D[1][n] * (ep.im - em.im);
          }
        }
#pragma unroll
        for (int q = 0; q < ZB; ++q) {
          wr[q] += kt[1] * gr[q];
          wi[q] += kt[1] * gi[q];
          if constexpr (MODE >= 2) acc[4] += er[q] * (double)gi[q] - ei[q] * (double)gr[q];
        }
      }
      {
        T gr[ZB], gi[ZB];
#pragma unroll
        for (int q = 0; q < ZB; ++q) gr[q] = gi[q] = 0;
#pragma unroll
        for (int n = 1; n <= 4; ++n) {
#pragma unroll
          for (int q = 0; q < ZB; ++q) {
            const cplx<T> ep = win[(j + n) % kColWin][q], em = win[(j + kColWin - n) % kColWin][q];
            vr[q] += C[0][n] * (ep.re + em.re);
            vi[q] += C[0][n] * (ep.im + em.im);
            gr[q] += D[0][n] * (ep.re - em.re);
            gi[q] += D[0][n] * (ep.im - em.im);
          }
        }
#pragma unroll
        for (int q = 0; q < ZB; ++q) {
          wr[q] += kt[0] * gr[q];
          wi[q] += kt[0] * gi[q];
          if constexpr (MODE >= 2) acc[3] += er[q] * (double)gi[q] - ei[q] * (double)gr[q];
        }
      }
      if constexpr (MODE >= 2) {
#pragma unroll
        for (int q = 0; q < ZB; ++q) {
          const double a2 = er[q] * er[q] + ei[q] * ei[q];
          acc[0] += a2;
          acc[1] += (double)V[cb][q] * a2;
          acc[2] += er[q] * (double)vr[q] + ei[q] * (double)vi[q];
        }
      }
      if constexpr (MODE != 2) {
        cplx<T> f[ZB];
#pragma unroll
        for (int q = 0; q < ZB; ++q) {
          const cplx<T> e = win[j][q];
          f[q].re = (V[cb][q] + A) * e.re - (T)0.5 * vr[q] + wi[q];
          f[q].im = (V[cb][q] + A) * e.im - (T)0.5 * vi[q] - wr[q];
        }
        if constexpr (MODE == 1) {
          // fold in the order of k_b_fold: z1 h0, z2 h1, z3 h2, z4 h3
#pragma unroll
          for (int q = 0; q < ZB; ++q) {
            cplx<T> z = zz[cb][q];
            const cplx<T> e = win[j][q];
            z.re += -z1i * a0[cb][q].im;
            z.im += z1i * a0[cb][q].re;
            z.re += z2r * a1[cb][q].re;
            z.im += z2r * a1[cb][q].im;
            z.re += -z3i * e.im;
            z.im += z3i * e.re;
            z.re += z4r * f[q].re;
            z.im += z4r * f[q].im;
            f[q] = z;
          }
        }
        store_run_rsrc<T, ZB>(ors, soff, f);
      }
    }
  }
}

// KB parts of one orbital, one wave per atom (lanes over the atom's sphere
// points, every orbital value gathered once per atom); kKbPass projectors of
// the atom per pass over its sphere, kKbChunks 64-point chunks of the sphere
// with their loads in flight together (indices past the sphere clamped to its
// first point and weighted by zero, so no branch separates loads from uses).
constexpr int kKbPass = 4;
constexpr int kKbChunks = 3;

// Projections of atom a (pseudo_pt phase 1, ARTED/common/hpsi.f90:183-193):
// U(il) = Hxyz iuV(il) sum_j uV(j,il) e^{i k.r_j} E(Jxyz(j)), phases ph.
template <typename T>
__device__ __forceinline__ void kb_project_atom(const KBTables& kb, double hxyz, int a, const cplx<T>* __restrict__ E,
                                                const cplx<T>* __restrict__ ph, cplx<T>* U, int lane) {
  const int j0 = kb.atom_start[a], m = kb.atom_start[a + 1] - j0;
  const int l1 = kb.atom_lma[a + 1];
  for (int lc = kb.atom_lma[a]; lc < l1; lc += kKbPass) {
    T v[2 * kKbPass];
#pragma unroll
    for (int k = 0; k < 2 * kKbPass; ++k) v[k] = 0;
    for (int jb = 0; jb < m; jb += kKbChunks * kWave) {
      int jj[kKbChunks];
      bool ok[kKbChunks];
      cplx<T> e[kKbChunks], p[kKbChunks];
#pragma unroll
      for (int h = 0; h < kKbChunks; ++h) {
        const int j = jb + h * kWave + lane;
        ok[h] = j < m;
        jj[h] = ok[h] ? j : 0;
        e[h] = E[kb.pidx[j0 + jj[h]]];
        p[h] = ph[j0 + jj[h]];
      }
#pragma unroll
      for (int h = 0; h < kKbChunks; ++h) {
        const T wr = p[h].re * e[h].re - p[h].im * e[h].im, wi = p[h].re * e[h].im + p[h].im * e[h].re;
#pragma unroll
        for (int l = 0; l < kKbPass; ++l) {
          const T u0 = (T)kb.uv[(size_t)min(lc + l, kb.Nlma - 1) * kb.Nps + jj[h]];
          const T u = (ok[h] && lc + l < l1) ? u0 : (T)0;
          v[2 * l] += u * wr;
          v[2 * l + 1] += u * wi;
        }
      }
    }
    wave_sums<2 * kKbPass>(v);
#pragma unroll
    for (int k = 0; k < 2 * kKbPass; ++k) {
      const int il = lc + k / 2;
      if (lane == wave_sums_lane(k) && il < l1) ((T*)U)[2 * il + (k & 1)] = v[0] * (T)(hxyz * kb.lma_sign[il]);
    }
  }
}

// dst(Jxyz(j)) += fac sum_il uV(j,il) e^{-i k.r_j} U(il) over atom a's sphere
// (pseudo_pt phase 2, ARTED/common/hpsi.f90:195-206); rs is dst's buffer
// resource (stores past the sphere are dropped).  Atoms whose spheres share
// a grid point must not run at the same time.
template <typename T>
__device__ __forceinline__ void kb_scatter_atom(const KBTables& kb, int a, const cplx<T>* __restrict__ ph,
                                                const cplx<T>* U, const cplx<T>* dst, __amdgpu_buffer_rsrc_t rs, T fac,
                                                int lane) {
  const int j0 = kb.atom_start[a], m = kb.atom_start[a + 1] - j0;
  const int l0 = kb.atom_lma[a], l1 = kb.atom_lma[a + 1];
  for (int jb = 0; jb < m; jb += kKbChunks * kWave) {
    int jj[kKbChunks], pt[kKbChunks];
    bool ok[kKbChunks];
    cplx<T> p[kKbChunks], d[kKbChunks];
#pragma unroll
    for (int h = 0; h < kKbChunks; ++h) {
      const int j = jb + h * kWave + lane;
      ok[h] = j < m;
      jj[h] = ok[h] ? j : 0;
      p[h] = ph[j0 + jj[h]];
      pt[h] = kb.pidx[j0 + jj[h]];
      d[h] = dst[pt[h]];
    }
#pragma unroll
    for (int h = 0; h < kKbChunks; ++h) {
      T fr = 0, fi = 0;
      for (int il = l0; il < l1; ++il) {
        const T u = (T)kb.uv[(size_t)il * kb.Nps + jj[h]];
        const cplx<T> w = U[il];
        fr += u * (p[h].re * w.re + p[h].im * w.im);
        fi += u * (p[h].re * w.im - p[h].im * w.re);
      }
      cplx<T> r[1] = {d[h]};
      r[0].re += fac * fr;
      r[0].im += fac * fi;
      store_run_rsrc<T, 1>(rs, ok[h] ? pt[h] * (int)sizeof(cplx<T>) : kDropOffset, r);
    }
  }
}

// kb_scatter_atom for an atom whose sphere no other atom touches and which has
// at most kKbPass projectors: the updated values are then final, so the
// projections of the next Taylor term (whose input dst is) are formed from
// them in the same pass, as kb_project_atom forms them, and replace the
// atom's U(il).
template <typename T>
__device__ __forceinline__ void kb_scatter_project_atom(const KBTables& kb, double hxyz, int a,
                                                        const cplx<T>* __restrict__ ph, cplx<T>* U, const cplx<T>* dst,
                                                        __amdgpu_buffer_rsrc_t rs, T fac, int lane) {
  const int j0 = kb.atom_start[a], m = kb.atom_start[a + 1] - j0;
  const int l0 = kb.atom_lma[a], l1 = kb.atom_lma[a + 1];
  T v[2 * kKbPass];
#pragma unroll
  for (int k = 0; k < 2 * kKbPass; ++k) v[k] = 0;
  for (int jb = 0; jb < m; jb += kKbChunks * kWave) {
    int jj[kKbChunks], pt[kKbChunks];
    bool ok[kKbChunks];
    cplx<T> p[kKbChunks], d[kKbChunks];
#pragma unroll
    for (int h = 0; h < kKbChunks; ++h) {
      const int j = jb + h * kWave + lane;
      ok[h] = j < m;
      jj[h] = ok[h] ? j : 0;
      p[h] = ph[j0 + jj[h]];
      pt[h] = kb.pidx[j0 + jj[h]];
      d[h] = dst[pt[h]];
    }
#pragma unroll
    for (int h = 0; h < kKbChunks; ++h) {
      T u[kKbPass];
      T fr = 0, fi = 0;
#pragma unroll
      for (int l = 0; l < kKbPass; ++l) {
        const int il = min(l0 + l, kb.Nlma - 1);
        const T u0 = (T)kb.uv[(size_t)il * kb.Nps + jj[h]];
        u[l] = l0 + l < l1 ? u0 : (T)0;
        const cplx<T> w = U[il];
        fr += u[l] * (p[h].re * w.re + p[h].im * w.im);
        fi += u[l] * (p[h].re * w.im - p[h].im * w.re);
      }
      cplx<T> r[1] = {d[h]};
      r[0].re += fac * fr;
      r[0].im += fac * fi;
      store_run_rsrc<T, 1>(rs, ok[h] ? pt[h] * (int)sizeof(cplx<T>) : kDropOffset, r);
      const T wr = p[h].re * r[0].re - p[h].im * r[0].im, wi = p[h].re * r[0].im + p[h].im * r[0].re;
#pragma unroll
      for (int l = 0; l < kKbPass; ++l) {
        const T um = ok[h] ? u[l] : (T)0;
        v[2 * l] += um * wr;
        v[2 * l + 1] += um * wi;
      }
    }
  }
  wave_sums<2 * kKbPass>(v);
#pragma unroll
  for (int k = 0; k < 2 * kKbPass; ++k) {
    const int il = l0 + k / 2;
    if (lane == wave_sums_lane(k) && il < l1) ((T*)U)[2 * il + (k & 1)] = v[0] * (T)(hxyz * kb.lma_sign[il]);
  }
}

// Nonlocal observables of atom a at the phases ph (current.f90:172-273 and
// total_energy.f90:191-199, as in observe_orbital): adds Jx, Jy, Jz, Enl.
// PROJ: from the same gathered values also the projections at the phases
// pp into U, as kb_project_atom forms them.
template <typename T, bool PROJ = false>
__device__ __forceinline__ void kb_observe_atom(const KBTables& kb, const Grid& g, double occ, int a,
                                                const cplx<T>* __restrict__ E, const cplx<T>* __restrict__ ph,
                                                double (&nl4)[4], int lane, const cplx<T>* __restrict__ pp = nullptr,
                                                cplx<T>* U = nullptr) {
  const int j0 = kb.atom_start[a], m = kb.atom_start[a + 1] - j0;
  const int l1 = kb.atom_lma[a + 1];
  for (int lc = kb.atom_lma[a]; lc < l1; lc += kKbPass) {
    // per projector: sum u t and sum r u t (t = e^{ik.r} psi), real and imaginary parts
    double v[8 * kKbPass];
    T vp[2 * kKbPass];
#pragma unroll
    for (int k = 0; k < 8 * kKbPass; ++k) v[k] = 0;
#pragma unroll
    for (int k = 0; k < 2 * kKbPass; ++k) vp[k] = 0;
    for (int jb = 0; jb < m; jb += kKbChunks * kWave) {
      int jj[kKbChunks];
      bool ok[kKbChunks];
      cplx<T> e[kKbChunks];
#pragma unroll
      for (int h = 0; h < kKbChunks; ++h) {
        const int j = jb + h * kWave + lane;
        ok[h] = j < m;
        jj[h] = ok[h] ? j : 0;
        e[h] = E[kb.pidx[j0 + jj[h]]];
      }
#pragma unroll
      for (int h = 0; h < kKbChunks; ++h) {
        const int jp = j0 + jj[h];
        const double px = kb.pos[3 * jp], py = kb.pos[3 * jp + 1], pz = kb.pos[3 * jp + 2];
        const double c = (double)ph[jp].re, sn = (double)ph[jp].im;
        const double tr0 = c * (double)e[h].re - sn * (double)e[h].im;
        const double ti0 = c * (double)e[h].im + sn * (double)e[h].re;
        T pwr = 0, pwi = 0;
        if constexpr (PROJ) {
          const cplx<T> q = pp[jp];
          pwr = q.re * e[h].re - q.im * e[h].im;
          pwi = q.re * e[h].im + q.im * e[h].re;
        }
#pragma unroll
        for (int l = 0; l < kKbPass; ++l) {
          const double u0 = kb.uv[(size_t)min(lc + l, kb.Nlma - 1) * kb.Nps + jj[h]];
          const double u = (ok[h] && lc + l < l1) ? u0 : 0.0;
          if constexpr (PROJ) {
            const T ut = (ok[h] && lc + l < l1) ? (T)u0 : (T)0;
            vp[2 * l] += ut * pwr;
            vp[2 * l + 1] += ut * pwi;
          }
          const double tr = u * tr0, ti = u * ti0;
          double* q = v + 8 * l;
          q[0] += tr;
          q[1] += ti;
          q[2] += px * tr;
          q[3] += px * ti;
          q[4] += py * tr;
          q[5] += py * ti;
          q[6] += pz * tr;
          q[7] += pz * ti;
        }
      }
    }
    if constexpr (PROJ) {
      wave_sums<2 * kKbPass>(vp);
#pragma unroll
      for (int k = 0; k < 2 * kKbPass; ++k) {
        const int il = lc + k / 2;
        if (lane == wave_sums_lane(k) && il < l1) ((T*)U)[2 * il + (k & 1)] = vp[0] * (T)(g.Hxyz * kb.lma_sign[il]);
      }
    }
    wave_sums<8 * kKbPass>(v);
    for (int l = 0; l < kKbPass && lc + l < l1; ++l) {
      const int il = lc + l;
      double q[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) q[k] = lane_read(v[0], wave_sums_lane(8 * l + k));
      const double sg = kb.lma_sign[il];
      const double Ur = q[0] * g.Hxyz * sg, Ui = q[1] * g.Hxyz * sg;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const double Xr = q[2 + 2 * d] * g.Hxyz, Xi = q[3 + 2 * d] * g.Hxyz;
        nl4[d] += occ / g.aLxyz * 2.0 * (Xr * Ui - Xi * Ur);
      }
      const double hr = q[0] * g.Hxyz, hi = q[1] * g.Hxyz;
      nl4[3] += occ * (hr * hr + hi * hi) * sg;
    }
  }
}

// Whole RT step of the orbitals o0 .. o0 + teams - 1 of system y.  The
// observables of the orbitals as they enter (those of the previous step, at
// k + A(t_n) for the launch of step n) are formed in the step's first sweep,
// which reads those orbitals anyway, and written to the rows; the first step
// of a run has none.  st.obs_only: only those observables (the launch that
// closes a run, after its last step).
template <typename T, int ZB>
__device__ __forceinline__ void sweep_item(const Sys<T>& y, const Step& st, int s, int o0, unsigned char* lds) {
  const Grid& g = y.g;
  const KBTables& kb = y.kb;
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
  const int tw = y.sw_tw, team = wv / tw, wt = wv - team * tw;
  const int tl = (int)threadIdx.x - team * tw * kWave;
  const int lane = (int)threadIdx.x & (kWave - 1);
  const int o = o0 + team;
  const bool act = o < y.NKB;
  const int PS = plane_slots<ZB>(g.NLy * g.NLz);
  const int nlma = kb.Nlma;
  double* red = (double*)(lds + 16);
  cplx<T>* pl = (cplx<T>*)(lds + kSweepHead) + (size_t)team * (2 * PS + nlma);
  cplx<T>* U = pl + 2 * PS;
  const size_t NL = (size_t)g.NL;
  const size_t ob = (size_t)(act ? o : 0) * NL;
  const int ik = (act ? o : 0) / y.NB;
  const int n = *st.step;
  const bool obs = st.obs_only || n > 0;
  // the observables' k: kAc0 + A(t_n)
  const double* acs = st.ac + (size_t)s * st.ac_stride;
  double kox, koy, koz;
  orbital_k(y.kac0, y.NK, ik, acs[3 * n], acs[3 * n + 1], acs[3 * n + 2], kox, koy, koz);
  const cplx<T>* phn = y.phn + (size_t)ik * kb.NPI;
  const double occ = act ? y.occ[o] : 0.0;
  // the observables' per-wave sums go to the reduction area as soon as they
  // are formed (not kept in registers through the step)
  auto flush_acc = [&](double (&acc)[6]) {
#pragma unroll
    for (int i = 0; i < 6; ++i) acc[i] = wave_sum(acc[i]);
    if (lane == 0) {
#pragma unroll
      for (int i = 0; i < 6; ++i) red[wv * kRedStride + i] = acc[i];
    }
  };
  const cplx<T>* phk = y.ph + (size_t)ik * kb.NPI;
  if (obs) {
    // nonlocal observables of the entering orbitals (the fourth sweep changes
    // them); in a step launch the same gathers give the first Taylor term's
    // KB projections
    double nl4[4] = {0, 0, 0, 0};  // nonlocal Jx, Jy, Jz, Enl of this wave's atoms
    if (act) {
      if (st.obs_only) {
        for (int a = wt; a < kb.NI; a += tw) kb_observe_atom<T>(kb, g, occ, a, y.zu + ob, phn, nl4, lane);
      } else {
        for (int a = wt; a < kb.NI; a += tw)
          kb_observe_atom<T, true>(kb, g, occ, a, y.zu + ob, phn, nl4, lane, phk, U);
      }
    }
    if (lane == 0) {
#pragma unroll
      for (int i = 0; i < 4; ++i) red[wv * kRedStride + 6 + i] = nl4[i];
    }
  }
  if (st.obs_only) {
    const SweepIO<T> io{y.zu + ob, nullptr, nullptr, nullptr, nullptr};
    double acc[6] = {0, 0, 0, 0, 0, 0};
    col_sweep<T, ZB, 2>(g, act, tl, pl, io, y.vloc, kox, koy, koz, st.dt, acc);
    flush_acc(acc);
  } else {
    double am[3], an[3];
    step_fields(st, s, am, an);
    double kx, ky, kz;
    orbital_k(y.kac0, y.NK, ik, am[0], am[1], am[2], kx, ky, kz);
    // no two atom spheres share a grid point and no atom has more than kKbPass
    // projectors: the back-projection of term n forms the projections of term
    // n+1 (kb_scatter_project_atom), so only term 0 projects on its own
    const bool fuse = y.ncolor == 1 && kb.max_lpa <= kKbPass;
    for (int term = 0; term < 4; ++term) {
      const cplx<T>* in = (term == 0 ? y.zu : y.h[term - 1]) + ob;
      // term 0: projected with the observables above when there are any
      if (act && ((term == 0 && !obs) || (term > 0 && !fuse)))
        for (int a = wt; a < kb.NI; a += tw) kb_project_atom<T>(kb, g.Hxyz, a, in, phk, U, lane);
      __syncthreads();
      SweepIO<T> io{in, (term < 3 ? y.h[term] : y.zu) + ob, y.zu + ob, y.h[0] + ob, y.h[1] + ob};
      double acc[6] = {0, 0, 0, 0, 0, 0};
      if (term == 0 && obs) {
        col_sweep<T, ZB, 3>(g, act, tl, pl, io, y.vloc, kx, ky, kz, st.dt, acc);
        flush_acc(acc);
      } else if (term < 3) {
        col_sweep<T, ZB, 0>(g, act, tl, pl, io, y.vloc, kx, ky, kz, st.dt, acc);
      } else {
        col_sweep<T, ZB, 1>(g, act, tl, pl, io, y.vloc, kx, ky, kz, st.dt, acc);
      }
      __syncthreads();
      // KB term: the atoms colour by colour (no two atoms of a colour share a
      // grid point), the atoms of a colour over the team's waves
      const T fac = term < 3 ? (T)1 : (T)(st.dt * st.dt * st.dt * st.dt / 24.0);
      const __amdgpu_buffer_rsrc_t rs = orbital_rsrc(io.out, (unsigned)(NL * sizeof(cplx<T>)));
      for (int c = 0; c < y.ncolor; ++c) {
        if (act) {
          if (fuse && term < 3) {
            for (int k = y.kb_cstart[c] + wt; k < y.kb_cstart[c + 1]; k += tw)
              kb_scatter_project_atom<T>(kb, g.Hxyz, y.kb_order[k], phk, U, io.out, rs, fac, lane);
          } else {
            for (int k = y.kb_cstart[c] + wt; k < y.kb_cstart[c + 1]; k += tw)
              kb_scatter_atom<T>(kb, y.kb_order[k], phk, U, io.out, rs, fac, lane);
          }
        }
        __syncthreads();
      }
    }
  }
  if (!obs) return;  // (the step's last barrier is behind)
  __syncthreads();
  if (act && tl == 0) {
    double v[kRedStride];
#pragma unroll
    for (int i = 0; i < kRedStride; ++i) v[i] = red[(team * tw) * kRedStride + i];
    for (int w = 1; w < tw; ++w)
#pragma unroll
      for (int i = 0; i < kRedStride; ++i) v[i] += red[(team * tw + w) * kRedStride + i];
    const double kk[3] = {kox, koy, koz};
    const double A0 = -(g.lap[0][0] + g.lap[1][0] + g.lap[2][0]) * 0.5;
    const double te = -0.5 * v[2];
    const double sc = g.Hxyz / g.aLxyz;
    double* r = y.rows + (size_t)o * 8;
    for (int d = 0; d < 3; ++d) r[d] = (occ * kk[d] * v[0] + v[3 + d] * occ) * sc + v[6 + d];
    r[3] = occ * (A0 * v[0] + te + kox * v[3] + koy * v[4] + koz * v[5]) * g.Hxyz +
           occ * (kox * kox + koy * koy + koz * koz) / 2.0;
    r[4] = v[9];
    r[5] = occ * v[1] * g.Hxyz;
    r[6] = occ * v[0] * g.Hxyz;
    r[7] = 0.0;
  }
  __syncthreads();
}

// Work items: (system, first orbital), costliest first; *next is zero at launch.
template <typename T, int ZBMAX>
__global__ void __launch_bounds__(sweep_waves<T>() * kWave) __attribute__((amdgpu_waves_per_eu(2))) k_b_sweep(const Sys<T>* __restrict__ S, Step st,
                                                           const int2* __restrict__ items, int nitems,
                                                           int* __restrict__ next) {
  extern __shared__ __align__(16) unsigned char lds[];
  int* slot = (int*)lds;
  for (;;) {
    if (threadIdx.x == 0) *slot = atomicAdd(next, 1);
    __syncthreads();
    const int it = __builtin_amdgcn_readfirstlane(*slot);
    __syncthreads();
    if (it >= nitems) break;
    const int2 wk = items[it];
    const int s = __builtin_amdgcn_readfirstlane(wk.x), o0 = __builtin_amdgcn_readfirstlane(wk.y);
    const Sys<T>& y = S[s];
    switch (y.sw_zb) {
      case 1: sweep_item<T, 1>(y, st, s, o0, lds); break;
      case 2: sweep_item<T, 2>(y, st, s, o0, lds); break;
      case 3: if constexpr (ZBMAX >= 3) sweep_item<T, 3>(y, st, s, o0, lds); break;
      default: if constexpr (ZBMAX >= 4) sweep_item<T, 4>(y, st, s, o0, lds); break;
    }
  }
}

__global__ void k_b_qreset(int* next) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *next = 0;
}

}  // namespace batch
}  // namespace salmon
