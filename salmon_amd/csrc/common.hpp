// Shared device types and helpers for the gfx950 ARTED kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace salmon {

// Interleaved complex (the Fortran complex(8)/complex(4) layout).  Aligned to
// its size so that a lane moves one value with a single 16-byte (complex128)
// or 8-byte (complex64) load.
template <typename T>
struct alignas(2 * sizeof(T)) cplx {
  T re, im;
};

// Grid constants, passed by value to every kernel.
struct Grid {
  int NLx, NLy, NLz, NL;
  double Hx, Hy, Hz, Hxyz;
  double aLx, aLy, aLz, aLxyz;
  double lap[3][5];  // lap(0:4)/H^2 per axis (x,y,z)
  double nab[3][5];  // nab(0:4)/H per axis, nab[.][0] unused
};

// Kleinman-Bylander tables in device memory.  "Packed" indices enumerate the
// sphere points of all atoms one atom after another (the reference's NPI
// vector, opt_variables.f90:233-265).
struct KBTables {
  int NI, Nps, Nlma, NPI, NT, NE;
  const int* atom_start;  // (NI+1) packed offset of each atom's sphere
  const int* pidx;        // (NPI) 0-based grid index of each packed point
  const double* pos;      // (NPI*3) unwrapped coordinates Lx*Hx - Jxx*aLx, ...
  const int* lma_atom;    // (Nlma) 0-based atom of each projector
  const double* lma_sign; // (Nlma) iuV as a real
  const double* uv;       // (Nlma*Nps) uV(j,ilma) at [ilma*Nps + j]
  const int* pt2t;        // (NL) touched-point number of a grid point, or -1
  const int* t_start;     // (NT+1) first entry of each touched point
  const int* e_lma;       // (NE) projector of each entry
  const int* e_jp;        // (NE) packed point of each entry
};

constexpr int kWave = 64;

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Deterministic sum of NV values over a 256-thread block; the result is
// valid in thread 0.  smem holds 4*NV doubles.
template <int NV>
__device__ __forceinline__ void block_sum256(double (&v)[NV], double* smem) {
#pragma unroll
  for (int i = 0; i < NV; ++i) v[i] = wave_sum(v[i]);
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) smem[w * NV + i] = v[i];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] = (smem[i] + smem[NV + i]) + (smem[2 * NV + i] + smem[3 * NV + i]);
  }
  __syncthreads();
}

__device__ __forceinline__ void sincos_t(double x, double* s, double* c) { sincos(x, s, c); }
__device__ __forceinline__ void sincos_t(float x, float* s, float* c) { sincosf(x, s, c); }

}  // namespace salmon
