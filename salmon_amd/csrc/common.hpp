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

// Wave-wide sum with DPP lane moves (VALU only; no LDS round trips):
// rotations by 8, 4, 2, 1 inside each 16-lane row leave every lane with its
// row sum; row_bcast:15 and row_bcast:31 then fold rows 0..3 into lane 63,
// whose value is broadcast with readlane.  The summation order is fixed.
namespace dpp {
constexpr int kRowRor1 = 0x121, kRowRor2 = 0x122, kRowRor4 = 0x124, kRowRor8 = 0x128;
constexpr int kRowBcast15 = 0x142, kRowBcast31 = 0x143;
}  // namespace dpp

template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_mov(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWS, 0xf, false));
}

template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_mov(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, ROWS, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROWS, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ float lane_read(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

__device__ __forceinline__ double lane_read(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// requires all 64 lanes active
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
  v += dpp_mov<dpp::kRowRor8, 0xf>(v);
  v += dpp_mov<dpp::kRowRor4, 0xf>(v);
  v += dpp_mov<dpp::kRowRor2, 0xf>(v);
  v += dpp_mov<dpp::kRowRor1, 0xf>(v);
  v += dpp_mov<dpp::kRowBcast15, 0xa>(v);
  v += dpp_mov<dpp::kRowBcast31, 0xc>(v);
  return lane_read(v, 63);
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
