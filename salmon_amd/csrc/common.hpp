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
  const int* atom_lma;    // (NI+1) first projector of each atom (projectors grouped by atom)
  int max_lpa;            // most projectors of one atom
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

// Wave sums of N values at once (N a power of two, 2 <= N <= 64), all 64
// lanes active.  Stage s (s = 0..5) pairs the values across lane bit 5 - s:
// the lanes with that bit 0 keep the pair sum of the even value of each pair,
// the lanes with it 1 that of the odd value, so each stage halves the number
// of values and the work per value.  Stages 0 and 1 move data with the
// permlane32/permlane16 swaps (halves and odd/even 16-lane rows), stages 2..5
// with DPP row_ror:8, row_half_mirror and quad_perm [2,3,0,1] / [1,0,3,2];
// once one value is left, the remaining stages are plain pair sums.  The
// total of value k ends in v[0] of the lanes whose bits 5, 4, ... equal bits
// 0, 1, ... of k (wave_sums_lane(k)).  The summation order is fixed.
namespace dpp {
constexpr int kRowHalfMirror = 0x141, kQuad2301 = 0x4e, kQuad1032 = 0xb1;
}  // namespace dpp

__device__ __forceinline__ void perm_swap(int stage, float& x, float& y) {
  const auto r = stage == 0 ? __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(y), false, false)
                            : __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(y), false, false);
  x = __uint_as_float(r[0]);
  y = __uint_as_float(r[1]);
}

__device__ __forceinline__ void perm_swap(int stage, double& x, double& y) {
  const unsigned long long bx = (unsigned long long)__double_as_longlong(x);
  const unsigned long long by = (unsigned long long)__double_as_longlong(y);
  const auto lo = stage == 0 ? __builtin_amdgcn_permlane32_swap((unsigned)bx, (unsigned)by, false, false)
                             : __builtin_amdgcn_permlane16_swap((unsigned)bx, (unsigned)by, false, false);
  const auto hi = stage == 0 ? __builtin_amdgcn_permlane32_swap((unsigned)(bx >> 32), (unsigned)(by >> 32), false, false)
                             : __builtin_amdgcn_permlane16_swap((unsigned)(bx >> 32), (unsigned)(by >> 32), false, false);
  x = __longlong_as_double((long long)(((unsigned long long)hi[0] << 32) | lo[0]));
  y = __longlong_as_double((long long)(((unsigned long long)hi[1] << 32) | lo[1]));
}

template <int CTRL, typename T>
__device__ __forceinline__ T dpp_pair(T p, T q, bool odd) {
  const T t = odd ? p : q, u = odd ? q : p;
  return u + dpp_mov<CTRL, 0xf>(t);
}

template <int S, typename T>
__device__ __forceinline__ T dpp_stage(T p, T q, bool odd) {
  if constexpr (S == 2) return dpp_pair<dpp::kRowRor8>(p, q, odd);
  if constexpr (S == 3) return dpp_pair<dpp::kRowHalfMirror>(p, q, odd);
  if constexpr (S == 4) return dpp_pair<dpp::kQuad2301>(p, q, odd);
  return dpp_pair<dpp::kQuad1032>(p, q, odd);
}

template <int S, int NV, int N, typename T>
__device__ __forceinline__ void wave_sums_stage(T (&v)[N], int lane) {
  if constexpr (S < 6) {
    if constexpr (NV > 1) {
#pragma unroll
      for (int i = 0; i < NV / 2; ++i) {
        T x = v[2 * i], y = v[2 * i + 1];
        if constexpr (S < 2) {
          perm_swap(S, x, y);
          v[i] = x + y;
        } else {
          v[i] = dpp_stage<S>(x, y, (lane >> (5 - S)) & 1);
        }
      }
    } else {
      T x = v[0], y = v[0];
      if constexpr (S < 2) {
        perm_swap(S, x, y);
        v[0] = x + y;
      } else {
        v[0] = dpp_stage<S>(x, x, false);
      }
    }
    wave_sums_stage<S + 1, (NV > 1 ? NV / 2 : 1)>(v, lane);
  }
}

template <int N, typename T>
__device__ __forceinline__ void wave_sums(T (&v)[N]) {
  static_assert(N >= 2 && N <= 64 && (N & (N - 1)) == 0, "N a power of two");
  wave_sums_stage<0, N>(v, (int)(threadIdx.x & (kWave - 1)));
}

__host__ __device__ constexpr int wave_sums_lane(int k) {
  return ((k & 1) << 5) | ((k & 2) << 3) | ((k & 4) << 1) | ((k & 8) >> 1) | ((k & 16) >> 3) | ((k & 32) >> 5);
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

// Workgroup barrier for exchanges through LDS only: it waits for this wave's
// LDS operations but not for its outstanding global loads and stores, which
// __syncthreads also waits for.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ __forceinline__ void sincos_t(double x, double* s, double* c) { sincos(x, s, c); }
__device__ __forceinline__ void sincos_t(float x, float* s, float* c) { sincosf(x, s, c); }

}  // namespace salmon
