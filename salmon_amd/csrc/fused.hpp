// LDS-resident RT step for grids whose orbital fits in LDS (16^3).
//
// One workgroup owns the bands [b0, b1) of one k point.  For each orbital it
// reads psi once from HBM into LDS, applies H four times out of LDS (stencil
// + local + KB), accumulates the Taylor-4 sum in registers
// (ARTED/RT/hamiltonian.f90:46-66,99-122), writes psi(t+dt) once, and then --
// with psi(t+dt) still in LDS -- evaluates the observables at k+A(t+dt):
// density (psi_rho.f90:130-144), current (RT/current.f90:172-273), kinetic
// and nonlocal energy (total_energy.f90:172-203) and the norm, all summed in
// FP64.  Per-workgroup sums and density partials are reduced afterwards in a
// fixed order, so results do not depend on scheduling.
//
// Thread layout: a thread owns PX = NX/XS consecutive x points of one (y,z)
// line; z is the fastest thread index, so every LDS and HBM access of a wave is
// to consecutive complex values.  y and z neighbours come from LDS, x
// neighbours from a register window (halo of 4 on each side, periodic).
// Two LDS buffers alternate between input and output of each H application.
// KB phases exp(i k.(r - T)) are built from per-axis tables
// exp(i k_a (l H_a - j L_a)), l = 0..N_a-1, j = -2..2, computed once per
// workgroup for the midpoint k (propagation) and for k+A(t+dt) (observables).
#pragma once

namespace salmon {

constexpr int kFusedMaxNPI = 1152;      // sphere points per cell (LDS correction array)
constexpr int kFusedMaxAtoms = 64;      // atoms per cell (LDS atom table)
constexpr int kFusedSpinMax = 1 << 14;  // bound of one KB handover wait (s_sleep 1 each, ~0.5 ms)
constexpr int kFusedMaxLmaPerAtom = 4;  // projectors per atom
constexpr int kLA = kFusedMaxLmaPerAtom;

template <typename T>
struct FusedArgs {
  cplx<T>* zu;
  const double* kac0;
  const double* occ;
  const T* vloc;
  const double* vloc64;
  int NK, NB, bpw, nbg;
  int prop;              // 1: Taylor step at k0 + am
  int obs;               // 1: observables of the input orbitals at k0 + an
  int rho_post;          // 1: the density of psi(t+dt), summed where it is stored
                         // (the observables sweep then leaves the density alone)
  int rho;               // 1: the observables sweep sums the density of its orbitals
                         // (a frozen-potential run needs only its last step's)
  double am[3], an[3];
  T z1i, z2r, z3i, z4r;
  T lapc[3][5];          // -lap(n)/2 per axis (propagation) -- uniform, kept in SGPRs
  T nab[3][5];           // nab(n) per axis
  // KB tables
  const int* pt_prim;    // (NL) first sphere point covering a grid point, or -1
  const int* ov_start;   // (NOV+1) grid points covered by several sphere points:
  const int* ov_jp;      //   entries [ov_start[o]+1, ov_start[o+1]) add into ov_jp[ov_start[o]]
  int NOV;
  const int* jp_atom;    // (NPI)
  const int* jp_j;       // (NPI) index within its atom's sphere
  const int* jp_grid;    // (NPI) grid index
  const int* jp_phase;   // (NPI) px | py << 10 | pz << 20 into the phase tables
  const double* jp_pos;  // (NPI*3) unwrapped coordinates
  const int* atom_start; // (NI+1)
  // projectors padded to kFusedMaxLmaPerAtom per atom (zero rows, zero sign
  // for an atom with fewer): row (at * kFusedMaxLmaPerAtom + l)
  const double* sgnp;    // (NI*kFusedMaxLmaPerAtom) iuV
  const T* uvp;          // (NI*kFusedMaxLmaPerAtom*Nps) uV(j, l) of atom at
  const double* uvp64;   // the same in FP64
  int NI, Nps, NPI;
  Grid g;
  double* wg_rows;       // (nwg*8)
  double* rho_part;      // (nwg*NL)
  double* res;           // the rows' sums (8), formed by the workgroup that finishes last; or null
  unsigned* cnt;         // workgroups finished (zero between launches)
};

template <typename T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return cplx<T>{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

template <typename T, int NX, int NY, int NZ>
__device__ __forceinline__ cplx<T> phase_at(const cplx<T>* ph, int code) {
  const int px = code & 1023, py = (code >> 10) & 1023, pz = (code >> 20) & 1023;
  return cmul(cmul(ph[px], ph[5 * NX + py]), ph[5 * NX + 5 * NY + pz]);
}

// Returns p unchanged but opaque to the optimiser, so loads through it are
// not hoisted out of the orbital and H loops (hoisted table values would stay
// live in registers for the whole kernel).
template <bool B>
struct BoolC {
  static constexpr bool value = B;
};

template <typename P>
__device__ __forceinline__ P* opaque(P* p) {
  asm volatile("" : "+s"(p));
  return p;
}

// x neighbour o (own-point offset, -4..PX+3) from the register window; o is a
// compile-time constant after unrolling, so the window stays in registers
template <typename T, int NX, int H, int XS, int W>
__device__ __forceinline__ cplx<T> win_at(const cplx<T> (&win)[W], int o) {
  if constexpr (XS > 1) {
    return win[H + o];
  } else {
    return win[(o + NX) % NX];
  }
}

// Neighbour sums of the stencil at own point i, added to lr/li and g:
//   l += sum_a sum_n lapc_a(n) (psi+ + psi-),   g_a += sum_n nab_a(n) (psi+ - psi-)
// (hpsi1_RT_stencil, stencil/F90/hpsi.f90:190, before the k factors); zb/yb
// are the LDS bases of the z and y neighbours (+1..+4, -1..-4), x neighbours
// come from the register window.
template <typename T, int NX, int NYZ, int H, int XS, int W>
__device__ __forceinline__ void nb_sums(const FusedArgs<T>& a, const cplx<T>* const (&zb)[8],
                                        const cplx<T>* const (&yb)[8], const cplx<T> (&win)[W], int i, T& lr,
                                        T& li, T (&gr)[3], T (&gi)[3]) {
#pragma unroll
  for (int nn = 1; nn <= 4; ++nn) {
    const cplx<T> zp = zb[nn - 1][i * NYZ];
    const cplx<T> zm = zb[nn + 3][i * NYZ];
    lr += a.lapc[2][nn] * (zp.re + zm.re);
    li += a.lapc[2][nn] * (zp.im + zm.im);
    gr[2] += a.nab[2][nn] * (zp.re - zm.re);
    gi[2] += a.nab[2][nn] * (zp.im - zm.im);
    const cplx<T> yp = yb[nn - 1][i * NYZ];
    const cplx<T> ym = yb[nn + 3][i * NYZ];
    lr += a.lapc[1][nn] * (yp.re + ym.re);
    li += a.lapc[1][nn] * (yp.im + ym.im);
    gr[1] += a.nab[1][nn] * (yp.re - ym.re);
    gi[1] += a.nab[1][nn] * (yp.im - ym.im);
    const cplx<T> xp = win_at<T, NX, H, XS, W>(win, i + nn);
    const cplx<T> xm = win_at<T, NX, H, XS, W>(win, i - nn);
    lr += a.lapc[0][nn] * (xp.re + xm.re);
    li += a.lapc[0][nn] * (xp.im + xm.im);
    gr[0] += a.nab[0][nn] * (xp.re - xm.re);
    gi[0] += a.nab[0][nn] * (xp.im - xm.im);
  }
}

// The same sums with the gradient term applied at once, as the reference's
// stencil does with D = kAc nab (common/hpsi.f90:122-124, stencil/F90/hpsi.f90:190):
//   h.re += sum lapc (psi+ + psi-).re + D (psi+ - psi-).im
//   h.im += sum lapc (psi+ + psi-).im - D (psi+ - psi-).re
template <typename T, int NX, int NYZ, int H, int XS, int W>
__device__ __forceinline__ void nb_sums_k(const FusedArgs<T>& a, const T (&D)[3][4], const cplx<T>* const (&zb)[8],
                                          const cplx<T>* const (&yb)[8], const cplx<T> (&win)[W], int i, T& lr,
                                          T& li) {
#pragma unroll
  for (int nn = 1; nn <= 4; ++nn) {
    const cplx<T> zp = zb[nn - 1][i * NYZ];
    const cplx<T> zm = zb[nn + 3][i * NYZ];
    lr += a.lapc[2][nn] * (zp.re + zm.re);
    li += a.lapc[2][nn] * (zp.im + zm.im);
    lr += D[2][nn - 1] * (zp.im - zm.im);
    li -= D[2][nn - 1] * (zp.re - zm.re);
    const cplx<T> yp = yb[nn - 1][i * NYZ];
    const cplx<T> ym = yb[nn + 3][i * NYZ];
    lr += a.lapc[1][nn] * (yp.re + ym.re);
    li += a.lapc[1][nn] * (yp.im + ym.im);
    lr += D[1][nn - 1] * (yp.im - ym.im);
    li -= D[1][nn - 1] * (yp.re - ym.re);
    const cplx<T> xp = win_at<T, NX, H, XS, W>(win, i + nn);
    const cplx<T> xm = win_at<T, NX, H, XS, W>(win, i - nn);
    lr += a.lapc[0][nn] * (xp.re + xm.re);
    li += a.lapc[0][nn] * (xp.im + xm.im);
    lr += D[0][nn - 1] * (xp.im - xm.im);
    li -= D[0][nn - 1] * (xp.re - xm.re);
  }
}

// LDS bases of the z and y neighbours of own point 0 (x0, iy, iz) in buf
template <typename T, int NY, int NZ>
__device__ __forceinline__ void nb_bases(const cplx<T>* buf, int x0, int iy, int iz, const cplx<T>* (&zb)[8],
                                         const cplx<T>* (&yb)[8]) {
#pragma unroll
  for (int nn = 1; nn <= 4; ++nn) {
    const cplx<T>* base = buf + x0 * NY * NZ;
    zb[nn - 1] = base + iy * NZ + (iz + nn) % NZ;
    zb[nn + 3] = base + iy * NZ + (iz - nn + NZ) % NZ;
    yb[nn - 1] = base + ((iy + nn) % NY) * NZ + iz;
    yb[nn + 3] = base + ((iy - nn + NY) % NY) * NZ + iz;
  }
}

// The complex64 kernel with one workgroup per CU (XS >= 2) has the LDS to
// hold per-workgroup sphere-point tables (phases at both k, grid index,
// projector values) and the density accumulator; the others read the tables
// from global memory and accumulate the density there.
template <typename T, int XS>
__host__ __device__ constexpr bool fused_lds_tables() {
  return sizeof(T) == 4 && XS >= 2;
}

// The complex64 kernel evaluates the observables of its input orbitals inside
// its first H application; the FP64 kernel, short of registers for that,
// evaluates those of psi(t+dt) after the step.
#ifndef SALMON_FUSED_DEFER
#define SALMON_FUSED_DEFER 1
#endif
template <typename T, int XS>
__host__ __device__ constexpr bool fused_defer() {
  return SALMON_FUSED_DEFER && sizeof(T) == 4;
}

// RP: the density of psi(t+dt) is summed where psi(t+dt) is stored (a.rho_post
// set by the host); a separate instantiation, so the plain step keeps its
// register budget.
template <typename T, int NX, int NY, int NZ, int XS, int WPE, bool OV, bool RP>
__global__ void __launch_bounds__(NY * NZ * XS, WPE) k_fused_step(FusedArgs<T> a) {
  constexpr int NYZ = NY * NZ, NT = NYZ * XS, NL = NX * NYZ;
  constexpr int PX = NX / XS;                 // own x points per thread
  constexpr int H = (XS > 1) ? 4 : 0;         // register halo on each side
  constexpr int W = (XS > 1) ? PX + 8 : NX;   // register window
  constexpr int NPH = 5 * (NX + NY + NZ);
  constexpr int NW = (NT + kWave - 1) / kWave;
  static_assert(NX % XS == 0, "x split");
  static_assert(XS == 1 || PX >= 4, "halo of 4 needs PX >= 4");

  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T>* bufA = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* bufB = bufA + NL;
  cplx<T>* phm = bufB + NL;      // phase tables at the midpoint k
  cplx<T>* phn = phm + NPH;      // phase tables at k + A(t+dt)
  cplx<T>* dps = phn + NPH;      // KB term at each sphere point of the current H
  constexpr bool LT = fused_lds_tables<T, XS>();
  // observables: of the input orbitals during the first H application
  // (DEFER), or of psi(t+dt) after the step
  constexpr bool DEFER = fused_defer<T, XS>();
  constexpr int MT = LT ? kFusedMaxNPI : 0;
  cplx<T>* em = dps + kFusedMaxNPI;  // e^{i km.(r_j - T_j)} per sphere point
  cplx<T>* en = em + MT;             // e^{i kn.(r_j - T_j)}
  T* ug = reinterpret_cast<T*>(en + MT);                    // uV(j, l) at [l * kFusedMaxNPI + jp]
  int* gix = reinterpret_cast<int*>(ug + kLA * MT);         // grid index of each sphere point
  double* rhol = reinterpret_cast<double*>(gix + MT);       // density of the workgroup's bands
  double* red = rhol + (LT ? NL : 0);
  // per atom: its first sphere point, and per projector Hxyz iuV (kfac, in
  // the propagation precision) and iuV (ksgn), so the KB loops read no global
  // memory between two barriers
  T* kfac = reinterpret_cast<T*>(red + 10 * NW);
  float* ksgn = reinterpret_cast<float*>(kfac + kFusedMaxAtoms * kLA);
  int* katom = reinterpret_cast<int*>(ksgn + kFusedMaxAtoms * kLA);
  // image offsets of each sphere point, (jx+2) + 5 (jy+2) + 25 (jz+2), for
  // its unwrapped position l H - j L in the nonlocal observables
  // KB handover: the KB waves' count of published applications, and a flag
  // set when a wait ran out
  int* kbcnt = katom + kFusedMaxAtoms + 1;
  int* kbfail = kbcnt + 1;
  unsigned char* gimg = reinterpret_cast<unsigned char*>(kbfail + 1);

  const int t = threadIdx.x;
  const int lane = t & (kWave - 1), wv = t / kWave;
  const int wvu = __builtin_amdgcn_readfirstlane(wv);  // the same, known to be wave-uniform
  const int xs = t / NYZ, yz = t - xs * NYZ;
  const int iy = yz / NZ, iz = yz - iy * NZ;
  const int x0 = xs * PX;
  const int ik = blockIdx.x / a.nbg;
  const int b0 = (blockIdx.x - ik * a.nbg) * a.bpw;
  const int b1 = min(a.NB, b0 + a.bpw);
  const Grid& g = a.g;

  double km[3], kn[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    km[d] = a.kac0[d * a.NK + ik] + a.am[d];
    kn[d] = a.kac0[d * a.NK + ik] + a.an[d];
  }
  // phase tables
  for (int q = t; q < NPH; q += NT) {
    int ax, n, r = q;
    if (r < 5 * NX) { ax = 0; n = NX; } else if (r < 5 * (NX + NY)) { ax = 1; n = NY; r -= 5 * NX; }
    else { ax = 2; n = NZ; r -= 5 * (NX + NY); }
    const int img = r / n, l = r - img * n;
    const double h = ax == 0 ? g.Hx : (ax == 1 ? g.Hy : g.Hz);
    const double L = ax == 0 ? g.aLx : (ax == 1 ? g.aLy : g.aLz);
    const double pos = l * h - (img - 2) * L;
    double s, c;
    sincos(km[ax] * pos, &s, &c);
    phm[q] = cplx<T>{(T)c, (T)s};
    sincos(kn[ax] * pos, &s, &c);
    phn[q] = cplx<T>{(T)c, (T)s};
  }
  // gradient coefficients at the midpoint k: D(a, n) = k_a nab_a(n)
  const T kmx = (T)km[0], kmy = (T)km[1], kmz = (T)km[2];
  T D[3][4];
#pragma unroll
  for (int d = 0; d < 3; ++d)
#pragma unroll
    for (int nn = 0; nn < 4; ++nn) D[d][nn] = (T)(km[d] * g.nab[d][nn + 1]);
  const T A = (T)((km[0] * km[0] + km[1] * km[1] + km[2] * km[2] - (g.lap[0][0] + g.lap[1][0] + g.lap[2][0])) * 0.5);
  // the (first) sphere point covering each own grid point, or the zero entry
  // NPI; two 16-bit indices per register
  int prim[(PX + 1) / 2];
#pragma unroll
  for (int i = 0; i < (PX + 1) / 2; ++i) prim[i] = 0;
#pragma unroll
  for (int i = 0; i < PX; ++i) {
    const int q = a.pt_prim[(x0 + i) * NYZ + yz];
    prim[i >> 1] |= (q >= 0 ? q : a.NPI) << (16 * (i & 1));
  }
  if (t == 0) dps[a.NPI] = cplx<T>{0, 0};
  const double A0 = -(g.lap[0][0] + g.lap[1][0] + g.lap[2][0]) * 0.5;
  // per-wave FP64 accumulators of the workgroup's observables, in LDS:
  // |psi|^2, V|psi|^2, kinetic, current stencil x/y/z, nonlocal current x/y/z, Enl
  if (lane < 10) red[wv * 10 + lane] = 0.0;
  if (t == 0) {
    *kbcnt = 0;
    *kbfail = 0;
  }
  for (int q = t; q <= a.NI; q += NT) katom[q] = a.atom_start[q];
  for (int q = t; q < a.NI * kLA; q += NT) {
    const double sg = a.sgnp[q];
    kfac[q] = (T)(g.Hxyz * sg);
    ksgn[q] = (float)sg;
  }
  // density partial of this workgroup: each thread owns its points, so the
  // slice is updated without atomics
  double* rho = a.rho_part + (size_t)blockIdx.x * NL;
  double* rho_acc = LT ? rhol : rho;
  if (LT || RP || a.rho) {
#pragma unroll
    for (int i = 0; i < PX; ++i) rho_acc[(x0 + i) * NYZ + yz] = 0.0;
  }
  __syncthreads();
  if constexpr (LT) {
    for (int q = t; q < a.NPI; q += NT) {
      const int code = a.jp_phase[q];
      em[q] = phase_at<T, NX, NY, NZ>(phm, code);
      en[q] = phase_at<T, NX, NY, NZ>(phn, code);
      gix[q] = a.jp_grid[q];
      gimg[q] = (unsigned char)((code & 1023) / NX + 5 * (((code >> 10) & 1023) / NY) + 25 * ((code >> 20) / NZ));
      const T* ur = a.uvp + (size_t)a.jp_atom[q] * kLA * a.Nps + a.jp_j[q];
#pragma unroll
      for (int l = 0; l < kLA; ++l) ug[l * kFusedMaxNPI + q] = ur[(size_t)l * a.Nps];
    }
    __syncthreads();
  }

  // Wave roles of one H application.
  // SEQ (complex64, and FP64 with overlapping spheres): every wave computes the
  // KB terms of an atom, a barrier (and the fold of the terms of points shared
  // by several spheres), then every wave runs the stencil of its points and
  // adds the KB term, the Taylor sum and the x window update at once.
  // AB (FP64 without overlapping spheres, whose KB loops read their tables
  // from global memory): one wave of each pair that shares a SIMD (waves w and
  // w+4, or 2j and 2j+1, whichever way the workgroup is placed: waves 0, 2,
  // 5, 7) computes the KB terms of two atoms and publishes them through an LDS
  // counter before its stencil, while its partner starts the stencil at once;
  // both add the KB terms once the counter shows every KB wave's terms, so the
  // latency-bound KB loops overlap stencil arithmetic.  Measured on the 16^3
  // step: AB is faster for FP64 and slower for complex64, whose stencil is
  // bound by LDS bandwidth as much as by VALU issue and pays for AB's extra
  // LDS round trip of the stencil part.
  constexpr bool AB = !OV && sizeof(T) == 8;
  static_assert(!AB || NW == 8, "KB wave pairing assumes 8 waves");
  const int order = ((wvu >> 2) ^ wvu) & 1;
  const bool kbw = !AB || order == 0;
  const int kat0 = AB ? (wvu >> 1) : wv, kats = AB ? NW / 2 : NW;
  int kseq = 0;  // KB publications expected so far
  // complex64 (SEQ): the potential of the thread's own points stays in
  // registers for the whole launch, so the sweeps issue no global loads
  constexpr bool VREG = sizeof(T) == 4 && !AB;
  T vf[VREG ? PX : 1];
  if constexpr (VREG) {
#pragma unroll
    for (int i = 0; i < PX; ++i) vf[i] = a.vloc[(x0 + i) * NYZ + yz];
  }
  for (int ib = b0; ib < b1; ++ib) {
    const size_t ikb = (size_t)ik * a.NB + ib;
    const double occ = a.occ[ikb];
    cplx<T>* zo = a.zu + ikb * NL;
    cplx<T> win[W], acc[PX];
#pragma unroll
    for (int i = 0; i < PX; ++i) {
      const int p = (x0 + i) * NYZ + yz;
      const cplx<T> v = zo[p];
      acc[i] = v;
      win[H + i] = v;
      bufA[p] = v;
    }
    lds_barrier();
    cplx<T>* in = bufA;
    cplx<T>* out = bufB;
#pragma unroll 1
    for (int n = 0; n < 4; ++n) {
      // the FP64 kernel's observables-only launch (RP instantiation, a.prop
      // unset): no H application; the observables below read the entry orbital
      if constexpr (RP && !DEFER) {
        if (!a.prop) break;
      }
      const bool obs = DEFER && n == 0 && a.obs;
      const int* jp_phase = opaque(a.jp_phase);
      const int* jp_grid = opaque(a.jp_grid);
      const T* uvt = opaque(a.uvp);
      auto kb_phase = [&]() __attribute__((always_inline)) {
      // KB term of pseudo_pt (hpsi.f90:176-206) per atom in one wave:
      //   uVpsi(l) = Hxyz iuV(l) sum_j uV(j,l) e^{ik.r_j} psi(J_j)       (projection)
      //   dps(j)   = e^{-ik.r_j} sum_l uV(j,l) uVpsi(l)                  (back-projection)
      // the wave sums leave uVpsi in every lane, so no barrier separates the
      // two.  The first two passes of 64 sphere points (every point when the
      // atom has at most 128) load together and keep their phase, orbital
      // value and projector values in registers for the back-projection and
      // the nonlocal observables; further points take a plain loop.
      for (int at = kat0; at < a.NI; at += kats) {
        const int j0 = katom[at], m = katom[at + 1] - j0;
        if (m == 0) continue;
        const T* ua = uvt + (size_t)at * kLA * a.Nps;
        T f[kLA];
#pragma unroll
        for (int l = 0; l < kLA; ++l) f[l] = kfac[at * kLA + l];
        cplx<T> e2[2], v2[2];
        T u2[2][kLA];
        int jp2[2];
        bool ok2[2];
#pragma unroll
        for (int s2 = 0; s2 < 2; ++s2) {
          const int j = lane + s2 * kWave;
          ok2[s2] = j < m;
          const int jc = ok2[s2] ? j : 0;  // an unused lane reads the atom's first point
          const int jp = j0 + jc;
          jp2[s2] = jp;
          if constexpr (LT) {
            e2[s2] = em[jp];
            v2[s2] = in[gix[jp]];
          } else {
            e2[s2] = phase_at<T, NX, NY, NZ>(phm, jp_phase[jp]);
            v2[s2] = in[jp_grid[jp]];
          }
#pragma unroll
          for (int l = 0; l < kLA; ++l) u2[s2][l] = LT ? ug[l * kFusedMaxNPI + jp] : ua[l * a.Nps + jc];
        }
        T q[2 * kLA];
#pragma unroll
        for (int l = 0; l < 2 * kLA; ++l) q[l] = 0;
#pragma unroll
        for (int s2 = 0; s2 < 2; ++s2) {
          if (ok2[s2]) {
            const cplx<T> w = cmul(e2[s2], v2[s2]);
#pragma unroll
            for (int l = 0; l < kLA; ++l) {
              q[2 * l] += u2[s2][l] * w.re;
              q[2 * l + 1] += u2[s2][l] * w.im;
            }
          }
        }
        for (int j = lane + 2 * kWave; j < m; j += kWave) {
          const int jp = j0 + j;
          cplx<T> w;
          if constexpr (LT) w = cmul(em[jp], in[gix[jp]]);
          else w = cmul(phase_at<T, NX, NY, NZ>(phm, jp_phase[jp]), in[jp_grid[jp]]);
#pragma unroll
          for (int l = 0; l < kLA; ++l) {
            const T u = LT ? ug[l * kFusedMaxNPI + jp] : ua[l * a.Nps + j];
            q[2 * l] += u * w.re;
            q[2 * l + 1] += u * w.im;
          }
        }
        wave_sums(q);
        T sr[kLA], si[kLA];
#pragma unroll
        for (int l = 0; l < kLA; ++l) {
          sr[l] = lane_read(q[0], wave_sums_lane(2 * l)) * f[l];
          si[l] = lane_read(q[0], wave_sums_lane(2 * l + 1)) * f[l];
        }
#pragma unroll
        for (int s2 = 0; s2 < 2; ++s2) {
          T tr = 0, ti = 0;
#pragma unroll
          for (int l = 0; l < kLA; ++l) {
            tr += u2[s2][l] * sr[l];
            ti += u2[s2][l] * si[l];
          }
          const cplx<T> e = e2[s2];
          if (ok2[s2]) dps[jp2[s2]] = cplx<T>{e.re * tr + e.im * ti, e.re * ti - e.im * tr};  // conj(e) * t
        }
        for (int j = lane + 2 * kWave; j < m; j += kWave) {
          const int jp = j0 + j;
          T tr = 0, ti = 0;
#pragma unroll
          for (int l = 0; l < kLA; ++l) {
            const T u = LT ? ug[l * kFusedMaxNPI + jp] : ua[l * a.Nps + j];
            tr += u * sr[l];
            ti += u * si[l];
          }
          const cplx<T> e = LT ? em[jp] : phase_at<T, NX, NY, NZ>(phm, jp_phase[jp]);
          dps[jp] = cplx<T>{e.re * tr + e.im * ti, e.re * ti - e.im * tr};  // conj(e) * t
        }
        if constexpr (DEFER && LT) {
          if (obs) {
            // nonlocal current and energy of this orbital (current.f90:247-269,
            // total_energy.f90:191-199) at k_obs, products and sums in FP64.
            // First the projections U_l = sum_j uV(j,l) e_j psi_j.  The current
            // sum_l sg_l Im(conj(U_l) X_l) of the first moments
            // X_l = sum_j r_j uV(j,l) e_j psi_j is linear in the point terms once
            // U is known: with w_j = e_j psi_j each point adds r_j c_j,
            //   c_j = Re w_j sum_l sg_l Im U_l uV(j,l) - Im w_j sum_l sg_l Re U_l uV(j,l),
            // so the wave sums 8 + 4 values instead of the 32 of (U, X).  uV and
            // the orbital values come from the registers above, the position
            // l H - j L from the grid point and its image offsets.
            const double hx = g.Hxyz, fo = occ / g.aLxyz * 2.0 * hx;
            double qu[2 * kLA];
#pragma unroll
            for (int l = 0; l < 2 * kLA; ++l) qu[l] = 0;
            cplx<T> w2[2];
#pragma unroll
            for (int s2 = 0; s2 < 2; ++s2) {
              w2[s2] = cmul(en[jp2[s2]], v2[s2]);
              if (ok2[s2]) {
#pragma unroll
                for (int l = 0; l < kLA; ++l) {
                  qu[2 * l] += (double)u2[s2][l] * (double)w2[s2].re;
                  qu[2 * l + 1] += (double)u2[s2][l] * (double)w2[s2].im;
                }
              }
            }
            for (int j = lane + 2 * kWave; j < m; j += kWave) {
              const int jp = j0 + j;
              const cplx<T> w = cmul(en[jp], in[gix[jp]]);
#pragma unroll
              for (int l = 0; l < kLA; ++l) {
                const double u = (double)ug[l * kFusedMaxNPI + jp];
                qu[2 * l] += u * (double)w.re;
                qu[2 * l + 1] += u * (double)w.im;
              }
            }
            wave_sums(qu);
            double P[kLA], Q[kLA], e9 = 0;
#pragma unroll
            for (int l = 0; l < kLA; ++l) {
              const double ur = lane_read(qu[0], wave_sums_lane(2 * l));
              const double ui = lane_read(qu[0], wave_sums_lane(2 * l + 1));
              const double sg = ksgn[at * kLA + l];
              P[l] = sg * ui;
              Q[l] = sg * ur;
              e9 += sg * (ur * ur + ui * ui);
            }
            double jm[4] = {0, 0, 0, 0};
            auto moment = [&](int jp, const cplx<T>& w, const T (&u)[kLA]) __attribute__((always_inline)) {
              const unsigned pg = (unsigned)gix[jp], c = gimg[jp];
              double sa = 0, sb = 0;
#pragma unroll
              for (int l = 0; l < kLA; ++l) {
                sa += P[l] * (double)u[l];
                sb += Q[l] * (double)u[l];
              }
              const double cj = (double)w.re * sa - (double)w.im * sb;
              jm[0] += ((int)(pg / NYZ) * g.Hx - ((int)(c % 5) - 2) * g.aLx) * cj;
              jm[1] += ((int)((pg / NZ) % NY) * g.Hy - ((int)((c / 5) % 5) - 2) * g.aLy) * cj;
              jm[2] += ((int)(pg % NZ) * g.Hz - ((int)(c / 25) - 2) * g.aLz) * cj;
            };
#pragma unroll
            for (int s2 = 0; s2 < 2; ++s2)
              if (ok2[s2]) moment(jp2[s2], w2[s2], u2[s2]);
            for (int j = lane + 2 * kWave; j < m; j += kWave) {
              const int jp = j0 + j;
              T u[kLA];
#pragma unroll
              for (int l = 0; l < kLA; ++l) u[l] = ug[l * kFusedMaxNPI + jp];
              moment(jp, cmul(en[jp], in[gix[jp]]), u);
            }
            wave_sums(jm);
            const double jx = lane_read(jm[0], wave_sums_lane(0));
            const double jy = lane_read(jm[0], wave_sums_lane(1));
            const double jz = lane_read(jm[0], wave_sums_lane(2));
            if (lane == 0) {
              red[wv * 10 + 6] += fo * hx * jx;
              red[wv * 10 + 7] += fo * hx * jy;
              red[wv * 10 + 8] += fo * hx * jz;
              red[wv * 10 + 9] += occ * hx * hx * e9;
            }
          }
        }
      }
      };  // kb_phase
      const T zr = (n == 1) ? a.z2r : ((n == 3) ? a.z4r : (T)0);
      const T zi = (n == 0) ? a.z1i : ((n == 2) ? a.z3i : (T)0);
      // stencil: h = (V+A) psi - 1/2 sum_a sum_n lap_n (psi+ + psi-) - i sum_a k_a sum_n nab_n (psi+ - psi-)
      // The stencil part of each result goes to LDS; the KB term and the
      // Taylor sum follow after the barrier that completes the KB terms
      // (finalize below).  The sweep loads its x window (own points and the
      // halo of 4) and forms its LDS bases itself, so nothing of it stays
      // live across the KB phase.
      auto sweep = [&](auto obs_c) __attribute__((always_inline)) {
        constexpr bool OBS = decltype(obs_c)::value;
        if constexpr (XS > 1) {
          // SEQ keeps its own points' inputs in the window from the previous
          // sweep (or the orbital's load) and reloads the halo only
#pragma unroll
          for (int i = 0; i < W; ++i)
            if (AB || i < H || i >= H + PX) win[i] = in[((x0 - H + i + NX) % NX) * NYZ + yz];
        } else {
#pragma unroll
          for (int i = 0; i < NX; ++i) win[i] = in[i * NYZ + yz];
        }
        // One LDS base address per neighbour direction; the x offset of each
        // own point is a compile-time constant that folds into the ds_read offset.
        const cplx<T>* zb[8];
        const cplx<T>* yb[8];
        nb_bases<T, NY, NZ>(in, x0, iy, iz, zb, yb);
        const T* vrow = opaque(a.vloc) + x0 * NYZ + yz;
        cplx<T> h[AB ? 1 : PX];
        // observables of the input orbital at k_obs, FP64 sums.  The
        // stencil's neighbour sums L = sum_a sum_n (-lap_n/2)(psi+ + psi-) and
        // G_a = sum_n nab_n (psi+ - psi-) give per point the kinetic term
        // Re conj(psi) L of total_energy_stencil (the centre term A0 |psi|^2
        // is added per orbital) and the stencil current Im conj(psi) G_a; the
        // gradient energy sum_i Re conj(psi)(-i k.G) is sum_a k_a current_a
        // per orbital.  Over the periodic grid, sum_i Im conj(psi_i)
        // (psi_{i+n} - psi_{i-n}) = 2 sum_i Im conj(psi_i) psi_{i+n}: the
        // central sums are twice the forward sums of current_stencil
        // (stencil/F90/current.f90:112-114).  L and G are in the propagation
        // precision, products and sums in FP64.
        double n2 = 0, vs = 0, te = 0, js[3] = {0, 0, 0};
        const double* v64 = opaque(a.vloc64) + x0 * NYZ + yz;
#pragma unroll
        for (int i = 0; i < PX; ++i) {
          const cplx<T> e0 = win[H + i];
          T vAi;
          if constexpr (VREG) vAi = vf[i] + A;
          else vAi = vrow[i * NYZ] + A;
          T lr, li;
          if constexpr (OBS) {
            T Lr = 0, Li = 0;
            T gr[3] = {0, 0, 0}, gi[3] = {0, 0, 0};
            nb_sums<T, NX, NYZ, H, XS, W>(a, zb, yb, win, i, Lr, Li, gr, gi);
            const double er = (double)e0.re, ei = (double)e0.im;
            const double a2 = er * er + ei * ei;
            n2 += a2;
            // complex64: the FP32 potential the propagation uses, promoted
            if constexpr (VREG) vs += (double)vf[i] * a2;
            else vs += v64[i * NYZ] * a2;
            if constexpr (!RP) rho_acc[(x0 + i) * NYZ + yz] += occ * a2;
            te += er * (double)Lr + ei * (double)Li;
#pragma unroll
            for (int d = 0; d < 3; ++d) js[d] += er * (double)gi[d] - ei * (double)gr[d];
            lr = vAi * e0.re + Lr + (kmx * gi[0] + kmy * gi[1] + kmz * gi[2]);
            li = vAi * e0.im + Li - (kmx * gr[0] + kmy * gr[1] + kmz * gr[2]);
          } else if constexpr (!DEFER) {
            lr = vAi * e0.re;
            li = vAi * e0.im;
            nb_sums_k<T, NX, NYZ, H, XS, W>(a, D, zb, yb, win, i, lr, li);
          } else {
            // the same arithmetic as the observables sweep, so that all four
            // H applications of a step use one Hamiltonian to the last bit
            T Lr = 0, Li = 0;
            T gr[3] = {0, 0, 0}, gi[3] = {0, 0, 0};
            nb_sums<T, NX, NYZ, H, XS, W>(a, zb, yb, win, i, Lr, Li, gr, gi);
            lr = vAi * e0.re + Lr + (kmx * gi[0] + kmy * gi[1] + kmz * gi[2]);
            li = vAi * e0.im + Li - (kmx * gr[0] + kmy * gr[1] + kmz * gr[2]);
          }
          if constexpr (AB) {
            out[(x0 + i) * NYZ + yz] = cplx<T>{lr, li};  // the KB term follows in finalize
          } else {
            // KB term of the sphere point covering this point (branch-free:
            // uncovered points read the zero entry dps[NPI]); the result goes
            // to LDS and into the Taylor sum, and replaces its input in the x
            // window once no later own point needs that input (a delay of 4
            // points; the first 4 of a periodic column wait to the end)
            const cplx<T> d = dps[(prim[i >> 1] >> (16 * (i & 1))) & 0xffff];
            h[i].re = lr + d.re;
            h[i].im = li + d.im;
            if (n < 3) out[(x0 + i) * NYZ + yz] = h[i];  // the fourth result only enters the Taylor sum
            if constexpr (sizeof(T) == 4) {
              // one of zr, zi is zero: two FMAs per part, the zero one adding
              // an exact zero (complex64; FP64 keeps the reference's rounding)
              acc[i].re = fmaf(zr, h[i].re, fmaf(-zi, h[i].im, acc[i].re));
              acc[i].im = fmaf(zr, h[i].im, fmaf(zi, h[i].re, acc[i].im));
            } else {
              acc[i].re += zr * h[i].re - zi * h[i].im;
              acc[i].im += zr * h[i].im + zi * h[i].re;
            }
            if constexpr (XS > 1) {
              if (i >= 4) win[H + i - 4] = h[i - 4];
            } else {
              if (i >= 8 && i - 4 < NX - 4) win[i - 4] = h[i - 4];
            }
          }
        }
        if constexpr (!AB) {
#pragma unroll
          for (int i = 0; i < PX; ++i) {
            if constexpr (XS > 1) {
              if (i >= PX - 4) win[H + i] = h[i];
            } else {
              if (i < 4 || i >= NX - 4) win[i] = h[i];
            }
          }
        }
        if constexpr (OBS) {
          // this orbital's sums into the wave's accumulators (one transposing
          // wave reduction of the six sums)
          double sv[8] = {n2, vs, te, js[0], js[1], js[2], 0.0, 0.0};
          wave_sums(sv);
          const double n2s = lane_read(sv[0], wave_sums_lane(0)), vss = lane_read(sv[0], wave_sums_lane(1));
          const double tes = lane_read(sv[0], wave_sums_lane(2)), j0 = lane_read(sv[0], wave_sums_lane(3));
          const double j1 = lane_read(sv[0], wave_sums_lane(4)), j2 = lane_read(sv[0], wave_sums_lane(5));
          if (lane == 0) {
            red[wv * 10 + 0] += occ * n2s;
            red[wv * 10 + 1] += occ * vss;
            red[wv * 10 + 2] += occ * (A0 * n2s + tes + kn[0] * j0 + kn[1] * j1 + kn[2] * j2);
            // reference: jx = jx + jxt*occ with jxt = 2*sum Im(...) over
            // forward neighbours (current.f90:206-210) = the central sum
            red[wv * 10 + 3] += occ * j0;
            red[wv * 10 + 4] += occ * j1;
            red[wv * 10 + 5] += occ * j2;
          }
        }
      };
      auto stencil_phase = [&]() __attribute__((always_inline)) {
        if constexpr (DEFER) {
          if (obs) sweep(BoolC<true>{});
          else sweep(BoolC<false>{});
        } else {
          sweep(BoolC<false>{});
        }
      };
      if constexpr (AB) {
        if (kbw) {
          kb_phase();
          // this wave's KB terms of this application are in LDS
          if (lane == 0) __hip_atomic_fetch_add(kbcnt, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
        stencil_phase();
        // wait for every KB wave's terms; the wait is bounded so that a wave
        // always reaches the barrier below (a wait that runs out poisons the
        // workgroup's sums instead of hanging)
        kseq += NW / 2;
        for (int spins = 0;; ++spins) {
          const int v = __builtin_amdgcn_readfirstlane(
              __hip_atomic_load(kbcnt, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP));
          if (v >= kseq) break;
          if (spins >= kFusedSpinMax) {
            if (lane == 0) *kbfail = 1;
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
      } else {
        kb_phase();
        lds_barrier();  // KB terms complete
      }
      if constexpr (OV) {
        // points inside several spheres (overlapping atoms): fold the other
        // sphere points' terms into the first, in table order
        for (int o = t; o < a.NOV; o += NT) {
          const int q0 = a.ov_start[o], q1 = a.ov_start[o + 1];
          cplx<T> s0 = dps[a.ov_jp[q0]];
          for (int q = q0 + 1; q < q1; ++q) {
            const cplx<T> d = dps[a.ov_jp[q]];
            s0.re += d.re;
            s0.im += d.im;
          }
          dps[a.ov_jp[q0]] = s0;
        }
        lds_barrier();
      }
      if constexpr (AB) {
        // finalize: result = stencil part + KB term of the sphere point
        // covering the point (branch-free: uncovered points read the zero
        // entry dps[NPI]); into LDS and the Taylor sum
#pragma unroll
        for (int i = 0; i < PX; ++i) {
          const int p = (x0 + i) * NYZ + yz;
          cplx<T> h = out[p];
          const cplx<T> d = dps[(prim[i >> 1] >> (16 * (i & 1))) & 0xffff];
          h.re += d.re;
          h.im += d.im;
          if (n < 3) out[p] = h;  // the fourth result only enters the Taylor sum
          acc[i].re += zr * h.re - zi * h.im;
          acc[i].im += zr * h.im + zi * h.re;
        }
      } else {
        stencil_phase();
      }
      lds_barrier();
      cplx<T>* tmp = in;
      in = out;
      out = tmp;
      if (!a.prop) break;  // observables only
    }
    if constexpr (DEFER) {
      // psi(t+dt) to HBM
      if (a.prop) {
#pragma unroll
        for (int i = 0; i < PX; ++i) zo[(x0 + i) * NYZ + yz] = acc[i];
        if constexpr (RP) {
#pragma unroll
          for (int i = 0; i < PX; ++i) {
            const double er = (double)acc[i].re, ei = (double)acc[i].im;
            rho_acc[(x0 + i) * NYZ + yz] += occ * (er * er + ei * ei);
          }
        }
      }
    } else {
    // psi(t+dt): to HBM, to the free LDS buffer and into the x window
#pragma unroll
    for (int i = 0; i < PX; ++i) {
      const int p = (x0 + i) * NYZ + yz;
      if (!RP || a.prop) zo[p] = acc[i];
      out[p] = acc[i];
      win[H + i] = acc[i];
    }
    if constexpr (RP) {
#pragma unroll
      for (int i = 0; i < PX; ++i) {
        const double er = (double)acc[i].re, ei = (double)acc[i].im;
        rho_acc[(x0 + i) * NYZ + yz] += occ * (er * er + ei * ei);
      }
    }
    // its observables unless the launch is the first half of an ETRS step
    // (a uniform branch: the barriers inside are taken by all threads or none)
    if (a.obs) {
    lds_barrier();
    if constexpr (XS > 1) {
#pragma unroll
      for (int i = 0; i < H; ++i) {
        win[i] = out[((x0 - H + i + NX) % NX) * NYZ + yz];
        win[H + PX + i] = out[((x0 + PX + i) % NX) * NYZ + yz];
      }
    }
    // observables at k + A(t+dt), FP64 sums.  The stencil's neighbour sums on
    // psi(t+dt), L = sum_a sum_n (-lap_n/2)(psi+ + psi-) and
    // G_a = sum_n nab_n (psi+ - psi-), give per point the kinetic term
    // Re conj(psi) L of total_energy_stencil (the centre term A0 |psi|^2 is
    // added per orbital) and the stencil current Im conj(psi) G_a; the
    // gradient energy sum_i Re conj(psi)(-i k.G) is sum_a k_a * current_a per
    // orbital.  Over the periodic grid, sum_i Im conj(psi_i)(psi_{i+n} -
    // psi_{i-n}) = 2 sum_i Im conj(psi_i) psi_{i+n}: the central sums are
    // twice the forward sums of current_stencil (stencil/F90/current.f90:112-114).
    // Products and sums are FP64; L and G are in the propagation precision.
    double n2 = 0, vs = 0, te = 0, js[3] = {0, 0, 0};
    {
      const cplx<T>* zb[8];
      const cplx<T>* yb[8];
      nb_bases<T, NY, NZ>(out, x0, iy, iz, zb, yb);
      const double* vrow = opaque(a.vloc64) + x0 * NYZ + yz;
      const __amdgpu_buffer_rsrc_t rrs = __builtin_amdgcn_make_buffer_rsrc(rho, (short)0, NL * (int)sizeof(double), 0x00020000);
#pragma unroll
      for (int i = 0; i < PX; ++i) {
        T lr = 0, li = 0;
        T gr[3] = {0, 0, 0}, gi[3] = {0, 0, 0};
        nb_sums<T, NX, NYZ, H, XS, W>(a, zb, yb, win, i, lr, li, gr, gi);
        const double er = (double)win[H + i].re, ei = (double)win[H + i].im;
        const double a2 = er * er + ei * ei;
        n2 += a2;
        vs += vrow[i * NYZ] * a2;
        if constexpr (!RP) {
          // through the buffer resource: with a.rho = 0 the offset is past the
          // slice, so the load returns 0 and the store is dropped, no memory access
          const int off = a.rho ? (int)(((x0 + i) * NYZ + yz) * sizeof(double)) : 0x7ffffff0;
          const auto ld = __builtin_amdgcn_raw_buffer_load_b64(rrs, off, 0, 0);
          const double v = __longlong_as_double(((long long)ld[1] << 32) | ld[0]) + occ * a2;
          const unsigned long long b = (unsigned long long)__double_as_longlong(v);
          typedef unsigned int u2 __attribute__((ext_vector_type(2)));
          __builtin_amdgcn_raw_buffer_store_b64(u2{(unsigned)b, (unsigned)(b >> 32)}, rrs, off, 0, 0);
        }
        te += er * (double)lr + ei * (double)li;
#pragma unroll
        for (int d = 0; d < 3; ++d) js[d] += er * (double)gi[d] - ei * (double)gr[d];
      }
    }
    {
      // this orbital's sums into the wave's accumulators
      const double n2s = wave_sum(n2), vss = wave_sum(vs), tes = wave_sum(te);
      const double j0 = wave_sum(js[0]), j1 = wave_sum(js[1]), j2 = wave_sum(js[2]);
      if (lane == 0) {
        red[wv * 10 + 0] += occ * n2s;
        red[wv * 10 + 1] += occ * vss;
        red[wv * 10 + 2] += occ * (A0 * n2s + tes + kn[0] * j0 + kn[1] * j1 + kn[2] * j2);
        // reference: jx = jx + jxt*occ with jxt = 2*sum Im(...) over forward
        // neighbours (current.f90:206-210) = the central sum
        red[wv * 10 + 3] += occ * j0;
        red[wv * 10 + 4] += occ * j1;
        red[wv * 10 + 5] += occ * j2;
      }
    }
    const int* jp_phase_n = opaque(a.jp_phase);
    const int* jp_grid_n = opaque(a.jp_grid);
    const double* uv64 = opaque(a.uvp64);
    const double* jp_pos = opaque(a.jp_pos);
    // nonlocal current and energy (current.f90:247-269, total_energy.f90:191-199):
    // one wave per atom; per projector l the projection U = sum_j uV e psi and
    // its first moments X_d = sum_j uV r_d e psi, all in FP64
    for (int at = wv; at < a.NI; at += NW) {
      const int j0 = katom[at], m = katom[at + 1] - j0;
      const double* ua = uv64 + (size_t)at * kLA * a.Nps;
      double q[8 * kLA];  // per l: Ur, Ui, Xr_x, Xi_x, Xr_y, Xi_y, Xr_z, Xi_z
#pragma unroll
      for (int l = 0; l < 8 * kLA; ++l) q[l] = 0;
      for (int j = lane; j < m; j += kWave) {
        const int jp = j0 + j;
        const cplx<T> e = phase_at<T, NX, NY, NZ>(phn, jp_phase_n[jp]);
        const cplx<T> w = cmul(e, out[jp_grid_n[jp]]);
        const double px = jp_pos[3 * jp], py = jp_pos[3 * jp + 1], pz = jp_pos[3 * jp + 2];
#pragma unroll
        for (int l = 0; l < kLA; ++l) {
          const double u = ua[l * a.Nps + j];
          const double tr = u * (double)w.re, ti = u * (double)w.im;
          q[8 * l + 0] += tr;
          q[8 * l + 1] += ti;
          q[8 * l + 2] += px * tr;
          q[8 * l + 3] += px * ti;
          q[8 * l + 4] += py * tr;
          q[8 * l + 5] += py * ti;
          q[8 * l + 6] += pz * tr;
          q[8 * l + 7] += pz * ti;
        }
      }
      wave_sums(q);
      const double hx = g.Hxyz, f = occ / g.aLxyz * 2.0 * hx;
#pragma unroll
      for (int l = 0; l < kLA; ++l) {
        double r[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) r[c] = lane_read(q[0], wave_sums_lane(8 * l + c));
        if (lane == 0) {
          const double sg = ksgn[at * kLA + l];
          const double Ur = r[0] * hx * sg, Ui = r[1] * hx * sg;
          red[wv * 10 + 6] += f * (r[2] * Ui - r[3] * Ur);
          red[wv * 10 + 7] += f * (r[4] * Ui - r[5] * Ur);
          red[wv * 10 + 8] += f * (r[6] * Ui - r[7] * Ur);
          red[wv * 10 + 9] += occ * (r[0] * r[0] + r[1] * r[1]) * hx * hx * sg;
        }
      }
    }
    }  // a.obs
    }
  }
  if constexpr (LT) {
    if (RP || a.rho) {
#pragma unroll
      for (int i = 0; i < PX; ++i) rho[(x0 + i) * NYZ + yz] = rhol[(x0 + i) * NYZ + yz];
    }
  }
  // per-workgroup sums of the per-wave accumulators, in wave order
  __syncthreads();
  if (t == 0) {
    double s[10];
#pragma unroll
    for (int i = 0; i < 10; ++i) {
      s[i] = 0;
      for (int w = 0; w < NW; ++w) s[i] += red[w * 10 + i];
    }
    double occsum = 0;
#pragma unroll 8
    for (int ib = b0; ib < b1; ++ib) occsum += a.occ[(size_t)ik * a.NB + ib];
    double* r = a.wg_rows + (size_t)blockIdx.x * 8;
    const double f = g.Hxyz / g.aLxyz;
#pragma unroll
    for (int d = 0; d < 3; ++d) r[d] = (kn[d] * s[0] + s[3 + d]) * f + s[6 + d];
    r[3] = s[2] * g.Hxyz + occsum * (kn[0] * kn[0] + kn[1] * kn[1] + kn[2] * kn[2]) / 2.0;
    r[4] = s[9];
    r[5] = s[1] * g.Hxyz;
    r[6] = s[0] * g.Hxyz;
    r[7] = 0.0;
    if (*kbfail) {
      for (int i = 0; i < 8; ++i) r[i] = __builtin_nan("");
    }
  }
  if (a.res) {
    // the workgroup that finishes last sums every workgroup's row in row
    // order (the same order whichever it is): no separate reduction launch
    __shared__ int last;
    if (t == 0) {
      const unsigned prev = __hip_atomic_fetch_add(a.cnt, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
      last = prev == gridDim.x - 1;
    }
    __syncthreads();
    if (last) {
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      for (int q = t; q < (int)gridDim.x; q += NT) {
#pragma unroll
        for (int i = 0; i < 8; ++i) v[i] += a.wg_rows[(size_t)q * 8 + i];
      }
      wave_sums(v);
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const double x = lane_read(v[0], wave_sums_lane(i));
        if (lane == 0) red[wv * 8 + i] = x;
      }
      __syncthreads();
      if (t < 8) {
        double acc = 0.0;
        for (int w = 0; w < NW; ++w) acc += red[w * 8 + t];
        a.res[t] = acc;
      }
      if (t == 0) __hip_atomic_store(a.cnt, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

}  // namespace salmon

// ---------------------------------------------------------------------------
// host side (included into salmon_gpu.hip after the context definition)
// ---------------------------------------------------------------------------
// x split of the 16^3 kernels: threads per workgroup = 256 * XS
// (build-time tunables, like the reference's stencil build options)
#ifndef SALMON_FUSED_XS64
#define SALMON_FUSED_XS64 2
#endif
#ifndef SALMON_FUSED_XS32
#define SALMON_FUSED_XS32 2
#endif
// waves per SIMD the register allocation targets: the FP64 kernel's 154 KiB of
// LDS allows one workgroup per CU, the complex64 kernel's 77 KiB two
#ifndef SALMON_FUSED_WPE64
#define SALMON_FUSED_WPE64 2
#endif
#ifndef SALMON_FUSED_WPE32
#define SALMON_FUSED_WPE32 2
#endif
constexpr int kFusedXS64 = SALMON_FUSED_XS64;
constexpr int kFusedXS32 = SALMON_FUSED_XS32;
constexpr int kFusedWPE64 = SALMON_FUSED_WPE64;
constexpr int kFusedWPE32 = SALMON_FUSED_WPE32;

struct FusedState {
  int shape = 0;  // 1: 16^3
  std::vector<void*> bufs;
  const int *pt_prim = nullptr, *ov_start = nullptr, *ov_jp = nullptr, *jp_atom = nullptr, *jp_j = nullptr,
            *jp_grid = nullptr, *jp_phase = nullptr;
  int NOV = 0;
  const double* jp_pos = nullptr;
  const double *sgnp = nullptr, *uvp64 = nullptr;
  const float* uvp32 = nullptr;
  double* wg_rows = nullptr;
  unsigned* cnt = nullptr;  // finished workgroups of the launch (the last sums the rows)
  double* rho_part = nullptr;
  int nwg_alloc = 0;
  bool tables = false;
};

static int fused_shape(const salmon::Grid& g) {
  if (g.NLx == 16 && g.NLy == 16 && g.NLz == 16) return 1;
  return 0;
}

static int fused_config(salmon_ctx* c) {
  if (!c->fz) c->fz = new FusedState();
  FusedState* f = (FusedState*)c->fz;
  f->shape = fused_shape(c->g);
  return f->shape;
}

static void fused_free_tables(FusedState* f) {
  for (void* p : f->bufs) (void)hipFree(p);
  f->bufs.clear();
  f->tables = false;
  f->sgnp = f->uvp64 = nullptr;
  f->uvp32 = nullptr;
}

static void fused_free(salmon_ctx* c) {
  FusedState* f = (FusedState*)c->fz;
  if (!f) return;
  fused_free_tables(f);
  if (f->wg_rows) (void)hipFree(f->wg_rows);
  if (f->rho_part) (void)hipFree(f->rho_part);
  if (f->cnt) (void)hipFree(f->cnt);
  delete f;
  c->fz = nullptr;
}

// Builds the per-point and per-sphere-point tables of the fused kernel from
// the streamed path's tables (host copies are rebuilt from the device ones).
static bool fused_ready(const salmon_ctx* c) {
  const FusedState* f = (const FusedState*)c->fz;
  return f && f->shape > 0 && f->tables && c->path == SALMON_PATH_AUTO;
}

static int fused_set_tables(salmon_ctx* c) {
  FusedState* f = (FusedState*)c->fz;
  if (!f || f->shape == 0) return SALMON_OK;
  fused_free_tables(f);
  const salmon::KBTables& kb = c->kb;
  const salmon::Grid& g = c->g;
  if (kb.NPI + 1 > salmon::kFusedMaxNPI || kb.NI > salmon::kFusedMaxAtoms)
    return SALMON_OK;  // the streamed path serves this case
  std::vector<int> start(kb.NI + 1), pidx(kb.NPI), lma_atom(kb.Nlma);
  std::vector<double> pos(3 * (size_t)kb.NPI), uv((size_t)kb.Nlma * kb.Nps), sign(kb.Nlma);
  if (hipMemcpy(start.data(), kb.atom_start, sizeof(int) * start.size(), hipMemcpyDeviceToHost) != hipSuccess ||
      (kb.NPI && hipMemcpy(pidx.data(), kb.pidx, sizeof(int) * pidx.size(), hipMemcpyDeviceToHost) != hipSuccess) ||
      (kb.NPI && hipMemcpy(pos.data(), kb.pos, sizeof(double) * pos.size(), hipMemcpyDeviceToHost) != hipSuccess) ||
      (kb.Nlma && hipMemcpy(lma_atom.data(), kb.lma_atom, sizeof(int) * lma_atom.size(), hipMemcpyDeviceToHost) != hipSuccess) ||
      (kb.Nlma && hipMemcpy(uv.data(), kb.uv, sizeof(double) * uv.size(), hipMemcpyDeviceToHost) != hipSuccess) ||
      (kb.Nlma && hipMemcpy(sign.data(), kb.lma_sign, sizeof(double) * sign.size(), hipMemcpyDeviceToHost) != hipSuccess))
    return fail(c, SALMON_ERR_HIP, "fused tables: device read-back failed");
  // projectors padded to kLA rows per atom, in their order within the atom
  constexpr int kLA = salmon::kLA;
  std::vector<int> nla(kb.NI, 0);
  std::vector<double> sgnp((size_t)kb.NI * kLA, 0.0), uvp64((size_t)kb.NI * kLA * kb.Nps, 0.0);
  for (int l = 0; l < kb.Nlma; ++l) {
    const int at = lma_atom[l];
    if (nla[at] == kLA) return SALMON_OK;  // more projectors per atom: the streamed path serves this case
    const size_t row = (size_t)at * kLA + nla[at]++;
    sgnp[row] = sign[l];
    for (int j = 0; j < kb.Nps; ++j) uvp64[row * kb.Nps + j] = uv[(size_t)l * kb.Nps + j];
  }
  std::vector<int> jp_atom(kb.NPI), jp_j(kb.NPI), jp_phase(kb.NPI);
  const int N[3] = {g.NLx, g.NLy, g.NLz};
  const double Hs[3] = {g.Hx, g.Hy, g.Hz}, Ls[3] = {g.aLx, g.aLy, g.aLz};
  for (int at = 0; at < kb.NI; ++at)
    for (int jp = start[at]; jp < start[at + 1]; ++jp) {
      jp_atom[jp] = at;
      jp_j[jp] = jp - start[at];
      const int p = pidx[jp];
      const int l[3] = {p / (g.NLy * g.NLz), (p / g.NLz) % g.NLy, p % g.NLz};
      int code = 0;
      for (int d = 0; d < 3; ++d) {
        // pos = l*H - img*L with img in -2..2
        const double img = (l[d] * Hs[d] - pos[3 * (size_t)jp + d]) / Ls[d];
        const int ii = (int)std::lround(img);
        if (ii < -2 || ii > 2) return fail(c, SALMON_ERR_ARG, "image offset outside -2..2");
        code |= ((ii + 2) * N[d] + l[d]) << (10 * d);
      }
      jp_phase[jp] = code;
    }
  std::vector<std::vector<int>> per(g.NL);
  for (int jp = 0; jp < kb.NPI; ++jp) per[pidx[jp]].push_back(jp);
  std::vector<int> pt_prim(g.NL, -1), ov_start(1, 0), ov_jp;
  for (int p = 0; p < g.NL; ++p) {
    if (per[p].empty()) continue;
    pt_prim[p] = per[p][0];
    if (per[p].size() > 1) {
      for (int jp : per[p]) ov_jp.push_back(jp);
      ov_start.push_back((int)ov_jp.size());
    }
  }
  f->NOV = (int)ov_start.size() - 1;
  std::vector<float> uvp32(uvp64.size());
  for (size_t i = 0; i < uvp64.size(); ++i) uvp32[i] = (float)uvp64[i];
  auto up = [&](const void* src, size_t bytes, void** dst) -> int {
    void* d = nullptr;
    if (hipMalloc(&d, bytes ? bytes : 4) != hipSuccess) return fail(c, SALMON_ERR_NOMEM, "fused tables");
    f->bufs.push_back(d);
    if (bytes && hipMemcpy(d, src, bytes, hipMemcpyHostToDevice) != hipSuccess)
      return fail(c, SALMON_ERR_HIP, "fused table upload");
    *dst = d;
    return SALMON_OK;
  };
  int r;
  if ((r = up(pt_prim.data(), sizeof(int) * pt_prim.size(), (void**)&f->pt_prim)) ||
      (r = up(ov_start.data(), sizeof(int) * ov_start.size(), (void**)&f->ov_start)) ||
      (r = up(ov_jp.data(), sizeof(int) * ov_jp.size(), (void**)&f->ov_jp)) ||
      (r = up(jp_atom.data(), sizeof(int) * jp_atom.size(), (void**)&f->jp_atom)) ||
      (r = up(jp_j.data(), sizeof(int) * jp_j.size(), (void**)&f->jp_j)) ||
      (r = up(pidx.data(), sizeof(int) * pidx.size(), (void**)&f->jp_grid)) ||
      (r = up(jp_phase.data(), sizeof(int) * jp_phase.size(), (void**)&f->jp_phase)) ||
      (r = up(sgnp.data(), sizeof(double) * sgnp.size(), (void**)&f->sgnp)) ||
      (r = up(uvp64.data(), sizeof(double) * uvp64.size(), (void**)&f->uvp64)) ||
      (r = up(uvp32.data(), sizeof(float) * uvp32.size(), (void**)&f->uvp32)))
    return r;
  f->jp_pos = kb.pos;
  f->tables = true;
  return SALMON_OK;
}

template <typename T, int NX, int NY, int NZ, int XS>
static size_t fused_lds_bytes() {
  constexpr bool LT = salmon::fused_lds_tables<T, XS>();
  constexpr size_t MT = LT ? salmon::kFusedMaxNPI : 0;
  return sizeof(salmon::cplx<T>) * (2 * NX * NY * NZ + 2 * 5 * (NX + NY + NZ) + salmon::kFusedMaxNPI + 2 * MT) +
         sizeof(T) * salmon::kLA * MT + sizeof(int) * MT + sizeof(double) * (LT ? NX * NY * NZ : 0) +
         sizeof(double) * 10 * ((NY * NZ * XS + 63) / 64) +
         (sizeof(T) + sizeof(float)) * salmon::kFusedMaxAtoms * salmon::kLA + sizeof(int) * (salmon::kFusedMaxAtoms + 3) +
         MT;  // gimg, one byte per sphere point
}

template <typename T, int NX, int NY, int NZ, int XS, int WPE, bool OV, bool RP>
static int fused_launch_rp(salmon_ctx* c, const salmon::FusedArgs<T>& args, int nwg) {
  auto kern = salmon::k_fused_step<T, NX, NY, NZ, XS, WPE, OV, RP>;
  const size_t lds = fused_lds_bytes<T, NX, NY, NZ, XS>();
  // set once per kernel instantiation (a thread-safe static initialisation)
  static const hipError_t attr =
      hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (attr != hipSuccess) return fail(c, SALMON_ERR_HIP, "hipFuncSetAttribute: %s", hipGetErrorString(attr));
  hipLaunchKernelGGL(kern, dim3(nwg), dim3(NY * NZ * XS), lds, c->stream, args);
  LAUNCHCHK(c);
  return SALMON_OK;
}

template <typename T, int NX, int NY, int NZ, int XS, int WPE, bool OV>
static int fused_launch(salmon_ctx* c, const salmon::FusedArgs<T>& args, int nwg) {
  // the self-consistent run's launches: steps that sum rho(t+dt), and the
  // FP64 kernel's observables-only launches
  if (args.rho_post || (sizeof(T) == 8 && !args.prop))
    return fused_launch_rp<T, NX, NY, NZ, XS, WPE, OV, true>(c, args, nwg);
  return fused_launch_rp<T, NX, NY, NZ, XS, WPE, OV, false>(c, args, nwg);
}

// One launch over all (k, band): with amid, the Taylor-4 step at kAc0 + amid;
// with aobs, the observables at kAc0 + aobs (density into d_rho, the 8
// result rows into res_dev) -- of the orbitals as they are on entry for the
// complex64 kernel (fused_defer), of psi(t+dt) for the FP64 kernel, or of
// the entry orbitals for an FP64 launch without amid (no density then).
// Either may be null (not both).
// rho_post (with amid): the density of psi(t+dt) into d_rho instead of that
// of the entry orbitals.  want_rho false: the observables leave d_rho as it is
// (a frozen-potential run sums the density of its last step only).
template <typename T>
// kbase: the device k points the launch adds amid / aobs to, kAc0 by default
// (the hamiltonian entry passes the caller's kAc(NK,3) with amid = 0)
static int fused_step(salmon_ctx* c, const double* amid, const double* aobs, double dt, double* res_dev,
                      bool rho_post = false, bool want_rho = true, const double* kbase = nullptr) {
  Region reg(c, amid ? SALMON_LOG_DT_EVOLVE : SALMON_LOG_CURRENT);
  FusedState* f = (FusedState*)c->fz;
  if (!f || !f->tables) return fail(c, SALMON_ERR_UNSUPPORTED, "fused tables not built");
  int ncu = 256;
  {
    int dev = 0;
    if (hipGetDevice(&dev) == hipSuccess) {
      int v = 0;
      if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0) ncu = v;
    }
  }
  // workgroups resident per CU: LDS allows 1 (FP64) or 2 (complex64), the
  // register target allows WPE*4/(threads/64)
  const int threads = 256 * (sizeof(T) == 8 ? kFusedXS64 : kFusedXS32);
  const int by_regs = std::max(1, (sizeof(T) == 8 ? kFusedWPE64 : kFusedWPE32) * 4 * 64 / threads);
  const int per_cu = std::min(sizeof(T) == 8 ? 1 : 2, by_regs);
  const int target = 2 * ncu * per_cu;
  int bpw = (c->NKB + target - 1) / target;
  bpw = std::max(1, std::min(bpw, c->NB));
  const int nbg = (c->NB + bpw - 1) / bpw;
  const int nwg = c->NK * nbg;
  if (nwg > f->nwg_alloc) {
    if (f->wg_rows) (void)hipFree(f->wg_rows);
    if (f->rho_part) (void)hipFree(f->rho_part);
    f->wg_rows = nullptr;
    f->rho_part = nullptr;
    int r;
    if ((r = dalloc(c, &f->wg_rows, 8 * (size_t)nwg)) || (r = dalloc(c, &f->rho_part, (size_t)nwg * c->g.NL))) return r;
    f->nwg_alloc = nwg;
  }
  salmon::FusedArgs<T> a{};
  a.zu = (salmon::cplx<T>*)c->d_zu;
  a.kac0 = kbase ? kbase : c->d_kac0;
  a.occ = c->d_occ;
  a.vloc = vloc_ptr<T>(c);
  a.vloc64 = c->d_vloc;
  a.NK = c->NK; a.NB = c->NB; a.bpw = bpw; a.nbg = nbg;
  a.prop = amid != nullptr;
  a.obs = aobs != nullptr;
  a.rho_post = rho_post && amid != nullptr;
  a.rho = want_rho;
  for (int d = 0; d < 3; ++d) {
    a.am[d] = amid ? amid[d] : 0.0;
    a.an[d] = aobs ? aobs[d] : 0.0;
  }
  a.z1i = (T)(-dt); a.z2r = (T)(-dt * dt / 2.0); a.z3i = (T)(dt * dt * dt / 6.0); a.z4r = (T)(dt * dt * dt * dt / 24.0);
  for (int d = 0; d < 3; ++d)
    for (int n = 0; n < 5; ++n) {
      a.lapc[d][n] = (T)(-0.5 * c->g.lap[d][n]);
      a.nab[d][n] = (T)c->g.nab[d][n];
    }
  a.pt_prim = f->pt_prim; a.ov_start = f->ov_start; a.ov_jp = f->ov_jp; a.NOV = f->NOV; a.jp_atom = f->jp_atom; a.jp_j = f->jp_j; a.jp_grid = f->jp_grid;
  a.jp_phase = f->jp_phase; a.jp_pos = f->jp_pos; a.atom_start = c->kb.atom_start;
  a.sgnp = f->sgnp;
  if constexpr (sizeof(T) == 8) a.uvp = (const T*)f->uvp64;
  else a.uvp = (const T*)f->uvp32;
  a.uvp64 = f->uvp64;
  a.NI = c->kb.NI; a.Nps = c->kb.Nps; a.NPI = c->kb.NPI;
  a.g = c->g;
  a.wg_rows = f->wg_rows;
  a.rho_part = f->rho_part;
  if (aobs && !f->cnt) {
    if (hipMalloc((void**)&f->cnt, sizeof(unsigned)) != hipSuccess) return fail(c, SALMON_ERR_NOMEM, "fused counter");
    HIPCHK(c, hipMemsetAsync(f->cnt, 0, sizeof(unsigned), c->stream));
  }
  a.res = aobs ? res_dev : nullptr;
  a.cnt = f->cnt;
  int r;
  // the overlap fold is compiled only into the instantiation for cells whose
  // atom spheres share grid points
  if constexpr (sizeof(T) == 8) {
    r = a.NOV > 0 ? fused_launch<T, 16, 16, 16, kFusedXS64, kFusedWPE64, true>(c, a, nwg)
                  : fused_launch<T, 16, 16, 16, kFusedXS64, kFusedWPE64, false>(c, a, nwg);
  } else {
    r = a.NOV > 0 ? fused_launch<T, 16, 16, 16, kFusedXS32, kFusedWPE32, true>(c, a, nwg)
                  : fused_launch<T, 16, 16, 16, kFusedXS32, kFusedWPE32, false>(c, a, nwg);
  }
  if (r || !(aobs || a.rho_post)) return r;
  // the FP64 observables-only launch sums no density (d_rho keeps rho(t+dt))
  if (sizeof(T) == 8 && !a.prop) return SALMON_OK;
  if (!a.rho && !a.rho_post) return SALMON_OK;
  return density_reduce(c, f->rho_part, nwg);
}
