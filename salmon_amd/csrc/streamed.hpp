// Streamed kernels: one pass over HBM per operator application, any grid.
//
// These serve every grid shape (the LDS-resident step kernel covers the
// grids whose orbital fits in LDS), the single-orbital hpsi entry and the
// ground-state routines.
#pragma once
#include "common.hpp"

namespace salmon {

// wave vector of orbital ikb: kAc0(ik,:) + a (a = A shift of the step)
__device__ __forceinline__ void orbital_k(const double* kac0, int NK, int ik, double ax, double ay, double az,
                                          double& kx, double& ky, double& kz) {
  kx = kac0[ik] + ax;
  ky = kac0[NK + ik] + ay;
  kz = kac0[2 * NK + ik] + az;
}

// ---------------------------------------------------------------------------
// KB projection, phase 1 of pseudo_pt (ARTED/common/hpsi.f90:183-193):
//   uVpsi(ilma) = Hxyz * iuV * sum_j conj(zproj(j,ilma)) * psi(Jxyz(j))
// with conj(zproj) = exp(i k.r_j) uV(j,ilma) (projector.f90:62,72) formed on
// the fly.  One wave per (projector, orbital); block = 4 waves.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void kb_project_wave(const cplx<T>* __restrict__ E, int ilma, double kx, double ky, double kz,
                                                const Grid& g, const KBTables& kb, cplx<T>* __restrict__ dst,
                                                const cplx<T>* __restrict__ ph = nullptr) {
  const int lane = threadIdx.x & (kWave - 1);
  const int a = kb.lma_atom[ilma];
  const int j0 = kb.atom_start[a], m = kb.atom_start[a + 1] - j0;
  // three 64-point chunks of the sphere with their loads in flight together:
  // indices past the sphere are clamped to its first point and weighted by
  // zero, so no branch separates the gathers from their uses
  constexpr int NC = 3;
  T sr = 0, si = 0;
  for (int jb = 0; jb < m; jb += NC * kWave) {
    int jj[NC];
    bool ok[NC];
    cplx<T> e[NC];
    T u[NC], c[NC], s[NC];
#pragma unroll
    for (int h = 0; h < NC; ++h) {
      const int j = jb + h * kWave + lane;
      ok[h] = j < m;
      jj[h] = ok[h] ? j : 0;
      e[h] = E[kb.pidx[j0 + jj[h]]];
      const T u0 = (T)kb.uv[(size_t)ilma * kb.Nps + jj[h]];
      u[h] = ok[h] ? u0 : (T)0;
    }
    if (ph) {
#pragma unroll
      for (int h = 0; h < NC; ++h) {
        c[h] = ph[j0 + jj[h]].re;
        s[h] = ph[j0 + jj[h]].im;
      }
    } else {
#pragma unroll
      for (int h = 0; h < NC; ++h) {
        const int jp = j0 + jj[h];
        const double kr = kx * kb.pos[3 * jp] + ky * kb.pos[3 * jp + 1] + kz * kb.pos[3 * jp + 2];
        sincos_t((T)kr, &s[h], &c[h]);
      }
    }
#pragma unroll
    for (int h = 0; h < NC; ++h) {
      // (u c + i u s) * (e.re + i e.im)
      sr += u[h] * (c[h] * e[h].re - s[h] * e[h].im);
      si += u[h] * (c[h] * e[h].im + s[h] * e[h].re);
    }
  }
  sr = wave_sum(sr);
  si = wave_sum(si);
  if (lane == 0) {
    const T f = (T)(g.Hxyz * kb.lma_sign[ilma]);
    *dst = cplx<T>{sr * f, si * f};
  }
}

template <typename T>
__global__ void __launch_bounds__(256) k_kb_project(const cplx<T>* __restrict__ psi, int orb0, int NB, int NK,
                                                    const double* __restrict__ kac0, double ax, double ay,
                                                    double az, Grid g, KBTables kb, cplx<T>* __restrict__ uvpsi,
                                                    const cplx<T>* __restrict__ ph) {
  const int ilma = blockIdx.x * 4 + (threadIdx.x / kWave);
  const int o = blockIdx.y;
  if (ilma >= kb.Nlma) return;
  const int ikb = orb0 + o;
  double kx, ky, kz;
  orbital_k(kac0, NK, ikb / NB, ax, ay, az, kx, ky, kz);
  kb_project_wave<T>(psi + (size_t)o * g.NL, ilma, kx, ky, kz, g, kb, uvpsi + (size_t)o * kb.Nlma + ilma,
                     ph ? ph + (size_t)(ikb / NB) * kb.NPI : nullptr);
}

// ---------------------------------------------------------------------------
// The same projection for the bands of one k point at once, on the matrix
// cores (the OpenACC build batches phase 1 of pseudo_pt over (ikb, ilma),
// ARTED/common/acc/hpsi.f90:72-93).  Per (atom, k point, tile of 16 bands,
// block of 8 projectors) one wave forms the real GEMM
//   C(16 bands x 16) = [psi_re | psi_im](16 x 2M) * [[w_re, w_im], [-w_im, w_re]](2M x 16)
// over the atom's M sphere points, w(j,l) = uV(j,l) e^{i k.(r_j - T_j)}, with
// columns 0..7 the real and 8..15 the imaginary parts of the block's
// projections: two v_mfma_*_16x16x4 per four sphere points.  Operand layout
// of the 16x16x4 MFMA: lane i holds A[i%16][i/16] and B[i/16][i%16]; its four
// accumulator values are rows row(i/16, r), r = 0..3, of column i%16, with
// row(g, r) = 4g + r for the f32 instruction and g + 4r for the f64 one (read
// back on gfx950 with one-hot operands).  Every index a lane reads is bounded
// by its own validity test, whatever the layout.
// ---------------------------------------------------------------------------
template <typename T>
struct MfmaAcc;
template <>
struct MfmaAcc<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ int row(int g, int r) { return g + 4 * r; }
  static __device__ __forceinline__ v4 fma(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};
template <>
struct MfmaAcc<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ int row(int g, int r) { return 4 * g + r; }
  static __device__ __forceinline__ v4 fma(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

// grid: (NI, band tiles of the chunk, projector blocks of 8); block: one wave.
// The chunk holds orbitals orb0 .. orb0+norb-1 (ikb = ik*NB + band); tile t
// covers k point k0 + t / tpk and bands 16*(t % tpk) .. +15.
template <typename T>
__global__ void __launch_bounds__(64) k_kb_project_mfma(const cplx<T>* __restrict__ psi, int orb0, int norb, int NB,
                                                        int NK, int k0, int tpk, const double* __restrict__ kac0,
                                                        double ax, double ay, double az, Grid g, KBTables kb,
                                                        cplx<T>* __restrict__ uvpsi, const cplx<T>* __restrict__ ph) {
  typedef typename MfmaAcc<T>::v4 v4;
  const int lane = threadIdx.x;
  const int a = blockIdx.x;
  const int ik = k0 + (int)blockIdx.y / tpk;
  const int b0 = ((int)blockIdx.y % tpk) * 16;
  const int l0 = kb.atom_lma[a] + 8 * (int)blockIdx.z, l1 = kb.atom_lma[a + 1];
  if (ik >= NK || l0 >= l1) return;
  const int j0 = kb.atom_start[a], m = kb.atom_start[a + 1] - j0;
  // this lane's A row (a band) and B column (a projector, real or imaginary part)
  const int row = lane & 15, kq = lane >> 4;
  const int band = b0 + row;
  const int o = ik * NB + band - orb0;  // chunk-local orbital
  const bool rvalid = band < NB && o >= 0 && o < norb;
  const cplx<T>* prow = psi + (size_t)(rvalid ? o : 0) * g.NL;
  const int col = lane & 15;
  const int lc = l0 + (col & 7);
  const bool cvalid = lc < l1;
  const double* ucol = kb.uv + (size_t)(cvalid ? lc : 0) * kb.Nps;
  const cplx<T>* phk = ph ? ph + (size_t)ik * kb.NPI : nullptr;
  double kx = 0, ky = 0, kz = 0;
  if (!ph) orbital_k(kac0, NK, ik, ax, ay, az, kx, ky, kz);
  v4 acc = {0, 0, 0, 0};
  for (int jj = 0; jj < m; jj += 4) {
    const int j = jj + kq;
    const bool jvalid = j < m;
    cplx<T> e{0, 0};
    if (rvalid && jvalid) e = prow[kb.pidx[j0 + j]];
    T wr = 0, wi = 0;
    if (cvalid && jvalid) {
      T sn, cs;
      if (phk) {
        cs = phk[j0 + j].re;
        sn = phk[j0 + j].im;
      } else {
        const int jp = j0 + j;
        const double kr = kx * kb.pos[3 * jp] + ky * kb.pos[3 * jp + 1] + kz * kb.pos[3 * jp + 2];
        sincos_t((T)kr, &sn, &cs);
      }
      const T u = (T)ucol[j];
      wr = u * cs;
      wi = u * sn;
    }
    // A = psi_re against [w_re | w_im], then A = psi_im against [-w_im | w_re]
    acc = MfmaAcc<T>::fma(e.re, col < 8 ? wr : wi, acc);
    acc = MfmaAcc<T>::fma(e.im, col < 8 ? -wi : wr, acc);
  }
  if (!cvalid) return;
  const T f = (T)(g.Hxyz * kb.lma_sign[lc]);
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int br = b0 + MfmaAcc<T>::row(kq, r);
    const int orow = ik * NB + br - orb0;
    if (br < NB && orow >= 0 && orow < norb) {
      T* dst = (T*)&uvpsi[(size_t)orow * kb.Nlma + lc];
      dst[col < 8 ? 0 : 1] = acc[r] * f;
    }
  }
}

// e^{i k.(r_j - T_j)} of every sphere point for every k point at kAc0 + a
// (NK, NPI): the phases of the KB projection and apply, shared by the four H
// applications of a Taylor step and by all bands of a k point.  FP64 sincos,
// rounded to the propagation precision.
template <typename T>
__global__ void __launch_bounds__(256) k_phase_table(const double* __restrict__ kac0, int NK, double ax, double ay,
                                                     double az, KBTables kb, cplx<T>* __restrict__ ph) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= (long)NK * kb.NPI) return;
  const int ik = (int)(i / kb.NPI), jp = (int)(i - (long)ik * kb.NPI);
  double kx, ky, kz;
  orbital_k(kac0, NK, ik, ax, ay, az, kx, ky, kz);
  double s, c;
  sincos(kx * kb.pos[3 * jp] + ky * kb.pos[3 * jp + 1] + kz * kb.pos[3 * jp + 2], &s, &c);
  ph[i] = cplx<T>{(T)c, (T)s};
}

// ---------------------------------------------------------------------------
// hpsi: stencil (hpsi1_RT_stencil, ARTED/stencil/F90/hpsi.f90:88-190) + local
// potential + KB apply (pseudo_pt phase 2, hpsi.f90:195-206, gathered per
// touched point through the transposed table as the OpenACC build does,
// common/acc/hpsi.f90:95-113).  One thread per grid point.
//   F = (V + A) E - 1/2 sum C_n (E[+n] + E[-n]) - i sum D_n (E[+n] - E[-n])
// with A = (|k|^2 - sum lap0)/2, C = lap/H^2, D = k.nab/H.
// tt (nullable) gets F - V E before the KB term (hpsi.f90:131-133).
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ cplx<T> hpsi_point(const cplx<T>* __restrict__ E, int p, double kx, double ky, double kz,
                                              const T* __restrict__ vloc, const Grid& g, const KBTables& kb,
                                              const cplx<T>* __restrict__ U, cplx<T>* __restrict__ ttp,
                                              const cplx<T>* __restrict__ ph = nullptr) {
  const T A = (T)((kx * kx + ky * ky + kz * kz - (g.lap[0][0] + g.lap[1][0] + g.lap[2][0])) * 0.5);
  const int NLx = g.NLx, NLy = g.NLy, NLz = g.NLz;
  const int iz = p % NLz, iy = (p / NLz) % NLy, ix = p / (NLz * NLy);
  const double kk[3] = {kx, ky, kz};
  T vr = 0, vi = 0, wr = 0, wi = 0;
  // z, y, x in the reference's order
#pragma unroll
  for (int ax3 = 2; ax3 >= 0; --ax3) {
    const int n_ax = ax3 == 2 ? NLz : (ax3 == 1 ? NLy : NLx);
    const int c_ax = ax3 == 2 ? iz : (ax3 == 1 ? iy : ix);
    const int stride = ax3 == 2 ? 1 : (ax3 == 1 ? NLz : NLz * NLy);
    const int base = p - c_ax * stride;
#pragma unroll
    for (int n = 1; n <= 4; ++n) {
      int cp = c_ax + n;
      if (cp >= n_ax) cp -= n_ax;
      int cm = c_ax - n;
      if (cm < 0) cm += n_ax;
      const cplx<T> ep = E[base + cp * stride], em = E[base + cm * stride];
      const T C = (T)g.lap[ax3][n], D = (T)(kk[ax3] * g.nab[ax3][n]);
      vr += C * (ep.re + em.re);
      vi += C * (ep.im + em.im);
      wr += D * (ep.re - em.re);
      wi += D * (ep.im - em.im);
    }
  }
  const cplx<T> e = E[p];
  const T V = vloc[p];
  cplx<T> f;
  f.re = (V + A) * e.re - (T)0.5 * vr + wi;
  f.im = (V + A) * e.im - (T)0.5 * vi - wr;
  if (ttp) *ttp = cplx<T>{f.re - V * e.re, f.im - V * e.im};
  const int t = kb.pt2t[p];
  if (t >= 0) {
    for (int q = kb.t_start[t]; q < kb.t_start[t + 1]; ++q) {
      const int il = kb.e_lma[q], jp = kb.e_jp[q];
      const int a = kb.lma_atom[il];
      T s, c;
      if (ph) {
        c = ph[jp].re;
        s = ph[jp].im;
      } else {
        const double kr = kx * kb.pos[3 * jp] + ky * kb.pos[3 * jp + 1] + kz * kb.pos[3 * jp + 2];
        sincos_t((T)kr, &s, &c);
      }
      const T u = (T)kb.uv[(size_t)il * kb.Nps + (jp - kb.atom_start[a])];
      // zproj * uVpsi = u (c - i s) (U.re + i U.im)
      const cplx<T> w = U[il];
      f.re += u * (c * w.re + s * w.im);
      f.im += u * (c * w.im - s * w.re);
    }
  }
  return f;
}

template <typename T>
__global__ void __launch_bounds__(256) k_hpsi(const cplx<T>* __restrict__ in, cplx<T>* __restrict__ out,
                                              cplx<T>* __restrict__ tt, int orb0, int NB, int NK,
                                              const double* __restrict__ kac0, double ax, double ay, double az,
                                              const T* __restrict__ vloc, Grid g, KBTables kb,
                                              const cplx<T>* __restrict__ uvpsi, const cplx<T>* __restrict__ ph) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  const int o = blockIdx.y;
  if (p >= g.NL) return;
  const int ikb = orb0 + o;
  double kx, ky, kz;
  orbital_k(kac0, NK, ikb / NB, ax, ay, az, kx, ky, kz);
  out[(size_t)o * g.NL + p] = hpsi_point<T>(in + (size_t)o * g.NL, p, kx, ky, kz, vloc, g, kb,
                                            uvpsi + (size_t)o * kb.Nlma, tt ? tt + (size_t)o * g.NL + p : nullptr,
                                            ph ? ph + (size_t)(ikb / NB) * kb.NPI : nullptr);
}

// zu += z1 h1 + z2 h2 + z3 h3 + z4 h4 (update, ARTED/RT/hamiltonian.f90:99-122)
template <typename T>
__global__ void __launch_bounds__(256) k_taylor_fold(cplx<T>* __restrict__ zu, const cplx<T>* __restrict__ h1,
                                                     const cplx<T>* __restrict__ h2, const cplx<T>* __restrict__ h3,
                                                     const cplx<T>* __restrict__ h4, size_t n, T z1i, T z2r, T z3i,
                                                     T z4r) {
  // zfac = (-i dt)^n/n!: z1 = (0, z1i), z2 = (z2r, 0), z3 = (0, z3i), z4 = (z4r, 0)
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  cplx<T> z = zu[i];
  const cplx<T> a = h1[i], b = h2[i], c = h3[i], d = h4[i];
  z.re += -z1i * a.im;
  z.im += z1i * a.re;
  z.re += z2r * b.re;
  z.im += z2r * b.im;
  z.re += -z3i * c.im;
  z.im += z3i * c.re;
  z.re += z4r * d.re;
  z.im += z4r * d.im;
  zu[i] = z;
}

// ---------------------------------------------------------------------------
// Observables of one orbital, all in FP64 (one 256-thread block per orbital):
//  current (RT/current.f90:172-273 + stencil/F90/current.f90:82-114):
//    J = (occ k sum|psi|^2 + occ 2 sum Im(conj psi sum_n nab_n psi[+n])) Hxyz/Omega
//        + occ/Omega 2 Im(conj(<x p|psi>) <p|psi> iuV) Hxyz^2
//  kinetic energy (total_energy.f90:183 + stencil/F90/total_energy.f90:93-180):
//    Ekin = occ Re sum conj(psi)(A0 psi - 1/2 sum C(psi+ + psi-) - i sum D(psi+ - psi-)) Hxyz
//           + occ |k|^2/2
//  nonlocal energy (total_energy.f90:191-199): Enl = occ sum iuV |Hxyz <p|psi>|^2
//  local energy of the frozen potential and the norm.
// out(ikb, 0..7) = Jx, Jy, Jz, Ekin, Enl, Eloc, Norm, 0
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void observe_orbital(const cplx<T>* __restrict__ E, double kx, double ky, double kz,
                                                double occ, const double* __restrict__ vloc, const Grid& g,
                                                const KBTables& kb, double* __restrict__ r,
                                                const cplx<T>* __restrict__ ph = nullptr) {
  __shared__ double smem[4 * 8];
  __shared__ double nl[4][4];
  const double kk[3] = {kx, ky, kz};
  const double A0 = -(g.lap[0][0] + g.lap[1][0] + g.lap[2][0]) * 0.5;
  const int NLx = g.NLx, NLy = g.NLy, NLz = g.NLz;
  // Per point the stencil's neighbour sums, in the orbitals' precision:
  //   L = sum_a sum_n lap_n (psi+ + psi-),  G_a = sum_n nab_n (psi+ - psi-);
  // Re conj(psi) L gives the kinetic term of total_energy_stencil
  // (total_energy.f90:172-190), Im conj(psi) G_a the stencil current; summed
  // over the periodic grid the central sum is twice the forward sum of
  // current_stencil (stencil/F90/current.f90:112-114), and the -i k.grad
  // energy is sum_a k_a current_a.  Products and sums in FP64.
  double n2 = 0, vs = 0, te = 0, jc[3] = {0, 0, 0};
  for (int p = threadIdx.x; p < g.NL; p += 256) {
    const int iz = p % NLz, iy = (p / NLz) % NLy, ix = p / (NLz * NLy);
    const cplx<T> e0 = E[p];
    const double er = (double)e0.re, ei = (double)e0.im;
    const double a2 = er * er + ei * ei;
    n2 += a2;
    vs += vloc[p] * a2;
    T Lr = 0, Li = 0, Gr[3] = {0, 0, 0}, Gi[3] = {0, 0, 0};
#pragma unroll
    for (int ax3 = 2; ax3 >= 0; --ax3) {
      const int n_ax = ax3 == 2 ? NLz : (ax3 == 1 ? NLy : NLx);
      const int c_ax = ax3 == 2 ? iz : (ax3 == 1 ? iy : ix);
      const int stride = ax3 == 2 ? 1 : (ax3 == 1 ? NLz : NLz * NLy);
      const int base = p - c_ax * stride;
#pragma unroll
      for (int n = 1; n <= 4; ++n) {
        int cp = c_ax + n;
        if (cp >= n_ax) cp -= n_ax;
        int cm = c_ax - n;
        if (cm < 0) cm += n_ax;
        const cplx<T> ep = E[base + cp * stride], em = E[base + cm * stride];
        const T C = (T)g.lap[ax3][n], nb = (T)g.nab[ax3][n];
        Lr += C * (ep.re + em.re);
        Li += C * (ep.im + em.im);
        Gr[ax3] += nb * (ep.re - em.re);
        Gi[ax3] += nb * (ep.im - em.im);
      }
    }
    te += er * (double)Lr + ei * (double)Li;
#pragma unroll
    for (int d = 0; d < 3; ++d) jc[d] += er * (double)Gi[d] - ei * (double)Gr[d];
  }
  // Re conj(psi) (A0 psi - L/2 - i sum_a k_a G_a) summed = A0 n2 - te/2 + sum_a k_a jc_a
  te = -0.5 * te;
  double v[6] = {n2, vs, te, jc[0], jc[1], jc[2]};
  block_sum256<6>(v, smem);
  // nonlocal: wave w takes projectors w, w+4, ...
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  double jn[3] = {0, 0, 0}, en = 0;
  for (int il = w; il < kb.Nlma; il += 4) {
    const int a = kb.lma_atom[il];
    const int j0 = kb.atom_start[a], m = kb.atom_start[a + 1] - j0;
    double ur = 0, ui = 0, xr[3] = {0, 0, 0}, xi[3] = {0, 0, 0};
    for (int j = lane; j < m; j += kWave) {
      const int jp = j0 + j;
      const double px = kb.pos[3 * jp], py = kb.pos[3 * jp + 1], pz = kb.pos[3 * jp + 2];
      double s, c;
      if (ph) {
        c = (double)ph[jp].re;
        s = (double)ph[jp].im;
      } else {
        sincos(kx * px + ky * py + kz * pz, &s, &c);
      }
      const double u = kb.uv[(size_t)il * kb.Nps + j];
      const cplx<T> e = E[kb.pidx[jp]];
      const double tr = u * (c * (double)e.re - s * (double)e.im);
      const double ti = u * (c * (double)e.im + s * (double)e.re);
      ur += tr;
      ui += ti;
      xr[0] += px * tr; xi[0] += px * ti;
      xr[1] += py * tr; xi[1] += py * ti;
      xr[2] += pz * tr; xi[2] += pz * ti;
    }
    ur = wave_sum(ur); ui = wave_sum(ui);
#pragma unroll
    for (int d = 0; d < 3; ++d) { xr[d] = wave_sum(xr[d]); xi[d] = wave_sum(xi[d]); }
    const double sg = kb.lma_sign[il];
    const double Ur = ur * g.Hxyz * sg, Ui = ui * g.Hxyz * sg;  // uVpsi (with iuV)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double Xr = xr[d] * g.Hxyz, Xi = xi[d] * g.Hxyz;
      jn[d] += occ / g.aLxyz * 2.0 * (Xr * Ui - Xi * Ur);  // Im(conj(X) U)
    }
    const double hr = ur * g.Hxyz, hi = ui * g.Hxyz;
    en += occ * (hr * hr + hi * hi) * sg;
  }
  if (lane == 0) { nl[w][0] = jn[0]; nl[w][1] = jn[1]; nl[w][2] = jn[2]; nl[w][3] = en; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double J[3], Enl = 0;
    for (int d = 0; d < 3; ++d) J[d] = (nl[0][d] + nl[1][d]) + (nl[2][d] + nl[3][d]);
    Enl = (nl[0][3] + nl[1][3]) + (nl[2][3] + nl[3][3]);
    const double s = g.Hxyz / g.aLxyz;
    for (int d = 0; d < 3; ++d) r[d] = (occ * kk[d] * v[0] + v[3 + d] * occ) * s + J[d];
    r[3] = occ * (A0 * v[0] + v[2] + kx * v[3] + ky * v[4] + kz * v[5]) * g.Hxyz +
           occ * (kx * kx + ky * ky + kz * kz) / 2.0;
    r[4] = Enl;
    r[5] = occ * v[1] * g.Hxyz;
    r[6] = occ * v[0] * g.Hxyz;
    r[7] = 0.0;
  }
}

template <typename T>
__global__ void __launch_bounds__(256) k_observe(const cplx<T>* __restrict__ zu, int orb0, int NB, int NK,
                                                 const double* __restrict__ kac0, double ax, double ay, double az,
                                                 const double* __restrict__ occv, const double* __restrict__ vloc,
                                                 Grid g, KBTables kb, double* __restrict__ out) {
  const int o = blockIdx.x;
  const int ikb = orb0 + o;
  double kx, ky, kz;
  orbital_k(kac0, NK, ikb / NB, ax, ay, az, kx, ky, kz);
  observe_orbital<T>(zu + (size_t)o * g.NL, kx, ky, kz, occv[ikb], vloc, g, kb, out + (size_t)ikb * 8);
}

// Deterministic sum of per-orbital rows (NKB, 8) -> dst[0..7]; one block.
__global__ void __launch_bounds__(256) k_sum_rows(const double* __restrict__ rows, int nrows,
                                                  double* __restrict__ dst) {
  __shared__ double smem[4 * 8];
  double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int r = threadIdx.x; r < nrows; r += 256) {
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] += rows[(size_t)r * 8 + i];
  }
  block_sum256<8>(v, smem);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < 8; ++i) dst[i] = v[i];
  }
}

// Density partials: part(c, p) = sum over orbitals of chunk c of occ |psi(p)|^2
template <typename T>
__global__ void __launch_bounds__(256) k_density_part(const cplx<T>* __restrict__ zu, int NKB, int NL, int per_chunk,
                                                      const double* __restrict__ occv, double* __restrict__ part) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  const int c = blockIdx.y;
  if (p >= NL) return;
  const int o0 = c * per_chunk, o1 = min(NKB, o0 + per_chunk);
  double r = 0;
  for (int o = o0; o < o1; ++o) {
    const cplx<T> e = zu[(size_t)o * NL + p];
    const double er = e.re, ei = e.im;
    r += occv[o] * (er * er + ei * ei);
  }
  part[(size_t)c * NL + p] = r;
}

__global__ void __launch_bounds__(256) k_density_final(const double* __restrict__ part, int nchunk, int NL,
                                                       double* __restrict__ rho) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= NL) return;
  double r = 0;
#pragma unroll 8
  for (int c = 0; c < nchunk; ++c) r += part[(size_t)c * NL + p];
  rho[p] = r;
}

// First stage of the density reduction: slice s of the nchunk partials,
// summed in chunk order -> part2(s, p).  The second stage (k_density_final over
// the slices) completes a fixed-order sum.
__global__ void __launch_bounds__(256) k_density_slices(const double* __restrict__ part, int nchunk, int per_slice,
                                                        int NL, double* __restrict__ part2) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  const int sl = blockIdx.y;
  if (p >= NL) return;
  const int c0 = sl * per_slice, c1 = min(nchunk, c0 + per_slice);
  double r = 0;
#pragma unroll 8
  for (int c = c0; c < c1; ++c) r += part[(size_t)c * NL + p];
  part2[(size_t)sl * NL + p] = r;
}

// <x|y> Hxyz over rows of (n, NL) pairs: dst(o) = sum conj(x) y * Hxyz (complex)
template <typename T>
__global__ void __launch_bounds__(256) k_dot_rows(const cplx<T>* __restrict__ x, const cplx<T>* __restrict__ y,
                                                  int NL, double Hxyz, double* __restrict__ dst) {
  __shared__ double smem[4 * 2];
  const cplx<T>* a = x + (size_t)blockIdx.x * NL;
  const cplx<T>* b = y + (size_t)blockIdx.x * NL;
  double v[2] = {0, 0};
  for (int p = threadIdx.x; p < NL; p += 256) {
    const double ar = a[p].re, ai = a[p].im, br = b[p].re, bi = b[p].im;
    v[0] += ar * br + ai * bi;
    v[1] += ar * bi - ai * br;
  }
  block_sum256<2>(v, smem);
  if (threadIdx.x == 0) { dst[2 * blockIdx.x] = v[0] * Hxyz; dst[2 * blockIdx.x + 1] = v[1] * Hxyz; }
}

template <typename S, typename D>
__global__ void __launch_bounds__(256) k_convert(const cplx<S>* __restrict__ src, cplx<D>* __restrict__ dst, size_t n) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) dst[i] = cplx<D>{(D)src[i].re, (D)src[i].im};
}

}  // namespace salmon
