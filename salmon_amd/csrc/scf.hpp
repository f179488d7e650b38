// Self-consistent local potential of the RT step (ARTED/RT/dt_evolve.f90:
// 185-207 with no_update_func = 'n'): from the density rho(t+dt), the Hartree
// potential (ARTED/common/Hartree.f90:20-128: rho(G) by one DFT pass per axis,
// V_H(G) = 4 pi rho(G)/|G|^2 without G = 0, back by three passes) and the
// PZ81 LDA exchange-correlation potential (src/xc/builtin_pz.f90:24-80 as
// called by src/xc/salmon_xc.f90:499-511), Vloc = V_H + Vpsl + V_xc.
// Orthorhombic cells.  FP64 throughout; the complex64 orbitals' potential copy
// is refreshed from the FP64 one.
#pragma once

namespace salmon {
namespace scf {

constexpr int kMaxN = 256;  // points per axis of the DFT passes (twiddle table rows)

// One DFT pass along `axis` over complex rows (one thread per output point):
// out(l) = sum_m in(m) e^{sgn 2 pi i l m / N}, twiddles tw[axis][j] =
// (cos, sin)(2 pi j / N) in the order the oracle sums.
__global__ void __launch_bounds__(256) k_dft_axis(const cplx<double>* __restrict__ in, cplx<double>* __restrict__ out,
                                                  Grid g, int axis, int sgn, const double* __restrict__ tw) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= g.NL) return;
  const int N = axis == 0 ? g.NLx : (axis == 1 ? g.NLy : g.NLz);
  const int es = axis == 0 ? g.NLy * g.NLz : (axis == 1 ? g.NLz : 1);
  const int l = (p / es) % N;
  const int base = p - l * es;
  const double* t = tw + (size_t)axis * 2 * kMaxN;
  double re = 0.0, im = 0.0;
  int j = 0;
  for (int m = 0; m < N; ++m) {
    const double c = t[2 * j], s = sgn * t[2 * j + 1];
    const cplx<double> v = in[base + m * es];
    re += c * v.re - s * v.im;
    im += c * v.im + s * v.re;
    j += l;
    if (j >= N) j -= N;
  }
  out[p] = cplx<double>{re, im};
}

// rho (real) into a complex row
__global__ void __launch_bounds__(256) k_real_to_cplx(const double* __restrict__ rho, cplx<double>* __restrict__ f,
                                                      int NL) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p < NL) f[p] = cplx<double>{rho[p], 0.0};
}

// F(G) *= 4 pi / |G|^2 / NL, G = 0 dropped
__global__ void __launch_bounds__(256) k_coulomb(cplx<double>* __restrict__ F, Grid g) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= g.NL) return;
  const int ix = p / (g.NLy * g.NLz), iy = (p / g.NLz) % g.NLy, iz = p % g.NLz;
  const int nx = ix < g.NLx - g.NLx / 2 ? ix : ix - g.NLx;
  const int ny = iy < g.NLy - g.NLy / 2 ? iy : iy - g.NLy;
  const int nz = iz < g.NLz - g.NLz / 2 ? iz : iz - g.NLz;
  const double pi = 3.141592653589793;
  const double gx = 2.0 * pi * nx / g.aLx, gy = 2.0 * pi * ny / g.aLy, gz = 2.0 * pi * nz / g.aLz;
  const double G2 = gx * gx + gy * gy + gz * gz;
  const double f = (nx == 0 && ny == 0 && nz == 0) ? 0.0 : 4.0 * pi / G2 / g.NL;
  F[p] = cplx<double>{f * F[p].re, f * F[p].im};
}

// PZ81 e_xc and its density derivative at the total density t (with the
// reference's + 1e-10 and its branch at rs = 1); returns v_xc.
__device__ __forceinline__ double pz_vxc(double rho) {
  const double pi = 3.141592653589793;
  const double gU = -0.1423, b1U = 1.0529, b2U = 0.3334, AU = 0.0311, BU = -0.048, CU = 0.002, DU = -0.0116;
  const double cst = 0.75 * pow(3.0 / (2.0 * pi), 2.0 / 3.0);
  const double t = rho + 1e-10;
  const double rs = pow(3.0 / (4.0 * pi * t), 1.0 / 3.0);
  double e = -cst / rs, de = e / (3.0 * t);
  if (rs > 1.0) {
    const double q = sqrt(rs), den = 1.0 + b1U * q + b2U * rs;
    e += gU / den;
    de += gU * (0.5 * b1U * q + b2U * rs) / (3.0 * t) / (den * den);
  } else {
    const double ln = log(rs);
    e += AU * ln + BU + CU * rs * ln + DU * rs;
    de -= rs / (3.0 * t) * (AU / rs + CU * (ln + 1.0) + DU);
  }
  return e + rho * de;
}

// Vloc = V_H + Vpsl + V_xc into the FP64 potential and its FP32 copy;
// vh/vxc (nullable) get the parts
__global__ void __launch_bounds__(256) k_vloc(const cplx<double>* __restrict__ vhc, const double* __restrict__ rho,
                                              const double* __restrict__ vpsl, int xc, int NL, double* __restrict__ v64,
                                              float* __restrict__ v32, double* __restrict__ vh, double* __restrict__ vxc) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= NL) return;
  const double h = vhc[p].re;
  const double x = xc ? pz_vxc(rho[p]) : 0.0;
  const double v = h + (vpsl ? vpsl[p] : 0.0) + x;
  v64[p] = v;
  if (v32) v32[p] = (float)v;
  if (vh) vh[p] = h;
  if (vxc) vxc[p] = x;
}

// ETRS potential extrapolation (dt_evolve.f90:247-251): Vnew = 3 V - 3 V_old1 +
// V_old2, the history shifts (old2 <- old1, old1 <- V) and V <- Vnew with its
// FP32 copy
__global__ void __launch_bounds__(256) k_vloc_extrapolate(double* __restrict__ v64, float* __restrict__ v32,
                                                          double* __restrict__ old1, double* __restrict__ old2,
                                                          int NL) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= NL) return;
  const double v = v64[p], o1 = old1[p];
  const double vn = 3.0 * v - 3.0 * o1 + old2[p];
  old2[p] = o1;
  old1[p] = v;
  v64[p] = vn;
  v32[p] = (float)vn;
}

}  // namespace scf
}  // namespace salmon
