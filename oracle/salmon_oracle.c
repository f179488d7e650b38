/*
 * CPU ORACLE for the ARTED real-time hot path.  TEST INFRASTRUCTURE ONLY:
 * it is the checker that the HIP library is compared against; nothing in the
 * product links or calls it.  Plain C99 + OpenMP, FP64 throughout.
 *
 * Each routine restates one reference routine (paths relative to the SALMON
 * source tree) and keeps its summation order:
 *   or_update_projector  ARTED/common/projector.f90:49-79
 *   or_hpsi              ARTED/common/hpsi.f90:98-207 (hpsi_omp_KB_base)
 *                        + ARTED/stencil/F90/hpsi.f90:27-198 (hpsi1_RT_stencil)
 *   or_hamiltonian       ARTED/RT/hamiltonian.f90:27-123 (Taylor-4)
 *   or_current           ARTED/RT/current.f90:78-292 (mode 'RT')
 *                        + ARTED/stencil/F90/current.f90:16-115
 *   or_density           ARTED/common/psi_rho.f90:40-206 (Sym=1)
 *   or_energy            ARTED/common/total_energy.f90:113-203 (Ekin, Enl)
 *                        + ARTED/stencil/F90/total_energy.f90:25-188
 *   or_gram_schmidt      ARTED/GS/Gram_Schmidt.f90:32-55
 *   or_sp_energy         ARTED/GS/sp_energy.f90:20-60
 *   or_cg                ARTED/GS/CG.f90:34-139 (CG_ompk, skip_gsortho='n');
 *                        the optional kinetic preconditioner has no reference
 *                        counterpart
 *   or_hartree           ARTED/common/Hartree.f90:20-128 (orthorhombic cell)
 *   or_pz                src/xc/builtin_pz.f90:24-80 (exc_cor_pz / PZxc) as
 *                        called by src/xc/salmon_xc.f90:499-511 (rho_s = rho/2)
 *   or_rt_run            ARTED/control/control_sc.f90:150-222 +
 *                        ARTED/RT/dt_evolve.f90:43-51,106-216 (middlepoint,
 *                        frozen potential: no_update_func='y')
 *
 * Layout (the reference's, Fortran column-major): orbital zu(NL,NB,NK)
 * complex(8) as interleaved (re,im); grid index i = ix*NLy*NLz+iy*NLz+iz,
 * z fastest; kAc(NK,3); occ(NB,NK); Jxyz/Jxx/Jyy/Jzz(Nps,NI);
 * uV(Nps,Nlma); zproj(Nps,Nlma).
 *
 * Reductions that the reference leaves to OpenMP (any order) are summed here
 * sequentially over (k, band), so the oracle is deterministic.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
  int NLx, NLy, NLz, NL;
  double Hx, Hy, Hz, Hxyz, aLx, aLy, aLz, aLxyz;
  double lapx[5], lapy[5], lapz[5], nabx[5], naby[5], nabz[5];
  const double *vloc;                 /* (NL) */
  int NI, Nps, Nlma;
  const int *Mps;                     /* (NI) */
  const int *Jxyz, *Jxx, *Jyy, *Jzz;  /* (Nps,NI), Jxyz 1-based */
  const int *a_tbl;                   /* (Nlma), 1-based */
  const double *uV;                   /* (Nps,Nlma) */
  const int *iuV;                     /* (Nlma) */
  int NK, NB;
  const double *occ;                  /* (NB,NK) */
} or_system;

typedef struct { double re, im; } zc;

static inline zc zmk(double r, double i) { zc z = {r, i}; return z; }
static inline zc zadd(zc a, zc b) { return zmk(a.re + b.re, a.im + b.im); }
static inline zc zsub(zc a, zc b) { return zmk(a.re - b.re, a.im - b.im); }
/* Fortran complex multiply (a.re*b.re - a.im*b.im, a.re*b.im + a.im*b.re) */
static inline zc zmul(zc a, zc b) { return zmk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
static inline zc zscale(double s, zc a) { return zmk(s * a.re, s * a.im); }
static inline zc zconj(zc a) { return zmk(a.re, -a.im); }
static inline double zabs2(zc a) { double h = hypot(a.re, a.im); return h * h; }

static inline int LX(const or_system *s, int i0) { return i0 / (s->NLy * s->NLz); }
static inline int LY(const or_system *s, int i0) { return (i0 / s->NLz) % s->NLy; }
static inline int LZ(const or_system *s, int i0) { return i0 % s->NLz; }

/* projector.f90:57-75: ekr = exp(i k.(r - T)), zproj = conj(ekr)*uV */
void or_update_projector(const or_system *s, const double k[3], double *zproj_out)
{
  zc *zproj = (zc *)zproj_out;
  for (int ilma = 0; ilma < s->Nlma; ++ilma) {
    int ia = s->a_tbl[ilma] - 1;
    for (int j = 0; j < s->Mps[ia]; ++j) {
      int i0 = s->Jxyz[ia * s->Nps + j] - 1;
      int ix = s->Jxx[ia * s->Nps + j], iy = s->Jyy[ia * s->Nps + j], iz = s->Jzz[ia * s->Nps + j];
      double kr = k[0] * (LX(s, i0) * s->Hx - ix * s->aLx) + k[1] * (LY(s, i0) * s->Hy - iy * s->aLy)
                + k[2] * (LZ(s, i0) * s->Hz - iz * s->aLz);
      zc ekr = zmk(cos(kr), sin(kr));
      zproj[ilma * s->Nps + j] = zscale(s->uV[ilma * s->Nps + j], zconj(ekr));
    }
  }
}

/* idx_lma(ia): number of projectors of the atoms before ia (opt_variables.f90:241-250) */
static void projector_offsets(const or_system *s, int *nproj, int *idx_lma, int *start)
{
  for (int a = 0; a < s->NI; ++a) nproj[a] = 0;
  for (int l = 0; l < s->Nlma; ++l) nproj[s->a_tbl[l] - 1]++;
  idx_lma[0] = 0;
  for (int a = 1; a < s->NI; ++a) idx_lma[a] = idx_lma[a - 1] + nproj[a - 1];
  int off = 0;
  for (int a = 0; a < s->NI; ++a) { start[a] = off; off += s->Mps[a]; }
}

#define IDX(ix, iy, iz) (((ix) * s->NLy + (iy)) * s->NLz + (iz))

/*
 * hpsi_omp_KB_base (hpsi.f90:98-207) on one orbital.  ttpsi (nullable) gets
 * the stencil result minus V*psi before the nonlocal term, as subtraction()
 * at hpsi.f90:131-133,144-160.
 */
void or_hpsi(const or_system *s, const double k[3], const double *zproj_in,
             const double *in_, double *out_, double *tt_)
{
  const zc *E = (const zc *)in_;
  zc *F = (zc *)out_;
  zc *TT = (zc *)tt_;
  const zc *zproj = (const zc *)zproj_in;
  const int NLx = s->NLx, NLy = s->NLy, NLz = s->NLz;
  double k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
  double A = (k2 - (s->lapx[0] + s->lapy[0] + s->lapz[0])) * 0.5;
  double C[13], D[13];
  for (int n = 1; n <= 4; ++n) {
    C[n] = s->lapx[n]; C[4 + n] = s->lapy[n]; C[8 + n] = s->lapz[n];
    D[n] = k[0] * s->nabx[n]; D[4 + n] = k[1] * s->naby[n]; D[8 + n] = k[2] * s->nabz[n];
  }
  /* hpsi1_RT_stencil, F90/hpsi.f90:88-193 (non-"optimized load" form) */
  for (int ix = 0; ix < NLx; ++ix)
    for (int iy = 0; iy < NLy; ++iy)
      for (int iz = 0; iz < NLz; ++iz) {
        zc v, w, t;
        /* z */
        v = zscale(C[9], zadd(E[IDX(ix, iy, (iz + 1) % NLz)], E[IDX(ix, iy, (iz - 1 + NLz) % NLz)]));
        v = zadd(v, zscale(C[10], zadd(E[IDX(ix, iy, (iz + 2) % NLz)], E[IDX(ix, iy, (iz - 2 + NLz) % NLz)])));
        v = zadd(v, zscale(C[11], zadd(E[IDX(ix, iy, (iz + 3) % NLz)], E[IDX(ix, iy, (iz - 3 + NLz) % NLz)])));
        v = zadd(v, zscale(C[12], zadd(E[IDX(ix, iy, (iz + 4) % NLz)], E[IDX(ix, iy, (iz - 4 + NLz) % NLz)])));
        w = zscale(D[9], zsub(E[IDX(ix, iy, (iz + 1) % NLz)], E[IDX(ix, iy, (iz - 1 + NLz) % NLz)]));
        w = zadd(w, zscale(D[10], zsub(E[IDX(ix, iy, (iz + 2) % NLz)], E[IDX(ix, iy, (iz - 2 + NLz) % NLz)])));
        w = zadd(w, zscale(D[11], zsub(E[IDX(ix, iy, (iz + 3) % NLz)], E[IDX(ix, iy, (iz - 3 + NLz) % NLz)])));
        w = zadd(w, zscale(D[12], zsub(E[IDX(ix, iy, (iz + 4) % NLz)], E[IDX(ix, iy, (iz - 4 + NLz) % NLz)])));
        /* y */
        t = zscale(C[5], zadd(E[IDX(ix, (iy + 1) % NLy, iz)], E[IDX(ix, (iy - 1 + NLy) % NLy, iz)]));
        t = zadd(t, zscale(C[6], zadd(E[IDX(ix, (iy + 2) % NLy, iz)], E[IDX(ix, (iy - 2 + NLy) % NLy, iz)])));
        t = zadd(t, zscale(C[7], zadd(E[IDX(ix, (iy + 3) % NLy, iz)], E[IDX(ix, (iy - 3 + NLy) % NLy, iz)])));
        t = zadd(t, zscale(C[8], zadd(E[IDX(ix, (iy + 4) % NLy, iz)], E[IDX(ix, (iy - 4 + NLy) % NLy, iz)])));
        v = zadd(t, v);
        t = zscale(D[5], zsub(E[IDX(ix, (iy + 1) % NLy, iz)], E[IDX(ix, (iy - 1 + NLy) % NLy, iz)]));
        t = zadd(t, zscale(D[6], zsub(E[IDX(ix, (iy + 2) % NLy, iz)], E[IDX(ix, (iy - 2 + NLy) % NLy, iz)])));
        t = zadd(t, zscale(D[7], zsub(E[IDX(ix, (iy + 3) % NLy, iz)], E[IDX(ix, (iy - 3 + NLy) % NLy, iz)])));
        t = zadd(t, zscale(D[8], zsub(E[IDX(ix, (iy + 4) % NLy, iz)], E[IDX(ix, (iy - 4 + NLy) % NLy, iz)])));
        w = zadd(t, w);
        /* x */
        t = zscale(C[1], zadd(E[IDX((ix + 1) % NLx, iy, iz)], E[IDX((ix - 1 + NLx) % NLx, iy, iz)]));
        t = zadd(t, zscale(C[2], zadd(E[IDX((ix + 2) % NLx, iy, iz)], E[IDX((ix - 2 + NLx) % NLx, iy, iz)])));
        t = zadd(t, zscale(C[3], zadd(E[IDX((ix + 3) % NLx, iy, iz)], E[IDX((ix - 3 + NLx) % NLx, iy, iz)])));
        t = zadd(t, zscale(C[4], zadd(E[IDX((ix + 4) % NLx, iy, iz)], E[IDX((ix - 4 + NLx) % NLx, iy, iz)])));
        v = zadd(t, v);
        t = zscale(D[1], zsub(E[IDX((ix + 1) % NLx, iy, iz)], E[IDX((ix - 1 + NLx) % NLx, iy, iz)]));
        t = zadd(t, zscale(D[2], zsub(E[IDX((ix + 2) % NLx, iy, iz)], E[IDX((ix - 2 + NLx) % NLx, iy, iz)])));
        t = zadd(t, zscale(D[3], zsub(E[IDX((ix + 3) % NLx, iy, iz)], E[IDX((ix - 3 + NLx) % NLx, iy, iz)])));
        t = zadd(t, zscale(D[4], zsub(E[IDX((ix + 4) % NLx, iy, iz)], E[IDX((ix - 4 + NLx) % NLx, iy, iz)])));
        w = zadd(t, w);
        /* F = B*E + A*E - 0.5*v - zI*w */
        int c = IDX(ix, iy, iz);
        double B = s->vloc[c];
        zc f = zadd(zscale(B, E[c]), zscale(A, E[c]));
        f = zsub(f, zscale(0.5, v));
        f = zsub(f, zmk(-w.im, w.re));
        F[c] = f;
      }
  if (TT) {
    for (int i = 0; i < s->NL; ++i) TT[i] = zsub(F[i], zscale(s->vloc[i], E[i]));
  }
  /* pseudo_pt, hpsi.f90:163-207 */
  int NI = s->NI;
  int *nproj = (int *)malloc(sizeof(int) * 3 * NI);
  int *idx_lma = nproj + NI, *start = nproj + 2 * NI;
  projector_offsets(s, nproj, idx_lma, start);
  int NPI = start[NI - 1] + s->Mps[NI - 1];
  zc *sp = (zc *)malloc(sizeof(zc) * 2 * NPI);
  zc *dp = sp + NPI;
  for (int a = 0; a < NI; ++a)
    for (int j = 0; j < s->Mps[a]; ++j) {
      sp[start[a] + j] = E[s->Jxyz[a * s->Nps + j] - 1];
      dp[start[a] + j] = zmk(0.0, 0.0);
    }
  for (int a = 0; a < NI; ++a)
    for (int ip = 0; ip < nproj[a]; ++ip) {
      int il = idx_lma[a] + ip;
      zc u = zmk(0.0, 0.0);
      for (int j = 0; j < s->Mps[a]; ++j) u = zadd(u, zmul(zconj(zproj[il * s->Nps + j]), sp[start[a] + j]));
      u = zscale(s->Hxyz, u);
      u = zscale((double)s->iuV[il], u);
      for (int j = 0; j < s->Mps[a]; ++j) dp[start[a] + j] = zadd(dp[start[a] + j], zmul(zproj[il * s->Nps + j], u));
    }
  for (int a = 0; a < NI; ++a)
    for (int j = 0; j < s->Mps[a]; ++j) {
      int i0 = s->Jxyz[a * s->Nps + j] - 1;
      F[i0] = zadd(F[i0], dp[start[a] + j]);
    }
  free(sp);
  free(nproj);
}

static inline const double *kvec(const double *kAc, int NK, int ik, double out[3])
{
  out[0] = kAc[ik]; out[1] = kAc[NK + ik]; out[2] = kAc[2 * NK + ik];
  return out;
}

/* hamiltonian.f90:46-49: zfac(1) = -i dt, zfac(n) = zfac(n-1)*(-i dt)/n */
static void taylor_coefficients(double dt, zc zfac[4])
{
  zc m = zmk(-0.0 * dt, -1.0 * dt);
  zfac[0] = m;
  for (int n = 1; n < 4; ++n) {
    zc t = zmul(zfac[n - 1], m);
    zfac[n] = zmk(t.re / (n + 1), t.im / (n + 1));
  }
}

/* hamiltonian.f90:53-73: Taylor-4 for every (k,band); kAc is the midpoint k+A */
void or_hamiltonian(const or_system *s, const double *kAc, double dt, double *zu_)
{
  zc zfac[4];
  taylor_coefficients(dt, zfac);
  const int NL = s->NL, NKB = s->NK * s->NB;
  size_t zsz = (size_t)s->Nps * s->Nlma;
  double *zproj = (double *)malloc(sizeof(zc) * zsz * s->NK);
  for (int ik = 0; ik < s->NK; ++ik) {
    double k[3];
    or_update_projector(s, kvec(kAc, s->NK, ik, k), zproj + 2 * zsz * ik);
  }
#pragma omp parallel
  {
    zc *h = (zc *)malloc(sizeof(zc) * 5 * NL);
#pragma omp for schedule(static)
    for (int ikb = 0; ikb < NKB; ++ikb) {
      int ik = ikb / s->NB;
      double k[3];
      kvec(kAc, s->NK, ik, k);
      zc *zu = (zc *)zu_ + (size_t)ikb * NL;
      const double *zp = zproj + 2 * zsz * ik;
      memcpy(h, zu, sizeof(zc) * NL);
      or_hpsi(s, k, zp, (double *)h, (double *)(h + NL), NULL);
      or_hpsi(s, k, zp, (double *)(h + NL), (double *)(h + 2 * NL), NULL);
      or_hpsi(s, k, zp, (double *)(h + 2 * NL), (double *)(h + 3 * NL), NULL);
      or_hpsi(s, k, zp, (double *)(h + 3 * NL), (double *)(h + 4 * NL), NULL);
      for (int i = 0; i < NL; ++i) {
        zc z = zu[i];
        z = zadd(z, zmul(zfac[0], h[NL + i]));
        z = zadd(z, zmul(zfac[1], h[2 * NL + i]));
        z = zadd(z, zmul(zfac[2], h[3 * NL + i]));
        z = zadd(z, zmul(zfac[3], h[4 * NL + i]));
        zu[i] = z;
      }
    }
    free(h);
  }
  free(zproj);
}

/* current_stencil, stencil/F90/current.f90:62-114: returns 2*sum Im(conj(E)*sum_n C_n E[+n]) per axis */
static void current_stencil(const or_system *s, const double C[13], const zc *E, double *Fx, double *Gy, double *Hz)
{
  const int NLx = s->NLx, NLy = s->NLy, NLz = s->NLz;
  double H = 0, G = 0, F = 0;
  for (int ix = 0; ix < NLx; ++ix)
    for (int iy = 0; iy < NLy; ++iy)
      for (int iz = 0; iz < NLz; ++iz) {
        zc w = zconj(E[IDX(ix, iy, iz)]), v;
        v = zscale(C[9], E[IDX(ix, iy, (iz + 1) % NLz)]);
        v = zadd(v, zscale(C[10], E[IDX(ix, iy, (iz + 2) % NLz)]));
        v = zadd(v, zscale(C[11], E[IDX(ix, iy, (iz + 3) % NLz)]));
        v = zadd(v, zscale(C[12], E[IDX(ix, iy, (iz + 4) % NLz)]));
        H += zmul(w, v).im;
        v = zscale(C[5], E[IDX(ix, (iy + 1) % NLy, iz)]);
        v = zadd(v, zscale(C[6], E[IDX(ix, (iy + 2) % NLy, iz)]));
        v = zadd(v, zscale(C[7], E[IDX(ix, (iy + 3) % NLy, iz)]));
        v = zadd(v, zscale(C[8], E[IDX(ix, (iy + 4) % NLy, iz)]));
        G += zmul(w, v).im;
        v = zscale(C[1], E[IDX((ix + 1) % NLx, iy, iz)]);
        v = zadd(v, zscale(C[2], E[IDX((ix + 2) % NLx, iy, iz)]));
        v = zadd(v, zscale(C[3], E[IDX((ix + 3) % NLx, iy, iz)]));
        v = zadd(v, zscale(C[4], E[IDX((ix + 4) % NLx, iy, iz)]));
        F += zmul(w, v).im;
      }
  *Hz = H * 2.0; *Gy = G * 2.0; *Fx = F * 2.0;
}

/* current(mode='RT'), RT/current.f90:106-273; jav(3) is the summed current */
void or_current(const or_system *s, const double *kAc, const double *zu_, double *jav)
{
  const int NL = s->NL, NKB = s->NK * s->NB;
  double C[13];
  for (int n = 1; n <= 4; ++n) { C[n] = s->nabx[n]; C[4 + n] = s->naby[n]; C[8 + n] = s->nabz[n]; }
  double IaLxyz = 1.0 / s->aLxyz;
  size_t zsz = (size_t)s->Nps * s->Nlma;
  double *zproj = (double *)malloc(sizeof(zc) * zsz * s->NK);
  for (int ik = 0; ik < s->NK; ++ik) {
    double k[3];
    or_update_projector(s, kvec(kAc, s->NK, ik, k), zproj + 2 * zsz * ik);
  }
  double *part = (double *)malloc(sizeof(double) * 3 * NKB);
#pragma omp parallel for schedule(static)
  for (int ikb = 0; ikb < NKB; ++ikb) {
    int ik = ikb / s->NB, ib = ikb % s->NB;
    const zc *z = (const zc *)zu_ + (size_t)ikb * NL;
    double occ = s->occ[ik * s->NB + ib];
    double k[3];
    kvec(kAc, s->NK, ik, k);
    /* init, :182-190 */
    double jt = 0.0;
    for (int i = 0; i < NL; ++i) jt = jt + z[i].re * z[i].re + z[i].im * z[i].im;
    double jx = occ * k[0] * jt, jy = occ * k[1] * jt, jz = occ * k[2] * jt;
    /* stencil, :203-210 */
    double jxt, jyt, jzt;
    current_stencil(s, C, z, &jxt, &jyt, &jzt);
    jx = jx + jxt * occ; jy = jy + jyt * occ; jz = jz + jzt * occ;
    /* pseudo_pt, :243-273 */
    const zc *zp = (const zc *)(zproj + 2 * zsz * ik);
    double nx = 0, ny = 0, nz = 0;
    for (int il = 0; il < s->Nlma; ++il) {
      int ia = s->a_tbl[il] - 1;
      zc u = zmk(0, 0), ux = zmk(0, 0), uy = zmk(0, 0), uz = zmk(0, 0);
      for (int j = 0; j < s->Mps[ia]; ++j) {
        int i0 = s->Jxyz[ia * s->Nps + j] - 1;
        double x = LX(s, i0) * s->Hx - s->Jxx[ia * s->Nps + j] * s->aLx;
        double y = LY(s, i0) * s->Hy - s->Jyy[ia * s->Nps + j] * s->aLy;
        double zz = LZ(s, i0) * s->Hz - s->Jzz[ia * s->Nps + j] * s->aLz;
        zc cp = zconj(zp[il * s->Nps + j]);
        u = zadd(u, zmul(cp, z[i0]));
        ux = zadd(ux, zmul(zscale(x, cp), z[i0]));
        uy = zadd(uy, zmul(zscale(y, cp), z[i0]));
        uz = zadd(uz, zmul(zscale(zz, cp), z[i0]));
      }
      u = zscale((double)s->iuV[il], zscale(s->Hxyz, u));
      ux = zscale(s->Hxyz, ux); uy = zscale(s->Hxyz, uy); uz = zscale(s->Hxyz, uz);
      nx = nx + occ * IaLxyz * 2 * zmul(zconj(ux), u).im;
      ny = ny + occ * IaLxyz * 2 * zmul(zconj(uy), u).im;
      nz = nz + occ * IaLxyz * 2 * zmul(zconj(uz), u).im;
    }
    part[3 * ikb + 0] = jx * s->Hxyz * IaLxyz + nx;
    part[3 * ikb + 1] = jy * s->Hxyz * IaLxyz + ny;
    part[3 * ikb + 2] = jz * s->Hxyz * IaLxyz + nz;
  }
  jav[0] = jav[1] = jav[2] = 0.0;
  for (int ikb = 0; ikb < NKB; ++ikb) {
    jav[0] += part[3 * ikb]; jav[1] += part[3 * ikb + 1]; jav[2] += part[3 * ikb + 2];
  }
  free(part);
  free(zproj);
}

/* psi_rho_RT with Sym=1 (psi_rho.f90:130-144): rho(i) = sum_{k,b} occ*|zu|^2 */
void or_density(const or_system *s, const double *zu_, double *rho)
{
  const int NL = s->NL;
  const zc *zu = (const zc *)zu_;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < NL; ++i) {
    double r = 0.0;
    for (int ik = 0; ik < s->NK; ++ik)
      for (int ib = 0; ib < s->NB; ++ib)
        r = r + (1.0 * s->occ[ik * s->NB + ib]) * zabs2(zu[((size_t)ik * s->NB + ib) * NL + i]);
    rho[i] = r;
  }
}

/* total_energy_stencil, stencil/F90/total_energy.f90:75-187 */
static zc energy_stencil(const or_system *s, double A, const double C[13], const double D[13], const zc *E)
{
  const int NLx = s->NLx, NLy = s->NLy, NLz = s->NLz;
  zc F = zmk(0, 0);
  for (int ix = 0; ix < NLx; ++ix)
    for (int iy = 0; iy < NLy; ++iy)
      for (int iz = 0; iz < NLz; ++iz) {
        zc v, w, t, z;
        z = zscale(A, E[IDX(ix, iy, iz)]);
        v = zscale(C[9], zadd(E[IDX(ix, iy, (iz + 1) % NLz)], E[IDX(ix, iy, (iz - 1 + NLz) % NLz)]));
        v = zadd(v, zscale(C[10], zadd(E[IDX(ix, iy, (iz + 2) % NLz)], E[IDX(ix, iy, (iz - 2 + NLz) % NLz)])));
        v = zadd(v, zscale(C[11], zadd(E[IDX(ix, iy, (iz + 3) % NLz)], E[IDX(ix, iy, (iz - 3 + NLz) % NLz)])));
        v = zadd(v, zscale(C[12], zadd(E[IDX(ix, iy, (iz + 4) % NLz)], E[IDX(ix, iy, (iz - 4 + NLz) % NLz)])));
        w = zscale(D[9], zsub(E[IDX(ix, iy, (iz + 1) % NLz)], E[IDX(ix, iy, (iz - 1 + NLz) % NLz)]));
        w = zadd(w, zscale(D[10], zsub(E[IDX(ix, iy, (iz + 2) % NLz)], E[IDX(ix, iy, (iz - 2 + NLz) % NLz)])));
        w = zadd(w, zscale(D[11], zsub(E[IDX(ix, iy, (iz + 3) % NLz)], E[IDX(ix, iy, (iz - 3 + NLz) % NLz)])));
        w = zadd(w, zscale(D[12], zsub(E[IDX(ix, iy, (iz + 4) % NLz)], E[IDX(ix, iy, (iz - 4 + NLz) % NLz)])));
        t = zscale(C[5], zadd(E[IDX(ix, (iy + 1) % NLy, iz)], E[IDX(ix, (iy - 1 + NLy) % NLy, iz)]));
        t = zadd(t, zscale(C[6], zadd(E[IDX(ix, (iy + 2) % NLy, iz)], E[IDX(ix, (iy - 2 + NLy) % NLy, iz)])));
        t = zadd(t, zscale(C[7], zadd(E[IDX(ix, (iy + 3) % NLy, iz)], E[IDX(ix, (iy - 3 + NLy) % NLy, iz)])));
        t = zadd(t, zscale(C[8], zadd(E[IDX(ix, (iy + 4) % NLy, iz)], E[IDX(ix, (iy - 4 + NLy) % NLy, iz)])));
        v = zadd(t, v);
        t = zscale(D[5], zsub(E[IDX(ix, (iy + 1) % NLy, iz)], E[IDX(ix, (iy - 1 + NLy) % NLy, iz)]));
        t = zadd(t, zscale(D[6], zsub(E[IDX(ix, (iy + 2) % NLy, iz)], E[IDX(ix, (iy - 2 + NLy) % NLy, iz)])));
        t = zadd(t, zscale(D[7], zsub(E[IDX(ix, (iy + 3) % NLy, iz)], E[IDX(ix, (iy - 3 + NLy) % NLy, iz)])));
        t = zadd(t, zscale(D[8], zsub(E[IDX(ix, (iy + 4) % NLy, iz)], E[IDX(ix, (iy - 4 + NLy) % NLy, iz)])));
        w = zadd(t, w);
        t = zscale(C[1], zadd(E[IDX((ix + 1) % NLx, iy, iz)], E[IDX((ix - 1 + NLx) % NLx, iy, iz)]));
        t = zadd(t, zscale(C[2], zadd(E[IDX((ix + 2) % NLx, iy, iz)], E[IDX((ix - 2 + NLx) % NLx, iy, iz)])));
        t = zadd(t, zscale(C[3], zadd(E[IDX((ix + 3) % NLx, iy, iz)], E[IDX((ix - 3 + NLx) % NLx, iy, iz)])));
        t = zadd(t, zscale(C[4], zadd(E[IDX((ix + 4) % NLx, iy, iz)], E[IDX((ix - 4 + NLx) % NLx, iy, iz)])));
        v = zadd(t, v);
        t = zscale(D[1], zsub(E[IDX((ix + 1) % NLx, iy, iz)], E[IDX((ix - 1 + NLx) % NLx, iy, iz)]));
        t = zadd(t, zscale(D[2], zsub(E[IDX((ix + 2) % NLx, iy, iz)], E[IDX((ix - 2 + NLx) % NLx, iy, iz)])));
        t = zadd(t, zscale(D[3], zsub(E[IDX((ix + 3) % NLx, iy, iz)], E[IDX((ix - 3 + NLx) % NLx, iy, iz)])));
        t = zadd(t, zscale(D[4], zsub(E[IDX((ix + 4) % NLx, iy, iz)], E[IDX((ix - 4 + NLx) % NLx, iy, iz)])));
        w = zadd(t, w);
        z = zsub(z, zscale(0.5, v));
        z = zsub(z, zmk(-w.im, w.re));
        F = zadd(F, zmul(zconj(E[IDX(ix, iy, iz)]), z));
      }
  return F;
}

/*
 * Wavefunction terms of Total_Energy_omp (total_energy.f90:113,172-203):
 * Ekin (with the |k+A|^2/2 term) and Enl.  eloc is the frozen-potential
 * local term sum_i V(i) rho(i) Hxyz of the synthetic model (the reference's
 * G-space Eloc/Eh/Eion/Exc are outside this path).
 */
void or_energy(const or_system *s, const double *kAc, const double *zu_, double *ekin, double *enl, double *eloc)
{
  const int NL = s->NL, NKB = s->NK * s->NB;
  double lap0_2 = -(s->lapx[0] + s->lapy[0] + s->lapz[0]) * 0.5;
  double *part = (double *)malloc(sizeof(double) * 2 * NKB);
  double *ekr = (double *)malloc(sizeof(zc) * (size_t)s->Nps * s->NI * s->NK);
  for (int ik = 0; ik < s->NK; ++ik)
    for (int ia = 0; ia < s->NI; ++ia)
      for (int j = 0; j < s->Mps[ia]; ++j) {
        int i0 = s->Jxyz[ia * s->Nps + j] - 1;
        double kr = kAc[ik] * (LX(s, i0) * s->Hx - s->Jxx[ia * s->Nps + j] * s->aLx)
                  + kAc[s->NK + ik] * (LY(s, i0) * s->Hy - s->Jyy[ia * s->Nps + j] * s->aLy)
                  + kAc[2 * s->NK + ik] * (LZ(s, i0) * s->Hz - s->Jzz[ia * s->Nps + j] * s->aLz);
        size_t o = 2 * (((size_t)ik * s->NI + ia) * s->Nps + j);
        ekr[o] = cos(kr); ekr[o + 1] = sin(kr);
      }
#pragma omp parallel for schedule(static)
  for (int ikb = 0; ikb < NKB; ++ikb) {
    int ik = ikb / s->NB, ib = ikb % s->NB;
    const zc *z = (const zc *)zu_ + (size_t)ikb * NL;
    double occ = s->occ[ik * s->NB + ib];
    double k[3];
    kvec(kAc, s->NK, ik, k);
    double C[13], D[13];
    for (int n = 1; n <= 4; ++n) {
      C[n] = s->lapx[n]; C[4 + n] = s->lapy[n]; C[8 + n] = s->lapz[n];
      D[n] = k[0] * s->nabx[n]; D[4 + n] = k[1] * s->naby[n]; D[8 + n] = k[2] * s->nabz[n];
    }
    zc tpsum = energy_stencil(s, lap0_2, C, D, z);
    part[2 * ikb] = (occ * tpsum.re) * s->Hxyz + occ * (k[0] * k[0] + k[1] * k[1] + k[2] * k[2]) / 2.0;
    double e = 0.0;
    for (int il = 0; il < s->Nlma; ++il) {
      int ia = s->a_tbl[il] - 1;
      zc u = zmk(0, 0);
      for (int j = 0; j < s->Mps[ia]; ++j) {
        size_t o = 2 * (((size_t)ik * s->NI + ia) * s->Nps + j);
        zc ek = zmk(ekr[o], ekr[o + 1]);
        u = zadd(u, zmul(zscale(s->uV[il * s->Nps + j], ek), z[s->Jxyz[ia * s->Nps + j] - 1]));
      }
      u = zscale(s->Hxyz, u);
      e = e + occ * zabs2(u) * s->iuV[il];
    }
    part[2 * ikb + 1] = e;
  }
  double ek = 0.0, en = 0.0;
  for (int ikb = 0; ikb < NKB; ++ikb) { ek += part[2 * ikb]; en += part[2 * ikb + 1]; }
  *ekin = ek; *enl = en;
  if (eloc) {
    double *rho = (double *)malloc(sizeof(double) * NL);
    or_density(s, zu_, rho);
    double el = 0.0;
    for (int i = 0; i < NL; ++i) el += s->vloc[i] * rho[i];
    *eloc = el * s->Hxyz;
    free(rho);
  }
  free(ekr);
  free(part);
}

/* total norm sum_{k,b} occ |psi|^2 Hxyz */
double or_norm(const or_system *s, const double *zu_)
{
  const zc *zu = (const zc *)zu_;
  double tot = 0.0;
  for (int ikb = 0; ikb < s->NK * s->NB; ++ikb) {
    double t = 0.0;
    for (int i = 0; i < s->NL; ++i) t += zu[(size_t)ikb * s->NL + i].re * zu[(size_t)ikb * s->NL + i].re
                                        + zu[(size_t)ikb * s->NL + i].im * zu[(size_t)ikb * s->NL + i].im;
    tot += s->occ[ikb] * t * s->Hxyz;
  }
  return tot;
}

/* Gram_Schmidt_ompk, GS/Gram_Schmidt.f90:41-51, on NB bands of every k */
void or_gram_schmidt(const or_system *s, double *zu_)
{
  const int NL = s->NL;
#pragma omp parallel for schedule(static)
  for (int ik = 0; ik < s->NK; ++ik) {
    zc *b0 = (zc *)zu_ + (size_t)ik * s->NB * NL;
    for (int ib = 0; ib < s->NB; ++ib) {
      zc *x = b0 + (size_t)ib * NL;
      for (int ibt = 0; ibt < ib; ++ibt) {
        zc *y = b0 + (size_t)ibt * NL;
        zc zov = zmk(0, 0);
        for (int i = 0; i < NL; ++i) zov = zadd(zov, zmul(zconj(y[i]), x[i]));
        zov = zscale(s->Hxyz, zov);
        for (int i = 0; i < NL; ++i) x[i] = zsub(x[i], zmul(y[i], zov));
      }
      double sn = 0.0;
      for (int i = 0; i < NL; ++i) sn += zabs2(x[i]);
      sn *= s->Hxyz;
      double r = sqrt(sn);
      for (int i = 0; i < NL; ++i) x[i] = zmk(x[i].re / r, x[i].im / r);
    }
  }
}

static zc zdot(const zc *a, const zc *b, int n)
{
  zc t = zmk(0, 0);
  for (int i = 0; i < n; ++i) t = zadd(t, zmul(zconj(a[i]), b[i]));
  return t;
}

static double znrm2(const zc *a, int n)
{
  double t = 0.0;
  for (int i = 0; i < n; ++i) t += zabs2(a[i]);
  return t;
}

/* sp_energy_omp, GS/sp_energy.f90:43-49: esp(NB,NK) = Re <psi|H|psi> Hxyz */
void or_sp_energy(const or_system *s, const double *kAc, const double *zu_, double *esp)
{
  const int NL = s->NL;
  size_t zsz = (size_t)s->Nps * s->Nlma;
#pragma omp parallel
  {
    double *zp = (double *)malloc(sizeof(zc) * zsz);
    zc *h = (zc *)malloc(sizeof(zc) * 2 * NL);
#pragma omp for schedule(static)
    for (int ikb = 0; ikb < s->NK * s->NB; ++ikb) {
      int ik = ikb / s->NB;
      double k[3];
      kvec(kAc, s->NK, ik, k);
      or_update_projector(s, k, zp);
      const zc *x = (const zc *)zu_ + (size_t)ikb * NL;
      or_hpsi(s, k, zp, (const double *)x, (double *)h, (double *)(h + NL));
      esp[ikb] = zdot(x, h, NL).re * s->Hxyz;
    }
    free(h);
    free(zp);
  }
}

/*
 * Kinetic-energy preconditioner K(g) (no reference counterpart: CG.f90 has
 * none).  Diagonal in the discrete plane-wave basis of the FD kinetic
 * operator: K = 1/(1 + t/Tx) with t the FD kinetic symbol of each plane
 * wave and Tx the band's kinetic energy <x|T|x> (a Teter-Payne-Allan-type
 * scale).  Applied by a naive DFT per axis (separable), adequate for an oracle.
 */
/*
 * Kinetic preconditioner (no reference counterpart; CG_ompk applies none,
 * GS/CG.f90:85-92): K g = F^-1[k(x) F g] with x = t(q)/Tx, t(q) the
 * finite-difference kinetic symbol -1/2 sum_axes (lap0 + 2 sum_m lap_m
 * cos(m 2 pi c/n)) and k(x) = f/(f + 16x^4), f = 27 + 18x + 12x^2 + 8x^3
 * (the Teter-Payne-Allan form).  F: unnormalised DFT along z, then y, then x;
 * F^-1: along x, y, z with 1/n per axis.  Twiddles come from the table
 * (cos, sin)(2 pi j/n) indexed by (c*m) mod n.
 */
static void dft_axis(const or_system *s, int ax, int sgn, const zc *in, zc *out)
{
  const int N[3] = {s->NLx, s->NLy, s->NLz};
  const int st[3] = {s->NLy * s->NLz, s->NLz, 1};
  const int n = N[ax], NL = s->NL;
  double *tc = (double *)malloc(sizeof(double) * 2 * n);
  for (int j = 0; j < n; ++j) {
    tc[2 * j] = cos(2.0 * M_PI * j / n);
    tc[2 * j + 1] = sin(2.0 * M_PI * j / n);
  }
  for (int i = 0; i < NL; ++i) {
    int c = (i / st[ax]) % n, base = i - c * st[ax];
    zc t = zmk(0, 0);
    for (int m = 0; m < n; ++m) {
      int j = (c * m) % n;
      t = zadd(t, zmul(zmk(tc[2 * j], sgn * tc[2 * j + 1]), in[base + m * st[ax]]));
    }
    out[i] = sgn > 0 ? zscale(1.0 / n, t) : t;
  }
  free(tc);
}

void or_kinetic_precondition(const or_system *s, double Tx, double *g_)
{
  zc *g = (zc *)g_;
  const int N[3] = {s->NLx, s->NLy, s->NLz};
  const double *lap[3] = {s->lapx, s->lapy, s->lapz};
  const int NL = s->NL;
  const int st[3] = {s->NLy * s->NLz, s->NLz, 1};
  zc *a = (zc *)malloc(sizeof(zc) * NL), *b = (zc *)malloc(sizeof(zc) * NL);
  double *symb[3];
  for (int ax = 0; ax < 3; ++ax) {
    symb[ax] = (double *)malloc(sizeof(double) * N[ax]);
    for (int c = 0; c < N[ax]; ++c) {
      double th = 2.0 * M_PI * c / N[ax];
      double sym = lap[ax][0];
      for (int m = 1; m <= 4; ++m) sym += 2.0 * lap[ax][m] * cos(m * th);
      symb[ax][c] = -0.5 * sym;
    }
  }
  dft_axis(s, 2, -1, g, a);
  dft_axis(s, 1, -1, a, b);
  dft_axis(s, 0, -1, b, a);
  double sc = Tx > 1e-12 ? Tx : 1e-12;
  for (int i = 0; i < NL; ++i) {
    double t = (symb[0][(i / st[0]) % N[0]] + symb[1][(i / st[1]) % N[1]]) + symb[2][(i / st[2]) % N[2]];
    double x = t / sc;
    double f = 27.0 + 18.0 * x + 12.0 * x * x + 8.0 * x * x * x;
    double kfac = f / (f + 16.0 * x * x * x * x);
    a[i] = zscale(kfac, a[i]);
  }
  dft_axis(s, 0, +1, a, b);
  dft_axis(s, 1, +1, b, a);
  dft_axis(s, 2, +1, a, g);
  for (int ax = 0; ax < 3; ++ax) free(symb[ax]);
  free(a);
  free(b);
}

/*
 * CG_ompk (GS/CG.f90:64-127) with skip_gsortho='n'.  The overlap used to
 * orthogonalise the starting band to lower bands is assigned to a real(8)
 * variable in the reference (CG.f90:46,70), so only its real part is
 * subtracted; that is kept.  precond=1 inserts the kinetic preconditioner
 * between the residual (CG.f90:85-90) and the search direction, and switches
 * the conjugation to the preconditioned (Polak-Ribiere-free) ratio
 * <Kg,g>/<Kg_old,g_old>.  iters_out(NB,NK) receives the iterations used.
 */
void or_cg(const or_system *s, const double *kAc, double *zu_, int ncg, int precond,
           double *esp_var, int *iters_out)
{
  const int NL = s->NL;
  const double Hxyz = s->Hxyz;
  size_t zsz = (size_t)s->Nps * s->Nlma;
#pragma omp parallel
  {
    double *zp = (double *)malloc(sizeof(zc) * zsz);
    zc *w = (zc *)malloc(sizeof(zc) * 9 * NL);
    zc *xk = w, *hxk = w + NL, *txk = w + 2 * NL, *gk = w + 3 * NL, *pk = w + 4 * NL, *pko = w + 5 * NL;
    zc *htp = w + 6 * NL, *ttp = w + 7 * NL, *kg = w + 8 * NL;
#pragma omp for schedule(static)
    for (int ik = 0; ik < s->NK; ++ik) {
      double k[3];
      kvec(kAc, s->NK, ik, k);
      or_update_projector(s, k, zp);
      zc *b0 = (zc *)zu_ + (size_t)ik * s->NB * NL;
      for (int ib = 0; ib < s->NB; ++ib) {
        zc *x = b0 + (size_t)ib * NL;
        for (int ibt = 0; ibt < ib; ++ibt) {
          double sr = zdot(b0 + (size_t)ibt * NL, x, NL).re * Hxyz;   /* real(8) :: s */
          for (int i = 0; i < NL; ++i) x[i] = zsub(x[i], zscale(sr, b0[(size_t)ibt * NL + i]));
        }
        double sc = 1.0 / sqrt(znrm2(x, NL) * Hxyz);
        for (int i = 0; i < NL; ++i) xk[i] = zscale(sc, x[i]);
        or_hpsi(s, k, zp, (double *)xk, (double *)hxk, (double *)txk);
        double xkHxk = zdot(xk, hxk, NL).re * Hxyz;
        double xkTxk = zdot(xk, txk, NL).re * Hxyz;
        double gkgk = 0.0;
        int it;
        for (it = 1; it <= ncg; ++it) {
          for (int i = 0; i < NL; ++i) gk[i] = zsub(hxk[i], zscale(xkHxk, xk[i]));
          for (int ibt = 0; ibt < ib; ++ibt) {
            zc zs = zscale(Hxyz, zdot(b0 + (size_t)ibt * NL, gk, NL));
            for (int i = 0; i < NL; ++i) gk[i] = zsub(gk[i], zmul(b0[(size_t)ibt * NL + i], zs));
          }
          double sg;
          zc *dir = gk;
          if (precond == 1) {
            memcpy(kg, gk, sizeof(zc) * NL);
            or_kinetic_precondition(s, xkTxk, (double *)kg);
            for (int ibt = 0; ibt < ib; ++ibt) {
              zc zs = zscale(Hxyz, zdot(b0 + (size_t)ibt * NL, kg, NL));
              for (int i = 0; i < NL; ++i) kg[i] = zsub(kg[i], zmul(b0[(size_t)ibt * NL + i], zs));
            }
            sg = zdot(kg, gk, NL).re * Hxyz;
            dir = kg;
          } else {
            sg = znrm2(gk, NL) * Hxyz;
          }
          if (it == 1) {
            memcpy(pk, dir, sizeof(zc) * NL);
          } else {
            double uk = sg / gkgk;
            for (int i = 0; i < NL; ++i) pk[i] = zadd(dir[i], zscale(uk, pk[i]));
          }
          gkgk = sg;
          zc zs = zscale(Hxyz, zdot(xk, pk, NL));
          for (int i = 0; i < NL; ++i) pko[i] = zsub(pk[i], zmul(xk[i], zs));
          double sn = 1.0 / sqrt(znrm2(pko, NL) * Hxyz);
          for (int i = 0; i < NL; ++i) pko[i] = zscale(sn, pko[i]);
          or_hpsi(s, k, zp, (double *)pko, (double *)htp, (double *)ttp);
          zc xkHpk = zscale(Hxyz, zdot(xk, htp, NL));
          double pkHpk = zdot(pko, htp, NL).re * Hxyz;
          double ev = 0.5 * ((xkHxk + pkHpk) - sqrt((xkHxk - pkHpk) * (xkHxk - pkHpk) + 4 * zabs2(xkHpk)));
          zc cx = zmk(xkHpk.re / (ev - xkHxk), xkHpk.im / (ev - xkHxk));
          double cp = 1.0 / sqrt(1.0 + zabs2(cx));
          cx = zscale(cp, cx);
          if (fabs(ev - xkHxk) < 1e-15) break;
          for (int i = 0; i < NL; ++i) {
            xk[i] = zadd(zmul(cx, xk[i]), zscale(cp, pko[i]));
            hxk[i] = zadd(zmul(cx, hxk[i]), zscale(cp, htp[i]));
            txk[i] = zadd(zmul(cx, txk[i]), zscale(cp, ttp[i]));
          }
          xkHxk = zdot(xk, hxk, NL).re * Hxyz;
          xkTxk = zdot(xk, txk, NL).re * Hxyz;
        }
        if (iters_out) iters_out[ik * s->NB + ib] = it > ncg ? ncg : it;
        sc = 1.0 / sqrt(znrm2(xk, NL) * Hxyz);
        for (int i = 0; i < NL; ++i) x[i] = zscale(sc, xk[i]);
        or_hpsi(s, k, zp, (double *)x, (double *)htp, (double *)ttp);
        double e = zdot(x, htp, NL).re * Hxyz;
        double r = 0.0;
        for (int i = 0; i < NL; ++i) r += zabs2(zsub(htp[i], zscale(e, x[i])));
        esp_var[ik * s->NB + ib] = sqrt(r * Hxyz) * s->occ[ik * s->NB + ib];
      }
    }
    free(w);
    free(zp);
  }
}

/*
 * The RT loop of tddft_sc (control_sc.f90:150-222) with the middlepoint
 * propagator and a frozen potential.  ac is A(t_n), n = 0..nsteps, as
 * (nsteps+1) rows of 3.  Step n: kAc = kAc0 + (A_n + A_{n+1})/2, Taylor-4
 * (dt_evolve.f90:47,181), density (:185), then kAc = kAc0 + A_{n+1}
 * (control_sc.f90:203-205), current (:209) and energy.  Outputs per step:
 * jt(nsteps,3) row-major, etot(nsteps) = Ekin+Enl+Eloc, norm(nsteps).
 * With energy_every > 1 the energy is computed on steps n with
 * (n+1) % energy_every == 0 only (others get 0).
 */
void or_rt_run(const or_system *s, const double *kAc0, const double *ac, int nsteps, double dt,
               double *zu, double *jt, double *etot, double *norm, double *rho, int energy_every)
{
  int NK = s->NK;
  double *kAc = (double *)malloc(sizeof(double) * 3 * NK);
  double *rtmp = rho ? rho : (double *)malloc(sizeof(double) * s->NL);
  for (int n = 0; n < nsteps; ++n) {
    for (int d = 0; d < 3; ++d)
      for (int ik = 0; ik < NK; ++ik)
        kAc[d * NK + ik] = kAc0[d * NK + ik] + 0.5 * (ac[3 * n + d] + ac[3 * (n + 1) + d]);
    or_hamiltonian(s, kAc, dt, zu);
    or_density(s, zu, rtmp);
    for (int d = 0; d < 3; ++d)
      for (int ik = 0; ik < NK; ++ik)
        kAc[d * NK + ik] = kAc0[d * NK + ik] + ac[3 * (n + 1) + d];
    if (jt) or_current(s, kAc, zu, jt + 3 * n);
    if (etot) {
      if (energy_every <= 1 || (n + 1) % energy_every == 0) {
        double ek, en, el;
        or_energy(s, kAc, zu, &ek, &en, &el);
        etot[n] = ek + en + el;
      } else {
        etot[n] = 0.0;
      }
    }
    if (norm) norm[n] = or_norm(s, zu);
  }
  if (!rho) free(rtmp);
  free(kAc);
}

/* Hartree potential of a density on the periodic grid: rho(G) = sum_r rho(r)
 * e^{-iG.r} / NL by one DFT pass per axis (z, y, x as the reference),
 * V_H(G) = 4 pi rho(G) / |G|^2 with the G = 0 term dropped, V_H(r) =
 * sum_G V_H(G) e^{iG.r}; G_a = 2 pi n_a / aL_a, n_a in -N_a/2 .. N_a-1-N_a/2. */
static void or_dft_axis(int NLx, int NLy, int NLz, int axis, int sgn, const zc *in, zc *out)
{
  const int N = axis == 0 ? NLx : (axis == 1 ? NLy : NLz);
  const int es = axis == 0 ? NLy * NLz : (axis == 1 ? NLz : 1);
  const int NL = NLx * NLy * NLz;
#pragma omp parallel for
  for (int p = 0; p < NL; ++p) {
    const int l = (p / es) % N;
    const int base = p - l * es;
    double re = 0.0, im = 0.0;
    for (int m = 0; m < N; ++m) {
      const double a = sgn * 2.0 * M_PI * (double)(((long)l * m) % N) / N;
      const double c = cos(a), sn = sin(a);
      const zc v = in[base + m * es];
      re += c * v.re - sn * v.im;
      im += c * v.im + sn * v.re;
    }
    out[p] = zmk(re, im);
  }
}

void or_hartree(int NLx, int NLy, int NLz, double aLx, double aLy, double aLz, const double *rho, double *vh)
{
  const int NL = NLx * NLy * NLz;
  zc *A = (zc *)malloc(sizeof(zc) * NL), *B = (zc *)malloc(sizeof(zc) * NL);
  for (int p = 0; p < NL; ++p) A[p] = zmk(rho[p], 0.0);
  or_dft_axis(NLx, NLy, NLz, 2, -1, A, B);
  or_dft_axis(NLx, NLy, NLz, 1, -1, B, A);
  or_dft_axis(NLx, NLy, NLz, 0, -1, A, B);
  for (int p = 0; p < NL; ++p) {
    const int ix = p / (NLy * NLz), iy = (p / NLz) % NLy, iz = p % NLz;
    const int nx = ix < NLx - NLx / 2 ? ix : ix - NLx;
    const int ny = iy < NLy - NLy / 2 ? iy : iy - NLy;
    const int nz = iz < NLz - NLz / 2 ? iz : iz - NLz;
    const double gx = 2.0 * M_PI * nx / aLx, gy = 2.0 * M_PI * ny / aLy, gz = 2.0 * M_PI * nz / aLz;
    const double G2 = gx * gx + gy * gy + gz * gz;
    const double f = (nx == 0 && ny == 0 && nz == 0) ? 0.0 : 4.0 * M_PI / G2 / NL;
    B[p] = zscale(f, B[p]);
  }
  or_dft_axis(NLx, NLy, NLz, 2, +1, B, A);
  or_dft_axis(NLx, NLy, NLz, 1, +1, A, B);
  or_dft_axis(NLx, NLy, NLz, 0, +1, B, A);
  for (int p = 0; p < NL; ++p) vh[p] = A[p].re;
  free(A);
  free(B);
}

/* Perdew-Zunger (1981) LDA exchange-correlation of the spin-unpolarised
 * density rho per point: e_xc(rho) and v_xc = e_xc + rho de_xc/drho, with the
 * reference's regularisation rho + 1e-10 and its branch at rs = 1. */
void or_pz(int n, const double *rho, double *vxc, double *exc)
{
  const double gU = -0.1423, b1U = 1.0529, b2U = 0.3334, AU = 0.0311, BU = -0.048, CU = 0.002, DU = -0.0116;
  const double cst = 0.75 * pow(3.0 / (2.0 * M_PI), 2.0 / 3.0);
  for (int i = 0; i < n; ++i) {
    const double t = rho[i] + 1e-10;
    const double rs = pow(3.0 / (4.0 * M_PI * t), 1.0 / 3.0);
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
    if (exc) exc[i] = e;
    vxc[i] = e + rho[i] * de;
  }
}

int or_num_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
