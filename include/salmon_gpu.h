/*
 * salmon_gpu.h -- C-ABI of the MI355X (gfx950) implementation of ARTED's
 * real-time electron-dynamics hot path.
 *
 * Conventions (those of the reference's Fortran module globals):
 *  - arrays are Fortran column-major; complex values are interleaved (re,im)
 *    doubles; the grid index is i = ix*NLy*NLz + iy*NLz + iz (z fastest),
 *    ARTED/preparation/init.f90:28-37;
 *  - orbitals zu(NL,NB,NK) complex(8), ARTED/modules/global_variables.f90 (zu_t);
 *  - kAc(NK,3) and kAc0(NK,3), occ(NB,NK), Jxyz/Jxx/Jyy/Jzz(Nps,NI) with Jxyz
 *    1-based, uV(Nps,Nlma), iuV(Nlma), a_tbl(Nlma) 1-based;
 *  - all pointers are host pointers.  Orbitals stay resident on the GPU
 *    between calls (the reference's OpenACC build does the same with
 *    "!$acc enter data", ARTED/control/control_sc.f90:133-149).
 *
 * Every entry returns 0 on success or a SALMON_ERR_* code; the message of the
 * last failure is salmon_gpu_last_error(ctx).  The reference has no status
 * codes and aborts through err_finalize (ARTED/main.f90:21-33); its Fortran
 * callers keep that behaviour by calling err_finalize on a non-zero status.
 *
 * A context owns one HIP stream on the current device.  Calls are ordered on
 * that stream; those that return host values synchronise it.  A context is
 * used by one host thread at a time.
 */
#ifndef SALMON_GPU_H
#define SALMON_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
  SALMON_OK = 0,
  SALMON_ERR_ARG = 1,         /* invalid argument or size */
  SALMON_ERR_HIP = 2,         /* a HIP runtime call failed */
  SALMON_ERR_STATE = 3,       /* required set-up call missing */
  SALMON_ERR_NOMEM = 4,       /* device allocation failed */
  SALMON_ERR_UNSUPPORTED = 5  /* option not available for this shape */
};

enum {
  SALMON_FP64 = 0,   /* complex128 orbitals, FP64 arithmetic */
  SALMON_MIXED = 1   /* complex64 stencil/KB/Taylor with FP32 accumulation;
                        FP64 norm, current, energy and density sums */
};

typedef struct salmon_ctx salmon_ctx;

/* Grid and finite-difference coefficients: lapx..lapz = lap(0:4)/H^2 and
 * nabx..nabz = nab(0:4)/H as produced by fd_coef
 * (ARTED/preparation/fd_coef.f90:75-80); nab*[0] is unused. */
typedef struct {
  int NLx, NLy, NLz;
  double aLx, aLy, aLz;
  double lapx[5], lapy[5], lapz[5];
  double nabx[5], naby[5], nabz[5];
} salmon_grid_t;

/* Creates a context.  Replaces the device side of the reference's
 * set-up (opt_vars_initialize_p2, ARTED/modules/opt_variables.f90:105-179). */
int salmon_gpu_init(const salmon_grid_t *grid, int precision, salmon_ctx **ctx);

/* Vloc(NL): the local potential read by hpsi1_RT_stencil as B
 * (ARTED/common/hpsi.f90:127; ARTED/RT/dt_evolve.f90:205-207). */
int salmon_gpu_set_potential(salmon_ctx *ctx, const double *vloc);

/* Kleinman-Bylander projector tables (ARTED/preparation/prep_ps.f90:81-208,
 * atom/pp/prep_pp.f90:263-347): Mps(NI), Jxyz/Jxx/Jyy/Jzz(Nps,NI),
 * a_tbl(Nlma), uV(Nps,Nlma), iuV(Nlma).  Projectors of one atom must be
 * consecutive in a_tbl (as idx_lma assumes, opt_variables.f90:241-250). */
int salmon_gpu_set_projectors(salmon_ctx *ctx, int NI, int Nps, const int *Mps, const int *Jxyz,
                              const int *Jxx, const int *Jyy, const int *Jzz, int Nlma,
                              const int *a_tbl, const double *uV, const int *iuV);

/* k points kAc0(NK,3) and occupations occ(NB,NK); NB = NBoccmax
 * (ARTED/preparation/init.f90:92-109, 283-331). */
int salmon_gpu_set_kpoints(salmon_ctx *ctx, int NK, int NB, const double *kAc0, const double *occ);

/* zu(NL,NB,NK) complex(8) host <-> device ("!$acc enter data" / "update self",
 * ARTED/control/control_sc.f90:133-149,215).  In SALMON_MIXED the device copy
 * is complex64. */
int salmon_gpu_put_orbitals(salmon_ctx *ctx, const double *zu);
int salmon_gpu_get_orbitals(salmon_ctx *ctx, double *zu);

/* H*psi for one orbital at wave vector kAc(3): hpsi_omp_KB_GS / _RT
 * (ARTED/common/hpsi.f90:33-96 -> hpsi_omp_KB_base :98-207).  in/out are
 * complex(8) (NL).  ttpsi (nullable) receives H*psi - Vloc*psi without the
 * nonlocal part (hpsi.f90:131-133). */
int salmon_gpu_hpsi(salmon_ctx *ctx, const double kAc[3], const double *in, double *out, double *ttpsi);

/* One 4th-order Taylor step of every resident orbital at kAc(NK,3)
 * (the midpoint k+A): subroutine hamiltonian(zu, flag_current),
 * ARTED/RT/hamiltonian.f90:27-123.  It takes rt_run's kernel path without
 * observables: on 16^3 the fused step, on other grids of at least 8 points
 * per axis the persistent sweep (SALMON_PATH_SWEEP), otherwise the streamed
 * kernels. */
int salmon_gpu_hamiltonian(salmon_ctx *ctx, const double *kAc, double dt);

/* Current jav(3) at kAc(NK,3): current_RT -> current('RT'),
 * ARTED/RT/current.f90:65-292, including the summation (:276-291).  The
 * current and the energies below come from one pass over the orbitals (one
 * observables-only launch of rt_run's kernel: the fused step on 16^3, the
 * persistent sweep on other grids of at least 8 points per axis, else the
 * streamed k_observe); a current or energy call at the
 * same kAc with nothing changed in between (no step, orbitals, potential,
 * tables or k points) reuses that pass. */
int salmon_gpu_current(salmon_ctx *ctx, const double *kAc, double jav[3]);

/* rho(NL) = sum occ |zu|^2: psi_rho_RT with Sym=1, ARTED/common/psi_rho.f90:40-206. */
int salmon_gpu_density(salmon_ctx *ctx, double *rho);

/* Wavefunction energies at kAc(NK,3): Ekin and Enl of Total_Energy_omp
 * (ARTED/common/total_energy.f90:113,172-203); eloc = sum V*rho*Hxyz of the
 * frozen local potential. */
int salmon_gpu_energy(salmon_ctx *ctx, const double *kAc, double *ekin, double *enl, double *eloc);

/* Total norm sum occ |zu|^2 Hxyz. */
int salmon_gpu_norm(salmon_ctx *ctx, double *norm);

/* Gram-Schmidt of the resident orbitals of every k, in FP64:
 * Gram_Schmidt_ompk, ARTED/GS/Gram_Schmidt.f90:32-55. */
int salmon_gpu_gram_schmidt(salmon_ctx *ctx);

/* esp(NB,NK) = <psi|H|psi> Hxyz at kAc(NK,3): sp_energy_omp, ARTED/GS/sp_energy.f90:20-60. */
int salmon_gpu_sp_energy(salmon_ctx *ctx, const double *kAc, double *esp);

/* Preconditioner applied to the CG residual.  CG_ompk applies none
 * (ARTED/GS/CG.f90:85-92); SALMON_PC_KINETIC multiplies the residual's
 * Fourier components by the Teter-Payne-Allan factor f/(f+16x^4),
 * f = 27+18x+12x^2+8x^3, x = (finite-difference kinetic symbol)/<x|T|x>, and
 * conjugates with the preconditioned ratio <Kg|g>/<Kg_old|g_old>; it has no
 * reference counterpart. */
enum { SALMON_PC_NONE = 0, SALMON_PC_KINETIC = 1, SALMON_PC_LEARNED = 2 };

/* One call of CG_ompk (ARTED/GS/CG.f90:34-139, skip_gsortho='n') on the
 * resident orbitals at kAc(NK,3): for each band in turn, the start vector is
 * orthogonalised to the lower bands and normalised, then up to iter_cg_max
 * orthogonalised CG iterations with the 2x2 Rayleigh-Ritz update, leaving at
 * |ev - <x|H|x>| < 1e-15.  esp_var(NB,NK) = occ ||Hx - <x|H|x> x|| sqrt(Hxyz)
 * (CG.f90:122-126); iters(NB,NK) (nullable) = iterations each band used.
 * The overlap of the start vector with the lower bands keeps only its real
 * part, as the reference's real(8) variable does (CG.f90:46,70).  The
 * coefficients of the reference's sequential Gram-Schmidt loops are formed
 * from one projection on all lower bands and a triangular solve with their
 * Gram matrix (equal up to rounding).  FP64 contexts, grid axes <= 64,
 * NB <= 1024. */
int salmon_gpu_cg(salmon_ctx *ctx, const double *kAc, int iter_cg_max, int precond, double *esp_var, int *iters);

/* Subspace diagonalisation diag_omp (ARTED/GS/diag.f90:95-130): za(i,j) =
 * <psi_i|H|psi_j> Hxyz hermitised as the reference does, its eigenvectors U
 * (ascending eigenvalues), zu <- zu U, esp(NB,NK) = the eigenvalues.
 * Eigenvector phases are those of a cyclic Jacobi solver (zheev's are
 * implementation-defined as well). */
int salmon_gpu_subspace_diag(salmon_ctx *ctx, const double *kAc, double *esp);

/* Sweeps of the ground-state loop with a frozen potential
 * (ARTED/control/ground_state.f90:136-140: Gram_Schmidt, diag_omp,
 * Gram_Schmidt, CG_omp(iter_cg_max), Gram_Schmidt), each followed by the band
 * energies <psi|H|psi> Hxyz (sp_energy_omp, ARTED/GS/sp_energy.f90:20-60),
 * until no band energy moves by tol or more between two sweeps, or
 * max_sweeps have run.  (The residual itself stalls near
 * sqrt(1e-15 * gap): CG_ompk leaves a band once its Rayleigh quotient moves by
 * less than 1e-15, CG.f90:42,114.)  Outputs (nullable): esp(NB,NK) = the
 * band energies after the last sweep, esp_var of its CG call, sweeps = sweeps
 * run, history(NB*NK, max_sweeps) = the band energies after each sweep. */
int salmon_gpu_ground_state(salmon_ctx *ctx, const double *kAc, int max_sweeps, int iter_cg_max, int precond,
                            double tol, double *esp, double *esp_var, int *sweeps, double *history);

/* ---- Learned CG preconditioner (no reference counterpart) ----------------
 * A three-layer periodic 3x3x3 convolution network without biases (channels
 * 2 -> 16 -> 16 -> 2, ReLU after the first two; weights W1(27,2,16),
 * W2(27,16,16), W3(27,16,2) as [tap][in][out], tap = (dx+1)*9+(dy+1)*3+(dz+1),
 * SALMON_PCNET_WEIGHTS floats in that order) maps the CG residual g, after
 * its phase is rotated so that sum g^2 is real and its norm scaled to
 * sqrt(NL), to a search direction (rotated back).  With SALMON_PC_LEARNED
 * each CG iteration is a steepest-descent step along that direction with the
 * 2x2 Rayleigh-Ritz of CG.f90:107-119 (no conjugation: the network is not a
 * fixed linear operator).  salmon_gpu_ground_state uses it for the first
 * sweeps and then falls back to SALMON_PC_KINETIC for the rest of the solve
 * (salmon_gpu_set_pc_handover).
 *
 * Training (one k point per context): salmon_gpu_pc_record makes the next CG
 * calls store each iteration's orthogonalised residual, iterate, band and
 * <x|H|x>; salmon_gpu_pc_set_targets takes the resident orbitals as the
 * converged bands Z and their energies as the targets lam.  The objective of
 * a sample is the fraction of the band's eigenvalue error removed by one
 * Rayleigh-Ritz step in span{x, p}, p the direction made orthogonal to x and
 * the lower bands Z_j (j < band) and normalised:
 *   loss = -mean(min((theta - ev)/(theta - lam), 1)),
 *   ev = (theta + c - sqrt((theta - c)^2 + 4|b|^2))/2,
 *   c = <p|H|p> Hxyz, b = <g|p> Hxyz, theta = <x|H|x>.
 * Adam (0.9, 0.999, 1e-8) updates FP32 weights. */
#define SALMON_PCNET_WEIGHTS (27 * 2 * 16 + 27 * 16 * 16 + 27 * 16 * 2)
typedef struct salmon_pcnet salmon_pcnet;
int salmon_gpu_pcnet_create(unsigned seed, salmon_pcnet **net);  /* He-normal initial weights */
int salmon_gpu_pcnet_set_weights(salmon_pcnet *net, const float *w);  /* also resets Adam */
int salmon_gpu_pcnet_get_weights(salmon_pcnet *net, float *w);
int salmon_gpu_pcnet_finalize(salmon_pcnet *net);
/* d(NL) = the learned direction of one residual g(NL) (complex, one k point),
 * computed as inside CG: BF16 weights and activations on the matrix cores
 * (v_mfma_f32_16x16x32_bf16, FP32 accumulation), scaled by |g|/sqrt(NL) */
int salmon_gpu_pcnet_apply(salmon_pcnet *net, salmon_ctx *ctx, const double *g, double *d);
/* network used by SALMON_PC_LEARNED on ctx (NULL detaches; not owned) */
int salmon_gpu_set_pcnet(salmon_ctx *ctx, salmon_pcnet *net);
/* No counterpart in the reference.  When salmon_gpu_ground_state with
 * SALMON_PC_LEARNED hands over to SALMON_PC_KINETIC: after at most
 * max_learned_sweeps sweeps (0: no limit), and earlier at the first sweep
 * (from the third on) that lowers the band-energy sum by less than
 * stall_ratio times the previous sweep's drop. */
int salmon_gpu_set_pc_handover(salmon_ctx *ctx, int max_learned_sweeps, double stall_ratio);
/* No counterpart in the reference.  CG and Gram-Schmidt calls run through a
 * graph cache (captured on the second call with the same arguments, replayed
 * after); enable = 0 runs them eagerly.  While a context's stream is being
 * captured, a HIP call of another host thread that touches the legacy stream
 * (a plain hipMemcpy, hipStreamCreate) fails, so programs that drive several
 * contexts from several threads disable the cache on every context. */
int salmon_gpu_set_graphs(salmon_ctx *ctx, int enable);
/* record up to max_samples residuals of the following CG calls, every
 * stride-th CG iteration (max_samples = 0 stops and frees the store) */
int salmon_gpu_pc_record(salmon_ctx *ctx, int max_samples, int stride);
int salmon_gpu_pc_samples(salmon_ctx *ctx, int *stored, int *valid);
int salmon_gpu_pc_get_sample(salmon_ctx *ctx, int i, double *g, double *x, int *band, double *theta, int *valid);
int salmon_gpu_pc_set_targets(salmon_ctx *ctx, const double *kAc);
/* loss and gradient (SALMON_PCNET_WEIGHTS doubles) of the mean objective over
 * the samples idx(nsamp), without an update */
int salmon_gpu_pcnet_gradient(salmon_pcnet *net, salmon_ctx *ctx, int nsamp, const int *idx, double *loss, double *grad);
/* nsteps Adam steps on batches of `batch` valid samples drawn with seed */
int salmon_gpu_pcnet_train(salmon_pcnet *net, salmon_ctx *ctx, int nsteps, int batch, double lr, unsigned seed,
                           double *mean_loss);

/* Device-resident RT loop: the time loop of tddft_sc
 * (ARTED/control/control_sc.f90:150-222) with the middlepoint propagator
 * (ARTED/RT/dt_evolve.f90:43-51,106-216) and a frozen potential.  ac holds
 * A(t_n) for n = 0..nsteps as (nsteps+1) rows of 3.  Step n propagates at
 * kAc0 + (A_n+A_{n+1})/2, computes the density, then current and energy at
 * kAc0 + A_{n+1}.  Outputs (nullable): jt (nsteps rows of 3), et(nsteps) =
 * Ekin+Enl+Eloc, norm(nsteps), rho(NL) = the density of the last step.
 * reortho_every > 0 applies an FP64 Gram-Schmidt after every reortho_every
 * steps. */
int salmon_gpu_rt_run(salmon_ctx *ctx, const double *ac, int nsteps, double dt, int reortho_every,
                      double *jt, double *et, double *norm, double *rho);

/* ---- Batched independent systems (no reference counterpart) --------------
 * A batch runs the RT loop of salmon_gpu_rt_run (middlepoint propagator,
 * frozen potentials, current/energy/norm every step) for several independent
 * set-up contexts at once -- different grids, band counts, k points and
 * atoms -- with every phase of a step one launch over all systems and the
 * step replayed from a captured HIP graph.  Contexts are referenced, not
 * copied: their resident orbitals advance; they must not be used elsewhere
 * while the batch runs and must outlive it.  All contexts share the batch's
 * precision, A(t) and dt.
 * Two step kernels: SALMON_BATCH_SWEEP, one persistent launch per step whose
 * workgroups take (system, orbitals) items from a device work counter and
 * run each orbital's whole step (grids of at least 8 points per axis whose x
 * plane splits into at most 256 columns of up to 4 z points in complex64, 512
 * columns of up to 2 z points in FP64), and SALMON_BATCH_PHASES,
 * one launch per phase over all systems (any grid).  SALMON_BATCH_AUTO (the
 * default) takes the sweep when every system fits it; salmon_gpu_batch_kernel
 * reports the kernel in use after tune (-1 before).  For the phase kernel
 * salmon_gpu_batch_tune times the stencil's z blocking (1, 2, 4 or 8 points
 * per thread along z) per grid shape and keeps the fastest (rt_run tunes on
 * first use); salmon_gpu_batch_zblock reports the choice.  Outputs:
 * jt(nsys, nsteps, 3), et(nsys, nsteps) = Ekin+Enl+Eloc, norm(nsys, nsteps),
 * row-major, each nullable. */
typedef struct salmon_batch salmon_batch;
enum { SALMON_BATCH_AUTO = 0, SALMON_BATCH_PHASES = 1, SALMON_BATCH_SWEEP = 2 };
int salmon_gpu_batch_create(int precision, salmon_batch **batch);
int salmon_gpu_batch_add(salmon_batch *batch, salmon_ctx *ctx);
/* SALMON_ERR_ARG from tune/rt_run if SALMON_BATCH_SWEEP is asked for and a system does not fit it */
int salmon_gpu_batch_set_kernel(salmon_batch *batch, int kernel);
int salmon_gpu_batch_kernel(const salmon_batch *batch);
int salmon_gpu_batch_tune(salmon_batch *batch);
int salmon_gpu_batch_zblock(salmon_batch *batch, int sys);
int salmon_gpu_batch_rt_run(salmon_batch *batch, const double *ac, int nsteps, double dt, double *jt, double *et,
                            double *norm);
/* The same with one A(t) per system, ac(nsys, nsteps+1, 3) row-major: the
 * macro-points of the multiscale Maxwell-TDDFT loop, each stepped from its
 * own Ac_m to Ac_new_m (replaces the per-macro-point dt_evolve_KB_MS and
 * current_RT_MS calls of ARTED/control/control_ms.f90:302-323; the
 * reference runs them one after another).  nsteps = 1 per call is the
 * coupled loop, where A(t+dt) depends on the currents. */
int salmon_gpu_batch_rt_run_ms(salmon_batch *batch, const double *ac, int nsteps, double dt, double *jt,
                               double *et, double *norm);
const char *salmon_gpu_batch_last_error(const salmon_batch *batch);
int salmon_gpu_batch_finalize(salmon_batch *batch);

/* Kernel path of the RT step.  SALMON_PATH_AUTO uses the LDS-resident fused
 * step where the grid has a compiled shape (16^3); on other grids of at least
 * 8 points per axis the persistent sweep kernel of the batches (one launch
 * per step; ETRS or the self-consistent potential: the Taylor-only
 * launch(es), density and potential, one observables launch per step; the
 * context as a batch of one on its own stream, its work buffers four times
 * the orbitals' size; a set-up that fails falls back); the streamed kernels
 * otherwise.  SALMON_PATH_STREAMED forces the streamed kernels (and frees
 * the sweep's work buffers),
 * SALMON_PATH_SWEEP the persistent sweep (rt_run then fails with
 * SALMON_ERR_UNSUPPORTED where it does not apply).  Like the reference's
 * build-time stencil selection (ARTED/stencil/CMakeLists.txt:1-52) the choice
 * changes the code path, not the results beyond rounding.  salmon_gpu_path
 * returns the path rt_run takes (SALMON_PATH_FUSED, SALMON_PATH_SWEEP or
 * SALMON_PATH_STREAMED).  The other entry points (hamiltonian, hpsi, the
 * observables) take the streamed kernels. */
enum { SALMON_PATH_AUTO = 0, SALMON_PATH_STREAMED = 1, SALMON_PATH_FUSED = 2, SALMON_PATH_SWEEP = 3 };
int salmon_gpu_set_path(salmon_ctx *ctx, int path);
int salmon_gpu_path(salmon_ctx *ctx);

/* Propagator of salmon_gpu_rt_run: the input variable propagator and its
 * dispatch in dt_evolve_KB (ARTED/RT/dt_evolve.f90:28-59).  MIDPOINT: one
 * Taylor-4 step over dt at kAc0 + (A_n + A_{n+1})/2 (the default).  ETRS
 * (dt_evolve_etrs_omp_KB, :218-348, with the frozen potential, whose
 * extrapolation 3V - 3V_old1 + V_old2 is then V): Taylor-4 over dt/2 at
 * kAc0 + A_n, then over dt/2 at kAc0 + A_{n+1}.  Observables are at
 * kAc0 + A_{n+1} either way.  A Fortran host running the reference's own
 * dt_evolve_KB reaches both through the hamiltonian drop-in instead. */
enum { SALMON_PROP_MIDPOINT = 0, SALMON_PROP_ETRS = 1 };
int salmon_gpu_set_propagator(salmon_ctx *ctx, int propagator);

/* Self-consistent local potential (dt_evolve.f90:185-207 with
 * no_update_func = 'n'): after each Taylor step of salmon_gpu_rt_run the
 * density rho(t+dt) gives Vloc = V_H + vpsl + V_xc -- the Hartree potential
 * of ARTED/common/Hartree.f90:20-128 (orthorhombic cell, up to 256 points per
 * axis) and, with SALMON_XC_PZ, the PZ81 LDA of src/xc/builtin_pz.f90 -- and
 * the observables are taken with the new potential.  salmon_gpu_set_scf sets
 * Vloc(0) from the resident orbitals' density at once; vpsl NULL returns to
 * the frozen potential set by salmon_gpu_set_potential (the last Vloc stays).
 * With SALMON_PROP_ETRS the second half step uses the extrapolated potential
 * 3 Vloc - 3 Vloc_old1 + Vloc_old2 (dt_evolve.f90:247-251), the history
 * starting at Vloc(0) when salmon_gpu_set_scf is called.  The energy's local term is then
 * sum Vloc rho Hxyz with the total Vloc (no Hartree double-counting
 * correction; the reference's G-space total energy is outside this path).
 * salmon_gpu_hartree_xc returns V_H and V_xc (xc = SALMON_XC_NONE: zeros) of
 * any density rho(NL) on the context's grid without changing its potential. */
enum { SALMON_XC_NONE = 0, SALMON_XC_PZ = 1 };
int salmon_gpu_set_scf(salmon_ctx *ctx, const double *vpsl, int xc);
int salmon_gpu_hartree_xc(salmon_ctx *ctx, const double *rho, int xc, double *vh, double *vxc);

/* salmon_gpu_hartree_xc without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight loops, hosts that forbid hipMalloc).  rho, vh and
 * vxc are DEVICE pointers of NL doubles (vh, vxc nullable); the work is
 * enqueued on the context's stream and the call returns without waiting.
 * Its DFT grids and twiddles live in one fixed device array per GPU (8 MiB,
 * held until the process ends, set up by salmon_gpu_init), so grids of at
 * most SALMON_NOALLOC_MAX_NL points are accepted (larger: SALMON_ERR_UNSUPPORTED).
 * Every context on a GPU shares that array: calls that overlap on one GPU
 * (several streams or threads) race and must be ordered by the caller. */
enum { SALMON_NOALLOC_MAX_NL = 64 * 64 * 64 };
int salmon_gpu_hartree_xc_noalloc(salmon_ctx *ctx, const double *rho_dev, int xc, double *vh_dev, double *vxc_dev);

/* Region timers with the reference's region numbers (misc/timer.f90:21-54;
 * timer_begin/timer_end :127-139, timer_get, timer_reset).  With the timers
 * on, each entry point times its GPU work per region: the stream time between
 * two events recorded around the region's launches, summed over calls.
 * Regions nest as in the reference (SALMON_LOG_DYNAMICS holds the
 * SALMON_LOG_DT_EVOLVE of its steps, which holds their SALMON_LOG_HPSI on the
 * streamed path; the fused 16^3 step is one launch and counts as
 * SALMON_LOG_DT_EVOLVE only).  The current and energy rows come from one
 * kernel and count as SALMON_LOG_CURRENT; the Hartree and exchange-correlation
 * potential from one chain and count as SALMON_LOG_HARTREE.  A region entered
 * while the context's stream is captured into a graph is not timed (the region
 * around the graph's replay is).  Every timed region is also a roctx range
 * named after the reference's constant ("LOG_HPSI", ...), for
 * rocprofv3 --marker-trace.  Timing adds a few microseconds per region. */
enum {
  SALMON_LOG_ALL = 0, SALMON_LOG_GROUND_STATE = 2, SALMON_LOG_DYNAMICS = 3,
  SALMON_LOG_CG = 10, SALMON_LOG_DIAG = 11, SALMON_LOG_SP_ENERGY = 12, SALMON_LOG_GRAM_SCHMIDT = 13,
  SALMON_LOG_DT_EVOLVE = 20, SALMON_LOG_HPSI = 21, SALMON_LOG_PSI_RHO = 22, SALMON_LOG_HARTREE = 23,
  SALMON_LOG_CURRENT = 25, SALMON_LOG_TOTAL_ENERGY = 26,
  SALMON_LOG_MAX = 41
};
/* on = 1 starts timing (the sums kept), 0 stops it. */
int salmon_gpu_timer_enable(salmon_ctx *ctx, int on);
/* timer_reset: all sums to zero (waits for the context's timed work). */
int salmon_gpu_timer_reset(salmon_ctx *ctx);
/* timer_get(id): seconds of GPU time and the number of timed calls of region
 * id (0 <= id < SALMON_LOG_MAX; waits for the context's timed work).
 * SALMON_LOG_ALL is the sum of the outermost regions' times. */
int salmon_gpu_timer_get(salmon_ctx *ctx, int id, double *seconds, int *calls);

/* Waits for all work of the context. */
int salmon_gpu_synchronize(salmon_ctx *ctx);

/* The context's HIP stream (a hipStream_t), for callers that order their own
 * work or timing events with the library's.  No reference counterpart (the
 * reference's OpenACC build uses the default queue). */
void *salmon_gpu_stream(salmon_ctx *ctx);

const char *salmon_gpu_errstr(int code);
/* with ctx NULL: the message of this thread's last failed salmon_gpu_init */
const char *salmon_gpu_last_error(const salmon_ctx *ctx);
int salmon_gpu_finalize(salmon_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
