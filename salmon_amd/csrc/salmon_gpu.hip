// C-ABI of the gfx950 implementation of ARTED's real-time hot path.
// The entry points are declared (with the reference routine each replaces)
// in include/salmon_gpu.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "../../include/salmon_gpu.h"
#include "common.hpp"
#include "streamed.hpp"

using namespace salmon;

// Region timers of a context (salmon_gpu_timer_*): per region the summed
// stream time and the number of timed calls; event pairs of finished regions
// wait in `pending` until the host asks for the sums.
struct RegionTimers {
  bool on = false;
  int depth = 0;  // regions open on the host
  double sec[SALMON_LOG_MAX] = {};
  int calls[SALMON_LOG_MAX] = {};
  struct Open {
    int id;
    bool outer;
    hipEvent_t b, e;
  };
  std::vector<Open> pending;
  std::vector<hipEvent_t> pool;
};

struct salmon_ctx {
  int precision = SALMON_FP64;
  int device = 0;
  hipStream_t stream = nullptr;
  Grid g{};
  std::string err;
  // potential
  double* d_vloc = nullptr;
  float* d_vloc_f = nullptr;
  bool have_v = false;
  // projectors
  bool have_kb = false;
  KBTables kb{};
  std::vector<void*> kb_bufs;
  // k points / orbitals
  int NK = 0, NB = 0, NKB = 0;
  double* d_kac0 = nullptr;   // (3*NK) as kAc0(NK,3)
  double* d_kac = nullptr;    // (3*NK) caller-given kAc
  double* d_occ = nullptr;    // (NKB)
  void* d_zu = nullptr;       // (NKB*NL) cplx<T>
  bool have_zu = false;
  // streamed work
  int chunk = 0;
  void* d_work[4] = {nullptr, nullptr, nullptr, nullptr};
  void* d_uvpsi = nullptr;
  int uvpsi_nlma = 0;         // projectors d_uvpsi is sized for
  bool graphs = true;         // CG / Gram-Schmidt through the graph cache
  // bumped whenever a device buffer that a launch may reference is freed or
  // replaced; a captured graph holds raw pointers and is dropped once the
  // generation it was captured at is gone
  unsigned long gen = 0;
  int propagator = SALMON_PROP_MIDPOINT;  // of rt_run
  // self-consistent potential (salmon_gpu_set_scf): Vloc = V_H[rho] + Vpsl + V_xc[rho]
  bool scf = false;
  int scf_xc = 0;
  double* d_vpsl = nullptr;
  double* d_scf_tw = nullptr;                 // DFT twiddles, 3 x kMaxN (cos, sin)
  void* d_scf_buf[2] = {nullptr, nullptr};    // two complex grids
  double* d_vold = nullptr;                   // ETRS potential history V_old1, V_old2 (2 NL)
  void* d_ph = nullptr;       // KB phases of the streamed Taylor step (NK, NPI)
  size_t ph_bytes = 0;
  double* d_rows = nullptr;   // (NKB*8)
  double* d_res = nullptr;    // rt_run's per-step results (8 per step), grown on demand
  size_t res_cap = 0;         // steps d_res holds
  double* d_sum = nullptr;    // (8)
  double* d_rho = nullptr;    // (NL)
  double* d_rho_part = nullptr;
  double* d_rho_part2 = nullptr;  // second stage of the density reduction
  int rho_chunks = 0;
  // single-orbital staging
  double* d_kone = nullptr;   // (3)
  double* d_oneocc = nullptr; // (1)
  void* d_stage[3] = {nullptr, nullptr, nullptr};
  cplx<double>* d_f64 = nullptr;  // FP64 scratch copy of the orbitals (mixed Gram-Schmidt)
  // fused step kernel configuration (0 = streamed path)
  int fused = -1;
  int path = SALMON_PATH_AUTO;
  void* fz = nullptr;  // state owned by the fused path
  void* gs = nullptr;  // state owned by the ground-state path
  void* pcnet = nullptr;  // network of the learned CG preconditioner (not owned)
  // learned -> kinetic handover of salmon_gpu_ground_state: at most
  // pc_max_sweeps learned sweeps (0: no limit), and earlier once a sweep lowers
  // the band-energy sum by less than pc_stall times the previous sweep's drop
  int pc_max_sweeps = 0;
  double pc_stall = 0.3;
  RegionTimers* tm = nullptr;  // salmon_gpu_timer_enable
  // the persistent sweep of rt_run off the fused shapes (SALMON_PATH_SWEEP):
  // a batch of this context alone, on the context's stream, rebuilt when the
  // buffer generation moves on; solo_fail marks a set-up that failed at
  // generation solo_gen (SALMON_PATH_AUTO then takes the streamed kernels)
  salmon_batch* solo = nullptr;
  unsigned long solo_gen = 0;
  bool solo_fail = false;
  // bumped by every entry point that may change the orbitals, potential,
  // tables, k points or kernel path; the observables sums of the last current
  // or energy call are reused by the next one at the same kAc and epoch (the
  // reference calls current_RT and Total_Energy_omp at one k after a step)
  std::vector<double> kac0_h;  // kAc0(NK,3) as set_kpoints gave it
  unsigned long epoch = 0;
  unsigned long obs_epoch = ~0ul;
  std::vector<double> obs_k;
  double obs_sums[8] = {};
};

static inline void touch(salmon_ctx* c) {
  if (c) ++c->epoch;
}

namespace {

// the reference's names of the timed regions (misc/timer.f90:21-54)
const char* log_name(int id) {
  switch (id) {
    case SALMON_LOG_GROUND_STATE: return "LOG_GROUND_STATE";
    case SALMON_LOG_DYNAMICS: return "LOG_DYNAMICS";
    case SALMON_LOG_CG: return "LOG_CG";
    case SALMON_LOG_DIAG: return "LOG_DIAG";
    case SALMON_LOG_SP_ENERGY: return "LOG_SP_ENERGY";
    case SALMON_LOG_GRAM_SCHMIDT: return "LOG_GRAM_SCHMIDT";
    case SALMON_LOG_DT_EVOLVE: return "LOG_DT_EVOLVE";
    case SALMON_LOG_HPSI: return "LOG_HPSI";
    case SALMON_LOG_PSI_RHO: return "LOG_PSI_RHO";
    case SALMON_LOG_HARTREE: return "LOG_HARTREE";
    case SALMON_LOG_CURRENT: return "LOG_CURRENT";
    case SALMON_LOG_TOTAL_ENERGY: return "LOG_TOTAL_ENERGY";
    default: return "LOG_OTHER";
  }
}

// Sums the finished regions' event pairs (waits for their end events).
void timers_resolve(salmon_ctx* c) {
  RegionTimers* t = c->tm;
  if (!t) return;
  for (const auto& o : t->pending) {
    float ms = 0.0f;
    if (hipEventSynchronize(o.e) == hipSuccess && hipEventElapsedTime(&ms, o.b, o.e) == hipSuccess) {
      t->sec[o.id] += 1e-3 * ms;
      t->calls[o.id] += 1;
      if (o.outer) {
        t->sec[SALMON_LOG_ALL] += 1e-3 * ms;
        t->calls[SALMON_LOG_ALL] += 1;
      }
    }
    t->pool.push_back(o.b);
    t->pool.push_back(o.e);
  }
  t->pending.clear();
}

// One timed region (timer_begin at construction, timer_end at scope exit):
// events on the context's stream around the launches enqueued in between, and
// a roctx range of the region's name.  Nothing is recorded with the timers
// off or while the stream is being captured.
class Region {
 public:
  Region(salmon_ctx* c, int id) : c_(c), id_(id) {
    RegionTimers* t = c->tm;
    if (!t || !t->on || id < 0 || id >= SALMON_LOG_MAX) return;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(c->stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return;
    b_ = take(t);
    e_ = take(t);
    if (!b_ || !e_ || hipEventRecord(b_, c->stream) != hipSuccess) {
      if (b_) t->pool.push_back(b_);
      if (e_) t->pool.push_back(e_);
      return;
    }
    outer_ = t->depth == 0;
    ++t->depth;
    roctxRangePushA(log_name(id));
    live_ = true;
  }
  ~Region() {
    if (!live_) return;
    RegionTimers* t = c_->tm;
    roctxRangePop();
    --t->depth;
    if (hipEventRecord(e_, c_->stream) == hipSuccess) {
      t->pending.push_back({id_, outer_, b_, e_});
    } else {
      t->pool.push_back(b_);
      t->pool.push_back(e_);
    }
    // bound the events held: a long timed run sums as it goes
    if (t->pending.size() >= 4096) timers_resolve(c_);
  }
  Region(const Region&) = delete;
  Region& operator=(const Region&) = delete;

 private:
  static hipEvent_t take(RegionTimers* t) {
    if (!t->pool.empty()) {
      hipEvent_t e = t->pool.back();
      t->pool.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    return hipEventCreate(&e) == hipSuccess ? e : nullptr;
  }
  salmon_ctx* c_;
  int id_;
  bool outer_ = false, live_ = false;
  hipEvent_t b_ = nullptr, e_ = nullptr;
};

int fail(salmon_ctx* c, int code, const char* fmt, ...) {
  if (c) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    c->err = buf;
  }
  return code;
}

#define HIPCHK(c, x)                                                                                 \
  do {                                                                                               \
    hipError_t e_ = (x);                                                                             \
    if (e_ != hipSuccess) return fail((c), SALMON_ERR_HIP, "%s: %s", #x, hipGetErrorString(e_));     \
  } while (0)

#define LAUNCHCHK(c)                                                                                 \
  do {                                                                                               \
    hipError_t e_ = hipGetLastError();                                                               \
    if (e_ != hipSuccess) return fail((c), SALMON_ERR_HIP, "launch: %s", hipGetErrorString(e_));    \
  } while (0)

template <typename X>
int dalloc(salmon_ctx* c, X** p, size_t n) {
  if (*p) {
    (void)hipFree(*p);
    *p = nullptr;
    if (c) ++c->gen;
  }
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void**)p, n * sizeof(X));
  if (e != hipSuccess) return fail(c, SALMON_ERR_NOMEM, "hipMalloc(%zu bytes): %s", n * sizeof(X), hipGetErrorString(e));
  return SALMON_OK;
}

void dfree(void** p) {
  if (*p) (void)hipFree(*p);
  *p = nullptr;
}

size_t elem_size(const salmon_ctx* c) { return c->precision == SALMON_MIXED ? sizeof(cplx<float>) : sizeof(cplx<double>); }

int ensure_work(salmon_ctx* c) {
  // the projection buffer follows the projector count (tables may be set after the first call)
  if (c->d_work[0]) {
    if (c->kb.Nlma <= c->uvpsi_nlma) return SALMON_OK;
    c->uvpsi_nlma = std::max(1, c->kb.Nlma);
    return dalloc(c, (char**)&c->d_uvpsi, (size_t)c->chunk * c->uvpsi_nlma * sizeof(cplx<double>));
  }
  // work buffers hold at most ~1 GiB each
  size_t per = (size_t)c->g.NL * elem_size(c);
  size_t cap = std::max<size_t>(1, ((size_t)1 << 30) / per);
  c->chunk = (int)std::min<size_t>((size_t)c->NKB, cap);
  for (int i = 0; i < 4; ++i) {
    hipError_t e = hipMalloc(&c->d_work[i], (size_t)c->chunk * per);
    if (e != hipSuccess) return fail(c, SALMON_ERR_NOMEM, "work buffer: %s", hipGetErrorString(e));
  }
  c->uvpsi_nlma = std::max(1, c->kb.Nlma);
  int r = dalloc(c, (char**)&c->d_uvpsi, (size_t)c->chunk * c->uvpsi_nlma * sizeof(cplx<double>));
  if (r) return r;
  return SALMON_OK;
}

int check_ready(salmon_ctx* c, bool need_zu) {
  if (!c) return SALMON_ERR_ARG;
  if (!c->have_v) return fail(c, SALMON_ERR_STATE, "salmon_gpu_set_potential not called");
  if (!c->have_kb) return fail(c, SALMON_ERR_STATE, "salmon_gpu_set_projectors not called");
  if (c->NKB == 0) return fail(c, SALMON_ERR_STATE, "salmon_gpu_set_kpoints not called");
  if (need_zu && !c->have_zu) return fail(c, SALMON_ERR_STATE, "salmon_gpu_put_orbitals not called");
  return SALMON_OK;
}

template <typename T>
const T* vloc_ptr(const salmon_ctx* c);
template <>
const double* vloc_ptr<double>(const salmon_ctx* c) { return c->d_vloc; }
template <>
const float* vloc_ptr<float>(const salmon_ctx* c) { return c->d_vloc_f; }

// bands per k point from which the KB projection runs on the matrix cores
// (16-band tiles; fewer bands leave most of a tile's rows empty)
constexpr int kMfmaMinBands = 8;

// H applied to nOrb orbitals (rows of NL) from `in` to `out` (optionally tt).
template <typename T>
int apply_h(salmon_ctx* c, const cplx<T>* in, cplx<T>* out, cplx<T>* tt, int orb0, int nOrb, int NB, int NK,
            const double* kbase, double ax, double ay, double az, void* uvbuf = nullptr,
            const cplx<T>* ph = nullptr) {
  Region reg(c, SALMON_LOG_HPSI);
  cplx<T>* uv = (cplx<T>*)(uvbuf ? uvbuf : c->d_uvpsi);
  if (c->kb.Nlma > 0 && NB >= kMfmaMinBands) {
    // the bands of each k point as the rows of one matrix-core GEMM per atom
    const int tpk = (NB + 15) / 16;
    const int k0 = orb0 / NB, k1 = (orb0 + nOrb - 1) / NB;
    const dim3 gm(c->kb.NI, (k1 - k0 + 1) * tpk, (c->kb.max_lpa + 7) / 8);
    hipLaunchKernelGGL(k_kb_project_mfma<T>, gm, dim3(64), 0, c->stream, in, orb0, nOrb, NB, NK, k0, tpk, kbase, ax,
                       ay, az, c->g, c->kb, uv, ph);
    LAUNCHCHK(c);
  } else if (c->kb.Nlma > 0) {
    dim3 gp((c->kb.Nlma + 3) / 4, nOrb);
    hipLaunchKernelGGL(k_kb_project<T>, gp, dim3(256), 0, c->stream, in, orb0, NB, NK, kbase, ax, ay, az, c->g,
                       c->kb, uv, ph);
    LAUNCHCHK(c);
  }
  dim3 gh((c->g.NL + 255) / 256, nOrb);
  hipLaunchKernelGGL(k_hpsi<T>, gh, dim3(256), 0, c->stream, in, out, tt, orb0, NB, NK, kbase, ax, ay, az,
                     vloc_ptr<T>(c), c->g, c->kb, (const cplx<T>*)uv, ph);
  LAUNCHCHK(c);
  return SALMON_OK;
}

template <typename T>
int taylor_streamed(salmon_ctx* c, const double* kbase, double ax, double ay, double az, double dt) {
  Region reg(c, SALMON_LOG_DT_EVOLVE);
  int r = ensure_work(c);
  if (r) return r;
  // zfac(n) = (-i dt)^n / n!  (hamiltonian.f90:46-49)
  const double z1i = -dt, z2r = -dt * dt / 2.0, z3i = dt * dt * dt / 6.0, z4r = dt * dt * dt * dt / 24.0;
  const size_t NL = c->g.NL;
  // KB phases of every k point at this step's k, shared by the four H applications
  const size_t nph = (size_t)c->NK * c->kb.NPI;
  cplx<T>* ph = nullptr;
  if (nph > 0) {
    if (nph * sizeof(cplx<T>) > c->ph_bytes) {
      if (c->d_ph) (void)hipFree(c->d_ph);
      c->d_ph = nullptr;
      c->ph_bytes = 0;
      ++c->gen;
      if (hipMalloc(&c->d_ph, nph * sizeof(cplx<T>)) != hipSuccess) return fail(c, SALMON_ERR_NOMEM, "phase table");
      c->ph_bytes = nph * sizeof(cplx<T>);
    }
    ph = (cplx<T>*)c->d_ph;
    hipLaunchKernelGGL(k_phase_table<T>, dim3((unsigned)((nph + 255) / 256)), dim3(256), 0, c->stream, kbase, c->NK, ax,
                       ay, az, c->kb, ph);
    LAUNCHCHK(c);
  }
  for (int o0 = 0; o0 < c->NKB; o0 += c->chunk) {
    const int n = std::min(c->chunk, c->NKB - o0);
    cplx<T>* zu = (cplx<T>*)c->d_zu + (size_t)o0 * NL;
    cplx<T>* h[4];
    for (int i = 0; i < 4; ++i) h[i] = (cplx<T>*)c->d_work[i];
    if ((r = apply_h<T>(c, zu, h[0], nullptr, o0, n, c->NB, c->NK, kbase, ax, ay, az, nullptr, ph))) return r;
    if ((r = apply_h<T>(c, h[0], h[1], nullptr, o0, n, c->NB, c->NK, kbase, ax, ay, az, nullptr, ph))) return r;
    if ((r = apply_h<T>(c, h[1], h[2], nullptr, o0, n, c->NB, c->NK, kbase, ax, ay, az, nullptr, ph))) return r;
    if ((r = apply_h<T>(c, h[2], h[3], nullptr, o0, n, c->NB, c->NK, kbase, ax, ay, az, nullptr, ph))) return r;
    const size_t tot = (size_t)n * NL;
    hipLaunchKernelGGL(k_taylor_fold<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, c->stream, zu, h[0], h[1],
                       h[2], h[3], tot, (T)z1i, (T)z2r, (T)z3i, (T)z4r);
    LAUNCHCHK(c);
  }
  return SALMON_OK;
}

template <typename T>
int observe_rows(salmon_ctx* c, const double* kbase, double ax, double ay, double az) {
  hipLaunchKernelGGL(k_observe<T>, dim3(c->NKB), dim3(256), 0, c->stream, (const cplx<T>*)c->d_zu, 0, c->NB, c->NK,
                     kbase, ax, ay, az, c->d_occ, c->d_vloc, c->g, c->kb, c->d_rows);
  LAUNCHCHK(c);
  return SALMON_OK;
}

// rho(p) = sum over nchunk partials (nchunk, NL), in a fixed order and with
// enough threads: slices of <= 16 partials, then the sum of the slices
constexpr int kRhoMaxSlices = 64;
int density_reduce(salmon_ctx* c, const double* part, int nchunk) {
  const int NL = c->g.NL;
  if (!c->d_rho_part2) {
    int r = dalloc(c, &c->d_rho_part2, (size_t)kRhoMaxSlices * NL);
    if (r) return r;
  }
  const int per = std::max(16, (nchunk + kRhoMaxSlices - 1) / kRhoMaxSlices);
  const int ns = (nchunk + per - 1) / per;
  hipLaunchKernelGGL(k_density_slices, dim3((NL + 255) / 256, ns), dim3(256), 0, c->stream, part, nchunk, per, NL,
                     c->d_rho_part2);
  LAUNCHCHK(c);
  hipLaunchKernelGGL(k_density_final, dim3((NL + 255) / 256), dim3(256), 0, c->stream, c->d_rho_part2, ns, NL,
                     c->d_rho);
  LAUNCHCHK(c);
  return SALMON_OK;
}

template <typename T>
int density_dev(salmon_ctx* c) {
  Region reg(c, SALMON_LOG_PSI_RHO);
  const int NL = c->g.NL;
  if (!c->d_rho_part) {
    c->rho_chunks = std::max(1, std::min(c->NKB, std::max(1, 4096 * 1024 / std::max(1, NL))));
    int r = dalloc(c, &c->d_rho_part, (size_t)c->rho_chunks * NL);
    if (r) return r;
  }
  const int per = (c->NKB + c->rho_chunks - 1) / c->rho_chunks;
  const int nch = (c->NKB + per - 1) / per;
  hipLaunchKernelGGL(k_density_part<T>, dim3((NL + 255) / 256, nch), dim3(256), 0, c->stream,
                     (const cplx<T>*)c->d_zu, c->NKB, NL, per, c->d_occ, c->d_rho_part);
  LAUNCHCHK(c);
  return density_reduce(c, c->d_rho_part, nch);
}

int gram_schmidt_multi_fwd(salmon_ctx* c, cplx<double>* zu);

// FP64 Gram-Schmidt of the resident orbitals (for complex64 orbitals on an
// FP64 copy)
int gram_schmidt_dev(salmon_ctx* c) {
  Region reg(c, SALMON_LOG_GRAM_SCHMIDT);
  const size_t n = (size_t)c->NKB * c->g.NL;
  if (c->precision == SALMON_FP64) return gram_schmidt_multi_fwd(c, (cplx<double>*)c->d_zu);
  if (!c->d_f64) {
    int r = dalloc(c, &c->d_f64, n);
    if (r) return r;
  }
  const unsigned nb = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL((k_convert<float, double>), dim3(nb), dim3(256), 0, c->stream, (const cplx<float>*)c->d_zu,
                     c->d_f64, n);
  LAUNCHCHK(c);
  int r = gram_schmidt_multi_fwd(c, c->d_f64);
  if (r) return r;
  hipLaunchKernelGGL((k_convert<double, float>), dim3(nb), dim3(256), 0, c->stream, (const cplx<double>*)c->d_f64,
                     (cplx<float>*)c->d_zu, n);
  LAUNCHCHK(c);
  return SALMON_OK;
}

int upload_kac(salmon_ctx* c, const double* kAc) {
  if (!kAc) return fail(c, SALMON_ERR_ARG, "kAc is NULL");
  HIPCHK(c, hipMemcpyAsync(c->d_kac, kAc, sizeof(double) * 3 * c->NK, hipMemcpyHostToDevice, c->stream));
  return SALMON_OK;
}

int read_sums(salmon_ctx* c, double out[8]) {
  hipLaunchKernelGGL(k_sum_rows, dim3(1), dim3(256), 0, c->stream, c->d_rows, c->NKB, c->d_sum);
  LAUNCHCHK(c);
  HIPCHK(c, hipMemcpyAsync(out, c->d_sum, sizeof(double) * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SALMON_OK;
}

}  // namespace

#include "fused.hpp"
#include "scf.hpp"
#include "gs.hpp"
#include "pc.hpp"
#include "batch_host.hpp"

namespace {
int gram_schmidt_multi_fwd(salmon_ctx* c, cplx<double>* zu) {
  int r = ensure_work(c);
  if (r || (r = gs_ensure(c))) return r;
  GsState* s = (GsState*)c->gs;
  const long key[4] = {(long)(uintptr_t)zu, c->NB, c->NK, 0};
  return graph_run(c, s->g_gs, key, [&]() { return gram_schmidt_multi(c, zu); });
}
}  // namespace

namespace {

// DFT twiddles (cos, sin)(2 pi j / N) of the three axes, kMaxN rows per axis.
void scf_twiddles(const Grid& g, double* tw) {
  const int N[3] = {g.NLx, g.NLy, g.NLz};
  for (size_t i = 0; i < (size_t)3 * 2 * salmon::scf::kMaxN; ++i) tw[i] = 0.0;
  for (int a = 0; a < 3; ++a)
    for (int j = 0; j < N[a]; ++j) {
      const double ang = 2.0 * M_PI * j / N[a];
      tw[(size_t)a * 2 * salmon::scf::kMaxN + 2 * j] = std::cos(ang);
      tw[(size_t)a * 2 * salmon::scf::kMaxN + 2 * j + 1] = std::sin(ang);
    }
}

// V_H and V_xc of the device density rho with the twiddles tw and two complex
// grids A, B (device): Hartree by per-axis DFTs and the Coulomb kernel, then
// Vloc = V_H + Vpsl + V_xc into the context's potential (set_v) and/or the
// parts into vh, vxc (device, nullable).  Launches only.
int scf_potential_on(salmon_ctx* c, const double* rho, bool set_v, int xc, double* vh, double* vxc, const double* tw,
                     cplx<double>* A, cplx<double>* B) {
  const Grid& g = c->g;
  const dim3 gr((g.NL + 255) / 256), bl(256);
  hipLaunchKernelGGL(salmon::scf::k_real_to_cplx, gr, bl, 0, c->stream, rho, A, g.NL);
  hipLaunchKernelGGL(salmon::scf::k_dft_axis, gr, bl, 0, c->stream, (const cplx<double>*)A, B, g, 2, -1, tw);
  hipLaunchKernelGGL(salmon::scf::k_dft_axis, gr, bl, 0, c->stream, (const cplx<double>*)B, A, g, 1, -1, tw);
  hipLaunchKernelGGL(salmon::scf::k_dft_axis, gr, bl, 0, c->stream, (const cplx<double>*)A, B, g, 0, -1, tw);
  hipLaunchKernelGGL(salmon::scf::k_coulomb, gr, bl, 0, c->stream, B, g);
  hipLaunchKernelGGL(salmon::scf::k_dft_axis, gr, bl, 0, c->stream, (const cplx<double>*)B, A, g, 2, +1, tw);
  hipLaunchKernelGGL(salmon::scf::k_dft_axis, gr, bl, 0, c->stream, (const cplx<double>*)A, B, g, 1, +1, tw);
  hipLaunchKernelGGL(salmon::scf::k_dft_axis, gr, bl, 0, c->stream, (const cplx<double>*)B, A, g, 0, +1, tw);
  LAUNCHCHK(c);
  if (set_v) {
    hipLaunchKernelGGL(salmon::scf::k_vloc, gr, bl, 0, c->stream, (const cplx<double>*)A, rho, (const double*)c->d_vpsl,
                       xc, g.NL, c->d_vloc, c->d_vloc_f, vh, vxc);
  } else {
    // parts only: Vpsl = 0 into a scratch potential (the second grid)
    double* pot = (double*)B;
    hipLaunchKernelGGL(salmon::scf::k_vloc, gr, bl, 0, c->stream, (const cplx<double>*)A, rho, (const double*)nullptr,
                       xc, g.NL, pot, (float*)nullptr, vh, vxc);
  }
  LAUNCHCHK(c);
  return SALMON_OK;
}

int scf_check_grid(salmon_ctx* c) {
  const Grid& g = c->g;
  if (g.NLx > salmon::scf::kMaxN || g.NLy > salmon::scf::kMaxN || g.NLz > salmon::scf::kMaxN)
    return fail(c, SALMON_ERR_UNSUPPORTED, "Hartree solve limited to %d points per axis", salmon::scf::kMaxN);
  return SALMON_OK;
}

// scf_potential_on with the context's own twiddles and grids (allocated on first use).
int scf_potential(salmon_ctx* c, const double* rho, bool set_v, int xc, double* vh, double* vxc) {
  Region reg(c, SALMON_LOG_HARTREE);
  const Grid& g = c->g;
  int r;
  if ((r = scf_check_grid(c))) return r;
  if (!c->d_scf_tw) {
    std::vector<double> tw((size_t)3 * 2 * salmon::scf::kMaxN);
    scf_twiddles(g, tw.data());
    if ((r = dalloc(c, &c->d_scf_tw, tw.size()))) return r;
    HIPCHK(c, hipMemcpy(c->d_scf_tw, tw.data(), sizeof(double) * tw.size(), hipMemcpyHostToDevice));
  }
  for (int i = 0; i < 2; ++i)
    if (!c->d_scf_buf[i] && (r = dalloc(c, (cplx<double>**)&c->d_scf_buf[i], (size_t)g.NL))) return r;
  return scf_potential_on(c, rho, set_v, xc, vh, vxc, c->d_scf_tw, (cplx<double>*)c->d_scf_buf[0],
                          (cplx<double>*)c->d_scf_buf[1]);
}

// One RT step n on resident orbitals (streamed path), A(t_n) = a0,
// A(t_{n+1}) = a1: middlepoint (dt_evolve.f90:43-51) Taylor-4 at
// kAc0 + (a0 + a1)/2, or ETRS (dt_evolve.f90:53-59,218-348 with a frozen
// potential, whose extrapolation 3V - 3V_old + V_old2 is then V) Taylor-4 over
// dt/2 at kAc0 + a0 and again at kAc0 + a1; then the density and the
// observables at kAc0 + a1 into res[0..7].
template <typename T>
int rt_step(salmon_ctx* c, const double a0[3], const double a1[3], double dt, double* res_dev) {
  int r;
  const double* anew = a1;
  if (c->propagator == SALMON_PROP_ETRS) {
    if ((r = taylor_streamed<T>(c, c->d_kac0, a0[0], a0[1], a0[2], 0.5 * dt))) return r;
    if ((r = taylor_streamed<T>(c, c->d_kac0, a1[0], a1[1], a1[2], 0.5 * dt))) return r;
  } else {
    const double amid[3] = {0.5 * (a0[0] + a1[0]), 0.5 * (a0[1] + a1[1]), 0.5 * (a0[2] + a1[2])};
    if ((r = taylor_streamed<T>(c, c->d_kac0, amid[0], amid[1], amid[2], dt))) return r;
  }
  if ((r = density_dev<T>(c))) return r;
  Region reg(c, SALMON_LOG_CURRENT);
  if ((r = observe_rows<T>(c, c->d_kac0, anew[0], anew[1], anew[2]))) return r;
  hipLaunchKernelGGL(k_sum_rows, dim3(1), dim3(256), 0, c->stream, c->d_rows, c->NKB, res_dev);
  LAUNCHCHK(c);
  return SALMON_OK;
}

// nsteps RT steps on the fused path.  The observables of step n (at
// kAc0 + A(t_{n+1}), of psi(t_{n+1})) are evaluated by the launch of step
// n+1 during its first H application, which reads psi(t_{n+1}) anyway; an
// observables-only launch closes the run, and one runs before each
// re-orthonormalisation, so every result is of the same orbitals as in the
// step-by-step order (step, observables, re-orthonormalisation).  With the
// potential frozen only the last step's density is kept (the run's rho
// output), so only the launch with the last step's observables sums it.
template <typename T>
int rt_run_fused(salmon_ctx* c, const double* ac, int nsteps, double dt, int reortho_every, double* d_res) {
  int r = SALMON_OK;
  const bool etrs = c->propagator == SALMON_PROP_ETRS;
  // the Taylor launches of step n: middlepoint, one at (A_n + A_{n+1})/2 over
  // dt; ETRS, one at A_n and one at A_{n+1}, each over dt/2 (see rt_step)
  auto kpoint = [&](int n, int part, double a[3]) {
    for (int d = 0; d < 3; ++d)
      a[d] = etrs ? ac[3 * (n + part) + d] : 0.5 * (ac[3 * n + d] + ac[3 * (n + 1) + d]);
  };
  const double dts = etrs ? 0.5 * dt : dt;
  if constexpr (!salmon::fused_defer<T, kFusedXS64>()) {
    // observables of psi(t+dt) in the (last) launch of the step itself
    for (int n = 0; n < nsteps && r == SALMON_OK; ++n) {
      double a[3], anew[3];
      for (int d = 0; d < 3; ++d) anew[d] = ac[3 * (n + 1) + d];
      if (etrs) {
        kpoint(n, 0, a);
        if ((r = fused_step<T>(c, a, nullptr, dts, nullptr, false, false))) break;
      }
      kpoint(n, 1, a);
      r = fused_step<T>(c, a, anew, dts, d_res + 8 * (size_t)n, false, n == nsteps - 1);
      if (r == SALMON_OK && reortho_every > 0 && (n + 1) % reortho_every == 0) r = gram_schmidt_dev(c);
    }
    return r;
  }
  double aobs[3];
  bool pending = false;
  for (int n = 0; n < nsteps && r == SALMON_OK; ++n) {
    double a[3];
    kpoint(n, 0, a);
    r = fused_step<T>(c, a, pending ? aobs : nullptr, dts, pending ? d_res + 8 * (size_t)(n - 1) : nullptr, false,
                      false);
    if (r == SALMON_OK && etrs) {
      kpoint(n, 1, a);
      r = fused_step<T>(c, a, nullptr, dts, nullptr, false, false);
    }
    for (int d = 0; d < 3; ++d) aobs[d] = ac[3 * (n + 1) + d];
    pending = true;
    if (r == SALMON_OK && reortho_every > 0 && (n + 1) % reortho_every == 0) {
      r = fused_step<T>(c, nullptr, aobs, dt, d_res + 8 * (size_t)n, false, n == nsteps - 1);
      pending = false;
      if (r == SALMON_OK) r = gram_schmidt_dev(c);
    }
  }
  if (r == SALMON_OK && pending) r = fused_step<T>(c, nullptr, aobs, dt, d_res + 8 * (size_t)(nsteps - 1), false, true);
  return r;
}

// nsteps RT steps with the self-consistent potential, in the reference's
// order (dt_evolve.f90:106-216 with no_update_func = 'n', then
// control_sc.f90:203-222): the Taylor-4 step with Vloc(t) (fused or
// streamed, no observables inside), rho(t+dt), Vloc(t+dt) = V_H + Vpsl + V_xc,
// then current and energy at kAc0 + A(t+dt) with the new potential.  ETRS
// (dt_evolve.f90:218-348): Taylor-4 over dt/2 at kAc0 + A_n with Vloc(t),
// the extrapolated potential 3 V - 3 V_old1 + V_old2, Taylor-4 over dt/2 at
// kAc0 + A_{n+1} with it, then the same update from rho(t+dt).
//
// On the fused path the step's launch also sums rho(t+dt) as it stores
// psi(t+dt).  The complex64 kernel's observables of step n are those it
// evaluates at the start of the next launch (or of an observables-only launch
// before a re-orthonormalisation and at the end), which reads psi(t_{n+1})
// with Vloc(t_{n+1}) already in place: the same order as above without a
// separate pass over the orbitals.
template <typename T>
int rt_run_scf(salmon_ctx* c, const double* ac, int nsteps, double dt, int reortho_every, double* d_res) {
  int r = SALMON_OK;
  const bool etrs = c->propagator == SALMON_PROP_ETRS;
  const bool fused = fused_ready(c);
  const bool defer = fused && salmon::fused_defer<T, kFusedXS64>();
  const dim3 gr((c->g.NL + 255) / 256), bl(256);
  double aobs[3];
  bool pending = false;  // observables of the previous step not yet evaluated (deferred form)
  // one Taylor launch; `last` marks the step's final one (it sums the density)
  auto taylor = [&](const double a[3], double h, bool last, int n) {
    if (!fused) return taylor_streamed<T>(c, c->d_kac0, a[0], a[1], a[2], h);
    const bool with_obs = defer && pending;
    const int q = fused_step<T>(c, a, with_obs ? aobs : nullptr, h, with_obs ? d_res + 8 * (size_t)(n - 1) : nullptr,
                                last);
    if (with_obs) pending = false;
    return q;
  };
  for (int n = 0; n < nsteps && r == SALMON_OK; ++n) {
    double amid[3];
    const double* anew = ac + 3 * (n + 1);
    if (etrs) {
      if ((r = taylor(ac + 3 * n, 0.5 * dt, false, n))) break;
      hipLaunchKernelGGL(salmon::scf::k_vloc_extrapolate, gr, bl, 0, c->stream, c->d_vloc, c->d_vloc_f, c->d_vold,
                         c->d_vold + c->g.NL, c->g.NL);
      LAUNCHCHK(c);
      r = taylor(anew, 0.5 * dt, true, n);
    } else {
      for (int d = 0; d < 3; ++d) amid[d] = 0.5 * (ac[3 * n + d] + anew[d]);
      r = taylor(amid, dt, true, n);
    }
    if (r || (!fused && (r = density_dev<T>(c)))) break;
    if ((r = scf_potential(c, c->d_rho, true, c->scf_xc, nullptr, nullptr))) break;
    if (defer) {
      for (int d = 0; d < 3; ++d) aobs[d] = anew[d];
      pending = true;
      if (reortho_every > 0 && (n + 1) % reortho_every == 0) {
        if ((r = fused_step<T>(c, nullptr, aobs, dt, d_res + 8 * (size_t)n))) break;
        pending = false;
        r = gram_schmidt_dev(c);
      }
      continue;
    }
    if (fused) {
      // FP64: an observables-only launch with Vloc(t+dt) in place
      if ((r = fused_step<T>(c, nullptr, anew, dt, d_res + 8 * (size_t)n))) break;
    } else {
      Region reg(c, SALMON_LOG_CURRENT);
      if ((r = observe_rows<T>(c, c->d_kac0, anew[0], anew[1], anew[2]))) break;
      hipLaunchKernelGGL(k_sum_rows, dim3(1), dim3(256), 0, c->stream, c->d_rows, c->NKB, d_res + 8 * (size_t)n);
      LAUNCHCHK(c);
    }
    if (reortho_every > 0 && (n + 1) % reortho_every == 0) r = gram_schmidt_dev(c);
  }
  if (r == SALMON_OK && pending) r = fused_step<T>(c, nullptr, aobs, dt, d_res + 8 * (size_t)(nsteps - 1));
  return r;
}

// The persistent sweep kernel of the batches run for this context alone:
// grids of at least 8 points per axis.
template <typename T>
bool sweep_eligible(const salmon_ctx* c) {
  return sweep_shape<T>(c->g)[0] > 0;
}

// The kernel path rt_run takes.
int rt_path(const salmon_ctx* c) {
  if (fused_ready(c)) return SALMON_PATH_FUSED;
  if (c->path == SALMON_PATH_STREAMED) return SALMON_PATH_STREAMED;
  if (c->path == SALMON_PATH_SWEEP) return SALMON_PATH_SWEEP;  // fails at run time where it does not apply
  const bool ok = c->precision == SALMON_FP64 ? sweep_eligible<double>(c) : sweep_eligible<float>(c);
  if (ok && !(c->solo_fail && c->solo_gen == c->gen)) return SALMON_PATH_SWEEP;
  return SALMON_PATH_STREAMED;
}

// The context's one-system batch, built (descriptors, work buffers, sweep
// plan) for the current buffers.  Runs no step.
template <typename T>
int solo_prepare(salmon_ctx* c) {
  if (!sweep_eligible<T>(c))
    return fail(c, SALMON_ERR_UNSUPPORTED, "persistent sweep: needs >= 8 points per axis");
  if (!c->solo) {
    salmon_batch* b = new salmon_batch();
    b->precision = c->precision;
    b->stream = c->stream;  // not owned: released before the batch is deleted
    b->kernel_req = SALMON_BATCH_SWEEP;
    b->ctx.push_back(c);
    c->solo = b;
    c->solo_gen = c->gen;
  }
  salmon_batch* b = c->solo;
  if (c->solo_gen != c->gen) {
    (void)hipStreamSynchronize(c->stream);
    batch_release(b);
    b->tuned = false;
    c->solo_gen = c->gen;
    c->solo_fail = false;
  }
  if (!b->tuned) {
    const int r = batch_tune<T>(b);
    if (r) {
      batch_release(b);  // frees what the failed set-up allocated
      b->tuned = false;
      c->solo_fail = true;
      (void)hipGetLastError();  // a failed allocation must not show up as a later launch's error
      return fail(c, r, "persistent sweep: %s", b->err.c_str());
    }
  }
  return SALMON_OK;
}

// nsteps steps through the persistent sweep, the observables closing each
// run (and each stretch before a re-orthonormalisation) as in the batches;
// the density of the last step's orbitals, taken (as psi_rho_RT follows the
// Taylor step) before the re-orthonormalisation that may end the run.
// Step by step on the persistent sweep, for ETRS and for the self-consistent
// potential (the order of rt_step and rt_run_scf): the Taylor-only launch over
// dt at kAc0 + (A_n + A_{n+1})/2, or ETRS's two over dt/2 at kAc0 + A_n and
// kAc0 + A_{n+1} (with the self-consistent potential the extrapolated one
// between them); the density (every step with the self-consistent
// potential, else after the last); the potential update from it; the
// observables-only launch at kAc0 + A_{n+1} with the potential then in place.
template <typename T>
int rt_run_sweep_steps(salmon_ctx* c, const double* ac, int nsteps, double dt, int reortho_every, double* jt,
                       double* et, double* norm) {
  salmon_batch* b = c->solo;
  const bool etrs = c->propagator == SALMON_PROP_ETRS;
  const size_t nk = (size_t)c->NK;
  if (c->kac0_h.size() != 3 * nk) return fail(c, SALMON_ERR_STATE, "salmon_gpu_set_kpoints not called");
  // one host table per upload of a step: the step's observables call
  // synchronises the stream before the tables are written again
  std::vector<double> k[2] = {std::vector<double>(3 * nk), std::vector<double>(3 * nk)};
  auto upload_at = [&](int slot, const double a[3]) {
    for (int d = 0; d < 3; ++d)
      for (size_t i = 0; i < nk; ++i) k[slot][d * nk + i] = c->kac0_h[d * nk + i] + a[d];
    return upload_kac(c, k[slot].data());
  };
  auto taylor = [&](double h) {
    const int q = batch_taylor<T>(b, &c->d_kac, h);
    return q ? fail(c, q, "persistent sweep: %s", b->err.c_str()) : SALMON_OK;
  };
  const dim3 gr((c->g.NL + 255) / 256), bl(256);
  int r;
  for (int n = 0; n < nsteps; ++n) {
    const double* a0 = ac + 3 * (size_t)n;
    const double* a1 = ac + 3 * (size_t)(n + 1);
    if (etrs) {
      if ((r = upload_at(0, a0)) || (r = taylor(0.5 * dt))) return r;
      if (c->scf) {
        hipLaunchKernelGGL(salmon::scf::k_vloc_extrapolate, gr, bl, 0, c->stream, c->d_vloc, c->d_vloc_f, c->d_vold,
                           c->d_vold + c->g.NL, c->g.NL);
        LAUNCHCHK(c);
      }
      if ((r = upload_at(1, a1)) || (r = taylor(0.5 * dt))) return r;
    } else {
      const double amid[3] = {0.5 * (a0[0] + a1[0]), 0.5 * (a0[1] + a1[1]), 0.5 * (a0[2] + a1[2])};
      if ((r = upload_at(0, amid)) || (r = taylor(dt)) || (r = upload_at(1, a1))) return r;
    }
    if ((c->scf || n == nsteps - 1) && (r = density_dev<T>(c))) return r;
    if (c->scf && (r = scf_potential(c, c->d_rho, true, c->scf_xc, nullptr, nullptr))) return r;
    double q[8];
    if ((r = batch_observe<T>(b, &c->d_kac, q))) return fail(c, r, "persistent sweep: %s", b->err.c_str());
    if (jt)
      for (int d = 0; d < 3; ++d) jt[3 * (size_t)n + d] = q[d];
    if (et) et[n] = q[3] + q[4] + q[5];
    if (norm) norm[n] = q[6];
    if (reortho_every > 0 && (n + 1) % reortho_every == 0 && (r = gram_schmidt_dev(c))) return r;
  }
  return SALMON_OK;
}

template <typename T>
int rt_run_sweep(salmon_ctx* c, const double* ac, int nsteps, double dt, int reortho_every, double* jt, double* et,
                 double* norm) {
  if (c->propagator == SALMON_PROP_ETRS || c->scf)
    return rt_run_sweep_steps<T>(c, ac, nsteps, dt, reortho_every, jt, et, norm);
  salmon_batch* b = c->solo;
  for (int n0 = 0; n0 < nsteps;) {
    const int seg = reortho_every > 0 ? std::min(reortho_every - n0 % reortho_every, nsteps - n0) : nsteps - n0;
    int r = batch_run<T>(b, ac + 3 * (size_t)n0, seg, dt, jt ? jt + 3 * (size_t)n0 : nullptr, et ? et + n0 : nullptr,
                         norm ? norm + n0 : nullptr, false);
    if (r) return fail(c, r, "persistent sweep: %s", b->err.c_str());
    n0 += seg;
    if (n0 == nsteps && (r = density_dev<T>(c))) return r;
    if (reortho_every > 0 && n0 % reortho_every == 0 && (r = gram_schmidt_dev(c))) return r;
  }
  return SALMON_OK;
}

// Fixed device scratch of salmon_gpu_hartree_xc_noalloc: the DFT twiddles
// (3 axes x kMaxN x (cos, sin)) followed by two complex grids of at most
// SALMON_NOALLOC_MAX_NL points.  In host code `scratch` only names the
// symbol; its device address on each GPU comes from hipGetSymbolAddress
// (noalloc_setup) and only that address is handed to kernels.
constexpr size_t kNoAllocTw = (size_t)3 * 2 * salmon::scf::kMaxN;
constexpr size_t kNoAllocDoubles = kNoAllocTw + (size_t)4 * SALMON_NOALLOC_MAX_NL;
static __device__ __align__(256) double scratch[kNoAllocDoubles];

constexpr int kNoAllocMaxDevices = 64;
std::mutex g_noalloc_mu;
double* g_noalloc_ptr[kNoAllocMaxDevices] = {};
hipError_t g_noalloc_err[kNoAllocMaxDevices] = {};

// Once per device, with that device current (salmon_gpu_init): the device
// address of this GPU's scratch copy.
void noalloc_setup(int device) {
  if (device < 0 || device >= kNoAllocMaxDevices) return;
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (g_noalloc_ptr[device]) return;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  g_noalloc_err[device] = e;
  if (e == hipSuccess) g_noalloc_ptr[device] = (double*)p;
}

double* noalloc_ptr(int device, hipError_t* err) {
  if (device < 0 || device >= kNoAllocMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  *err = g_noalloc_err[device];
  return g_noalloc_ptr[device];
}

// (cos, sin)(2 pi j / N_a) for j < N_a, zero rows beyond, as scf_twiddles
// lays them out
__global__ void __launch_bounds__(256) k_noalloc_twiddles(double* tw, int nx, int ny, int nz) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= 3 * salmon::scf::kMaxN) return;
  const int a = q / salmon::scf::kMaxN, j = q - a * salmon::scf::kMaxN;
  const int n = a == 0 ? nx : (a == 1 ? ny : nz);
  double s = 0.0, c = 0.0;
  if (j < n) sincospi(2.0 * j / n, &s, &c);
  tw[2 * q] = c;
  tw[2 * q + 1] = s;
}

}  // namespace

extern "C" {

const char* salmon_gpu_errstr(int code) {
  switch (code) {
    case SALMON_OK: return "ok";
    case SALMON_ERR_ARG: return "invalid argument";
    case SALMON_ERR_HIP: return "HIP runtime error";
    case SALMON_ERR_STATE: return "missing set-up call";
    case SALMON_ERR_NOMEM: return "device memory allocation failed";
    case SALMON_ERR_UNSUPPORTED: return "unsupported option";
    default: return "unknown error";
  }
}

// message of this thread's last failed salmon_gpu_init (no context holds it)
static thread_local std::string tl_init_err;

const char* salmon_gpu_last_error(const salmon_ctx* c) {
  if (c) return c->err.c_str();
  return tl_init_err.empty() ? "no context" : tl_init_err.c_str();
}

int salmon_gpu_init(const salmon_grid_t* grid, int precision, salmon_ctx** out) {
  if (!grid || !out) return SALMON_ERR_ARG;
  *out = nullptr;
  tl_init_err.clear();
  // the stencil wraps each neighbour index once, as the reference's periodic tables
  // modx/mody/modz do (ARTED/modules/opt_variables.f90:155-167): Nd = 4 points per axis or more
  if (grid->NLx < 4 || grid->NLy < 4 || grid->NLz < 4) return SALMON_ERR_ARG;
  if (precision != SALMON_FP64 && precision != SALMON_MIXED) return SALMON_ERR_ARG;
  if (!(grid->aLx > 0 && grid->aLy > 0 && grid->aLz > 0)) return SALMON_ERR_ARG;
  salmon_ctx* c = new salmon_ctx();
  c->precision = precision;
  hipError_t e = hipGetDevice(&c->device);
  if (e == hipSuccess) {
    e = hipStreamCreate(&c->stream);
    if (e != hipSuccess) tl_init_err = std::string("hipStreamCreate: ") + hipGetErrorString(e);
  } else {
    tl_init_err = std::string("hipGetDevice: ") + hipGetErrorString(e);
  }
  if (e != hipSuccess) {
    delete c;
    return SALMON_ERR_HIP;
  }
  Grid& g = c->g;
  g.NLx = grid->NLx; g.NLy = grid->NLy; g.NLz = grid->NLz;
  g.NL = g.NLx * g.NLy * g.NLz;
  g.aLx = grid->aLx; g.aLy = grid->aLy; g.aLz = grid->aLz;
  g.aLxyz = g.aLx * g.aLy * g.aLz;
  g.Hx = g.aLx / g.NLx; g.Hy = g.aLy / g.NLy; g.Hz = g.aLz / g.NLz;
  g.Hxyz = g.Hx * g.Hy * g.Hz;
  for (int n = 0; n < 5; ++n) {
    g.lap[0][n] = grid->lapx[n]; g.lap[1][n] = grid->lapy[n]; g.lap[2][n] = grid->lapz[n];
    g.nab[0][n] = grid->nabx[n]; g.nab[1][n] = grid->naby[n]; g.nab[2][n] = grid->nabz[n];
  }
  int r;
  if ((r = dalloc(c, &c->d_vloc, g.NL)) || (r = dalloc(c, &c->d_vloc_f, g.NL)) || (r = dalloc(c, &c->d_rho, g.NL)) ||
      (r = dalloc(c, &c->d_kone, 3)) || (r = dalloc(c, &c->d_oneocc, 1)) || (r = dalloc(c, &c->d_sum, 8))) {
    salmon_gpu_finalize(c);
    return r;
  }
  for (int i = 0; i < 3; ++i) {
    if (hipMalloc(&c->d_stage[i], (size_t)g.NL * sizeof(cplx<double>)) != hipSuccess) {
      salmon_gpu_finalize(c);
      return SALMON_ERR_NOMEM;
    }
  }
  const double one = 1.0;
  if (hipMemcpy(c->d_oneocc, &one, sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
    salmon_gpu_finalize(c);
    return SALMON_ERR_HIP;
  }
  c->fused = fused_config(c);
  noalloc_setup(c->device);
  *out = c;
  return SALMON_OK;
}

int salmon_gpu_set_potential(salmon_ctx* c, const double* vloc) {
  touch(c);
  if (!c) return SALMON_ERR_ARG;
  if (!vloc) return fail(c, SALMON_ERR_ARG, "vloc is NULL");
  const int NL = c->g.NL;
  std::vector<float> vf(NL);
  for (int i = 0; i < NL; ++i) vf[i] = (float)vloc[i];
  HIPCHK(c, hipMemcpy(c->d_vloc, vloc, sizeof(double) * NL, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->d_vloc_f, vf.data(), sizeof(float) * NL, hipMemcpyHostToDevice));
  c->have_v = true;
  return SALMON_OK;
}

int salmon_gpu_set_projectors(salmon_ctx* c, int NI, int Nps, const int* Mps, const int* Jxyz, const int* Jxx,
                              const int* Jyy, const int* Jzz, int Nlma, const int* a_tbl, const double* uV,
                              const int* iuV) {
  touch(c);
  if (!c) return SALMON_ERR_ARG;
  if (NI < 0 || Nps < 0 || Nlma < 0) return fail(c, SALMON_ERR_ARG, "negative table size");
  if ((NI > 0 && (!Mps || !Jxyz || !Jxx || !Jyy || !Jzz)) || (Nlma > 0 && (!a_tbl || !uV || !iuV)))
    return fail(c, SALMON_ERR_ARG, "projector table pointer is NULL");
  const Grid& g = c->g;
  std::vector<int> start(NI + 1, 0);
  for (int a = 0; a < NI; ++a) {
    if (Mps[a] < 0 || Mps[a] > Nps) return fail(c, SALMON_ERR_ARG, "Mps(%d)=%d outside 0..Nps=%d", a + 1, Mps[a], Nps);
    start[a + 1] = start[a] + Mps[a];
  }
  const int NPI = start[NI];
  std::vector<int> pidx(NPI);
  std::vector<double> pos(3 * (size_t)NPI);
  for (int a = 0; a < NI; ++a)
    for (int j = 0; j < Mps[a]; ++j) {
      const int i = Jxyz[(size_t)a * Nps + j];
      if (i < 1 || i > g.NL) return fail(c, SALMON_ERR_ARG, "Jxyz(%d,%d)=%d outside 1..NL", j + 1, a + 1, i);
      const int i0 = i - 1, jp = start[a] + j;
      const int lx = i0 / (g.NLy * g.NLz), ly = (i0 / g.NLz) % g.NLy, lz = i0 % g.NLz;
      pidx[jp] = i0;
      pos[3 * jp + 0] = lx * g.Hx - Jxx[(size_t)a * Nps + j] * g.aLx;
      pos[3 * jp + 1] = ly * g.Hy - Jyy[(size_t)a * Nps + j] * g.aLy;
      pos[3 * jp + 2] = lz * g.Hz - Jzz[(size_t)a * Nps + j] * g.aLz;
    }
  std::vector<int> lma_atom(Nlma);
  std::vector<double> sign(Nlma);
  for (int l = 0; l < Nlma; ++l) {
    const int a = a_tbl[l] - 1;
    if (a < 0 || a >= NI) return fail(c, SALMON_ERR_ARG, "a_tbl(%d)=%d outside 1..NI", l + 1, a_tbl[l]);
    if (l > 0 && a < lma_atom[l - 1]) return fail(c, SALMON_ERR_ARG, "a_tbl must group projectors by atom in order");
    lma_atom[l] = a;
    sign[l] = (double)iuV[l];
  }
  // first projector of each atom (atoms without projectors: an empty range)
  std::vector<int> atom_lma(NI + 1, Nlma);
  for (int l = Nlma - 1; l >= 0; --l) atom_lma[lma_atom[l]] = l;
  for (int a = NI - 1; a >= 0; --a) atom_lma[a] = std::min(atom_lma[a], atom_lma[a + 1]);
  int max_lpa = 0;
  for (int a = 0; a < NI; ++a) max_lpa = std::max(max_lpa, atom_lma[a + 1] - atom_lma[a]);
  // transposed table: per touched grid point, its (projector, packed point) entries
  std::vector<std::vector<std::pair<int, int>>> per(g.NL);
  for (int l = 0; l < Nlma; ++l) {
    const int a = lma_atom[l];
    for (int j = 0; j < Mps[a]; ++j) per[pidx[start[a] + j]].push_back({start[a] + j, l});
  }
  std::vector<int> pt2t(g.NL, -1), tstart(1, 0), elma, ejp;
  int NT = 0;
  for (int p = 0; p < g.NL; ++p) {
    if (per[p].empty()) continue;
    std::sort(per[p].begin(), per[p].end());
    pt2t[p] = NT++;
    for (auto& e : per[p]) {
      ejp.push_back(e.first);
      elma.push_back(e.second);
    }
    tstart.push_back((int)ejp.size());
  }
  // the old tables, projection and work buffers go: drop graphs that captured them
  (void)hipStreamSynchronize(c->stream);
  ++c->gen;
  for (void* p : c->kb_bufs) (void)hipFree(p);
  c->kb_bufs.clear();
  auto up = [&](const void* src, size_t bytes, void** dst) -> int {
    void* d = nullptr;
    if (hipMalloc(&d, bytes ? bytes : 4) != hipSuccess) return fail(c, SALMON_ERR_NOMEM, "projector tables");
    c->kb_bufs.push_back(d);
    if (bytes && hipMemcpy(d, src, bytes, hipMemcpyHostToDevice) != hipSuccess)
      return fail(c, SALMON_ERR_HIP, "projector table upload");
    *dst = d;
    return SALMON_OK;
  };
  KBTables kb{};
  kb.NI = NI; kb.Nps = Nps; kb.Nlma = Nlma; kb.NPI = NPI; kb.NT = NT; kb.NE = (int)ejp.size();
  kb.max_lpa = max_lpa;
  int r;
  if ((r = up(start.data(), sizeof(int) * start.size(), (void**)&kb.atom_start)) ||
      (r = up(pidx.data(), sizeof(int) * pidx.size(), (void**)&kb.pidx)) ||
      (r = up(pos.data(), sizeof(double) * pos.size(), (void**)&kb.pos)) ||
      (r = up(lma_atom.data(), sizeof(int) * lma_atom.size(), (void**)&kb.lma_atom)) ||
      (r = up(atom_lma.data(), sizeof(int) * atom_lma.size(), (void**)&kb.atom_lma)) ||
      (r = up(sign.data(), sizeof(double) * sign.size(), (void**)&kb.lma_sign)) ||
      (r = up(uV, sizeof(double) * (size_t)Nps * Nlma, (void**)&kb.uv)) ||
      (r = up(pt2t.data(), sizeof(int) * pt2t.size(), (void**)&kb.pt2t)) ||
      (r = up(tstart.data(), sizeof(int) * tstart.size(), (void**)&kb.t_start)) ||
      (r = up(elma.data(), sizeof(int) * elma.size(), (void**)&kb.e_lma)) ||
      (r = up(ejp.data(), sizeof(int) * ejp.size(), (void**)&kb.e_jp)))
    return r;
  c->kb = kb;
  c->have_kb = true;
  dfree(&c->d_uvpsi);
  for (int i = 0; i < 4; ++i) dfree(&c->d_work[i]);
  int fr = fused_set_tables(c);
  if (fr) return fr;
  return SALMON_OK;
}

int salmon_gpu_set_kpoints(salmon_ctx* c, int NK, int NB, const double* kAc0, const double* occ) {
  touch(c);
  if (!c) return SALMON_ERR_ARG;
  if (NK < 1 || NB < 1 || !kAc0 || !occ) return fail(c, SALMON_ERR_ARG, "NK, NB must be >= 1 with kAc0, occ given");
  if ((size_t)NK * NB * c->g.NL > ((size_t)1 << 40)) return fail(c, SALMON_ERR_ARG, "orbital set too large");
  (void)hipStreamSynchronize(c->stream);
  ++c->gen;  // orbitals, k tables and work buffers are replaced
  c->NK = NK; c->NB = NB; c->NKB = NK * NB;
  int r;
  if ((r = dalloc(c, &c->d_kac0, 3 * (size_t)NK)) || (r = dalloc(c, &c->d_kac, 3 * (size_t)NK)) ||
      (r = dalloc(c, &c->d_occ, (size_t)c->NKB)) || (r = dalloc(c, &c->d_rows, 8 * (size_t)c->NKB)))
    return r;
  HIPCHK(c, hipMemcpy(c->d_kac0, kAc0, sizeof(double) * 3 * NK, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->d_occ, occ, sizeof(double) * c->NKB, hipMemcpyHostToDevice));
  c->kac0_h.assign(kAc0, kAc0 + 3 * (size_t)NK);
  dfree(&c->d_zu);
  dfree((void**)&c->d_f64);
  dfree((void**)&c->d_rho_part);
  dfree(&c->d_uvpsi);
  for (int i = 0; i < 4; ++i) dfree(&c->d_work[i]);
  c->have_zu = false;
  void* z = nullptr;
  HIPCHK(c, hipMalloc(&z, (size_t)c->NKB * c->g.NL * elem_size(c)));
  c->d_zu = z;
  return SALMON_OK;
}

int salmon_gpu_put_orbitals(salmon_ctx* c, const double* zu) {
  touch(c);
  if (!c) return SALMON_ERR_ARG;
  if (!c->d_zu) return fail(c, SALMON_ERR_STATE, "salmon_gpu_set_kpoints not called");
  if (!zu) return fail(c, SALMON_ERR_ARG, "zu is NULL");
  const size_t n = (size_t)c->NKB * c->g.NL;
  if (c->precision == SALMON_FP64) {
    HIPCHK(c, hipMemcpy(c->d_zu, zu, n * sizeof(cplx<double>), hipMemcpyHostToDevice));
  } else {
    // complex(8) rows are converted to complex64 on the device, a block of rows at a time
    const size_t blk = (size_t)c->g.NL, rows = std::max<size_t>(1, std::min<size_t>(c->NKB, (64u << 20) / (blk * 16)));
    cplx<double>* st = nullptr;
    HIPCHK(c, hipMalloc((void**)&st, rows * blk * sizeof(cplx<double>)));
    int r = SALMON_OK;
    for (size_t o = 0; o < (size_t)c->NKB && r == SALMON_OK; o += rows) {
      const size_t m = std::min(rows, (size_t)c->NKB - o) * blk;
      hipError_t e = hipMemcpyAsync(st, zu + 2 * o * blk, m * sizeof(cplx<double>), hipMemcpyHostToDevice, c->stream);
      if (e == hipSuccess) {
        hipLaunchKernelGGL((k_convert<double, float>), dim3((unsigned)((m + 255) / 256)), dim3(256), 0, c->stream, st,
                           (cplx<float>*)c->d_zu + o * blk, m);
        e = hipGetLastError();
      }
      if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
      if (e != hipSuccess) r = fail(c, SALMON_ERR_HIP, "put_orbitals: %s", hipGetErrorString(e));
    }
    (void)hipFree(st);
    if (r) return r;
  }
  c->have_zu = true;
  return SALMON_OK;
}

int salmon_gpu_get_orbitals(salmon_ctx* c, double* zu) {
  if (!c) return SALMON_ERR_ARG;
  if (!c->have_zu) return fail(c, SALMON_ERR_STATE, "no resident orbitals");
  if (!zu) return fail(c, SALMON_ERR_ARG, "zu is NULL");
  const size_t n = (size_t)c->NKB * c->g.NL;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (c->precision == SALMON_FP64) {
    HIPCHK(c, hipMemcpy(zu, c->d_zu, n * sizeof(cplx<double>), hipMemcpyDeviceToHost));
  } else {
    const size_t blk = (size_t)c->g.NL, rows = std::max<size_t>(1, std::min<size_t>(c->NKB, (64u << 20) / (blk * 16)));
    cplx<double>* st = nullptr;
    HIPCHK(c, hipMalloc((void**)&st, rows * blk * sizeof(cplx<double>)));
    int r = SALMON_OK;
    for (size_t o = 0; o < (size_t)c->NKB && r == SALMON_OK; o += rows) {
      const size_t m = std::min(rows, (size_t)c->NKB - o) * blk;
      hipLaunchKernelGGL((k_convert<float, double>), dim3((unsigned)((m + 255) / 256)), dim3(256), 0, c->stream,
                         (const cplx<float>*)c->d_zu + o * blk, st, m);
      hipError_t e = hipGetLastError();
      if (e == hipSuccess) e = hipMemcpyAsync(zu + 2 * o * blk, st, m * sizeof(cplx<double>), hipMemcpyDeviceToHost, c->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
      if (e != hipSuccess) r = fail(c, SALMON_ERR_HIP, "get_orbitals: %s", hipGetErrorString(e));
    }
    (void)hipFree(st);
    if (r) return r;
  }
  return SALMON_OK;
}

int salmon_gpu_hpsi(salmon_ctx* c, const double kAc[3], const double* in, double* out, double* ttpsi) {
  int r = check_ready(c, false);
  if (r) return r;
  if (!kAc || !in || !out) return fail(c, SALMON_ERR_ARG, "NULL argument");
  if ((r = ensure_work(c))) return r;
  const size_t NL = c->g.NL;
  HIPCHK(c, hipMemcpyAsync(c->d_kone, kAc, 3 * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_stage[0], in, NL * sizeof(cplx<double>), hipMemcpyHostToDevice, c->stream));
  if (c->precision == SALMON_FP64) {
    if ((r = apply_h<double>(c, (const cplx<double>*)c->d_stage[0], (cplx<double>*)c->d_stage[1],
                             ttpsi ? (cplx<double>*)c->d_stage[2] : nullptr, 0, 1, 1, 1, c->d_kone, 0, 0, 0)))
      return r;
  } else {
    cplx<float>* w0 = (cplx<float>*)c->d_work[0];
    cplx<float>* w1 = (cplx<float>*)c->d_work[1];
    cplx<float>* w2 = (cplx<float>*)c->d_work[2];
    const unsigned nb = (unsigned)((NL + 255) / 256);
    hipLaunchKernelGGL((k_convert<double, float>), dim3(nb), dim3(256), 0, c->stream,
                       (const cplx<double>*)c->d_stage[0], w0, NL);
    LAUNCHCHK(c);
    if ((r = apply_h<float>(c, w0, w1, ttpsi ? w2 : nullptr, 0, 1, 1, 1, c->d_kone, 0, 0, 0))) return r;
    hipLaunchKernelGGL((k_convert<float, double>), dim3(nb), dim3(256), 0, c->stream, w1,
                       (cplx<double>*)c->d_stage[1], NL);
    LAUNCHCHK(c);
    if (ttpsi) {
      hipLaunchKernelGGL((k_convert<float, double>), dim3(nb), dim3(256), 0, c->stream, w2,
                         (cplx<double>*)c->d_stage[2], NL);
      LAUNCHCHK(c);
    }
  }
  HIPCHK(c, hipMemcpyAsync(out, c->d_stage[1], NL * sizeof(cplx<double>), hipMemcpyDeviceToHost, c->stream));
  if (ttpsi)
    HIPCHK(c, hipMemcpyAsync(ttpsi, c->d_stage[2], NL * sizeof(cplx<double>), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SALMON_OK;
}

int salmon_gpu_hamiltonian(salmon_ctx* c, const double* kAc, double dt) {
  touch(c);
  int r = check_ready(c, true);
  if (r) return r;
  if ((r = upload_kac(c, kAc))) return r;
  if (fused_ready(c)) {
    // the fused step without observables or density: Taylor-4 at kAc
    const double zero[3] = {0.0, 0.0, 0.0};
    return c->precision == SALMON_FP64 ? fused_step<double>(c, zero, nullptr, dt, nullptr, false, false, c->d_kac)
                                       : fused_step<float>(c, zero, nullptr, dt, nullptr, false, false, c->d_kac);
  }
  if (rt_path(c) == SALMON_PATH_SWEEP) {
    // the persistent sweep without observables at kAc; a set-up that fails
    // leaves SALMON_PATH_AUTO with the streamed kernels
    r = c->precision == SALMON_FP64 ? solo_prepare<double>(c) : solo_prepare<float>(c);
    if (r == SALMON_OK) {
      Region reg(c, SALMON_LOG_DT_EVOLVE);
      r = c->precision == SALMON_FP64 ? batch_taylor<double>(c->solo, &c->d_kac, dt)
                                      : batch_taylor<float>(c->solo, &c->d_kac, dt);
      return r ? fail(c, r, "persistent sweep: %s", c->solo->err.c_str()) : SALMON_OK;
    }
    if (c->path != SALMON_PATH_AUTO) return r;
  }
  if (c->precision == SALMON_FP64) return taylor_streamed<double>(c, c->d_kac, 0, 0, 0, dt);
  return taylor_streamed<float>(c, c->d_kac, 0, 0, 0, dt);
}

// Sums of the observables rows of the resident orbitals at k = kbase (J,
// Ekin, Enl, Eloc, norm, as k_sum_rows lays them out): one observables-only
// launch of rt_run's kernel (the fused step on 16^3, the persistent sweep on
// other grids of at least 8 points per axis), otherwise k_observe per orbital.
static int observe_sums(salmon_ctx* c, const double* kbase, double out[8]) {
  int r;
  if (fused_ready(c)) {
    const double zero[3] = {0.0, 0.0, 0.0};
    r = c->precision == SALMON_FP64 ? fused_step<double>(c, nullptr, zero, 0.0, c->d_sum, false, false, kbase)
                                    : fused_step<float>(c, nullptr, zero, 0.0, c->d_sum, false, false, kbase);
    if (r) return r;
    HIPCHK(c, hipMemcpyAsync(out, c->d_sum, sizeof(double) * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SALMON_OK;
  }
  if (rt_path(c) == SALMON_PATH_SWEEP) {
    // one observables-only persistent sweep launch; a set-up that fails
    // leaves SALMON_PATH_AUTO with k_observe
    r = c->precision == SALMON_FP64 ? solo_prepare<double>(c) : solo_prepare<float>(c);
    if (r == SALMON_OK) {
      const double* const* kb = kbase == c->d_kac0 ? &c->d_kac0 : &c->d_kac;
      r = c->precision == SALMON_FP64 ? batch_observe<double>(c->solo, kb, out) : batch_observe<float>(c->solo, kb, out);
      return r ? fail(c, r, "persistent sweep: %s", c->solo->err.c_str()) : SALMON_OK;
    }
    if (c->path != SALMON_PATH_AUTO) return r;
  }
  r = c->precision == SALMON_FP64 ? observe_rows<double>(c, kbase, 0, 0, 0) : observe_rows<float>(c, kbase, 0, 0, 0);
  if (r) return r;
  return read_sums(c, out);
}

// observe_sums at the caller's kAc(NK,3), or the sums the previous current or
// energy call formed at the same kAc with nothing changed since
static int observe_sums_at(salmon_ctx* c, const double* kAc, double out[8]) {
  const size_t nk = 3 * (size_t)c->NK;
  if (kAc && c->obs_epoch == c->epoch && c->obs_k.size() == nk && std::equal(kAc, kAc + nk, c->obs_k.begin())) {
    std::copy(c->obs_sums, c->obs_sums + 8, out);
    return SALMON_OK;
  }
  int r;
  if ((r = upload_kac(c, kAc))) return r;
  if ((r = observe_sums(c, c->d_kac, out))) return r;
  c->obs_k.assign(kAc, kAc + nk);
  std::copy(out, out + 8, c->obs_sums);
  c->obs_epoch = c->epoch;
  return SALMON_OK;
}

int salmon_gpu_current(salmon_ctx* c, const double* kAc, double jav[3]) {
  int r = check_ready(c, true);
  if (r) return r;
  Region reg(c, SALMON_LOG_CURRENT);
  if (!jav) return fail(c, SALMON_ERR_ARG, "jav is NULL");
  double s[8];
  if ((r = observe_sums_at(c, kAc, s))) return r;
  jav[0] = s[0]; jav[1] = s[1]; jav[2] = s[2];
  return SALMON_OK;
}

int salmon_gpu_energy(salmon_ctx* c, const double* kAc, double* ekin, double* enl, double* eloc) {
  int r = check_ready(c, true);
  if (r) return r;
  Region reg(c, SALMON_LOG_TOTAL_ENERGY);
  double s[8];
  if ((r = observe_sums_at(c, kAc, s))) return r;
  if (ekin) *ekin = s[3];
  if (enl) *enl = s[4];
  if (eloc) *eloc = s[5];
  return SALMON_OK;
}

int salmon_gpu_norm(salmon_ctx* c, double* norm) {
  int r = check_ready(c, true);
  if (r) return r;
  if (!norm) return fail(c, SALMON_ERR_ARG, "norm is NULL");
  double s[8];
  if ((r = observe_sums(c, c->d_kac0, s))) return r;
  *norm = s[6];
  return SALMON_OK;
}

int salmon_gpu_density(salmon_ctx* c, double* rho) {
  int r = check_ready(c, true);
  if (r) return r;
  if (!rho) return fail(c, SALMON_ERR_ARG, "rho is NULL");
  r = c->precision == SALMON_FP64 ? density_dev<double>(c) : density_dev<float>(c);
  if (r) return r;
  HIPCHK(c, hipMemcpyAsync(rho, c->d_rho, sizeof(double) * c->g.NL, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SALMON_OK;
}

int salmon_gpu_gram_schmidt(salmon_ctx* c) {
  touch(c);
  // needs the orbitals only (no potential or projectors)
  if (!c) return SALMON_ERR_ARG;
  if (c->NKB == 0) return fail(c, SALMON_ERR_STATE, "salmon_gpu_set_kpoints not called");
  if (!c->have_zu) return fail(c, SALMON_ERR_STATE, "salmon_gpu_put_orbitals not called");
  return gram_schmidt_dev(c);
}

int salmon_gpu_sp_energy(salmon_ctx* c, const double* kAc, double* esp) {
  int r = check_ready(c, true);
  if (r) return r;
  Region reg(c, SALMON_LOG_SP_ENERGY);
  if (!esp) return fail(c, SALMON_ERR_ARG, "esp is NULL");
  if ((r = upload_kac(c, kAc))) return r;
  if ((r = ensure_work(c))) return r;
  const size_t NL = c->g.NL;
  std::vector<double> tmp;
  double* d_dots = c->d_rows;  // (NKB*2) fits in NKB*8
  for (int o0 = 0; o0 < c->NKB; o0 += c->chunk) {
    const int n = std::min(c->chunk, c->NKB - o0);
    if (c->precision == SALMON_FP64) {
      const cplx<double>* zu = (const cplx<double>*)c->d_zu + (size_t)o0 * NL;
      if ((r = apply_h<double>(c, zu, (cplx<double>*)c->d_work[0], nullptr, o0, n, c->NB, c->NK, c->d_kac, 0, 0, 0)))
        return r;
      hipLaunchKernelGGL(k_dot_rows<double>, dim3(n), dim3(256), 0, c->stream, zu, (const cplx<double>*)c->d_work[0],
                         c->g.NL, c->g.Hxyz, d_dots + 2 * (size_t)o0);
    } else {
      const cplx<float>* zu = (const cplx<float>*)c->d_zu + (size_t)o0 * NL;
      if ((r = apply_h<float>(c, zu, (cplx<float>*)c->d_work[0], nullptr, o0, n, c->NB, c->NK, c->d_kac, 0, 0, 0)))
        return r;
      hipLaunchKernelGGL(k_dot_rows<float>, dim3(n), dim3(256), 0, c->stream, zu, (const cplx<float>*)c->d_work[0],
                         c->g.NL, c->g.Hxyz, d_dots + 2 * (size_t)o0);
    }
    LAUNCHCHK(c);
  }
  tmp.resize(2 * (size_t)c->NKB);
  HIPCHK(c, hipMemcpyAsync(tmp.data(), d_dots, sizeof(double) * tmp.size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int o = 0; o < c->NKB; ++o) esp[o] = tmp[2 * (size_t)o];
  return SALMON_OK;
}

static int gs_check(salmon_ctx* c, const double* kAc) {
  int r = check_ready(c, true);
  if (r) return r;
  if (c->precision != SALMON_FP64) return fail(c, SALMON_ERR_UNSUPPORTED, "the ground-state path runs on FP64 contexts");
  if (c->g.NLx > kDftMax || c->g.NLy > kDftMax || c->g.NLz > kDftMax)
    return fail(c, SALMON_ERR_UNSUPPORTED, "grid axes longer than %d", kDftMax);
  if (c->NB > kCgMaxNB) return fail(c, SALMON_ERR_UNSUPPORTED, "more than %d bands", kCgMaxNB);
  if ((r = upload_kac(c, kAc))) return r;
  if ((r = ensure_work(c))) return r;
  return gs_ensure(c);
}

static int check_precond(salmon_ctx* c, int precond) {
  if (precond != SALMON_PC_NONE && precond != SALMON_PC_KINETIC && precond != SALMON_PC_LEARNED)
    return fail(c, SALMON_ERR_ARG, "unknown preconditioner %d", precond);
  if (precond == SALMON_PC_LEARNED && !c->pcnet)
    return fail(c, SALMON_ERR_STATE, "learned preconditioner selected but no network set (salmon_gpu_set_pcnet)");
  return SALMON_OK;
}

static int cg_results(salmon_ctx* c, double* esp_var, int* iters) {
  GsState* s = (GsState*)c->gs;
  if (esp_var)
    HIPCHK(c, hipMemcpyAsync(esp_var, s->esp_var, sizeof(double) * c->NKB, hipMemcpyDeviceToHost, c->stream));
  if (iters) HIPCHK(c, hipMemcpyAsync(iters, s->iters, sizeof(int) * c->NKB, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SALMON_OK;
}

int salmon_gpu_cg(salmon_ctx* c, const double* kAc, int iter_cg_max, int precond, double* esp_var, int* iters) {
  touch(c);
  int r = gs_check(c, kAc);
  if (r) return r;
  Region reg(c, SALMON_LOG_CG);
  if (iter_cg_max < 0) return fail(c, SALMON_ERR_ARG, "iter_cg_max < 0");
  if ((r = check_precond(c, precond))) return r;
  if ((r = cg_call(c, iter_cg_max, precond))) return r;
  return cg_results(c, esp_var, iters);
}

int salmon_gpu_subspace_diag(salmon_ctx* c, const double* kAc, double* esp) {
  touch(c);
  int r = gs_check(c, kAc);
  if (r) return r;
  Region reg(c, SALMON_LOG_DIAG);
  return subspace_diag(c, esp);
}

int salmon_gpu_ground_state(salmon_ctx* c, const double* kAc, int max_sweeps, int iter_cg_max, int precond,
                            double tol, double* esp, double* esp_var, int* sweeps, double* history) {
  touch(c);
  int r = gs_check(c, kAc);
  if (r) return r;
  Region reg(c, SALMON_LOG_GROUND_STATE);
  if (max_sweeps < 1 || iter_cg_max < 1) return fail(c, SALMON_ERR_ARG, "max_sweeps and iter_cg_max must be >= 1");
  if ((r = check_precond(c, precond))) return r;
  std::vector<double> e0(c->NKB, 0.0), e1(c->NKB), var(c->NKB);
  int n = 0, pc = precond;
  double prev_drop = 0.0, prev_sum = 0.0;
  for (n = 1; n <= max_sweeps; ++n) {
    // ground_state.f90:136-140
    if ((r = gram_schmidt_dev(c))) return r;
    if ((r = subspace_diag(c, nullptr))) return r;
    if ((r = gram_schmidt_dev(c))) return r;
    if ((r = cg_call(c, iter_cg_max, pc))) return r;
    if ((r = gram_schmidt_dev(c))) return r;
    if ((r = cg_results(c, var.data(), nullptr))) return r;
    if ((r = salmon_gpu_sp_energy(c, kAc, e1.data()))) return r;
    double change = 0.0, sum = 0.0;
    for (int o = 0; o < c->NKB; ++o) {
      change = std::max(change, std::fabs(e1[o] - e0[o]));
      sum += e1[o];
    }
    // the learned direction leads for at most pc_max_sweeps sweeps, and only
    // while each sweep lowers the band-energy sum by at least pc_stall of the
    // previous sweep's drop; then the kinetic preconditioner takes over for
    // the rest of the solve
    if (pc == SALMON_PC_LEARNED) {
      const double drop = prev_sum - sum;
      if (n > 2 && !(drop >= c->pc_stall * prev_drop)) pc = SALMON_PC_KINETIC;
      if (c->pc_max_sweeps > 0 && n >= c->pc_max_sweeps) pc = SALMON_PC_KINETIC;
      prev_drop = drop;
    }
    prev_sum = sum;
    if (history) std::copy(e1.begin(), e1.end(), history + (size_t)(n - 1) * c->NKB);
    e0.swap(e1);
    if (n > 1 && change < tol) break;
  }
  if (sweeps) *sweeps = std::min(n, max_sweeps);
  if (esp_var) std::copy(var.begin(), var.end(), esp_var);
  if (esp) std::copy(e0.begin(), e0.end(), esp);
  return SALMON_OK;
}

int salmon_gpu_pcnet_create(unsigned seed, salmon_pcnet** out) {
  if (!out) return SALMON_ERR_ARG;
  *out = nullptr;
  salmon_pcnet* n = new salmon_pcnet();
  int r = 0;
  std::vector<void*> own;
  r |= pc_alloc(own, &n->w, NW);
  r |= pc_alloc(own, &n->m, NW);
  r |= pc_alloc(own, &n->v, NW);
  r |= pc_alloc(own, &n->grad, NW);
  r |= pc_alloc(own, &n->bt, (size_t)kBt1 + kBt2 + kBt3);
  n->owned = own;
  if (r) {
    salmon_gpu_pcnet_finalize(n);
    return SALMON_ERR_NOMEM;
  }
  // He-normal initial weights (fan-in 27*cin; the output layer at half the variance)
  std::mt19937_64 rng(seed);
  std::normal_distribution<double> nd(0.0, 1.0);
  std::vector<float> w(NW);
  for (int i = 0; i < NW; ++i) {
    const double sd = i < NW1 ? std::sqrt(2.0 / (27 * C0)) : (i < NW1 + NW2 ? std::sqrt(2.0 / (27 * C1)) : std::sqrt(1.0 / (27 * C2)));
    w[i] = (float)(nd(rng) * sd);
  }
  if (hipMemcpy(n->w, w.data(), sizeof(float) * NW, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(n->m, 0, sizeof(float) * NW) != hipSuccess || hipMemset(n->v, 0, sizeof(float) * NW) != hipSuccess) {
    salmon_gpu_pcnet_finalize(n);
    return SALMON_ERR_HIP;
  }
  *out = n;
  return SALMON_OK;
}

int salmon_gpu_pcnet_set_weights(salmon_pcnet* n, const float* w) {
  if (!n || !w) return SALMON_ERR_ARG;
  if (hipMemcpy(n->w, w, sizeof(float) * NW, hipMemcpyHostToDevice) != hipSuccess) return SALMON_ERR_HIP;
  if (hipMemset(n->m, 0, sizeof(float) * NW) != hipSuccess || hipMemset(n->v, 0, sizeof(float) * NW) != hipSuccess)
    return SALMON_ERR_HIP;
  n->step = 0;
  n->packed = false;
  return SALMON_OK;
}

int salmon_gpu_pcnet_get_weights(salmon_pcnet* n, float* w) {
  if (!n || !w) return SALMON_ERR_ARG;
  return hipMemcpy(w, n->w, sizeof(float) * NW, hipMemcpyDeviceToHost) == hipSuccess ? SALMON_OK : SALMON_ERR_HIP;
}

int salmon_gpu_pcnet_finalize(salmon_pcnet* n) {
  if (!n) return SALMON_ERR_ARG;
  (void)hipDeviceSynchronize();
  pcnet_free_buffers(n);
  for (void* b : n->owned) (void)hipFree(b);
  delete n;
  return SALMON_OK;
}

int salmon_gpu_pcnet_apply(salmon_pcnet* n, salmon_ctx* c, const double* g, double* d) {
  touch(c);
  if (!n || !g || !d) return c ? fail(c, SALMON_ERR_ARG, "NULL argument") : SALMON_ERR_ARG;
  int r = check_ready(c, false);
  if (r) return r;
  if (c->NK != 1) return fail(c, SALMON_ERR_UNSUPPORTED, "pcnet_apply on a context with one k point");
  if ((r = ensure_work(c)) || (r = gs_ensure(c))) return r;
  GsState* s = (GsState*)c->gs;
  const size_t NL = c->g.NL;
  HIPCHK(c, hipMemcpyAsync(s->w[W_G], g, sizeof(cplx<double>) * NL, hipMemcpyHostToDevice, c->stream));
  if ((r = pc_pack(c, n)) || (r = pc_infer(c, s, n, s->w[W_G], s->w[W_KG]))) return r;
  HIPCHK(c, hipMemcpyAsync(d, s->w[W_KG], sizeof(cplx<double>) * NL, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SALMON_OK;
}

int salmon_gpu_set_scf(salmon_ctx* c, const double* vpsl, int xc) {
  touch(c);
  if (!c) return SALMON_ERR_ARG;
  if (!vpsl) {
    c->scf = false;
    return SALMON_OK;
  }
  if (xc != SALMON_XC_NONE && xc != SALMON_XC_PZ) return fail(c, SALMON_ERR_ARG, "xc %d is neither NONE nor PZ", xc);
  int r = check_ready(c, true);
  if (r) return r;
  if ((r = dalloc(c, &c->d_vpsl, (size_t)c->g.NL))) return r;
  HIPCHK(c, hipMemcpy(c->d_vpsl, vpsl, sizeof(double) * c->g.NL, hipMemcpyHostToDevice));
  c->scf_xc = xc;
  // Vloc(0) from the resident orbitals' density; the ETRS history starts at
  // V_old1 = V_old2 = Vloc(0) (control_sc.f90:82)
  if ((r = c->precision == SALMON_FP64 ? density_dev<double>(c) : density_dev<float>(c))) return r;
  if ((r = scf_potential(c, c->d_rho, true, xc, nullptr, nullptr))) return r;
  if (!c->d_vold && (r = dalloc(c, &c->d_vold, 2 * (size_t)c->g.NL))) return r;
  HIPCHK(c, hipMemcpyAsync(c->d_vold, c->d_vloc, sizeof(double) * c->g.NL, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_vold + c->g.NL, c->d_vloc, sizeof(double) * c->g.NL, hipMemcpyDeviceToDevice,
                           c->stream));
  c->scf = true;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SALMON_OK;
}

int salmon_gpu_hartree_xc(salmon_ctx* c, const double* rho, int xc, double* vh, double* vxc) {
  touch(c);
  if (!c || !rho) return c ? fail(c, SALMON_ERR_ARG, "rho is NULL") : SALMON_ERR_ARG;
  if (xc != SALMON_XC_NONE && xc != SALMON_XC_PZ) return fail(c, SALMON_ERR_ARG, "xc %d is neither NONE nor PZ", xc);
  const size_t NL = c->g.NL;
  double *d_in = nullptr, *d_vh = nullptr, *d_vx = nullptr;
  int r;
  if ((r = dalloc(c, &d_in, 3 * NL))) return r;
  d_vh = d_in + NL;
  d_vx = d_in + 2 * NL;
  HIPCHK(c, hipMemcpy(d_in, rho, sizeof(double) * NL, hipMemcpyHostToDevice));
  r = scf_potential(c, d_in, false, xc, d_vh, d_vx);
  if (r == SALMON_OK) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (vh) HIPCHK(c, hipMemcpy(vh, d_vh, sizeof(double) * NL, hipMemcpyDeviceToHost));
    if (vxc) HIPCHK(c, hipMemcpy(vxc, d_vx, sizeof(double) * NL, hipMemcpyDeviceToHost));
  }
  (void)hipFree(d_in);
  return r;
}

int salmon_gpu_hartree_xc_noalloc(salmon_ctx* c, const double* rho_dev, int xc, double* vh_dev, double* vxc_dev) {
  touch(c);
  if (!c || !rho_dev) return c ? fail(c, SALMON_ERR_ARG, "rho_dev is NULL") : SALMON_ERR_ARG;
  if (xc != SALMON_XC_NONE && xc != SALMON_XC_PZ) return fail(c, SALMON_ERR_ARG, "xc %d is neither NONE nor PZ", xc);
  const Grid& g = c->g;
  if ((size_t)g.NL > (size_t)SALMON_NOALLOC_MAX_NL)
    return fail(c, SALMON_ERR_UNSUPPORTED, "grid of %d points exceeds the allocation-free limit of %d", g.NL,
                (int)SALMON_NOALLOC_MAX_NL);
  int r;
  if ((r = scf_check_grid(c))) return r;
  hipError_t e = hipSuccess;
  double* base = noalloc_ptr(c->device, &e);
  if (!base)
    return fail(c, SALMON_ERR_HIP, "device scratch not set up on device %d: %s", c->device, hipGetErrorString(e));
  double* tw = base;
  cplx<double>* A = (cplx<double>*)(base + kNoAllocTw);
  cplx<double>* B = A + g.NL;
  hipLaunchKernelGGL(k_noalloc_twiddles, dim3((3 * salmon::scf::kMaxN + 255) / 256), dim3(256), 0, c->stream, tw,
                     g.NLx, g.NLy, g.NLz);
  LAUNCHCHK(c);
  return scf_potential_on(c, rho_dev, false, xc, vh_dev, vxc_dev, tw, A, B);
}

int salmon_gpu_set_propagator(salmon_ctx* c, int propagator) {
  touch(c);
  if (!c) return SALMON_ERR_ARG;
  if (propagator != SALMON_PROP_MIDPOINT && propagator != SALMON_PROP_ETRS)
    return fail(c, SALMON_ERR_ARG, "propagator %d is neither SALMON_PROP_MIDPOINT nor SALMON_PROP_ETRS", propagator);
  c->propagator = propagator;
  return SALMON_OK;
}

int salmon_gpu_set_graphs(salmon_ctx* c, int enable) {
  if (!c) return SALMON_ERR_ARG;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  c->graphs = enable != 0;
  return SALMON_OK;
}

int salmon_gpu_set_pcnet(salmon_ctx* c, salmon_pcnet* n) {
  touch(c);
  if (!c) return SALMON_ERR_ARG;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  ++c->gen;  // a captured learned CG reads the previous network's weights
  c->pcnet = n;
  return SALMON_OK;
}

int salmon_gpu_set_pc_handover(salmon_ctx* c, int max_learned_sweeps, double stall_ratio) {
  if (!c) return SALMON_ERR_ARG;
  if (max_learned_sweeps < 0 || !(stall_ratio >= 0.0))
    return fail(c, SALMON_ERR_ARG, "max_learned_sweeps < 0 or stall_ratio not >= 0");
  c->pc_max_sweeps = max_learned_sweeps;
  c->pc_stall = stall_ratio;
  return SALMON_OK;
}

int salmon_gpu_pc_record(salmon_ctx* c, int max_samples, int stride) {
  touch(c);
  int r = check_ready(c, true);
  if (r) return r;
  if (c->NK != 1) return fail(c, SALMON_ERR_UNSUPPORTED, "training samples are recorded for one k point");
  if (max_samples < 0 || stride < 1) return fail(c, SALMON_ERR_ARG, "max_samples < 0 or stride < 1");
  if ((r = ensure_work(c)) || (r = gs_ensure(c))) return r;
  GsState* s = (GsState*)c->gs;
  for (void* p : s->smp_bufs) (void)hipFree(p);
  s->smp_bufs.clear();
  s->smp_cap = s->smp_n = 0;
  s->smp_stride = stride;
  s->smp_seen = 0;
  s->recording = max_samples > 0;
  if (!s->recording) return SALMON_OK;
  const size_t NL = c->g.NL;
  int e = 0;
  e |= pc_alloc(s->smp_bufs, &s->smp_g, (size_t)max_samples * NL);
  e |= pc_alloc(s->smp_bufs, &s->smp_x, (size_t)max_samples * NL);
  e |= pc_alloc(s->smp_bufs, &s->smp_theta, (size_t)max_samples);
  e |= pc_alloc(s->smp_bufs, &s->smp_ib, (size_t)max_samples);
  e |= pc_alloc(s->smp_bufs, &s->smp_valid, (size_t)max_samples);
  if (e) {
    s->recording = false;
    return fail(c, SALMON_ERR_NOMEM, "sample store of %d residuals", max_samples);
  }
  s->smp_cap = max_samples;
  return SALMON_OK;
}

int salmon_gpu_pc_samples(salmon_ctx* c, int* stored, int* valid) {
  if (!c) return SALMON_ERR_ARG;
  GsState* s = (GsState*)c->gs;
  int r;
  if (!s) {
    if (stored) *stored = 0;
    if (valid) *valid = 0;
    return SALMON_OK;
  }
  if ((r = pc_sync_samples(c, s))) return r;
  if (stored) *stored = s->smp_n;
  if (valid) {
    int v = 0;
    for (int x : s->h_valid) v += x;
    *valid = v;
  }
  return SALMON_OK;
}

int salmon_gpu_pc_get_sample(salmon_ctx* c, int i, double* g, double* x, int* ib, double* theta, int* valid) {
  if (!c || !c->gs) return SALMON_ERR_ARG;
  GsState* s = (GsState*)c->gs;
  if (i < 0 || i >= s->smp_n) return fail(c, SALMON_ERR_ARG, "sample %d of %d", i, s->smp_n);
  int r = pc_sync_samples(c, s);
  if (r) return r;
  const size_t NL = c->g.NL;
  if (g) HIPCHK(c, hipMemcpy(g, s->smp_g + (size_t)i * NL, sizeof(cplx<double>) * NL, hipMemcpyDeviceToHost));
  if (x) HIPCHK(c, hipMemcpy(x, s->smp_x + (size_t)i * NL, sizeof(cplx<double>) * NL, hipMemcpyDeviceToHost));
  if (ib) *ib = s->h_ib[i];
  if (theta) *theta = s->h_theta[i];
  if (valid) *valid = s->h_valid[i];
  return SALMON_OK;
}

int salmon_gpu_pc_set_targets(salmon_ctx* c, const double* kAc) {
  touch(c);
  int r = gs_check(c, kAc);
  if (r) return r;
  if (c->NK != 1) return fail(c, SALMON_ERR_UNSUPPORTED, "training targets are set for one k point");
  GsState* s = (GsState*)c->gs;
  const size_t n = (size_t)c->NB * c->g.NL;
  if (!s->Zc && (r = gs_alloc(c, s, &s->Zc, n))) return r;
  HIPCHK(c, hipMemcpyAsync(s->Zc, c->d_zu, sizeof(cplx<double>) * n, hipMemcpyDeviceToDevice, c->stream));
  s->lam.resize(c->NB);
  return salmon_gpu_sp_energy(c, kAc, s->lam.data());
}

static int pc_train_check(salmon_pcnet* n, salmon_ctx* c) {
  if (!n || !c) return SALMON_ERR_ARG;
  GsState* s = (GsState*)c->gs;
  if (!s || s->smp_n == 0) return fail(c, SALMON_ERR_STATE, "no recorded samples (salmon_gpu_pc_record, then CG)");
  if (!s->Zc || (int)s->lam.size() != c->NB) return fail(c, SALMON_ERR_STATE, "salmon_gpu_pc_set_targets not called");
  return pc_sync_samples(c, s);
}

int salmon_gpu_pcnet_gradient(salmon_pcnet* n, salmon_ctx* c, int nsamp, const int* idx, double* loss, double* grad) {
  touch(c);
  int r = pc_train_check(n, c);
  if (r) return r;
  if (nsamp < 1 || !idx) return fail(c, SALMON_ERR_ARG, "no samples given");
  std::vector<int> v(idx, idx + nsamp);
  return pc_step(c, n, v, false, 0.0f, loss, grad);
}

int salmon_gpu_pcnet_train(salmon_pcnet* n, salmon_ctx* c, int nsteps, int batch, double lr, unsigned seed,
                           double* mean_loss) {
  touch(c);
  int r = pc_train_check(n, c);
  if (r) return r;
  if (nsteps < 0 || batch < 1) return fail(c, SALMON_ERR_ARG, "nsteps < 0 or batch < 1");
  GsState* s = (GsState*)c->gs;
  std::vector<int> pool;
  for (int i = 0; i < s->smp_n; ++i)
    if (s->h_valid[i]) pool.push_back(i);
  if (pool.empty()) return fail(c, SALMON_ERR_STATE, "no valid samples");
  std::mt19937 rng(seed);
  std::uniform_int_distribution<int> pick(0, (int)pool.size() - 1);
  double tot = 0.0;
  for (int k = 0; k < nsteps; ++k) {
    std::vector<int> idx(batch);
    for (int& i : idx) i = pool[pick(rng)];
    double l = 0.0;
    if ((r = pc_step(c, n, idx, true, (float)lr, &l, nullptr))) return r;
    tot += l;
  }
  if (mean_loss) *mean_loss = nsteps > 0 ? tot / nsteps : 0.0;
  return SALMON_OK;
}

int salmon_gpu_rt_run(salmon_ctx* c, const double* ac, int nsteps, double dt, int reortho_every, double* jt, double* et,
                      double* norm, double* rho) {
  touch(c);
  int r = check_ready(c, true);
  if (r) return r;
  if (!ac || nsteps < 0) return fail(c, SALMON_ERR_ARG, "ac is NULL or nsteps < 0");
  Region reg(c, SALMON_LOG_DYNAMICS);
  if (nsteps == 0) return SALMON_OK;
  if (rt_path(c) == SALMON_PATH_SWEEP) {
    // set-up failures (grid, LDS or memory) before any step: SALMON_PATH_AUTO
    // takes the streamed kernels instead
    r = c->precision == SALMON_FP64 ? solo_prepare<double>(c) : solo_prepare<float>(c);
    if (r == SALMON_OK) {
      r = c->precision == SALMON_FP64 ? rt_run_sweep<double>(c, ac, nsteps, dt, reortho_every, jt, et, norm)
                                      : rt_run_sweep<float>(c, ac, nsteps, dt, reortho_every, jt, et, norm);
      if (r == SALMON_OK && rho) {
        const hipError_t e = hipMemcpy(rho, c->d_rho, sizeof(double) * c->g.NL, hipMemcpyDeviceToHost);
        if (e != hipSuccess) r = fail(c, SALMON_ERR_HIP, "rt_run density: %s", hipGetErrorString(e));
      }
      (void)hipStreamSynchronize(c->stream);
      return r;
    }
    if (c->path != SALMON_PATH_AUTO) return r;
    r = SALMON_OK;
  }
  // the per-step results stay allocated between runs, sized for at least
  // 4096 steps (256 KiB), so runs of the usual length allocate nothing
  if ((size_t)nsteps > c->res_cap) {
    const size_t cap = std::max<size_t>((size_t)nsteps, std::max<size_t>(4096, 2 * c->res_cap));
    (void)hipStreamSynchronize(c->stream);
    if ((r = dalloc(c, &c->d_res, 8 * cap))) return r;
    c->res_cap = cap;
  }
  double* d_res = c->d_res;
  if (c->scf) {
    r = c->precision == SALMON_FP64 ? rt_run_scf<double>(c, ac, nsteps, dt, reortho_every, d_res)
                                    : rt_run_scf<float>(c, ac, nsteps, dt, reortho_every, d_res);
  } else if (fused_ready(c)) {
    r = c->precision == SALMON_FP64 ? rt_run_fused<double>(c, ac, nsteps, dt, reortho_every, d_res)
                                    : rt_run_fused<float>(c, ac, nsteps, dt, reortho_every, d_res);
  }
  for (int n = 0; n < nsteps && r == SALMON_OK && !fused_ready(c) && !c->scf; ++n) {  // streamed
    const double* a0 = ac + 3 * n;
    const double* a1 = ac + 3 * (n + 1);
    r = c->precision == SALMON_FP64 ? rt_step<double>(c, a0, a1, dt, d_res + 8 * (size_t)n)
                                    : rt_step<float>(c, a0, a1, dt, d_res + 8 * (size_t)n);
    if (r == SALMON_OK && reortho_every > 0 && (n + 1) % reortho_every == 0) r = gram_schmidt_dev(c);
  }
  if (r == SALMON_OK) {
    std::vector<double> res(8 * (size_t)nsteps);
    hipError_t e = hipMemcpyAsync(res.data(), d_res, sizeof(double) * res.size(), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
      r = fail(c, SALMON_ERR_HIP, "rt_run results: %s", hipGetErrorString(e));
    } else {
      for (int n = 0; n < nsteps; ++n) {
        const double* q = &res[8 * (size_t)n];
        if (jt) { jt[3 * n] = q[0]; jt[3 * n + 1] = q[1]; jt[3 * n + 2] = q[2]; }
        if (et) et[n] = q[3] + q[4] + q[5];
        if (norm) norm[n] = q[6];
      }
      if (rho) {
        e = hipMemcpy(rho, c->d_rho, sizeof(double) * c->g.NL, hipMemcpyDeviceToHost);
        if (e != hipSuccess) r = fail(c, SALMON_ERR_HIP, "rt_run density: %s", hipGetErrorString(e));
      }
    }
  }
  (void)hipStreamSynchronize(c->stream);
  return r;
}

int salmon_gpu_batch_create(int precision, salmon_batch** out) {
  if (!out) return SALMON_ERR_ARG;
  *out = nullptr;
  if (precision != SALMON_FP64 && precision != SALMON_MIXED) return SALMON_ERR_ARG;
  salmon_batch* b = new salmon_batch();
  b->precision = precision;
  if (hipStreamCreate(&b->stream) != hipSuccess) {
    delete b;
    return SALMON_ERR_HIP;
  }
  *out = b;
  return SALMON_OK;
}

int salmon_gpu_batch_add(salmon_batch* b, salmon_ctx* c) {
  touch(c);
  if (!b) return SALMON_ERR_ARG;
  if (!c) return bfail(b, SALMON_ERR_ARG, "context is NULL");
  int r = check_ready(c, true);
  if (r) return bfail(b, r, "context %zu not set up: %s", b->ctx.size(), c->err.c_str());
  if (c->precision != b->precision) return bfail(b, SALMON_ERR_ARG, "context precision differs from the batch's");
  if (c->g.NLz > 1 << 20) return bfail(b, SALMON_ERR_ARG, "grid too large");
  (void)hipStreamSynchronize(c->stream);
  batch_release(b);
  b->tuned = false;
  b->ctx.push_back(c);
  return SALMON_OK;
}

int salmon_gpu_batch_set_kernel(salmon_batch* b, int kernel) {
  if (!b) return SALMON_ERR_ARG;
  if (kernel != SALMON_BATCH_AUTO && kernel != SALMON_BATCH_PHASES && kernel != SALMON_BATCH_SWEEP)
    return bfail(b, SALMON_ERR_ARG, "unknown batch kernel %d", kernel);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  batch_release(b);
  b->tuned = false;
  b->kernel_req = kernel;
  return SALMON_OK;
}

int salmon_gpu_batch_kernel(const salmon_batch* b) {
  if (!b || !b->ready || !b->tuned) return -1;
  return b->sweep ? SALMON_BATCH_SWEEP : SALMON_BATCH_PHASES;
}

int salmon_gpu_batch_tune(salmon_batch* b) {
  if (b)
    for (salmon_ctx* c : b->ctx) touch(c);
  if (!b) return SALMON_ERR_ARG;
  if (b->ctx.empty()) return bfail(b, SALMON_ERR_STATE, "empty batch");
  return b->precision == SALMON_FP64 ? batch_tune<double>(b) : batch_tune<float>(b);
}

int salmon_gpu_batch_zblock(salmon_batch* b, int i) {
  if (!b || i < 0 || i >= (int)b->zb.size() || !b->tuned) return -1;
  return b->zb[i];
}

int salmon_gpu_batch_rt_run(salmon_batch* b, const double* ac, int nsteps, double dt, double* jt, double* et,
                            double* norm) {
  if (b)
    for (salmon_ctx* c : b->ctx) touch(c);
  if (!b) return SALMON_ERR_ARG;
  if (b->ctx.empty()) return bfail(b, SALMON_ERR_STATE, "empty batch");
  if (!ac || nsteps < 0) return bfail(b, SALMON_ERR_ARG, "ac is NULL or nsteps < 0");
  if (nsteps == 0) return SALMON_OK;
  return b->precision == SALMON_FP64 ? batch_run<double>(b, ac, nsteps, dt, jt, et, norm, false)
                                     : batch_run<float>(b, ac, nsteps, dt, jt, et, norm, false);
}

int salmon_gpu_batch_rt_run_ms(salmon_batch* b, const double* ac, int nsteps, double dt, double* jt, double* et,
                               double* norm) {
  if (b)
    for (salmon_ctx* c : b->ctx) touch(c);
  if (!b) return SALMON_ERR_ARG;
  if (b->ctx.empty()) return bfail(b, SALMON_ERR_STATE, "empty batch");
  if (!ac || nsteps < 0) return bfail(b, SALMON_ERR_ARG, "ac is NULL or nsteps < 0");
  if (nsteps == 0) return SALMON_OK;
  return b->precision == SALMON_FP64 ? batch_run<double>(b, ac, nsteps, dt, jt, et, norm, true)
                                     : batch_run<float>(b, ac, nsteps, dt, jt, et, norm, true);
}

const char* salmon_gpu_batch_last_error(const salmon_batch* b) { return b ? b->err.c_str() : "no batch"; }

int salmon_gpu_batch_finalize(salmon_batch* b) {
  if (!b) return SALMON_ERR_ARG;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  batch_release(b);
  if (b->d_step) (void)hipFree(b->d_step);
  if (b->d_next) (void)hipFree(b->d_next);
  if (b->d_ac) (void)hipFree(b->d_ac);
  if (b->d_res) (void)hipFree(b->d_res);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return SALMON_OK;
}

int salmon_gpu_synchronize(salmon_ctx* c) {
  if (!c) return SALMON_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SALMON_OK;
}

int salmon_gpu_timer_enable(salmon_ctx* c, int on) {
  if (!c) return SALMON_ERR_ARG;
  if (on && !c->tm) c->tm = new RegionTimers();
  if (c->tm) c->tm->on = on != 0;
  return SALMON_OK;
}

int salmon_gpu_timer_reset(salmon_ctx* c) {
  if (!c) return SALMON_ERR_ARG;
  if (!c->tm) return SALMON_OK;
  timers_resolve(c);
  for (int i = 0; i < SALMON_LOG_MAX; ++i) {
    c->tm->sec[i] = 0.0;
    c->tm->calls[i] = 0;
  }
  return SALMON_OK;
}

int salmon_gpu_timer_get(salmon_ctx* c, int id, double* seconds, int* calls) {
  if (!c) return SALMON_ERR_ARG;
  if (id < 0 || id >= SALMON_LOG_MAX) return fail(c, SALMON_ERR_ARG, "timer region %d outside 0..%d", id, SALMON_LOG_MAX - 1);
  if (!seconds && !calls) return fail(c, SALMON_ERR_ARG, "seconds and calls are NULL");
  timers_resolve(c);
  if (seconds) *seconds = c->tm ? c->tm->sec[id] : 0.0;
  if (calls) *calls = c->tm ? c->tm->calls[id] : 0;
  return SALMON_OK;
}

int salmon_gpu_set_path(salmon_ctx* c, int path) {
  touch(c);
  if (!c) return SALMON_ERR_ARG;
  if (path != SALMON_PATH_AUTO && path != SALMON_PATH_STREAMED && path != SALMON_PATH_SWEEP)
    return fail(c, SALMON_ERR_ARG, "unknown path %d", path);
  c->path = path;
  if (path == SALMON_PATH_STREAMED && c->solo) {
    // the persistent sweep's work buffers (four times the orbitals) go
    (void)hipStreamSynchronize(c->stream);
    batch_release(c->solo);
    c->solo->tuned = false;
  }
  return SALMON_OK;
}

int salmon_gpu_path(salmon_ctx* c) {
  if (!c) return -1;
  return rt_path(c);
}

void* salmon_gpu_stream(salmon_ctx* c) { return c ? (void*)c->stream : nullptr; }

int salmon_gpu_finalize(salmon_ctx* c) {
  if (!c) return SALMON_ERR_ARG;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->solo) {
    c->solo->stream = nullptr;  // the context's own stream, destroyed below
    (void)salmon_gpu_batch_finalize(c->solo);
    c->solo = nullptr;
  }
  fused_free(c);
  gs_free(c);
  dfree((void**)&c->d_vloc); dfree((void**)&c->d_vloc_f); dfree((void**)&c->d_rho); dfree((void**)&c->d_rho_part);
  dfree((void**)&c->d_rho_part2);
  dfree((void**)&c->d_kac0); dfree((void**)&c->d_kac); dfree((void**)&c->d_occ); dfree((void**)&c->d_rows);
  dfree((void**)&c->d_res);
  dfree((void**)&c->d_sum); dfree((void**)&c->d_kone); dfree((void**)&c->d_oneocc); dfree(&c->d_zu);
  dfree((void**)&c->d_f64); dfree(&c->d_uvpsi); dfree(&c->d_ph);
  dfree((void**)&c->d_vpsl); dfree((void**)&c->d_vold); dfree((void**)&c->d_scf_tw); dfree(&c->d_scf_buf[0]); dfree(&c->d_scf_buf[1]);
  for (int i = 0; i < 4; ++i) dfree(&c->d_work[i]);
  for (int i = 0; i < 3; ++i) dfree(&c->d_stage[i]);
  for (void* p : c->kb_bufs) (void)hipFree(p);
  if (c->tm) {
    timers_resolve(c);
    for (hipEvent_t e : c->tm->pool) (void)hipEventDestroy(e);
    delete c->tm;
  }
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return SALMON_OK;
}

}  // extern "C"
