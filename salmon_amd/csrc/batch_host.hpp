// Host side of the batched step: a batch references set-up contexts (their
// orbitals, potentials, tables and k points stay where the contexts put them),
// owns per-system work buffers, tunes the stencil's z blocking per grid shape
// and runs the time loop as replays of one captured step.
#pragma once
#include <array>

#include <algorithm>

#include "batch_sweep.hpp"

using namespace salmon::batch;

struct salmon_batch {
  int precision = SALMON_FP64;
  hipStream_t stream = nullptr;
  std::string err;
  std::vector<salmon_ctx*> ctx;
  std::vector<void*> bufs;        // work buffers (rebuilt by prepare)
  void* d_sys = nullptr;          // Sys<T>[n]
  std::vector<int> zb;            // chosen z blocking per system
  std::vector<long> first[kPhases + 1];
  unsigned short* d_owner[kPhases] = {};  // block -> system tables (in bufs)
  long zcls[4][2] = {};                   // stencil blocks of the z-blocking classes 1, 2, 4, 8: (first, count)
  bool ready = false, tuned = false;
  int* d_step = nullptr;
  double* d_ac = nullptr;
  double* d_res = nullptr;
  size_t cap_steps = 0;
  size_t cap_ac = 0;              // doubles in d_ac
  long max_ph = 0;                // largest NK * NPI over the systems
  // persistent sweep (one launch per step) or one launch per phase
  int kernel_req = SALMON_BATCH_AUTO;
  bool sweep = false;
  int2* d_items = nullptr;        // (system, first orbital), costliest first (in bufs)
  int nitems = 0;
  int* d_next = nullptr;          // work counter of the sweep launch
  unsigned sweep_grid = 0;
  size_t sweep_lds = 0;
  // one Taylor step at caller-given k points (batch_taylor): descriptors whose
  // kac0 is the given table, a zero A(t) table and a zero step counter (in bufs)
  void* d_sys_k = nullptr;
  const double* const* sys_k_base = nullptr;
  double* d_zero_ac = nullptr;
  int* d_zero_step = nullptr;
  hipGraphExec_t exec = nullptr;
  double graph_dt = 0.0;
  long graph_stride = 0;          // the A(t) table stride the graph was captured with
};

static int bfail(salmon_batch* b, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  b->err = buf;
  return code;
}

#define BHIP(b, x)                                                                                   \
  do {                                                                                               \
    hipError_t e_ = (x);                                                                             \
    if (e_ != hipSuccess) return bfail((b), SALMON_ERR_HIP, "%s: %s", #x, hipGetErrorString(e_));    \
  } while (0)

static void batch_release(salmon_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  b->exec = nullptr;
  for (void* p : b->bufs) (void)hipFree(p);
  b->bufs.clear();
  b->d_sys = nullptr;
  b->d_sys_k = nullptr;
  b->sys_k_base = nullptr;
  b->d_zero_ac = nullptr;
  b->d_zero_step = nullptr;
  for (int ph = 0; ph < kPhases; ++ph) b->d_owner[ph] = nullptr;
  b->d_items = nullptr;
  b->nitems = 0;
  b->sweep = false;
  b->ready = false;
}

// Column length (points along z per thread) and team size (waves) of the
// persistent sweep for one grid: the candidate that leaves the fewest team
// threads without a column, the longer column on a tie; {0, 0} if no
// candidate fits a workgroup or the grid is too small for single periodic wraps.
template <typename T>
static std::array<int, 2> sweep_shape(const Grid& g) {
  constexpr int zbmax = sizeof(T) == 4 ? 4 : 2;
  std::array<int, 2> best{0, 0};
  if (g.NLx < 8 || g.NLy < 8 || g.NLz < 8) return best;
  const int P = g.NLy * g.NLz;
  double bu = -1.0;
  for (int zb = 1; zb <= zbmax; ++zb) {
    if (g.NLz % zb) continue;
    const int ncol = P / zb;
    int tw = 1;
    while (tw < sweep_waves<T>() && tw * kWave < ncol) tw *= 2;
    if (tw * kWave < ncol) continue;
    const double u = (double)ncol / (tw * kWave);
    if (u > bu + 1e-12 || (u > bu - 1e-12 && zb > best[0])) {
      bu = u;
      best = {zb, tw};
    }
  }
  return best;
}

// The persistent sweep's work list, LDS size and grid; b->sweep stays false
// (one launch per phase) if the phase kernel is asked for or, under
// SALMON_BATCH_AUTO, a system does not fit the sweep.
template <typename T>
static int sweep_plan(salmon_batch* b, std::vector<Sys<T>>& S) {
  b->sweep = false;
  if (b->kernel_req == SALMON_BATCH_PHASES) return SALMON_OK;
  const bool must = b->kernel_req == SALMON_BATCH_SWEEP;
  std::vector<std::array<long, 3>> items;  // (cost, system, first orbital)
  size_t lds = 0;
  for (size_t i = 0; i < S.size(); ++i) {
    S[i].kb_order = S[i].kb_cstart = nullptr;
    S[i].ncolor = 0;
  }
  for (size_t i = 0; i < S.size(); ++i) {
    const auto sh = sweep_shape<T>(S[i].g);
    if (!sh[0]) {
      if (must) return bfail(b, SALMON_ERR_ARG, "system %zu (%dx%dx%d) does not fit the persistent sweep", i,
                             S[i].g.NLx, S[i].g.NLy, S[i].g.NLz);
      return SALMON_OK;
    }
    S[i].sw_zb = sh[0];
    S[i].sw_tw = sh[1];
    const int teams = sweep_waves<T>() / sh[1];
    const int zb = sh[0], cs = zb % 2 == 0 ? zb + 1 : zb;  // col_stride<zb>()
    const size_t PS = ((size_t)S[i].g.NLy * S[i].g.NLz / zb + 1) * cs;  // plane_slots<zb>(P)
    lds = std::max(lds, (size_t)kSweepHead + (size_t)teams * (2 * PS + S[i].kb.Nlma) * sizeof(cplx<T>));
    for (int o = 0; o < S[i].NKB; o += teams)
      items.push_back({(long)S[i].g.NL * std::min(teams, S[i].NKB - o), (long)i, (long)o});
  }
  // colours of the atoms of every system: greedy, an atom takes the first
  // colour none of whose atoms shares a grid point with it
  for (size_t i = 0; i < S.size(); ++i) {
    const KBTables& kb = S[i].kb;
    if (kb.NI <= 0 || kb.Nlma <= 0) continue;
    std::vector<int> ast(kb.NI + 1), pidx(std::max(1, kb.NPI));
    BHIP(b, hipMemcpy(ast.data(), kb.atom_start, sizeof(int) * (kb.NI + 1), hipMemcpyDeviceToHost));
    if (kb.NPI > 0) BHIP(b, hipMemcpy(pidx.data(), kb.pidx, sizeof(int) * kb.NPI, hipMemcpyDeviceToHost));
    std::vector<std::vector<char>> used;
    std::vector<int> colour(kb.NI);
    for (int a = 0; a < kb.NI; ++a) {
      int c = 0;
      for (;; ++c) {
        if (c == (int)used.size()) used.emplace_back(S[i].g.NL, 0);
        bool clash = false;
        for (int jp = ast[a]; jp < ast[a + 1] && !clash; ++jp) clash = used[c][pidx[jp]] != 0;
        if (!clash) break;
      }
      for (int jp = ast[a]; jp < ast[a + 1]; ++jp) used[c][pidx[jp]] = 1;
      colour[a] = c;
    }
    const int nc = (int)used.size();
    std::vector<int> tab;  // kb_order (NI), then kb_cstart (nc + 1)
    std::vector<int> cst(nc + 1, 0);
    for (int c = 0; c < nc; ++c) {
      cst[c] = (int)tab.size();
      for (int a = 0; a < kb.NI; ++a)
        if (colour[a] == c) tab.push_back(a);
    }
    cst[nc] = (int)tab.size();
    tab.insert(tab.end(), cst.begin(), cst.end());
    int* d = nullptr;
    if (hipMalloc((void**)&d, sizeof(int) * tab.size()) != hipSuccess)
      return bfail(b, SALMON_ERR_NOMEM, "batch atom colours");
    b->bufs.push_back(d);
    BHIP(b, hipMemcpy(d, tab.data(), sizeof(int) * tab.size(), hipMemcpyHostToDevice));
    S[i].kb_order = d;
    S[i].kb_cstart = d + kb.NI;
    S[i].ncolor = nc;
  }
  if (lds > 65536 || items.empty()) {
    if (must) return bfail(b, SALMON_ERR_ARG, "the persistent sweep needs %zu bytes of LDS", lds);
    return SALMON_OK;
  }
  std::stable_sort(items.begin(), items.end(), [](const std::array<long, 3>& a, const std::array<long, 3>& c) {
    return a[0] > c[0];
  });
  std::vector<int2> it(items.size());
  for (size_t k = 0; k < items.size(); ++k) it[k] = int2{(int)items[k][1], (int)items[k][2]};
  if (hipMalloc((void**)&b->d_items, sizeof(int2) * it.size()) != hipSuccess)
    return bfail(b, SALMON_ERR_NOMEM, "batch work list");
  b->bufs.push_back(b->d_items);
  BHIP(b, hipMemcpy(b->d_items, it.data(), sizeof(int2) * it.size(), hipMemcpyHostToDevice));
  if (!b->d_next) BHIP(b, hipMalloc((void**)&b->d_next, sizeof(int)));
  BHIP(b, hipMemset(b->d_next, 0, sizeof(int)));
  int dev = 0, ncu = 0, per = 0;
  BHIP(b, hipGetDevice(&dev));
  BHIP(b, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  const void* kern = nullptr;
  if constexpr (sizeof(T) == 4)
    kern = (const void*)k_b_sweep<T, 4>;
  else
    kern = (const void*)k_b_sweep<T, 2>;
  BHIP(b, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, sweep_waves<T>() * kWave, lds));
  if (per < 1) {
    if (must) return bfail(b, SALMON_ERR_ARG, "the persistent sweep does not fit a compute unit");
    return SALMON_OK;
  }
  b->nitems = (int)it.size();
  b->sweep_lds = lds;
  b->sweep_grid = (unsigned)std::min<long>((long)ncu * per, (long)it.size());
  b->sweep = true;
  return SALMON_OK;
}

template <typename T>
static int batch_build(salmon_batch* b) {
  const int n = (int)b->ctx.size();
  std::vector<Sys<T>> S(n);
  b->max_ph = 0;
  long acc[kPhases] = {0, 0, 0, 0, 0};
  for (int i = 0; i < n; ++i) {
    salmon_ctx* c = b->ctx[i];
    Sys<T>& y = S[i];
    y.g = c->g;
    y.kb = c->kb;
    y.zu = (cplx<T>*)c->d_zu;
    y.vloc = vloc_ptr<T>(c);
    y.vloc64 = c->d_vloc;
    y.kac0 = c->d_kac0;
    y.occ = c->d_occ;
    y.NK = c->NK;
    y.NB = c->NB;
    y.NKB = c->NKB;
    y.zb = b->zb[i];
    y.sw_zb = y.sw_tw = 0;
    const size_t nv = (size_t)c->NKB * c->g.NL;
    for (int t = 0; t < 4; ++t) {
      if (hipMalloc((void**)&y.h[t], nv * sizeof(cplx<T>)) != hipSuccess)
        return bfail(b, SALMON_ERR_NOMEM, "batch work buffers (system %d)", i);
      b->bufs.push_back(y.h[t]);
    }
    if (hipMalloc((void**)&y.uv, (size_t)c->NKB * std::max(1, c->kb.Nlma) * sizeof(cplx<T>)) != hipSuccess ||
        hipMalloc((void**)&y.ph, (size_t)c->NK * std::max(1, c->kb.NPI) * sizeof(cplx<T>)) != hipSuccess ||
        hipMalloc((void**)&y.phn, (size_t)c->NK * std::max(1, c->kb.NPI) * sizeof(cplx<T>)) != hipSuccess ||
        hipMalloc((void**)&y.kbv, (size_t)c->NKB * std::max(1, c->kb.NT) * sizeof(cplx<T>)) != hipSuccess ||
        hipMalloc((void**)&y.rows, (size_t)c->NKB * 8 * sizeof(double)) != hipSuccess)
      return bfail(b, SALMON_ERR_NOMEM, "batch buffers (system %d)", i);
    b->bufs.push_back(y.uv);
    b->bufs.push_back(y.ph);
    b->bufs.push_back(y.phn);
    b->bufs.push_back(y.kbv);
    b->bufs.push_back(y.rows);
    b->max_ph = std::max(b->max_ph, (long)c->NK * c->kb.NPI);
    y.pblk = (std::max(1, c->kb.Nlma) + 3) / 4;
    y.hblk = (c->g.NL / y.zb + 255) / 256;
    y.kblk = (c->kb.NT + 255) / 256;
    const long nb[kPhases] = {(long)c->NKB * y.pblk, 0, (long)((nv + 255) / 256), (long)c->NKB,
                              (long)c->NKB * y.kblk};
    for (int ph = 0; ph < kPhases; ++ph) {
      if (ph == 1) continue;
      y.first[ph] = acc[ph];
      acc[ph] += nb[ph];
    }
    y.first[kPhases] = 0;
  }
  // stencil blocks class by class, so that one launch per z blocking covers
  // every system of the class
  const int zbs[4] = {1, 2, 4, 8};
  for (int k = 0; k < 4; ++k) {
    b->zcls[k][0] = acc[1];
    for (int i = 0; i < n; ++i) {
      if (S[i].zb != zbs[k]) continue;
      S[i].first[1] = acc[1];
      acc[1] += (long)S[i].NKB * S[i].hblk;
    }
    b->zcls[k][1] = acc[1] - b->zcls[k][0];
  }
  for (int ph = 0; ph < kPhases; ++ph) {
    b->first[ph].assign(1, acc[ph]);
    std::vector<unsigned short> own((size_t)acc[ph]);
    for (int i = 0; i < n; ++i) {
      const long cnt = ph == 0 ? (long)S[i].NKB * S[i].pblk
                     : ph == 1 ? (long)S[i].NKB * S[i].hblk
                     : ph == 2 ? (long)(((size_t)S[i].NKB * S[i].g.NL + 255) / 256)
                     : ph == 3 ? (long)S[i].NKB : (long)S[i].NKB * S[i].kblk;
      std::fill(own.begin() + S[i].first[ph], own.begin() + S[i].first[ph] + cnt, (unsigned short)i);
    }
    if (hipMalloc((void**)&b->d_owner[ph], sizeof(unsigned short) * std::max<size_t>(1, own.size())) != hipSuccess)
      return bfail(b, SALMON_ERR_NOMEM, "batch block tables");
    b->bufs.push_back(b->d_owner[ph]);
    BHIP(b, hipMemcpy(b->d_owner[ph], own.data(), sizeof(unsigned short) * own.size(), hipMemcpyHostToDevice));
  }
  int r = sweep_plan<T>(b, S);
  if (r) return r;
  if (hipMalloc(&b->d_sys, sizeof(Sys<T>) * n) != hipSuccess) return bfail(b, SALMON_ERR_NOMEM, "batch descriptors");
  b->bufs.push_back(b->d_sys);
  BHIP(b, hipMemcpy(b->d_sys, S.data(), sizeof(Sys<T>) * n, hipMemcpyHostToDevice));
  b->ready = true;
  return SALMON_OK;
}

// One step's launches (captured by batch_run).
template <typename T>
static int batch_enqueue_step(salmon_batch* b, double dt, long ac_stride) {
  const int n = (int)b->ctx.size();
  const Sys<T>* S = (const Sys<T>*)b->d_sys;
  const Step st{b->d_ac, b->d_step, dt, {b->d_owner[0], b->d_owner[1], b->d_owner[2], b->d_owner[3], b->d_owner[4]},
                ac_stride};
  if (b->sweep) {
    // the step's launch also forms the previous step's observables (at
    // A(t_n)), whose rows the sum then takes
    if (b->max_ph > 0) {
      hipLaunchKernelGGL(k_b_phase<T>, dim3((unsigned)((b->max_ph + 255) / 256), n), dim3(256), 0, b->stream, S, st, 0);
      hipLaunchKernelGGL(k_b_phase<T>, dim3((unsigned)((b->max_ph + 255) / 256), n), dim3(256), 0, b->stream, S, st, 2);
    }
    hipLaunchKernelGGL(k_b_qreset, dim3(1), dim3(64), 0, b->stream, b->d_next);
    if constexpr (sizeof(T) == 4)
      hipLaunchKernelGGL((k_b_sweep<T, 4>), dim3(b->sweep_grid), dim3(sweep_waves<T>() * kWave), b->sweep_lds, b->stream, S, st,
                         (const int2*)b->d_items, b->nitems, b->d_next);
    else
      hipLaunchKernelGGL((k_b_sweep<T, 2>), dim3(b->sweep_grid), dim3(sweep_waves<T>() * kWave), b->sweep_lds, b->stream, S, st,
                         (const int2*)b->d_items, b->nitems, b->d_next);
    hipLaunchKernelGGL(k_b_sum<T>, dim3(n), dim3(256), 0, b->stream, S, n, b->d_step, b->d_res, 1);
    hipLaunchKernelGGL(k_b_tick, dim3(1), dim3(64), 0, b->stream, b->d_step);
    BHIP(b, hipGetLastError());
    return SALMON_OK;
  }
  if (b->max_ph > 0)
    hipLaunchKernelGGL(k_b_phase<T>, dim3((unsigned)((b->max_ph + 255) / 256), n), dim3(256), 0, b->stream, S, st, 0);
  for (int t = 0; t < 4; ++t) {
    hipLaunchKernelGGL(k_b_project<T>, dim3((unsigned)b->first[0][0]), dim3(256), 0, b->stream, S, n, t, st);
    if (b->first[4][0] > 0)
      hipLaunchKernelGGL(k_b_kbapply<T>, dim3((unsigned)b->first[4][0]), dim3(256), 0, b->stream, S, st);
    if (b->zcls[0][1]) hipLaunchKernelGGL((k_b_hpsi<T, 1>), dim3((unsigned)b->zcls[0][1]), dim3(256), 0, b->stream, S, t, st, b->zcls[0][0]);
    if (b->zcls[1][1]) hipLaunchKernelGGL((k_b_hpsi<T, 2>), dim3((unsigned)b->zcls[1][1]), dim3(256), 0, b->stream, S, t, st, b->zcls[1][0]);
    if (b->zcls[2][1]) hipLaunchKernelGGL((k_b_hpsi<T, 4>), dim3((unsigned)b->zcls[2][1]), dim3(256), 0, b->stream, S, t, st, b->zcls[2][0]);
    if (b->zcls[3][1]) hipLaunchKernelGGL((k_b_hpsi<T, 8>), dim3((unsigned)b->zcls[3][1]), dim3(256), 0, b->stream, S, t, st, b->zcls[3][0]);
  }
  hipLaunchKernelGGL(k_b_fold<T>, dim3((unsigned)b->first[2][0]), dim3(256), 0, b->stream, S, n, st);
  // the observables' phases at k + A(t+dt) (the step's projections are done)
  if (b->max_ph > 0)
    hipLaunchKernelGGL(k_b_phase<T>, dim3((unsigned)((b->max_ph + 255) / 256), n), dim3(256), 0, b->stream, S, st, 1);
  hipLaunchKernelGGL(k_b_observe<T>, dim3((unsigned)b->first[3][0]), dim3(256), 0, b->stream, S, n, st);
  hipLaunchKernelGGL(k_b_sum<T>, dim3(n), dim3(256), 0, b->stream, S, n, b->d_step, b->d_res, 0);
  hipLaunchKernelGGL(k_b_tick, dim3(1), dim3(64), 0, b->stream, b->d_step);
  BHIP(b, hipGetLastError());
  return SALMON_OK;
}

// Descriptors of the batch whose k tables are kbase[i] (system i's device
// kAc(NK,3) table, in place of its kAc0 + A), a zero A(t) table and a zero
// step counter, built on the first call for a kbase array and kept with the
// work buffers.
template <typename T>
static int batch_k_descriptors(salmon_batch* b, const double* const* kbase) {
  const int n = (int)b->ctx.size();
  int r;
  if (!b->tuned && (r = batch_tune<T>(b))) return r;
  if (!b->sweep) return bfail(b, SALMON_ERR_UNSUPPORTED, "the persistent sweep is not set up");
  if (b->d_sys_k && b->sys_k_base == kbase) return SALMON_OK;
  if (!b->d_sys_k) {
    if (hipMalloc(&b->d_sys_k, sizeof(Sys<T>) * n) != hipSuccess ||
        hipMalloc((void**)&b->d_zero_ac, sizeof(double) * 6) != hipSuccess ||
        hipMalloc((void**)&b->d_zero_step, sizeof(int)) != hipSuccess)
      return bfail(b, SALMON_ERR_NOMEM, "batch k descriptors");
    b->bufs.push_back(b->d_sys_k);
    b->bufs.push_back(b->d_zero_ac);
    b->bufs.push_back(b->d_zero_step);
    BHIP(b, hipMemsetAsync(b->d_zero_ac, 0, sizeof(double) * 6, b->stream));
    BHIP(b, hipMemsetAsync(b->d_zero_step, 0, sizeof(int), b->stream));
  }
  std::vector<Sys<T>> S(n);
  BHIP(b, hipMemcpyAsync(S.data(), b->d_sys, sizeof(Sys<T>) * n, hipMemcpyDeviceToHost, b->stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  for (int i = 0; i < n; ++i) S[i].kac0 = kbase[i];
  BHIP(b, hipMemcpyAsync(b->d_sys_k, S.data(), sizeof(Sys<T>) * n, hipMemcpyHostToDevice, b->stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  b->sys_k_base = kbase;
  return SALMON_OK;
}

template <typename T>
static void batch_sweep_launch(salmon_batch* b, const Sys<T>* S, const Step& st) {
  if constexpr (sizeof(T) == 4)
    hipLaunchKernelGGL((k_b_sweep<T, 4>), dim3(b->sweep_grid), dim3(sweep_waves<T>() * kWave), b->sweep_lds, b->stream,
                       S, st, (const int2*)b->d_items, b->nitems, b->d_next);
  else
    hipLaunchKernelGGL((k_b_sweep<T, 2>), dim3(b->sweep_grid), dim3(sweep_waves<T>() * kWave), b->sweep_lds, b->stream,
                       S, st, (const int2*)b->d_items, b->nitems, b->d_next);
}

// One Taylor-4 step (hamiltonian.f90:27-123) of every orbital of the batch at
// the k points kbase[i] over dt, without observables: the persistent sweep
// with the step counter at 0 (so no observables are formed) and A = 0.
template <typename T>
static int batch_taylor(salmon_batch* b, const double* const* kbase, double dt) {
  const int n = (int)b->ctx.size();
  int r;
  if ((r = batch_k_descriptors<T>(b, kbase))) return r;
  const Sys<T>* S = (const Sys<T>*)b->d_sys_k;
  const Step st{b->d_zero_ac, b->d_zero_step, dt,
                {b->d_owner[0], b->d_owner[1], b->d_owner[2], b->d_owner[3], b->d_owner[4]}, 0L};
  if (b->max_ph > 0)
    hipLaunchKernelGGL(k_b_phase<T>, dim3((unsigned)((b->max_ph + 255) / 256), n), dim3(256), 0, b->stream, S, st, 0);
  hipLaunchKernelGGL(k_b_qreset, dim3(1), dim3(64), 0, b->stream, b->d_next);
  batch_sweep_launch<T>(b, S, st);
  BHIP(b, hipGetLastError());
  return SALMON_OK;
}

// The observables rows' sums of every system's orbitals at the k points
// kbase[i] (current.f90, total_energy.f90 terms and the norm, as k_b_sum lays
// them out) into out (nsys x 8): one observables-only sweep launch.
template <typename T>
static int batch_observe(salmon_batch* b, const double* const* kbase, double* out) {
  const int n = (int)b->ctx.size();
  int r;
  if ((r = batch_k_descriptors<T>(b, kbase))) return r;
  if (b->cap_steps < 1 || !b->d_res) {
    if (b->d_res) (void)hipFree(b->d_res);
    b->d_res = nullptr;
    BHIP(b, hipMalloc((void**)&b->d_res, sizeof(double) * 8 * (size_t)n));
    b->cap_steps = 1;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);  // the step graph holds the old results pointer
    b->exec = nullptr;
  }
  const Sys<T>* S = (const Sys<T>*)b->d_sys_k;
  const Step st{b->d_zero_ac, b->d_zero_step, 0.0,
                {b->d_owner[0], b->d_owner[1], b->d_owner[2], b->d_owner[3], b->d_owner[4]}, 0L, 1};
  if (b->max_ph > 0)
    hipLaunchKernelGGL(k_b_phase<T>, dim3((unsigned)((b->max_ph + 255) / 256), n), dim3(256), 0, b->stream, S, st, 2);
  hipLaunchKernelGGL(k_b_qreset, dim3(1), dim3(64), 0, b->stream, b->d_next);
  batch_sweep_launch<T>(b, S, st);
  hipLaunchKernelGGL(k_b_sum<T>, dim3(n), dim3(256), 0, b->stream, S, n, b->d_zero_step, b->d_res, 0);
  BHIP(b, hipGetLastError());
  BHIP(b, hipMemcpyAsync(out, b->d_res, sizeof(double) * 8 * (size_t)n, hipMemcpyDeviceToHost, b->stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  return SALMON_OK;
}

// After the last step of a run on the persistent sweep: the observables of
// the final orbitals (at A(t_nsteps), the device counter then holding nsteps)
// into the run's last result row.
template <typename T>
static int batch_close_sweep(salmon_batch* b, double dt, long ac_stride) {
  const int n = (int)b->ctx.size();
  const Sys<T>* S = (const Sys<T>*)b->d_sys;
  const Step st{b->d_ac, b->d_step, dt, {b->d_owner[0], b->d_owner[1], b->d_owner[2], b->d_owner[3], b->d_owner[4]},
                ac_stride, 1};
  if (b->max_ph > 0)
    hipLaunchKernelGGL(k_b_phase<T>, dim3((unsigned)((b->max_ph + 255) / 256), n), dim3(256), 0, b->stream, S, st, 2);
  hipLaunchKernelGGL(k_b_qreset, dim3(1), dim3(64), 0, b->stream, b->d_next);
  if constexpr (sizeof(T) == 4)
    hipLaunchKernelGGL((k_b_sweep<T, 4>), dim3(b->sweep_grid), dim3(sweep_waves<T>() * kWave), b->sweep_lds, b->stream,
                       S, st, (const int2*)b->d_items, b->nitems, b->d_next);
  else
    hipLaunchKernelGGL((k_b_sweep<T, 2>), dim3(b->sweep_grid), dim3(sweep_waves<T>() * kWave), b->sweep_lds, b->stream,
                       S, st, (const int2*)b->d_items, b->nitems, b->d_next);
  hipLaunchKernelGGL(k_b_sum<T>, dim3(n), dim3(256), 0, b->stream, S, n, b->d_step, b->d_res, 1);
  BHIP(b, hipGetLastError());
  return SALMON_OK;
}

// Times the stencil launch of system i alone for each z blocking that divides NLz.
template <typename T>
static int batch_tune_one(salmon_batch* b, int i, int* best) {
  salmon_ctx* c = b->ctx[i];
  std::vector<Sys<T>> one(1);
  BHIP(b, hipMemcpy(one.data(), (Sys<T>*)b->d_sys + i, sizeof(Sys<T>), hipMemcpyDeviceToHost));
  Sys<T>* d1 = nullptr;
  BHIP(b, hipMalloc((void**)&d1, sizeof(Sys<T>)));
  hipEvent_t e0, e1;
  BHIP(b, hipEventCreate(&e0));
  BHIP(b, hipEventCreate(&e1));
  float bestt = 1e30f;
  *best = 1;
  // the single-system launch below addresses blocks 0.. of a one-system table
  std::vector<unsigned short> zeros((size_t)one[0].NKB * ((c->g.NL + 255) / 256) + 1, 0);
  unsigned short* d_zero = nullptr;
  BHIP(b, hipMalloc((void**)&d_zero, sizeof(unsigned short) * zeros.size()));
  BHIP(b, hipMemcpy(d_zero, zeros.data(), sizeof(unsigned short) * zeros.size(), hipMemcpyHostToDevice));
  const Step st{b->d_ac, b->d_step, 0.0, {d_zero, d_zero, d_zero, d_zero, d_zero}, 0};
  for (int zb : {1, 2, 4, 8}) {
    if (c->g.NLz % zb) continue;
    Sys<T> y = one[0];
    y.zb = zb;
    y.hblk = (c->g.NL / zb + 255) / 256;
    y.first[1] = 0;
    BHIP(b, hipMemcpy(d1, &y, sizeof(Sys<T>), hipMemcpyHostToDevice));
    if ((long)y.NK * y.kb.NPI > 0)
      hipLaunchKernelGGL(k_b_phase<T>, dim3((unsigned)(((long)y.NK * y.kb.NPI + 255) / 256), 1), dim3(256), 0, b->stream,
                         (const Sys<T>*)d1, st, 0);
    const unsigned nblk = (unsigned)((long)y.NKB * y.hblk);
    float tmin = 1e30f;
    for (int rep = 0; rep < 4; ++rep) {
      BHIP(b, hipEventRecord(e0, b->stream));
      switch (zb) {
        case 1: hipLaunchKernelGGL((k_b_hpsi<T, 1>), dim3(nblk), dim3(256), 0, b->stream, (const Sys<T>*)d1, 0, st, 0L); break;
        case 2: hipLaunchKernelGGL((k_b_hpsi<T, 2>), dim3(nblk), dim3(256), 0, b->stream, (const Sys<T>*)d1, 0, st, 0L); break;
        case 4: hipLaunchKernelGGL((k_b_hpsi<T, 4>), dim3(nblk), dim3(256), 0, b->stream, (const Sys<T>*)d1, 0, st, 0L); break;
        default: hipLaunchKernelGGL((k_b_hpsi<T, 8>), dim3(nblk), dim3(256), 0, b->stream, (const Sys<T>*)d1, 0, st, 0L); break;
      }
      BHIP(b, hipEventRecord(e1, b->stream));
      BHIP(b, hipEventSynchronize(e1));
      float ms = 0;
      BHIP(b, hipEventElapsedTime(&ms, e0, e1));
      if (rep > 0) tmin = std::min(tmin, ms);
    }
    if (tmin < bestt) {
      bestt = tmin;
      *best = zb;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(d1);
  (void)hipFree(d_zero);
  return SALMON_OK;
}

template <typename T>
static int batch_tune(salmon_batch* b) {
  // build with z blocking 1 first (work buffers and descriptors for timing)
  batch_release(b);
  b->zb.assign(b->ctx.size(), 1);
  int r = batch_build<T>(b);
  if (r) return r;
  if (!b->d_step) {
    BHIP(b, hipMalloc((void**)&b->d_step, sizeof(int)));
    BHIP(b, hipMemset(b->d_step, 0, sizeof(int)));
  }
  if (!b->d_ac) {
    BHIP(b, hipMalloc((void**)&b->d_ac, sizeof(double) * 6));
    BHIP(b, hipMemset(b->d_ac, 0, sizeof(double) * 6));
    b->cap_ac = 6;
  }
  BHIP(b, hipMemset(b->d_step, 0, sizeof(int)));
  if (b->sweep) {  // the persistent sweep chooses its column shapes without timing
    b->tuned = true;
    return SALMON_OK;
  }
  std::vector<std::array<int, 4>> shapes;  // (NLx, NLy, NLz, zb)
  std::vector<int> zb(b->ctx.size(), 1);
  for (size_t i = 0; i < b->ctx.size(); ++i) {
    const Grid& g = b->ctx[i]->g;
    int found = -1;
    for (size_t j = 0; j < shapes.size(); ++j)
      if (shapes[j][0] == g.NLx && shapes[j][1] == g.NLy && shapes[j][2] == g.NLz) found = (int)j;
    if (found < 0) {
      int best = 1;
      if ((r = batch_tune_one<T>(b, (int)i, &best))) return r;
      shapes.push_back({g.NLx, g.NLy, g.NLz, best});
      found = (int)shapes.size() - 1;
    }
    zb[i] = shapes[found][3];
  }
  batch_release(b);
  b->zb = zb;
  b->tuned = true;
  return batch_build<T>(b);
}

// nsteps steps; ac is one A(t) table (nsteps+1, 3) for every system, or
// (per_sys) one per system, (nsys, nsteps+1, 3)
template <typename T>
static int batch_run(salmon_batch* b, const double* ac, int nsteps, double dt, double* jt, double* et, double* norm,
                     bool per_sys) {
  const int n = (int)b->ctx.size();
  int r;
  if (!b->tuned && (r = batch_tune<T>(b))) return r;
  const size_t need_ac = (size_t)3 * (nsteps + 1) * (per_sys ? n : 1);
  const long stride = per_sys ? 3L * (nsteps + 1) : 0L;
  if (need_ac > b->cap_ac || !b->d_ac) {
    if (b->d_ac) (void)hipFree(b->d_ac);
    b->d_ac = nullptr;
    BHIP(b, hipMalloc((void**)&b->d_ac, sizeof(double) * need_ac));
    b->cap_ac = need_ac;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);  // the graph holds the old table pointer
    b->exec = nullptr;
  }
  if ((size_t)nsteps > b->cap_steps || !b->d_res) {
    if (b->d_res) (void)hipFree(b->d_res);
    b->d_res = nullptr;
    BHIP(b, hipMalloc((void**)&b->d_res, sizeof(double) * 8 * (size_t)n * nsteps));
    b->cap_steps = nsteps;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
  }
  BHIP(b, hipMemcpyAsync(b->d_ac, ac, sizeof(double) * need_ac, hipMemcpyHostToDevice, b->stream));
  BHIP(b, hipMemsetAsync(b->d_step, 0, sizeof(int), b->stream));
  if (b->exec && (b->graph_dt != dt || b->graph_stride != stride)) {
    (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
  }
  if (!b->exec) {
    hipGraph_t graph = nullptr;
    BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeRelaxed));
    r = batch_enqueue_step<T>(b, dt, stride);
    const hipError_t e = hipStreamEndCapture(b->stream, &graph);
    if (r) {
      if (graph) (void)hipGraphDestroy(graph);
      return r;
    }
    if (e != hipSuccess) return bfail(b, SALMON_ERR_HIP, "step capture: %s", hipGetErrorString(e));
    const hipError_t ei = hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) return bfail(b, SALMON_ERR_HIP, "step graph: %s", hipGetErrorString(ei));
    b->graph_dt = dt;
    b->graph_stride = stride;
  }
  for (int s = 0; s < nsteps; ++s) BHIP(b, hipGraphLaunch(b->exec, b->stream));
  if (b->sweep && (r = batch_close_sweep<T>(b, dt, stride))) return r;
  std::vector<double> res(8 * (size_t)n * nsteps);
  BHIP(b, hipMemcpyAsync(res.data(), b->d_res, sizeof(double) * res.size(), hipMemcpyDeviceToHost, b->stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  for (int i = 0; i < n; ++i)
    for (int s = 0; s < nsteps; ++s) {
      const double* q = &res[((size_t)s * n + i) * 8];
      if (jt)
        for (int d = 0; d < 3; ++d) jt[((size_t)i * nsteps + s) * 3 + d] = q[d];
      if (et) et[(size_t)i * nsteps + s] = q[3] + q[4] + q[5];
      if (norm) norm[(size_t)i * nsteps + s] = q[6];
    }
  return SALMON_OK;
}
