// Host side of the learned CG preconditioner: the network object (FP32
// master weights, Adam moments), its inference on the CG residuals of all k
// points, the training samples recorded on a context, and the training step.
// Included by salmon_gpu.hip after gs.hpp.
#pragma once
#include <random>

#include "pcnet.hpp"

using namespace salmon::pc;

struct salmon_pcnet {
  float* w = nullptr;   // (NW) W1 | W2 | W3
  float* m = nullptr;   // Adam moments
  float* v = nullptr;
  double* grad = nullptr;
  long step = 0;
  std::string err;
  // training buffers, sized for n samples of NL points
  int cap_n = 0, cap_NL = 0, cap_NB = 0, cap_nlma = 0;
  std::vector<void*> bufs;
  float *u = nullptr, *a1 = nullptr, *a2 = nullptr, *y = nullptr, *gy = nullptr, *ga1 = nullptr, *ga2 = nullptr;
  float* wpart = nullptr;
  size_t wpart_n = 0;
  cplx<double> *g = nullptr, *x = nullptr, *d = nullptr, *q = nullptr, *p = nullptr, *hp = nullptr, *gq = nullptr;
  cplx<double>* coef = nullptr;
  double *stats = nullptr, *rpart = nullptr, *dpart = nullptr, *nu = nullptr, *co = nullptr, *theta = nullptr,
         *lam = nullptr;
  int* ib = nullptr;
  void* uv = nullptr;
  std::vector<void*> owned;  // w, m, v, grad, bt
  unsigned short* bt = nullptr;  // BF16 weights of the three layers, [n][kk] per layer
  bool packed = false;           // bt holds the current w
};

constexpr int kBt1 = 16 * kpad<C0>(), kBt2 = 16 * kpad<C1>(), kBt3 = 16 * kpad<C2>();

static int pc_fail(salmon_pcnet* n, int code, const char* msg) {
  if (n) n->err = msg;
  return code;
}

static void pcnet_free_buffers(salmon_pcnet* n) {
  for (void* b : n->bufs) (void)hipFree(b);
  n->bufs.clear();
  n->cap_n = n->cap_NL = n->cap_NB = n->cap_nlma = 0;
}

template <typename X>
static int pc_alloc(std::vector<void*>& bufs, X** p, size_t n) {
  if (n == 0) n = 1;
  if (hipMalloc((void**)p, n * sizeof(X)) != hipSuccess) return SALMON_ERR_NOMEM;
  bufs.push_back((void*)*p);
  return SALMON_OK;
}

static int pcnet_buffers(salmon_pcnet* n, int ns, int NL, int NB, int Nlma) {
  if (ns <= n->cap_n && NL == n->cap_NL && NB <= n->cap_NB && Nlma <= n->cap_nlma) return SALMON_OK;
  pcnet_free_buffers(n);
  const size_t v = (size_t)ns * NL;
  const int nchunk = (NL + kRedChunk - 1) / kRedChunk;
  int r = 0;
  r |= pc_alloc(n->bufs, &n->u, v * C0);
  r |= pc_alloc(n->bufs, &n->a1, v * C1);
  r |= pc_alloc(n->bufs, &n->a2, v * C2);
  r |= pc_alloc(n->bufs, &n->y, v * C3);
  r |= pc_alloc(n->bufs, &n->gy, v * C3);
  r |= pc_alloc(n->bufs, &n->ga1, v * C1);
  r |= pc_alloc(n->bufs, &n->ga2, v * C2);
  n->wpart_n = (size_t)27 * ns * ((NL + kWChunk - 1) / kWChunk) * 256;
  r |= pc_alloc(n->bufs, &n->wpart, n->wpart_n);
  for (cplx<double>** b : {&n->g, &n->x, &n->d, &n->q, &n->p, &n->hp, &n->gq}) r |= pc_alloc(n->bufs, b, v);
  r |= pc_alloc(n->bufs, &n->coef, (size_t)ns * (NB + 1));
  r |= pc_alloc(n->bufs, &n->stats, (size_t)4 * ns);
  r |= pc_alloc(n->bufs, &n->rpart, (size_t)3 * ns * nchunk);
  r |= pc_alloc(n->bufs, &n->dpart, (size_t)2 * ns * (NB + 1) * nchunk);
  r |= pc_alloc(n->bufs, &n->nu, (size_t)ns);
  r |= pc_alloc(n->bufs, &n->co, (size_t)6 * ns);
  r |= pc_alloc(n->bufs, &n->theta, (size_t)ns);
  r |= pc_alloc(n->bufs, &n->lam, (size_t)ns);
  r |= pc_alloc(n->bufs, &n->ib, (size_t)ns);
  r |= pc_alloc(n->bufs, (cplx<double>**)&n->uv, (size_t)ns * std::max(1, Nlma));
  if (r) {
    pcnet_free_buffers(n);
    return pc_fail(n, SALMON_ERR_NOMEM, "pcnet training buffers");
  }
  n->cap_n = ns;
  n->cap_NL = NL;
  n->cap_NB = NB;
  n->cap_nlma = std::max(1, Nlma);
  return SALMON_OK;
}

// y = net(u) for n rows; a1, a2 keep the activations.
static int pc_forward(hipStream_t st, const float* W, const Grid& g, int n, const float* u, float* a1, float* a2, float* y) {
  const dim3 gr((g.NL + kBlk - 1) / kBlk, n);
  hipLaunchKernelGGL((k_conv_fwd<C0, C1, true>), gr, dim3(kBlk), 0, st, u, W, g, a1);
  hipLaunchKernelGGL((k_conv_fwd<C1, C2, true>), gr, dim3(kBlk), 0, st, (const float*)a1, W + NW1, g, a2);
  hipLaunchKernelGGL((k_conv_fwd<C2, C3, false>), gr, dim3(kBlk), 0, st, (const float*)a2, W + NW1 + NW2, g, y);
  return hipGetLastError() == hipSuccess ? SALMON_OK : SALMON_ERR_HIP;
}

// Canonicalised input of n residual rows (stride NL).
static int pc_canon(hipStream_t st, const cplx<double>* g, int n, int NL, double* part, double* stats, float* u) {
  const int nchunk = (NL + kRedChunk - 1) / kRedChunk;
  hipLaunchKernelGGL(k_pc_stats, dim3(nchunk, n), dim3(kBlk), 0, st, g, NL, part);
  hipLaunchKernelGGL(k_pc_stats_fin, dim3((n + 63) / 64), dim3(64), 0, st, (const double*)part, nchunk, n, NL, stats);
  hipLaunchKernelGGL(k_pc_input, dim3((NL + kBlk - 1) / kBlk, n), dim3(kBlk), 0, st, g, (const double*)stats, NL, u);
  return hipGetLastError() == hipSuccess ? SALMON_OK : SALMON_ERR_HIP;
}

// q = v - sum_z z <z|v> h over the batch's lower bands and x.
static int pc_project(hipStream_t st, const Batch& B, salmon_pcnet* n, const cplx<double>* v, cplx<double>* out) {
  const int nchunk = (B.NL + kRedChunk - 1) / kRedChunk;
  hipLaunchKernelGGL(k_pc_dots, dim3(nchunk, B.NB + 1, B.n), dim3(kBlk), 0, st, B, v, n->dpart);
  hipLaunchKernelGGL(k_pc_dots_fin, dim3((B.n * (B.NB + 1) + 63) / 64), dim3(64), 0, st, B, (const double*)n->dpart,
                     nchunk, n->coef);
  hipLaunchKernelGGL(k_pc_sub, dim3((B.NL + kBlk - 1) / kBlk, B.n), dim3(kBlk), 0, st, B, v,
                     (const cplx<double>*)n->coef, out);
  return hipGetLastError() == hipSuccess ? SALMON_OK : SALMON_ERR_HIP;
}

template <int CI, int CO>
static int pc_wgrad(hipStream_t st, salmon_pcnet* n, const Grid& g, int ns, const float* ain, const float* gout,
                    double* grad) {
  const int per = (g.NL + kWChunk - 1) / kWChunk;
  const int nblk = ns * per;
  hipLaunchKernelGGL((k_conv_bwd_w<CI, CO>), dim3(nblk, 27), dim3(kBlk), 0, st, ain, gout, g, n->wpart);
  hipLaunchKernelGGL(k_conv_bwd_w_sum, dim3((27 * CI * CO + kBlk - 1) / kBlk), dim3(kBlk), 0, st,
                     (const float*)n->wpart, nblk, CI * CO, grad);
  return hipGetLastError() == hipSuccess ? SALMON_OK : SALMON_ERR_HIP;
}

// BF16 copy of the network's current weights for the matrix-core inference.
static int pc_pack(salmon_ctx* c, salmon_pcnet* net) {
  if (net->packed) return SALMON_OK;
  const hipStream_t st = c->stream;
  hipLaunchKernelGGL((k_pc_pack_bf16<C0, C1>), dim3((kBt1 + 255) / 256), dim3(256), 0, st, (const float*)net->w, net->bt);
  hipLaunchKernelGGL((k_pc_pack_bf16<C1, C2>), dim3((kBt2 + 255) / 256), dim3(256), 0, st, (const float*)(net->w + NW1),
                     net->bt + kBt1);
  hipLaunchKernelGGL((k_pc_pack_bf16<C2, C3>), dim3((kBt3 + 255) / 256), dim3(256), 0, st,
                     (const float*)(net->w + NW1 + NW2), net->bt + kBt1 + kBt2);
  LAUNCHCHK(c);
  net->packed = true;
  return SALMON_OK;
}

// Learned direction of n residual rows g (stride NL) into out, on the matrix
// cores with BF16 weights and activations.
static int pc_infer(salmon_ctx* c, GsState* s, salmon_pcnet* net, const cplx<double>* g, cplx<double>* out) {
  PcInfer& I = s->inf;
  const int NK = s->NK, NL = s->NL;
  const hipStream_t st = c->stream;
  if (I.NK != NK || I.NL != NL) {
    for (void* b : I.bufs) (void)hipFree(b);
    if (!I.bufs.empty()) ++c->gen;
    I.bufs.clear();
    const size_t v = (size_t)NK * NL;
    int r = 0;
    r |= pc_alloc(I.bufs, &I.ub, v * C0);
    r |= pc_alloc(I.bufs, &I.a1b, v * C1);
    r |= pc_alloc(I.bufs, &I.a2b, v * C2);
    r |= pc_alloc(I.bufs, &I.y, v * C3);
    r |= pc_alloc(I.bufs, &I.stats, (size_t)4 * NK);
    r |= pc_alloc(I.bufs, &I.part, (size_t)3 * NK * ((NL + kRedChunk - 1) / kRedChunk));
    if (r) return fail(c, SALMON_ERR_NOMEM, "learned preconditioner buffers");
    I.NK = NK;
    I.NL = NL;
  }
  if (!net->packed) return fail(c, SALMON_ERR_STATE, "learned preconditioner weights not packed");
  const int nchunk = (NL + kRedChunk - 1) / kRedChunk;
  hipLaunchKernelGGL(k_pc_stats, dim3(nchunk, NK), dim3(kBlk), 0, st, g, NL, I.part);
  hipLaunchKernelGGL(k_pc_stats_fin, dim3((NK + 63) / 64), dim3(64), 0, st, (const double*)I.part, nchunk, NK, NL, I.stats);
  hipLaunchKernelGGL(k_pc_input_bf16, dim3((NL + kBlk - 1) / kBlk, NK), dim3(kBlk), 0, st, g, (const double*)I.stats, NL,
                     I.ub);
  const dim3 gm((NL + 63) / 64, NK);
  hipLaunchKernelGGL((k_conv_mfma<C0, C1, true, true>), gm, dim3(kBlk), 0, st, (const unsigned short*)I.ub,
                     (const unsigned short*)net->bt, c->g, (void*)I.a1b);
  hipLaunchKernelGGL((k_conv_mfma<C1, C2, true, true>), gm, dim3(kBlk), 0, st, (const unsigned short*)I.a1b,
                     (const unsigned short*)(net->bt + kBt1), c->g, (void*)I.a2b);
  hipLaunchKernelGGL((k_conv_mfma<C2, C3, false, false>), gm, dim3(kBlk), 0, st, (const unsigned short*)I.a2b,
                     (const unsigned short*)(net->bt + kBt1 + kBt2), c->g, (void*)I.y);
  hipLaunchKernelGGL(k_pc_output, dim3((NL + kBlk - 1) / kBlk, NK), dim3(kBlk), 0, st, (const float*)I.y,
                     (const double*)I.stats, NL, 1, out);
  LAUNCHCHK(c);
  return SALMON_OK;
}

// Host mirrors of the recorded samples' scalars.
static int pc_sync_samples(salmon_ctx* c, GsState* s) {
  s->h_ib.resize(s->smp_n);
  s->h_theta.resize(s->smp_n);
  s->h_valid.resize(s->smp_n);
  if (s->smp_n == 0) return SALMON_OK;
  HIPCHK(c, hipMemcpyAsync(s->h_ib.data(), s->smp_ib, sizeof(int) * s->smp_n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(s->h_theta.data(), s->smp_theta, sizeof(double) * s->smp_n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(s->h_valid.data(), s->smp_valid, sizeof(int) * s->smp_n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SALMON_OK;
}

// One evaluation of the objective and its gradient over the samples idx;
// with update, one Adam step (lr) follows.  loss = mean over the samples.
static int pc_step(salmon_ctx* c, salmon_pcnet* n, const std::vector<int>& idx, bool update, float lr, double* loss,
                   double* grad_out) {
  GsState* s = (GsState*)c->gs;
  const int B = (int)idx.size(), NL = s->NL, NB = s->NB;
  const hipStream_t st = c->stream;
  int r = pcnet_buffers(n, B, NL, NB, c->kb.Nlma);
  if (r) return fail(c, r, "pcnet training buffers");
  std::vector<int> ib(B);
  std::vector<double> th(B), lam(B);
  for (int i = 0; i < B; ++i) {
    const int j = idx[i];
    if (j < 0 || j >= s->smp_n) return fail(c, SALMON_ERR_ARG, "sample index %d out of range", j);
    ib[i] = s->h_ib[j];
    th[i] = s->h_theta[j];
    lam[i] = s->lam[ib[i]];
    HIPCHK(c, hipMemcpyAsync(n->g + (size_t)i * NL, s->smp_g + (size_t)j * NL, sizeof(cplx<double>) * NL,
                             hipMemcpyDeviceToDevice, st));
    HIPCHK(c, hipMemcpyAsync(n->x + (size_t)i * NL, s->smp_x + (size_t)j * NL, sizeof(cplx<double>) * NL,
                             hipMemcpyDeviceToDevice, st));
  }
  HIPCHK(c, hipMemcpyAsync(n->ib, ib.data(), sizeof(int) * B, hipMemcpyHostToDevice, st));
  HIPCHK(c, hipMemcpyAsync(n->theta, th.data(), sizeof(double) * B, hipMemcpyHostToDevice, st));
  HIPCHK(c, hipMemcpyAsync(n->lam, lam.data(), sizeof(double) * B, hipMemcpyHostToDevice, st));
  const Batch Bt{B, NL, NB, c->g.Hxyz, s->Zc, n->x, n->g, n->ib, n->theta, n->lam};
  const int nchunk = (NL + kRedChunk - 1) / kRedChunk;
  const dim3 gp((NL + kBlk - 1) / kBlk, B);
  if (pc_canon(st, n->g, B, NL, n->rpart, n->stats, n->u)) return fail(c, SALMON_ERR_HIP, "pc canon");
  if (pc_forward(st, n->w, c->g, B, n->u, n->a1, n->a2, n->y)) return fail(c, SALMON_ERR_HIP, "pc forward");
  hipLaunchKernelGGL(k_pc_output, gp, dim3(kBlk), 0, st, (const float*)n->y, (const double*)n->stats, NL, 0, n->d);
  LAUNCHCHK(c);
  if (pc_project(st, Bt, n, n->d, n->q)) return fail(c, SALMON_ERR_HIP, "pc project");
  hipLaunchKernelGGL(k_pc_norm_part, dim3(nchunk, B), dim3(kBlk), 0, st, (const cplx<double>*)n->q, NL, n->rpart);
  hipLaunchKernelGGL(k_pc_normalise, gp, dim3(kBlk), 0, st, (const cplx<double>*)n->q, (const double*)n->rpart, nchunk,
                     NL, c->g.Hxyz, n->p, n->nu);
  LAUNCHCHK(c);
  if ((r = apply_h<double>(c, n->p, n->hp, nullptr, 0, B, B, c->NK, c->d_kac0, 0, 0, 0, n->uv))) return r;
  hipLaunchKernelGGL(k_pc_bc_part, dim3(nchunk, B), dim3(kBlk), 0, st, Bt, (const cplx<double>*)n->p,
                     (const cplx<double>*)n->hp, n->rpart);
  hipLaunchKernelGGL(k_pc_objective, dim3((B + 63) / 64), dim3(64), 0, st, Bt, (const double*)n->rpart, nchunk, n->co);
  hipLaunchKernelGGL(k_pc_gradq, gp, dim3(kBlk), 0, st, Bt, (const cplx<double>*)n->p, (const cplx<double>*)n->hp,
                     (const double*)n->co, (const double*)n->nu, n->gq);
  LAUNCHCHK(c);
  if (pc_project(st, Bt, n, n->gq, n->d)) return fail(c, SALMON_ERR_HIP, "pc project gradient");
  hipLaunchKernelGGL(k_pc_gy, gp, dim3(kBlk), 0, st, (const cplx<double>*)n->d, (const double*)n->stats, NL, n->gy);
  LAUNCHCHK(c);
  const Grid& g = c->g;
  const float* W = n->w;
  if (pc_wgrad<C2, C3>(st, n, g, B, n->a2, n->gy, n->grad + NW1 + NW2)) return fail(c, SALMON_ERR_HIP, "pc wgrad 3");
  hipLaunchKernelGGL((k_conv_bwd_data<C2, C3>), gp, dim3(kBlk), 0, st, (const float*)n->gy, (const float*)n->a2,
                     W + NW1 + NW2, g, n->ga2);
  if (pc_wgrad<C1, C2>(st, n, g, B, n->a1, n->ga2, n->grad + NW1)) return fail(c, SALMON_ERR_HIP, "pc wgrad 2");
  hipLaunchKernelGGL((k_conv_bwd_data<C1, C2>), gp, dim3(kBlk), 0, st, (const float*)n->ga2, (const float*)n->a1,
                     W + NW1, g, n->ga1);
  if (pc_wgrad<C0, C1>(st, n, g, B, n->u, n->ga1, n->grad)) return fail(c, SALMON_ERR_HIP, "pc wgrad 1");
  LAUNCHCHK(c);
  std::vector<double> co(6 * (size_t)B);
  HIPCHK(c, hipMemcpyAsync(co.data(), n->co, sizeof(double) * co.size(), hipMemcpyDeviceToHost, st));
  if (grad_out) HIPCHK(c, hipMemcpyAsync(grad_out, n->grad, sizeof(double) * NW, hipMemcpyDeviceToHost, st));
  if (update) {
    n->step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - std::pow(b1, (float)n->step), c2 = 1.0f - std::pow(b2, (float)n->step);
    hipLaunchKernelGGL(k_adam, dim3((NW + kBlk - 1) / kBlk), dim3(kBlk), 0, st, n->w, n->m, n->v,
                       (const double*)n->grad, NW, lr, b1, b2, 1e-8f, c1, c2);
    LAUNCHCHK(c);
    n->packed = false;
  }
  HIPCHK(c, hipStreamSynchronize(st));
  double l = 0.0;
  for (int i = 0; i < B; ++i) l += co[6 * (size_t)i];
  if (loss) *loss = l / B;
  return SALMON_OK;
}

// One CG call through the graph cache (eager while samples are recorded).
static int cg_call(salmon_ctx* c, int ncg, int precond) {
  GsState* s = (GsState*)c->gs;
  salmon_pcnet* net = precond == SALMON_PC_LEARNED ? (salmon_pcnet*)c->pcnet : nullptr;
  int r;
  if (net && (r = pc_pack(c, net))) return r;
  if (s->recording && s->smp_n < s->smp_cap) return cg_run(c, ncg, precond);
  const long key[4] = {ncg, precond, (long)(uintptr_t)net, (long)(uintptr_t)c->d_zu};
  return graph_run(c, s->g_cg, key, [&]() { return cg_run(c, ncg, precond); });
}
