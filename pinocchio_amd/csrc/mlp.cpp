// mlp.cpp -- C ABI of the forward-dynamics surrogate (rbd_mlp_*): parameter storage,
// training step (forward, backward, Adam), inference, and the RNEA-residual gate with
// exact-ABA fallback.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/rbd.h"
#include "k_mlp.hpp"
#include "k_sample.hpp"
#include "kernels_common.hpp"
#include "ops.hpp"
#include "runtime.hpp"

using namespace rbd;
using namespace rbd::mlp;

namespace {

constexpr int kBM = 128;     // row tile of the products other than the forward layers
constexpr int kPadRows = 256; // batch padding unit: the forward layers' row tile (kNtBM)
constexpr int kMaxSplits = 32;

int64_t roundUp(int64_t x, int64_t m) { return (x + m - 1) / m * m; }
unsigned blocksFor(int64_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

template<class T> struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  T* get(size_t want) {
    if (n < want) {
      if (p) HIP_CHECK(hipFree(p));
      p = nullptr;
      HIP_CHECK(hipMalloc(&p, want * sizeof(T)));
      n = want;
    }
    return p;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

// splitmix64 stream for the initial weights
struct Rng {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

}  // namespace

struct rbd_mlp {
  rbd_ctx* ctx = nullptr;
  const rbd_model* model = nullptr;
  int nin = 0, nout = 0, H = 0, L = 0;  // L hidden layers -> L + 1 linear layers
  int kin = 0, kout = 0;                // padded input / output widths
  FeatMap fm{};
  std::vector<int> inR, outR, inP, outP;         // real and padded layer widths
  std::vector<int64_t> offW, offB;              // offsets in the parameter block
  int64_t total = 0;                            // padded parameter count
  int64_t nparams = 0;                          // real parameter count
  int64_t step = 0;                             // Adam steps taken
  DevBuf<float> P, m1, m2, G, norm, lossAcc;    // norm: xmean, xinv, ymean, ystd, yinv
  DevBuf<bf16> Pb, Pf;  // bf16 copy; its fragment-major copy for the fused inference kernel
  bool fragStale = true; // Pf lags Pb
  // activations / workspace (grown with the batch)
  DevBuf<bf16> X, act, dOut, dZ;
  DevBuf<float> Z, out, Y, tauChk, cq, cv, ct, cdd, rel;
  DevBuf<unsigned int> count;
  DevBuf<int64_t> idx;
  std::vector<float> hostNorm;

  ~rbd_mlp() {
    for (DevBuf<float>* b : {&P, &m1, &m2, &G, &norm, &lossAcc, &Z, &out, &Y, &tauChk, &cq, &cv, &ct, &cdd, &rel})
      b->release();
    for (DevBuf<bf16>* b : {&Pb, &Pf, &X, &act, &dOut, &dZ}) b->release();
    count.release();
    idx.release();
  }
  float* xmean() { return norm.p; }
  float* xinv() { return norm.p + kMaxFeat; }
  float* ymean() { return norm.p + 2 * kMaxFeat; }
  float* ystd() { return norm.p + 3 * kMaxFeat; }
  float* yinv() { return norm.p + 4 * kMaxFeat; }
  bf16* actL(int l, int64_t Bp) { return act.p + (int64_t)l * Bp * H; }  // l = 0 .. L-1
  float* ZL(int l, int64_t Bp) { return Z.p + (int64_t)l * Bp * H; }
};

namespace {

void setDev(rbd_ctx* ctx) {
  RBD_CHECK_ARG(!ctx->capturing, "capture: the surrogate's calls cannot be recorded");
  HIP_CHECK(hipSetDevice(ctx->device));
}

template<int BM, int BN, bool AK, bool BK, int EPI> void gemm(hipStream_t st, const GemmArgs& g, int splits = 1) {
  if (g.M % BM || g.N % BN || g.K % kStage || g.kChunk % kStage || (int64_t)g.kChunk * splits < g.K)
    throw Error(RBD_EINVAL, "internal: product shape does not fit its tiles");
  hipLaunchKernelGGL((k_gemm<BM, BN, AK, BK, EPI>), dim3(g.M / BM, g.N / BN, splits), dim3(kGemmThreads), 0, st, g);
}

// forward product over 256 x 128 tiles (hidden layers)
template<int EPI> void gemmNt(hipStream_t st, const GemmArgs& g) {
  if (g.M % kNtBM || g.N % kNtBN || g.K % kNtBK || g.kChunk != g.K)
    throw Error(RBD_EINVAL, "internal: product shape does not fit its tiles");
  hipLaunchKernelGGL((k_gemm_nt<EPI>), dim3((g.M / kNtBM) * (g.N / kNtBN)), dim3(kGemmThreads), 0, st, g);
}

GemmArgs gargs(const bf16* A, int64_t lda, const bf16* B, int64_t ldb, int M, int N, int K) {
  GemmArgs g{};
  g.A = A;
  g.lda = lda;
  g.B = B;
  g.ldb = ldb;
  g.M = M;
  g.N = N;
  g.K = K;
  g.kChunk = K;
  return g;
}

// grow the batch-dependent buffers for Bp padded rows
void ensure(rbd_mlp* m, int64_t Bp, bool train) {
  m->X.get((size_t)Bp * m->kin);
  m->act.get((size_t)Bp * m->H * m->L);
  m->out.get((size_t)Bp * m->kout);
  if (train) {
    m->Z.get((size_t)Bp * m->H * m->L);
    m->Y.get((size_t)Bp * m->kout);
    m->dOut.get((size_t)Bp * m->kout);
    m->dZ.get((size_t)Bp * m->H * 2);
    m->G.get((size_t)kMaxSplits * m->total);
  }
}

void checkVec(const void* p, int64_t ld, int64_t B, const char* what) {
  if (B > 0 && !p) throw Error(RBD_EINVAL, std::string("null ") + what);
  if (B > 0 && ld < B) throw Error(RBD_EINVAL, std::string(what) + ": leading dimension below the batch size");
}

// features of B configurations into m->X (rows [B, Bp) zeroed)
void features(rbd_mlp* m, hipStream_t st, int64_t B, int64_t Bp, const float* q, int64_t ldq, const float* v,
              int64_t ldv, const float* tau, int64_t ldt, float* Xraw) {
  if (B > 0)
    hipLaunchKernelGGL(k_features, dim3((unsigned)((B + 63) / 64)), dim3(64), 64 * (m->fm.ldx + 2) * sizeof(bf16), st,
                       m->fm, B, q, ldq, v, ldv, tau, ldt, m->xmean(), m->xinv(), m->X.p, Xraw);
  if (Bp > B) hipLaunchKernelGGL(k_zero_rows, dim3(blocksFor((Bp - B) * m->kin)), dim3(256), 0, st, m->X.p,
                                 (int64_t)m->kin, B, Bp);
}

// The fused inference kernel takes hidden widths 128, 256 and 512 (the activations of its
// 128 configurations stay in LDS) and at most 64 padded outputs.
bool fusedInference(const rbd_mlp* m) {
  return (m->H == 128 || m->H == 256 || m->H == 512) && m->kout <= kInfMaxOut && m->L + 1 <= kInfMaxLayers;
}

template<int H> void launchInfer(hipStream_t st, const InferArgs& a, size_t lds, unsigned grid) {
  constexpr int R = inferRows(H);
  // dynamic LDS above the default 64 KB per workgroup needs the kernel attribute (set once)
  static const hipError_t attr =
      hipFuncSetAttribute((const void*)&k_mlp_infer<H, R>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  if (attr != hipSuccess) throw Error(RBD_EDEVICE, "fused inference: the device refuses 160 KB of LDS per workgroup");
  hipLaunchKernelGGL((k_mlp_infer<H, R>), dim3(grid), dim3(kInfThreads), lds, st, a);
}

// ddq of B configurations through the fused kernel (no workspace)
void infer(rbd_mlp* m, hipStream_t st, int64_t B, const float* q, int64_t ldq, const float* v, int64_t ldv,
           const float* tau, int64_t ldt, float* ddq, int64_t ldd) {
  InferArgs a{};
  a.fm = m->fm;
  a.B = B;
  a.q = q;
  a.ldq = ldq;
  a.v = v;
  a.ldv = ldv;
  a.tau = tau;
  a.ldt = ldt;
  a.xmean = m->xmean();
  a.xinv = m->xinv();
  a.ymean = m->ymean();
  a.ystd = m->ystd();
  if (m->fragStale) {  // refresh the fragment-major weights (stream-ordered after the update)
    bf16* F = m->Pf.get(m->total);
    for (int l = 0; l <= m->L; ++l) {
      const int64_t n = (int64_t)m->outP[l] * m->inP[l];
      hipLaunchKernelGGL(k_pack_fragments, dim3(blocksFor(n)), dim3(256), 0, st, m->Pb.p + m->offW[l], F + m->offW[l],
                         m->outP[l], m->inP[l]);
    }
    m->fragStale = false;
  }
  a.Wb = m->Pf.p;
  a.P = m->P.p;
  for (int l = 0; l <= m->L; ++l) {
    a.offW[l] = m->offW[l];
    a.offB[l] = m->offB[l];
  }
  a.L = m->L;
  a.kin = m->kin;
  a.kout = m->kout;
  a.nout = m->nout;
  a.pitch = std::max(m->kin, m->H) + 8;
  a.ddq = ddq;
  a.ldd = ldd;
  const int R = inferRows(m->H);
  const size_t lds = (size_t)R * a.pitch * sizeof(bf16) + (size_t)m->L * m->H * sizeof(float);
  if (lds > 160 * 1024 || m->kin % 64 || m->kin > kMaxLdx) throw Error(RBD_EINVAL, "internal: network does not fit the fused kernel");
  const unsigned grid = (unsigned)((B + R - 1) / R);
  if (m->H == 512) launchInfer<512>(st, a, lds, grid);
  else if (m->H == 256) launchInfer<256>(st, a, lds, grid);
  else launchInfer<128>(st, a, lds, grid);
}

// forward pass over Bp rows of m->X; pre-activations kept when training
void forward(rbd_mlp* m, hipStream_t st, int64_t Bp, bool keepZ) {
  const int M = (int)Bp;
  const bf16* in = m->X.p;
  int64_t ldin = m->kin;
  for (int l = 0; l <= m->L; ++l) {
    GemmArgs g = gargs(in, ldin, m->Pb.p + m->offW[l], m->inP[l], M, m->outP[l], m->inP[l]);
    g.bias = m->P.p + m->offB[l];
    if (l < m->L) {
      g.Cb = m->actL(l, Bp);
      g.ldc = m->H;
      g.Zout = keepZ ? m->ZL(l, Bp) : nullptr;
      g.ldz = m->H;
      gemmNt<EPI_SILU>(st, g);
      in = g.Cb;
      ldin = m->H;
    } else {
      g.Cf = m->out.p;
      g.ldc = m->kout;
      gemm<kBM, 32, true, true, EPI_LINEAR>(st, g);
    }
  }
}

int splitsFor(int64_t Bp) {
  int s = kMaxSplits;
  while (s > 1 && (Bp / s) % kStage) s /= 2;
  return s;
}

// backward pass (after forward with keepZ and k_loss_grad): split partials of every
// parameter's gradient in m->G (split stride m->total); returns the split count
int backward(rbd_mlp* m, hipStream_t st, int64_t Bp) {
  const int S = splitsFor(Bp);
  const int chunk = (int)(Bp / S);
  const int M = (int)Bp;
  // gradient w.r.t. the pre-activation of layer l's output: dOut for l = L, else dZ ping-pong
  const bf16* dcur = m->dOut.p;
  int64_t lddcur = m->kout;
  for (int l = m->L; l >= 0; --l) {
    const bf16* actIn = l == 0 ? m->X.p : m->actL(l - 1, Bp);
    const int64_t ldIn = l == 0 ? m->kin : m->H;
    // weight gradient: dW_l(o, i) = sum_b d(b, o) act(b, i)
    GemmArgs g = gargs(dcur, lddcur, actIn, ldIn, m->outP[l], m->inP[l], M);
    g.kChunk = chunk;
    g.Cf = m->G.p + m->offW[l];
    g.ldc = m->inP[l];
    g.zstride = m->total;
    if (l == m->L) gemm<32, 128, false, false, EPI_PARTIAL>(st, g, S);
    else if (l == 0) gemm<kBM, 32, false, false, EPI_PARTIAL>(st, g, S);
    else gemm<kBM, 128, false, false, EPI_PARTIAL>(st, g, S);
    // bias gradient
    hipLaunchKernelGGL(k_colsum, dim3(blocksFor(m->outP[l]), S), dim3(256), 0, st, dcur, lddcur, m->outP[l],
                       (int64_t)chunk, m->G.p + m->offB[l], m->total);
    if (l == 0) break;
    // input gradient through SiLU: dZ_{l-1}(b, i) = sum_o d(b, o) W_l(o, i) * silu'(Z_{l-1}(b, i))
    bf16* dnext = m->dZ.p + (int64_t)(l & 1) * Bp * m->H;
    GemmArgs h = gargs(dcur, lddcur, m->Pb.p + m->offW[l], m->inP[l], M, m->inP[l], m->outP[l]);
    h.Z = m->ZL(l - 1, Bp);
    h.ldz = m->H;
    h.Cb = dnext;
    h.ldc = m->H;
    gemm<kBM, 128, true, false, EPI_SILU_GRAD>(st, h);
    dcur = dnext;
    lddcur = m->H;
  }
  return S;
}

// loss gradient of B real rows against targets ddq (batch-fastest); accumulates the
// mean squared error (normalised units) into *lossAcc
void lossGrad(rbd_mlp* m, hipStream_t st, int64_t B, int64_t Bp, const float* ddq, int64_t ldd, float* lossAcc) {
  hipLaunchKernelGGL(k_targets, dim3(blocksFor(Bp * m->kout)), dim3(256), 0, st, m->nout, m->kout, B, Bp, ddq, ldd,
                     m->ymean(), m->yinv(), m->Y.p);
  hipLaunchKernelGGL(k_loss_grad, dim3(blocksFor(Bp * m->kout)), dim3(256), 0, st, m->nout, m->kout, B, Bp,
                     m->out.p, m->Y.p, m->dOut.p, lossAcc);
}

void uploadParamsBf16(rbd_mlp* m, hipStream_t st) {
  hipLaunchKernelGGL(k_to_bf16, dim3(blocksFor(m->total)), dim3(256), 0, st, m->total, m->P.p, m->Pb.p);
  m->fragStale = true;
}

// packed (unpadded, per layer W row-major then b) <-> padded block
void packedToPadded(const rbd_mlp* m, const float* src, std::vector<float>& dst) {
  dst.assign(m->total, 0.f);
  int64_t o = 0;
  for (int l = 0; l <= m->L; ++l) {
    for (int r = 0; r < m->outR[l]; ++r)
      for (int c = 0; c < m->inR[l]; ++c) dst[m->offW[l] + (int64_t)r * m->inP[l] + c] = src[o++];
    for (int r = 0; r < m->outR[l]; ++r) dst[m->offB[l] + r] = src[o++];
  }
}
void paddedToPacked(const rbd_mlp* m, const float* src, float* dst) {
  int64_t o = 0;
  for (int l = 0; l <= m->L; ++l) {
    for (int r = 0; r < m->outR[l]; ++r)
      for (int c = 0; c < m->inR[l]; ++c) dst[o++] = src[m->offW[l] + (int64_t)r * m->inP[l] + c];
    for (int r = 0; r < m->outR[l]; ++r) dst[o++] = src[m->offB[l] + r];
  }
}

void checkMlp(rbd_ctx* ctx, rbd_mlp* m) {
  RBD_CHECK_ARG(ctx && m, "null argument");
  RBD_CHECK_ARG(m->ctx == ctx, "surrogate belongs to another context");
}

}  // namespace

extern "C" {

int rbd_mlp_create(rbd_ctx* ctx, const rbd_model* model, int32_t hidden, int32_t nhidden, uint64_t seed,
                   rbd_mlp** out) {
  return guarded([&] {
    RBD_CHECK_ARG(ctx && model && out, "null argument");
    RBD_CHECK_ARG(model->ctx == ctx, "model belongs to another context");
    RBD_CHECK_ARG(hidden >= 128 && hidden % 128 == 0 && hidden <= 4096, "hidden width must be a multiple of 128");
    RBD_CHECK_ARG(nhidden >= 1 && nhidden <= 8, "between 1 and 8 hidden layers");
    RBD_CHECK_ARG(model->ops && model->ops->rnea_f32 && model->ops->aba_f32,
                  "the surrogate needs the model's RNEA and ABA kernels");
    auto m = std::make_unique<rbd_mlp>();
    m->ctx = ctx;
    m->model = model;
    FeatMap& fm = m->fm;
    fm.nq = model->nq;
    fm.nv = model->nv;
    fm.hasFF = 0;
    fm.nrev = 0;
    for (int j = 1; j < model->njoints; ++j) {
      const int t = model->jtype[j];
      if (t == RBD_JOINT_FREEFLYER) {
        RBD_CHECK_ARG(j == 1 && model->parents[j] == 0, "a free flyer is supported as the root joint only");
        fm.hasFF = 1;
      } else {
        RBD_CHECK_ARG(fm.nrev < 96, "too many revolute joints for the feature map");
        fm.revq[fm.nrev++] = model->iq[j];
      }
    }
    fm.nfeat = 2 * fm.nrev + (fm.hasFF ? 13 : 0) + 2 * model->nv;
    RBD_CHECK_ARG(fm.nfeat <= kMaxFeat && (int)roundUp(fm.nfeat, kNtBK) <= kMaxLdx, "too many input features");
    m->nin = fm.nfeat;
    m->nout = model->nv;
    m->H = hidden;
    m->L = nhidden;
    m->kin = (int)roundUp(m->nin, kNtBK);  // padded widths: whole LDS stages of the products
    m->kout = (int)roundUp(m->nout, kStage);
    RBD_CHECK_ARG(m->nout <= kMaxFeat, "too many outputs");
    fm.ldx = m->kin;
    for (int l = 0; l <= m->L; ++l) {
      m->inR.push_back(l == 0 ? m->nin : m->H);
      m->inP.push_back(l == 0 ? m->kin : m->H);
      m->outR.push_back(l == m->L ? m->nout : m->H);
      m->outP.push_back(l == m->L ? m->kout : m->H);
    }
    for (int l = 0; l <= m->L; ++l) {
      m->offW.push_back(m->total);
      m->total += (int64_t)m->outP[l] * m->inP[l];
      m->offB.push_back(m->total);
      m->total += roundUp(m->outP[l], 8);
      m->nparams += (int64_t)m->outR[l] * m->inR[l] + m->outR[l];
    }
    setDev(ctx);
    // initial weights and biases U(-1/sqrt(fan_in), 1/sqrt(fan_in))
    std::vector<float> packed(m->nparams);
    Rng rng{seed};
    {
      int64_t o = 0;
      for (int l = 0; l <= m->L; ++l) {
        const double bound = 1.0 / std::sqrt((double)m->inR[l]);
        const int64_t n = (int64_t)m->outR[l] * m->inR[l] + m->outR[l];
        for (int64_t k = 0; k < n; ++k) packed[o++] = (float)((2.0 * rng.uniform() - 1.0) * bound);
      }
    }
    std::vector<float> padded;
    packedToPadded(m.get(), packed.data(), padded);
    m->P.get(m->total);
    m->m1.get(m->total);
    m->m2.get(m->total);
    m->Pb.get(m->total);
    m->norm.get(5 * kMaxFeat);
    m->lossAcc.get(1);
    m->count.get(1);
    HIP_CHECK(hipMemcpy(m->P.p, padded.data(), m->total * sizeof(float), hipMemcpyHostToDevice));
    std::vector<float> zeros(m->total, 0.f);
    HIP_CHECK(hipMemcpy(m->m1.p, zeros.data(), m->total * sizeof(float), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(m->m2.p, zeros.data(), m->total * sizeof(float), hipMemcpyHostToDevice));
    m->hostNorm.assign(5 * kMaxFeat, 0.f);
    for (int k = 0; k < kMaxFeat; ++k) {
      m->hostNorm[kMaxFeat + k] = 1.f;      // xinv
      m->hostNorm[3 * kMaxFeat + k] = 1.f;  // ystd
      m->hostNorm[4 * kMaxFeat + k] = 1.f;  // yinv
    }
    HIP_CHECK(hipMemcpy(m->norm.p, m->hostNorm.data(), 5 * kMaxFeat * sizeof(float), hipMemcpyHostToDevice));
    uploadParamsBf16(m.get(), ctx->stream);
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    *out = m.release();
  });
}

int rbd_mlp_destroy(rbd_mlp* m) {
  return guarded([&] {
    if (!m) return;
    (void)hipSetDevice(m->ctx->device);
    (void)hipStreamSynchronize(m->ctx->stream);
    delete m;
  });
}

int rbd_mlp_dims(const rbd_mlp* m, int32_t* nin, int32_t* nout, int32_t* hidden, int32_t* nhidden, int64_t* nparams) {
  return guarded([&] {
    RBD_CHECK_ARG(m, "null surrogate");
    if (nin) *nin = m->nin;
    if (nout) *nout = m->nout;
    if (hidden) *hidden = m->H;
    if (nhidden) *nhidden = m->L;
    if (nparams) *nparams = m->nparams;
  });
}

int rbd_mlp_set_normalization(rbd_mlp* m, const float* xmean, const float* xstd, const float* ymean,
                              const float* ystd) {
  return guarded([&] {
    RBD_CHECK_ARG(m && xmean && xstd && ymean && ystd, "null argument");
    for (int k = 0; k < m->nin; ++k) RBD_CHECK_ARG(xstd[k] > 0.f, "feature scales must be positive");
    for (int k = 0; k < m->nout; ++k) RBD_CHECK_ARG(ystd[k] > 0.f, "output scales must be positive");
    for (int k = 0; k < m->nin; ++k) {
      m->hostNorm[k] = xmean[k];
      m->hostNorm[kMaxFeat + k] = 1.f / xstd[k];
    }
    for (int k = 0; k < m->nout; ++k) {
      m->hostNorm[2 * kMaxFeat + k] = ymean[k];
      m->hostNorm[3 * kMaxFeat + k] = ystd[k];
      m->hostNorm[4 * kMaxFeat + k] = 1.f / ystd[k];
    }
    setDev(m->ctx);
    HIP_CHECK(hipStreamSynchronize(m->ctx->stream));
    HIP_CHECK(hipMemcpy(m->norm.p, m->hostNorm.data(), 5 * kMaxFeat * sizeof(float), hipMemcpyHostToDevice));
  });
}

int rbd_mlp_get_params(rbd_mlp* m, float* packed, int64_t n) {
  return guarded([&] {
    RBD_CHECK_ARG(m && packed, "null argument");
    RBD_CHECK_ARG(n == m->nparams, "parameter count mismatch");
    setDev(m->ctx);
    HIP_CHECK(hipStreamSynchronize(m->ctx->stream));
    std::vector<float> padded(m->total);
    HIP_CHECK(hipMemcpy(padded.data(), m->P.p, m->total * sizeof(float), hipMemcpyDeviceToHost));
    paddedToPacked(m, padded.data(), packed);
  });
}

int rbd_mlp_set_params(rbd_mlp* m, const float* packed, int64_t n) {
  return guarded([&] {
    RBD_CHECK_ARG(m && packed, "null argument");
    RBD_CHECK_ARG(n == m->nparams, "parameter count mismatch");
    std::vector<float> padded;
    packedToPadded(m, packed, padded);
    setDev(m->ctx);
    HIP_CHECK(hipStreamSynchronize(m->ctx->stream));
    HIP_CHECK(hipMemcpy(m->P.p, padded.data(), m->total * sizeof(float), hipMemcpyHostToDevice));
    std::vector<float> zeros(m->total, 0.f);
    HIP_CHECK(hipMemcpy(m->m1.p, zeros.data(), m->total * sizeof(float), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(m->m2.p, zeros.data(), m->total * sizeof(float), hipMemcpyHostToDevice));
    m->step = 0;
    uploadParamsBf16(m, m->ctx->stream);
    HIP_CHECK(hipStreamSynchronize(m->ctx->stream));
  });
}

int rbd_mlp_features(rbd_ctx* ctx, rbd_mlp* m, int64_t B, const float* q, int64_t ldq, const float* v, int64_t ldv,
                     const float* tau, int64_t ldt, float* Xraw, uint32_t flags) {
  return guarded([&] {
    checkMlp(ctx, m);
    RBD_CHECK_ARG(B >= 0, "negative batch");
    checkVec(q, ldq, B, "q");
    checkVec(v, ldv, B, "v");
    checkVec(tau, ldt, B, "tau");
    RBD_CHECK_ARG(B == 0 || Xraw, "null output");
    if (B == 0) return;
    setDev(ctx);
    const int64_t Bp = roundUp(B, kPadRows);
    ensure(m, Bp, false);
    features(m, ctx->stream, B, Bp, q, ldq, v, ldv, tau, ldt, Xraw);
    HIP_CHECK(hipGetLastError());
    if (!(flags & RBD_ASYNC)) HIP_CHECK(hipStreamSynchronize(ctx->stream));
  });
}

int rbd_mlp_predict(rbd_ctx* ctx, rbd_mlp* m, int64_t B, const float* q, int64_t ldq, const float* v, int64_t ldv,
                    const float* tau, int64_t ldt, float* ddq, int64_t ldd, uint32_t flags) {
  return guarded([&] {
    checkMlp(ctx, m);
    RBD_CHECK_ARG(B >= 0, "negative batch");
    checkVec(q, ldq, B, "q");
    checkVec(v, ldv, B, "v");
    checkVec(tau, ldt, B, "tau");
    checkVec(ddq, ldd, B, "ddq");
    if (B == 0) return;
    setDev(ctx);
    if (fusedInference(m) && !(flags & RBD_MLP_LAYERED)) {
      infer(m, ctx->stream, B, q, ldq, v, ldv, tau, ldt, ddq, ldd);
    } else {
      const int64_t Bp = roundUp(B, kPadRows);
      ensure(m, Bp, false);
      features(m, ctx->stream, B, Bp, q, ldq, v, ldv, tau, ldt, nullptr);
      forward(m, ctx->stream, Bp, false);
      hipLaunchKernelGGL(k_outputs, dim3(blocksFor(B * m->nout)), dim3(256), 0, ctx->stream, m->nout, m->kout, B,
                         m->out.p, m->ymean(), m->ystd(), ddq, ldd);
    }
    HIP_CHECK(hipGetLastError());
    if (!(flags & RBD_ASYNC)) HIP_CHECK(hipStreamSynchronize(ctx->stream));
  });
}

// One optimiser step on a minibatch: loss and gradients of every parameter, then Adam.
int rbd_mlp_train_step(rbd_ctx* ctx, rbd_mlp* m, int64_t B, const float* q, int64_t ldq, const float* v, int64_t ldv,
                       const float* tau, int64_t ldt, const float* ddq, int64_t ldd, float lr, float* loss,
                       uint32_t flags) {
  return guarded([&] {
    checkMlp(ctx, m);
    RBD_CHECK_ARG(B > 0, "empty minibatch");
    checkVec(q, ldq, B, "q");
    checkVec(v, ldv, B, "v");
    checkVec(tau, ldt, B, "tau");
    checkVec(ddq, ldd, B, "ddq");
    RBD_CHECK_ARG(lr > 0.f, "learning rate must be positive");
    setDev(ctx);
    hipStream_t st = ctx->stream;
    const int64_t Bp = roundUp(B, kPadRows);
    ensure(m, Bp, true);
    features(m, st, B, Bp, q, ldq, v, ldv, tau, ldt, nullptr);
    forward(m, st, Bp, true);
    lossGrad(m, st, B, Bp, ddq, ldd, loss);
    const int S = backward(m, st, Bp);
    m->step += 1;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = (float)(1.0 / (1.0 - std::pow((double)b1, (double)m->step)));
    const float c2 = (float)(1.0 / (1.0 - std::pow((double)b2, (double)m->step)));
    hipLaunchKernelGGL(k_adam, dim3(blocksFor(m->total)), dim3(256), 0, st, m->total, m->P.p, m->m1.p, m->m2.p, m->G.p,
                       S, m->total, lr, b1, b2, eps, c1, c2, m->Pb.p);
    m->fragStale = true;
    HIP_CHECK(hipGetLastError());
    if (!(flags & RBD_ASYNC)) HIP_CHECK(hipStreamSynchronize(st));
  });
}

// Loss and gradient of every parameter (packed layout, host arrays) without an update.
int rbd_mlp_gradients(rbd_ctx* ctx, rbd_mlp* m, int64_t B, const float* q, int64_t ldq, const float* v, int64_t ldv,
                      const float* tau, int64_t ldt, const float* ddq, int64_t ldd, float* grad, int64_t n,
                      float* loss) {
  return guarded([&] {
    checkMlp(ctx, m);
    RBD_CHECK_ARG(B > 0, "empty minibatch");
    checkVec(q, ldq, B, "q");
    checkVec(v, ldv, B, "v");
    checkVec(tau, ldt, B, "tau");
    checkVec(ddq, ldd, B, "ddq");
    RBD_CHECK_ARG(grad && n == m->nparams, "gradient buffer of nparams floats expected");
    setDev(ctx);
    hipStream_t st = ctx->stream;
    const int64_t Bp = roundUp(B, kPadRows);
    ensure(m, Bp, true);
    hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(64), 0, st, m->lossAcc.p, (int64_t)1, 0.f);
    features(m, st, B, Bp, q, ldq, v, ldv, tau, ldt, nullptr);
    forward(m, st, Bp, true);
    lossGrad(m, st, B, Bp, ddq, ldd, m->lossAcc.p);
    const int S = backward(m, st, Bp);
    float* sum = m->Y.p;  // Y is consumed; reuse it when large enough, else a fresh buffer
    DevBuf<float> tmp;
    if ((int64_t)m->Y.n < m->total) sum = tmp.get(m->total);
    hipLaunchKernelGGL(k_sum_partials, dim3(blocksFor(m->total)), dim3(256), 0, st, m->total, m->G.p, S, m->total,
                       sum);
    HIP_CHECK(hipGetLastError());
    std::vector<float> padded(m->total);
    HIP_CHECK(hipMemcpyAsync(padded.data(), sum, m->total * sizeof(float), hipMemcpyDeviceToHost, st));
    float lh = 0.f;
    HIP_CHECK(hipMemcpyAsync(&lh, m->lossAcc.p, sizeof(float), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    tmp.release();
    paddedToPacked(m, padded.data(), grad);
    if (loss) *loss = lh;
  });
}

// Surrogate forward dynamics with exact fallback: ddq = MLP(q, v, tau); configurations
// whose RNEA residual ||rnea(q, v, ddq) - tau|| exceeds tol ||tau|| are recomputed with
// ABA.  Synchronous (the fallback count sizes the ABA launch).
int rbd_mlp_forward_dynamics(rbd_ctx* ctx, rbd_mlp* m, int64_t B, const float* q, int64_t ldq, const float* v,
                             int64_t ldv, const float* tau, int64_t ldt, float tol, float* ddq, int64_t ldd,
                             int64_t* nfallback, unsigned char* accepted, float* residual) {
  return guarded([&] {
    checkMlp(ctx, m);
    RBD_CHECK_ARG(B >= 0, "negative batch");
    checkVec(q, ldq, B, "q");
    checkVec(v, ldv, B, "v");
    checkVec(tau, ldt, B, "tau");
    checkVec(ddq, ldd, B, "ddq");
    RBD_CHECK_ARG(tol >= 0.f, "negative tolerance");
    if (nfallback) *nfallback = 0;
    if (B == 0) return;
    setDev(ctx);
    hipStream_t st = ctx->stream;
    const rbd_model* md = m->model;
    const int nq = md->nq, nv = md->nv;
    if (fusedInference(m)) {
      infer(m, st, B, q, ldq, v, ldv, tau, ldt, ddq, ldd);
    } else {
      const int64_t Bp = roundUp(B, kPadRows);
      ensure(m, Bp, false);
      features(m, st, B, Bp, q, ldq, v, ldv, tau, ldt, nullptr);
      forward(m, st, Bp, false);
      hipLaunchKernelGGL(k_outputs, dim3(blocksFor(B * nv)), dim3(256), 0, st, nv, m->kout, B, m->out.p, m->ymean(),
                         m->ystd(), ddq, ldd);
    }
    float* tc = m->tauChk.get((size_t)nv * B);
    md->ops->rnea_f32(st, md->gen.get(), md->mc32, B, CView<float>{q, ldq, 1}, CView<float>{v, ldv, 1}, CView<float>{ddq, ldd, 1},
                      View<float>{tc, B, 1});
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, m->count.p, (int64_t)1, 0u);
    float* cq = m->cq.get((size_t)nq * B);
    float* cv = m->cv.get((size_t)nv * B);
    float* ct = m->ct.get((size_t)nv * B);
    int64_t* idx = m->idx.get((size_t)B);
    hipLaunchKernelGGL(k_gate, dim3(blocksFor(B)), dim3(256), 0, st, nq, nv, B, tc, B, tau, ldt, q, ldq, v, ldv, tol,
                       m->count.p, idx, cq, cv, ct, B, accepted, residual);
    unsigned int cnt = 0;
    HIP_CHECK(hipMemcpyAsync(&cnt, m->count.p, sizeof(cnt), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    if (cnt > 0) {
      float* cdd = m->cdd.get((size_t)nv * B);
      md->ops->aba_f32(st, md->gen.get(), md->mc32, (int64_t)cnt, CView<float>{cq, B, 1}, CView<float>{cv, B, 1},
                       CView<float>{ct, B, 1}, View<float>{cdd, B, 1});
      hipLaunchKernelGGL(k_scatter, dim3(blocksFor((int64_t)cnt * nv)), dim3(256), 0, st, nv, (int64_t)cnt, idx, cdd,
                         B, ddq, ldd);
    }
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(st));
    if (nfallback) *nfallback = cnt;
  });
}

// Seeded batch inputs generated on the device, draw for draw the stream of the host
// sampler: per joint in order, free-flyer translation U(-r, r)^3 and a uniform unit
// quaternion, revolute angles U(-pi, pi); then v ~ N(0, 1), a ~ N(0, 1), tau ~ N(0, s).
int rbd_sample_inputs(rbd_ctx* ctx, const rbd_model* model, int64_t B, uint64_t seed, double tau_sigma,
                      double base_range, float* q, int64_t ldq, float* v, int64_t ldv, float* a, int64_t lda,
                      float* tau, int64_t ldt, uint32_t flags) {
  return guarded([&] {
    RBD_CHECK_ARG(ctx && model, "null argument");
    RBD_CHECK_ARG(model->ctx == ctx, "model belongs to another context");
    RBD_CHECK_ARG(B >= 0, "negative batch");
    checkVec(q, ldq, B, "q");
    checkVec(v, ldv, B, "v");
    checkVec(tau, ldt, B, "tau");
    if (a) checkVec(a, lda, B, "a");
    if (B == 0) return;
    setDev(ctx);
    hipStream_t st = ctx->stream;
    uint64_t pos = 0;
    const double pi = 3.141592653589793;
    for (int j = 1; j < model->njoints; ++j) {
      const int iq = model->iq[j];
      if (model->jtype[j] == RBD_JOINT_FREEFLYER) {
        hipLaunchKernelGGL(sample::k_uniform, dim3(blocksFor(3 * B)), dim3(256), 0, st, seed, pos, 3, B, -base_range,
                           base_range, q + (int64_t)iq * ldq, ldq);
        pos += 3 * (uint64_t)B;
        hipLaunchKernelGGL(sample::k_quaternion, dim3(blocksFor(B)), dim3(256), 0, st, seed, pos, B,
                           q + (int64_t)(iq + 3) * ldq, ldq);
        pos += 3 * (uint64_t)B;
      } else {
        hipLaunchKernelGGL(sample::k_uniform, dim3(blocksFor(B)), dim3(256), 0, st, seed, pos, 1, B, -pi, pi,
                           q + (int64_t)iq * ldq, ldq);
        pos += (uint64_t)B;
      }
    }
    const int nv = model->nv;
    const uint64_t pairs = ((uint64_t)nv * B + 1) / 2;
    float* outs[3] = {v, a, tau};
    const int64_t lds[3] = {ldv, lda, ldt};
    const double sig[3] = {1.0, 1.0, tau_sigma};
    for (int k = 0; k < 3; ++k) {
      if (outs[k])
        hipLaunchKernelGGL(sample::k_normal, dim3(blocksFor((int64_t)nv * B)), dim3(256), 0, st, seed, pos, nv, B,
                           sig[k], outs[k], lds[k]);
      pos += 2 * pairs;
    }
    HIP_CHECK(hipGetLastError());
    if (!(flags & RBD_ASYNC)) HIP_CHECK(hipStreamSynchronize(st));
  });
}

}  // extern "C"
