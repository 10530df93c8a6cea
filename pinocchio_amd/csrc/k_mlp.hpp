// k_mlp.hpp -- kernels of the forward-dynamics surrogate: a multilayer perceptron
// (features -> H -> H -> H -> nv, SiLU) trained and evaluated with BF16 matrix-core
// products and FP32 accumulation.
//
// Activations are sample-major (row b holds one configuration's vector, stride ld);
// weights are [out, in] row-major like a linear layer (y = W x + b).  Every product is
// C(m, n) = sum_k A(m, k) B(n, k) over LDS tiles fed to v_mfma_f32_16x16x32_bf16; the
// operands may be k-contiguous (forward: activations x weights) or m/n-contiguous
// (backward: weight gradients reduce over the batch, input gradients over the outputs),
// in which case the tile is transposed on its way into LDS.  Epilogues fuse the bias,
// the activation, its derivative, and split-K partial sums.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels_common.hpp"

namespace rbd {
namespace mlp {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kBK = 32;       // k extent of one LDS stage (= one MFMA's k)
constexpr int kStage = kBK;   // shapes are padded to whole stages
constexpr int kLdk = kBK + 8; // LDS row pitch in bf16 (80 B: rows stay 16-B aligned)
constexpr int kGemmThreads = 256;

enum Epi : int {
  EPI_SILU = 0,       // Cb = bf16(silu(acc + bias)), optionally Zout = acc + bias (fp32)
  EPI_LINEAR = 1,     // Cf = acc + bias (fp32)
  EPI_SILU_GRAD = 2,  // Cb = bf16(acc * silu'(Z))
  EPI_PARTIAL = 3     // Cf[z] = acc (fp32 split-K partial, z = blockIdx.z)
};

struct GemmArgs {
  const bf16* A;
  int64_t lda;  // A(m, k) = AK ? A[m * lda + k] : A[k * lda + m]
  const bf16* B;
  int64_t ldb;  // B(n, k) = BK ? B[n * ldb + k] : B[k * ldb + n]
  int M, N, K;  // multiples of the tile extents and of kBK
  int kChunk;   // k range of one blockIdx.z (split-K); multiple of kBK
  const float* bias;
  const float* Z;  // EPI_SILU_GRAD: pre-activation, [M, ldz]
  float* Zout;     // EPI_SILU: pre-activation out (may be null)
  int64_t ldz;
  bf16* Cb;
  float* Cf;
  int64_t ldc;
  int64_t zstride;  // EPI_PARTIAL: distance between the partials of consecutive splits
};

// logistic function with the hardware reciprocal (1 ulp; an IEEE division costs about ten
// more instructions per activation)
__device__ __forceinline__ float sigmoidf(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }

// Stage ROWS x kBK operand values (row r, k) into s[r * kLdk + k].
template<int ROWS, bool KCONTIG>
__device__ __forceinline__ void loadTile(bf16* __restrict__ s, const bf16* __restrict__ g, int64_t ld, int r0, int k0,
                                         int tid) {
  if constexpr (KCONTIG) {
    // 16-byte chunks along k: 4 per row
    for (int c = tid; c < ROWS * (kBK / 8); c += kGemmThreads) {
      const int r = c / (kBK / 8), kc = (c % (kBK / 8)) * 8;
      *reinterpret_cast<uint4*>(s + r * kLdk + kc) =
          *reinterpret_cast<const uint4*>(g + (int64_t)(r0 + r) * ld + k0 + kc);
    }
  } else {
    // 16-byte chunks along r, transposed into LDS
    for (int c = tid; c < kBK * (ROWS / 8); c += kGemmThreads) {
      const int k = c / (ROWS / 8), rc = (c % (ROWS / 8)) * 8;
      const uint4 w = *reinterpret_cast<const uint4*>(g + (int64_t)(k0 + k) * ld + r0 + rc);
      const bf16* e = reinterpret_cast<const bf16*>(&w);
#pragma unroll
      for (int j = 0; j < 8; ++j) s[(rc + j) * kLdk + k] = e[j];
    }
  }
}

// C tile of BM x BN per workgroup; four waves in a 2 x 2 grid, each wave BM/2 x BN/2 as
// (BM/32) x (BN/32) MFMA blocks of 16 x 16.
template<int BM, int BN, bool AK, bool BK, int EPI>
__global__ __launch_bounds__(kGemmThreads) void k_gemm(GemmArgs g) {
  constexpr int WM = BM / 2, WN = BN / 2, TM = WM / 16, TN = WN / 16;
  static_assert(TM >= 1 && TN >= 1, "tile too small");
  __shared__ __attribute__((aligned(16))) bf16 sA[BM * kLdk];
  __shared__ __attribute__((aligned(16))) bf16 sB[BN * kLdk];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int wm = wave >> 1, wn = wave & 1;
  const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
  const int kBeg = blockIdx.z * g.kChunk;
  const int kEnd = min(g.K, kBeg + g.kChunk);
  f32x4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

  const int lr = lane & 15, lk = (lane >> 4) * 8;
  for (int k0 = kBeg; k0 < kEnd; k0 += kBK) {
    loadTile<BM, AK>(sA, g.A, g.lda, m0, k0, tid);
    loadTile<BN, BK>(sB, g.B, g.ldb, n0, k0, tid);
    __syncthreads();
    bf16x8 af[TM], bfr[TN];
#pragma unroll
    for (int i = 0; i < TM; ++i) af[i] = *reinterpret_cast<const bf16x8*>(sA + (wm * WM + i * 16 + lr) * kLdk + lk);
#pragma unroll
    for (int j = 0; j < TN; ++j) bfr[j] = *reinterpret_cast<const bf16x8*>(sB + (wn * WN + j * 16 + lr) * kLdk + lk);
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }

  // lane holds C(row = 4 * (lane / 16) + e, col = lane % 16) of each 16 x 16 block
#pragma unroll
  for (int i = 0; i < TM; ++i) {
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int col = n0 + wn * WN + j * 16 + lr;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int row = m0 + wm * WM + i * 16 + (lane >> 4) * 4 + e;
        const float a = acc[i][j][e];
        if constexpr (EPI == EPI_SILU) {
          const float z = a + g.bias[col];
          if (g.Zout) g.Zout[(int64_t)row * g.ldz + col] = z;
          g.Cb[(int64_t)row * g.ldc + col] = (bf16)(z * sigmoidf(z));
        } else if constexpr (EPI == EPI_LINEAR) {
          g.Cf[(int64_t)row * g.ldc + col] = a + g.bias[col];
        } else if constexpr (EPI == EPI_SILU_GRAD) {
          const float z = g.Z[(int64_t)row * g.ldz + col];
          const float s = sigmoidf(z);
          g.Cb[(int64_t)row * g.ldc + col] = (bf16)(a * s * (1.0f + z * (1.0f - s)));
        } else {
          g.Cf[(int64_t)blockIdx.z * g.zstride + (int64_t)row * g.ldc + col] = a;
        }
      }
    }
  }
}

// ---- forward product: both operands k-contiguous ---------------------------------------
// C(m, n) = sum_k A[m * lda + k] B[n * ldb + k] (activations x weights of a linear layer).
// A 256 x 128 tile per workgroup; four waves in a 2 x 2 grid, each 128 x 64 as 4 x 2
// blocks of v_mfma_f32_32x32x16_bf16 (128 accumulators per lane), so every operand value a
// wave reads from LDS feeds 64 or 128 products.  Stages of 64 k: while the matrix cores
// work on stage s from LDS, the loads of stage s + 1 are in flight into registers (32
// products per wave to cover their latency), parked into LDS after the stage's barrier.
// Launch: (M / 256) * (N / 128) workgroups on a one-dimensional grid; K a multiple of 64.
constexpr int kNtBM = 256, kNtBN = 128, kNtBK = 64, kNtLdk = kNtBK + 8;
typedef float f32x16 __attribute__((ext_vector_type(16)));

template<int EPI>
__global__ __launch_bounds__(kGemmThreads) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_gemm_nt(GemmArgs g) {
  constexpr int BM = kNtBM, BN = kNtBN, WM = BM / 2, WN = BN / 2, TM = WM / 32;
  __shared__ __attribute__((aligned(16))) bf16 sA[BM * kNtLdk];
  __shared__ __attribute__((aligned(16))) bf16 sB[BN * kNtLdk];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int wm = wave >> 1, wn = wave & 1;
  // One-dimensional grid over the tiles, column tile fastest.  Consecutive workgroups go
  // to different XCDs (round robin over 8), each with its own L2: remap so that every XCD
  // owns a contiguous run of tiles, where the column tiles of one row tile run side by
  // side and read their A rows from HBM once.
  const int ncol = g.N / BN, ntiles = gridDim.x;
  int t = blockIdx.x;
  if (ntiles % 8 == 0) t = (t % 8) * (ntiles / 8) + t / 8;
  const int m0 = (t / ncol) * BM, n0 = (t % ncol) * BN;
  const int nst = g.K / kNtBK;
  // thread tid stages the 16-byte chunks of rows tid / 8 + 32 i, k offset (tid % 8) * 8
  const bf16* gA = g.A + (int64_t)(m0 + (tid >> 3)) * g.lda + (tid & 7) * 8;
  const bf16* gB = g.B + (int64_t)(n0 + (tid >> 3)) * g.ldb + (tid & 7) * 8;
  const int64_t stA = 32 * g.lda, stB = 32 * g.ldb;
  const int so = (tid >> 3) * kNtLdk + (tid & 7) * 8, ss = 32 * kNtLdk;
  uint4 ra0, ra1, ra2, ra3, ra4, ra5, ra6, ra7, rb0, rb1, rb2, rb3;
#define RBD_NT_FETCH(k0)                                                                                      \
  ra0 = *reinterpret_cast<const uint4*>(gA + (k0));                                                           \
  ra1 = *reinterpret_cast<const uint4*>(gA + stA + (k0));                                                     \
  ra2 = *reinterpret_cast<const uint4*>(gA + 2 * stA + (k0));                                                 \
  ra3 = *reinterpret_cast<const uint4*>(gA + 3 * stA + (k0));                                                 \
  ra4 = *reinterpret_cast<const uint4*>(gA + 4 * stA + (k0));                                                 \
  ra5 = *reinterpret_cast<const uint4*>(gA + 5 * stA + (k0));                                                 \
  ra6 = *reinterpret_cast<const uint4*>(gA + 6 * stA + (k0));                                                 \
  ra7 = *reinterpret_cast<const uint4*>(gA + 7 * stA + (k0));                                                 \
  rb0 = *reinterpret_cast<const uint4*>(gB + (k0));                                                           \
  rb1 = *reinterpret_cast<const uint4*>(gB + stB + (k0));                                                     \
  rb2 = *reinterpret_cast<const uint4*>(gB + 2 * stB + (k0));                                                 \
  rb3 = *reinterpret_cast<const uint4*>(gB + 3 * stB + (k0));
#define RBD_NT_PARK                                                                                           \
  *reinterpret_cast<uint4*>(&sA[so]) = ra0;                                                                   \
  *reinterpret_cast<uint4*>(&sA[so + ss]) = ra1;                                                              \
  *reinterpret_cast<uint4*>(&sA[so + 2 * ss]) = ra2;                                                          \
  *reinterpret_cast<uint4*>(&sA[so + 3 * ss]) = ra3;                                                          \
  *reinterpret_cast<uint4*>(&sA[so + 4 * ss]) = ra4;                                                          \
  *reinterpret_cast<uint4*>(&sA[so + 5 * ss]) = ra5;                                                          \
  *reinterpret_cast<uint4*>(&sA[so + 6 * ss]) = ra6;                                                          \
  *reinterpret_cast<uint4*>(&sA[so + 7 * ss]) = ra7;                                                          \
  *reinterpret_cast<uint4*>(&sB[so]) = rb0;                                                                   \
  *reinterpret_cast<uint4*>(&sB[so + ss]) = rb1;                                                              \
  *reinterpret_cast<uint4*>(&sB[so + 2 * ss]) = rb2;                                                          \
  *reinterpret_cast<uint4*>(&sB[so + 3 * ss]) = rb3;
  f32x16 acc[TM][2];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int e = 0; e < 16; ++e) acc[i][0][e] = acc[i][1][e] = 0.f;

  // lane: row (or column) lane % 32 of a 32-block, k half (lane / 32) * 8 of a 16-k step
  const int lr = lane & 31, lk = (lane >> 5) * 8;
  RBD_NT_FETCH(0)
  for (int s = 0; s < nst; ++s) {
    RBD_NT_PARK
    __syncthreads();
    if (s + 1 < nst) {
      RBD_NT_FETCH((s + 1) * kNtBK)
    }
#pragma unroll
    for (int kk = 0; kk < kNtBK; kk += 16) {
      const bf16x8 b0 = *reinterpret_cast<const bf16x8*>(&sB[(wn * WN + lr) * kNtLdk + kk + lk]);
      const bf16x8 b1 = *reinterpret_cast<const bf16x8*>(&sB[(wn * WN + 32 + lr) * kNtLdk + kk + lk]);
#pragma unroll
      for (int i = 0; i < TM; ++i) {
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(&sA[(wm * WM + i * 32 + lr) * kNtLdk + kk + lk]);
        acc[i][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b0, acc[i][0], 0, 0, 0);
        acc[i][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b1, acc[i][1], 0, 0, 0);
      }
    }
    __syncthreads();
  }
#undef RBD_NT_FETCH
#undef RBD_NT_PARK

  // lane holds C(row = 8 (e / 4) + 4 (lane / 32) + e % 4, col = lane % 32) of each 32 x 32 block
  const float bias0 = g.bias[n0 + wn * WN + lr], bias1 = g.bias[n0 + wn * WN + 32 + lr];
  if constexpr (EPI == EPI_SILU) {
    // bf16 activations leave through LDS: each wave turns one 32 x 64 block row at a time
    // into 16-byte row segments (the LDS tiles are free after the last barrier)
    constexpr int P = WN + 8;  // staging pitch in bf16
    bf16* stg = sA + wave * 32 * P;
    const bool keepZ = g.Zout != nullptr;
#pragma unroll
    for (int i = 0; i < TM; ++i) {
#pragma unroll
      for (int j = 0; j < 2; ++j) {
#pragma unroll
        for (int e = 0; e < 16; ++e) {
          const int r = (e >> 2) * 8 + (lane >> 5) * 4 + (e & 3);
          const float z = acc[i][j][e] + (j ? bias1 : bias0);
          if (keepZ) g.Zout[(int64_t)(m0 + wm * WM + i * 32 + r) * g.ldz + n0 + wn * WN + j * 32 + lr] = z;
          stg[r * P + j * 32 + lr] = (bf16)(z * sigmoidf(z));
        }
      }
      __syncthreads();
#pragma unroll
      for (int c = 0; c < 4; ++c) {  // 32 rows x 8 segments of 8 bf16
        const int idx = c * 64 + lane, r = idx >> 3, sg = (idx & 7) * 8;
        *reinterpret_cast<uint4*>(g.Cb + (int64_t)(m0 + wm * WM + i * 32 + r) * g.ldc + n0 + wn * WN + sg) =
            *reinterpret_cast<const uint4*>(&stg[r * P + sg]);
      }
      __syncthreads();
    }
  } else {
    static_assert(EPI == EPI_LINEAR, "forward epilogues only");
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int e = 0; e < 16; ++e) {
          const int row = m0 + wm * WM + i * 32 + (e >> 2) * 8 + (lane >> 5) * 4 + (e & 3);
          g.Cf[(int64_t)row * g.ldc + n0 + wn * WN + j * 32 + lr] = acc[i][j][e] + (j ? bias1 : bias0);
        }
  }
}

// ---- features ------------------------------------------------------------------------
// Input vector of one configuration (feature order):
//   sin q_r, cos q_r for every revolute joint r       2 * nrev
//   free-flyer quaternion (qx, qy, qz, qw)            4   (if the root is a free flyer)
//   its rotation matrix, row major                    9   (if the root is a free flyer)
//   v, tau                                            2 * nv
// normalised as (x - mean) / std and zero-padded to the product's k extent.
constexpr int kMaxFeat = 256;
struct FeatMap {
  int nq, nv, nrev, hasFF, nfeat, ldx;  // ldx: padded row length
  int revq[96];                          // q index of each revolute joint
};

__device__ __forceinline__ void quatRot9(float x, float y, float z, float w, float* r) {
  // unit-quaternion rotation matrix (Eigen's Quaternion::toRotationMatrix formula)
  const float tx = 2 * x, ty = 2 * y, tz = 2 * z;
  const float twx = tx * w, twy = ty * w, twz = tz * w;
  const float txx = tx * x, txy = ty * x, txz = tz * x;
  const float tyy = ty * y, tyz = tz * y, tzz = tz * z;
  r[0] = 1 - (tyy + tzz); r[1] = txy - twz;       r[2] = txz + twy;
  r[3] = txy + twz;       r[4] = 1 - (txx + tzz); r[5] = tyz - twx;
  r[6] = txz - twy;       r[7] = tyz + twx;       r[8] = 1 - (txx + tyy);
}

// X[b, f] = bf16((x_f - mean_f) * invstd_f), zero for f >= nfeat.  One thread per
// configuration (its q, v, tau reads are coalesced across the wave); a 64-configuration
// tile of feature rows is assembled in LDS and written out as one contiguous block.
constexpr int kMaxLdx = 256;
__global__ __launch_bounds__(64) void k_features(FeatMap fm, int64_t B, const float* q, int64_t ldq, const float* v,
                                                 int64_t ldv, const float* tau, int64_t ldt, const float* mean,
                                                 const float* invstd, bf16* X, float* Xraw) {
  // row pitch ldx + 2: an odd number of 4-byte words, so the per-thread row writes of a
  // wave fall in distinct LDS banks (dynamic LDS of 64 (ldx + 2) bf16)
  extern __shared__ __attribute__((aligned(16))) bf16 tile[];
  const int t = threadIdx.x;
  const int64_t b0 = (int64_t)blockIdx.x * 64, b = b0 + t;
  const bool live = b < B;
  const int ldx = fm.ldx, pitch = ldx + 2;
  bf16* row = tile + t * pitch;
  auto put = [&](int f, float raw) {
    if (Xraw) Xraw[b * fm.nfeat + f] = raw;
    row[f] = (bf16)((raw - mean[f]) * invstd[f]);
  };
  if (live) {
    int f = 0;
    for (int r = 0; r < fm.nrev; ++r) {
      float sn, cs;
      sincosT(q[fm.revq[r] * ldq + b], &sn, &cs);
      put(r, sn);
      put(fm.nrev + r, cs);
    }
    f = 2 * fm.nrev;
    if (fm.hasFF) {
      const float qx = q[3 * ldq + b], qy = q[4 * ldq + b], qz = q[5 * ldq + b], qw = q[6 * ldq + b];
      float r9[9];
      quatRot9(qx, qy, qz, qw, r9);
      put(f, qx);
      put(f + 1, qy);
      put(f + 2, qz);
      put(f + 3, qw);
#pragma unroll
      for (int k = 0; k < 9; ++k) put(f + 4 + k, r9[k]);
      f += 13;
    }
    for (int k = 0; k < fm.nv; ++k) put(f + k, v[k * ldv + b]);
    f += fm.nv;
    for (int k = 0; k < fm.nv; ++k) put(f + k, tau[k * ldt + b]);
    f += fm.nv;
    for (; f < ldx; ++f) row[f] = (bf16)0.f;
  }
  __syncthreads();
  // the tile's rows are consecutive in X: copy it out as 4-byte pairs, coalesced
  const int64_t rows = B - b0 < 64 ? B - b0 : 64;
  const int half = ldx / 2;
  uint32_t* dst = reinterpret_cast<uint32_t*>(X + b0 * ldx);
  const uint32_t* src = reinterpret_cast<const uint32_t*>(tile);
  for (int c = t; c < rows * half; c += 64) {
    const int r = c / half, k = c - r * half;
    dst[c] = src[r * (half + 1) + k];
  }
}

// ---- fused inference --------------------------------------------------------------------
// The whole network for R configurations per workgroup (inferRows below): input features,
// every hidden layer and the output layer, with the activations resident in LDS (row per configuration,
// bf16, pitch max(kin, H) + 8) and the weights streamed from L2 straight into the matrix
// cores' operand registers -- no activation leaves the chip.  Products are computed
// transposed, C(n, b) = sum_k W(n, k) X(b, k) (v_mfma_f32_32x32x16_bf16, A = 32 weight
// rows, B = 32 configurations), so a lane holds four consecutive outputs of one
// configuration: the SiLU epilogue writes them back to LDS as one 8-byte store, and the
// output layer's stores run along b, coalesced in the batch-fastest ddq.  Eight waves; in
// a hidden layer wave w owns 64 outputs (column group w % (H / 64)) of R / WR
// configurations (WR = 8 / (H / 64) row groups).  k runs in 16-steps from 0 to K, the
// bias and SiLU are those of the layer-by-layer kernels.  LDS: the activation rows
// (R x pitch bf16), then the hidden layers' biases (L x H fp32).
constexpr int kInfThreads = 512, kInfMaxOut = 64, kInfMaxLayers = 9;
constexpr int kFragElems = 512;  // one 32 x 16 weight fragment
// configurations per workgroup: 64 for hidden widths 256 and 512 (two workgroups per CU,
// so that one's VALU phases -- features, SiLU epilogues, output -- overlap the other's
// products; four waves per SIMD, 128 registers), 128 for width 128 (32-row blocks)
constexpr int inferRows(int H) { return H >= 256 ? 64 : 128; }
// weight prefetch distance in 16-k steps: 64 k with 256 registers, 32 k with 128
template<int R> constexpr int kInfStageSteps = R == 128 ? 4 : 2;

// Fragment-major copy of a layer's weights W[N][K] (N % 32 == 0, K % 16 == 0): fragment
// (n / 32, k / 16), fragments of one 32-row block consecutive along k, holds its 512 values
// in the order the matrix cores take them: lane (n % 32) + 32 ((k % 16) / 8), then k % 8.
__global__ void k_pack_fragments(const bf16* __restrict__ W, bf16* __restrict__ F, int N, int K) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)N * K) return;
  const int n = (int)(i / K), k = (int)(i - (int64_t)n * K);
  const int64_t frag = (int64_t)(n / 32) * (K / 16) + k / 16;
  F[frag * kFragElems + ((n % 32) + 32 * ((k % 16) / 8)) * 8 + (k % 8)] = W[i];
}

struct InferArgs {
  FeatMap fm;
  int64_t B;
  const float* q;
  int64_t ldq;
  const float* v;
  int64_t ldv;
  const float* tau;
  int64_t ldt;
  const float *xmean, *xinv, *ymean, *ystd;
  const bf16* Wb;  // fragment-major bf16 weights (k_pack_fragments), layer l at offW[l]
  const float* P;  // fp32 parameter block (biases)
  int64_t offW[kInfMaxLayers], offB[kInfMaxLayers];
  int L, kin, kout, nout, pitch;
  float* ddq;
  int64_t ldd;
};

// features of the tile's R configurations into X[s * pitch + f] (bf16, normalised;
// padding features zero, configurations past the batch all zero).  The 512 threads form
// 512 / R groups of R (one configuration each), cycling through four roles -- sin / cos of
// the revolute joints, the free flyer's 13 and the zero padding, v, tau -- with a role's
// inputs split between its groups when there are eight.
template<int R>
__device__ __forceinline__ void inferFeatures(const InferArgs& a, bf16* __restrict__ X, int64_t b0) {
  constexpr int NREP = kInfThreads / R / 4;  // groups per role
  const FeatMap& fm = a.fm;
  const int tid = threadIdx.x, s = tid & (R - 1);
  const int part = __builtin_amdgcn_readfirstlane(tid / R);
  const int role = part & 3, rep = part >> 2;
  const int64_t b = b0 + s;
  const bool live = b < a.B;
  bf16* row = X + s * a.pitch;
  auto put = [&](int f, float raw) { row[f] = live ? (bf16)((raw - a.xmean[f]) * a.xinv[f]) : (bf16)0.f; };
  auto split = [&](int n, int& lo, int& hi) {
    lo = n * rep / NREP;
    hi = n * (rep + 1) / NREP;
  };
  const int fv = 2 * fm.nrev + (fm.hasFF ? 13 : 0);
  // inputs are read eight at a time, all eight requests in flight before the first is used
  if (role == 0) {
    int lo, hi;
    split(fm.nrev, lo, hi);
    for (int r0 = lo; r0 < hi; r0 += 8) {
      float x[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) x[j] = (live && r0 + j < hi) ? a.q[fm.revq[r0 + j] * a.ldq + b] : 0.f;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        if (r0 + j < hi) {
          float sn, cs;
          sincosT(x[j], &sn, &cs);
          put(r0 + j, sn);
          put(fm.nrev + r0 + j, cs);
        }
      }
    }
  } else if (role == 1) {
    if (fm.hasFF && rep == 0) {
      float qx = 0.f, qy = 0.f, qz = 0.f, qw = 1.f;
      if (live) {
        qx = a.q[3 * a.ldq + b];
        qy = a.q[4 * a.ldq + b];
        qz = a.q[5 * a.ldq + b];
        qw = a.q[6 * a.ldq + b];
      }
      float r9[9];
      quatRot9(qx, qy, qz, qw, r9);
      const int f = 2 * fm.nrev;
      put(f, qx);
      put(f + 1, qy);
      put(f + 2, qz);
      put(f + 3, qw);
#pragma unroll
      for (int k = 0; k < 9; ++k) put(f + 4 + k, r9[k]);
    }
    if (rep == NREP - 1)
      for (int f = fm.nfeat; f < a.kin; ++f) row[f] = (bf16)0.f;
  } else {
    const float* src = role == 2 ? a.v : a.tau;
    const int64_t ld = role == 2 ? a.ldv : a.ldt;
    const int f0 = role == 2 ? fv : fv + fm.nv;
    int lo, hi;
    split(fm.nv, lo, hi);
    for (int k0 = lo; k0 < hi; k0 += 8) {
      float x[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) x[j] = (live && k0 + j < hi) ? src[(k0 + j) * ld + b] : 0.f;
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if (k0 + j < hi) put(f0 + k0 + j, x[j]);
    }
  }
}

// acc[j][r] += W(n0 + 32 j + i, k) X(r0 + 32 r + c, k) over k in [0, K): A operand from
// the fragment-major weights in global memory (one load = one contiguous 1 KB fragment),
// B operand from the LDS activations.  K is a compile-time constant so that the k loop
// unrolls completely: the weights of the next 64-k stage are requested before the current
// stage's products (nothing loop-carried for the compiler to sink back to the point of use).
template<int RB, int K, int SK>
__device__ __forceinline__ void inferProduct(f32x16 (&acc)[2][RB], const bf16* __restrict__ X, int pitch,
                                             const bf16* __restrict__ W, int n0, int r0, int lane) {
  static_assert(K % (16 * SK) == 0, "whole stages of SK k-steps");
  constexpr int NST = K / (16 * SK), KS = K / 16;
  const int lr = lane & 31, lk = (lane >> 5) * 8;
  const bf16* w0 = W + (int64_t)(n0 / 32) * KS * kFragElems + lane * 8;
  const bf16* w1 = w0 + (int64_t)KS * kFragElems;
  const bf16* x = X + (r0 + lr) * pitch + lk;
  bf16x8 wr[NST][SK][2];
#pragma unroll
  for (int kk = 0; kk < SK; ++kk) {
    wr[0][kk][0] = *reinterpret_cast<const bf16x8*>(w0 + kk * kFragElems);
    wr[0][kk][1] = *reinterpret_cast<const bf16x8*>(w1 + kk * kFragElems);
  }
#pragma unroll
  for (int st = 0; st < NST; ++st) {
    if (st + 1 < NST) {
#pragma unroll
      for (int kk = 0; kk < SK; ++kk) {
        wr[st + 1][kk][0] = *reinterpret_cast<const bf16x8*>(w0 + ((st + 1) * SK + kk) * kFragElems);
        wr[st + 1][kk][1] = *reinterpret_cast<const bf16x8*>(w1 + ((st + 1) * SK + kk) * kFragElems);
      }
    }
    // keep those requests ahead of this stage's products (the scheduler would otherwise
    // move them next to their use and expose the L2 latency at every step)
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int kk = 0; kk < SK; ++kk) {
      const int k = (st * SK + kk) * 16;
      bf16x8 xb[RB];
#pragma unroll
      for (int r = 0; r < RB; ++r) xb[r] = *reinterpret_cast<const bf16x8*>(x + r * 32 * pitch + k);
#pragma unroll
      for (int r = 0; r < RB; ++r) {
        acc[0][r] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wr[st][kk][0], xb[r], acc[0][r], 0, 0, 0);
        acc[1][r] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wr[st][kk][1], xb[r], acc[1][r], 0, 0, 0);
      }
    }
  }
}

// one hidden layer: products, then (after every wave has read the input) bias + SiLU back
// into the activation rows
template<int H, int K, int R>
__device__ __forceinline__ void inferLayer(bf16* __restrict__ X, int pitch, const bf16* __restrict__ W,
                                           const float* __restrict__ bias, int lane, int wave) {
  constexpr int WC = H / 64, WR = 8 / WC, RB = R / WR / 32;
  static_assert(RB >= 1, "32-row blocks");
  const int n0 = (wave % WC) * 64, r0 = (wave / WC) * (R / WR);
  const int lr = lane & 31, lh = (lane >> 5) * 4;
  f32x16 acc[2][RB];
#pragma unroll
  for (int j = 0; j < 2; ++j)
#pragma unroll
    for (int r = 0; r < RB; ++r)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[j][r][e] = 0.f;
  inferProduct<RB, K, kInfStageSteps<R>>(acc, X, pitch, W, n0, r0, lane);
  __syncthreads();  // every wave has read the layer's input
#pragma unroll
  for (int j = 0; j < 2; ++j) {
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const int n = n0 + j * 32 + g * 8 + lh;  // this lane's outputs n .. n + 3
      const float4 bv = *reinterpret_cast<const float4*>(bias + n);
#pragma unroll
      for (int r = 0; r < RB; ++r) {
        bf16 o[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const float z = acc[j][r][4 * g + t] + (t == 0 ? bv.x : t == 1 ? bv.y : t == 2 ? bv.z : bv.w);
          o[t] = (bf16)(z * sigmoidf(z));
        }
        *reinterpret_cast<uint2*>(X + (r0 + r * 32 + lr) * pitch + n) = *reinterpret_cast<const uint2*>(o);
      }
    }
  }
  __syncthreads();
}

// output layer: block (nb, rb) of 32 outputs x 32 configurations per wave (nb < kout / 32),
// then ddq = (acc + bias) * std + mean stored along b.  The block's weights are requested
// eight fragments at a time ahead of their products, and the k steps feed four
// independent accumulator chains (summed at the end): a few MFMA latencies, not H / 16.
template<int H, int R>
__device__ __forceinline__ void inferOutput(const InferArgs& a, const bf16* __restrict__ X, int64_t b0, int lane,
                                            int wave) {
  constexpr int NS = H / 16, NRB = R / 32;
  const int nb = wave / NRB, rb = wave % NRB;
  if (nb * 32 >= a.kout) return;
  const int lr = lane & 31, lh = (lane >> 5) * 4;
  const bf16* W = a.Wb + a.offW[a.L] + (int64_t)nb * NS * kFragElems + lane * 8;
  const bf16* x = X + (rb * 32 + lr) * a.pitch + (lane >> 5) * 8;
  const float* bias = a.P + a.offB[a.L];
  float cb[16], cs[16], cm[16];  // bias, std and mean of this lane's outputs (n < kout: in range)
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    const int n = nb * 32 + (e >> 2) * 8 + lh + (e & 3);
    cb[e] = bias[n];
    cs[e] = a.ystd[n];
    cm[e] = a.ymean[n];
  }
  f32x16 acc[4];
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int e = 0; e < 16; ++e) acc[c][e] = 0.f;
#pragma unroll
  for (int k0 = 0; k0 < NS; k0 += 8) {  // eight weight fragments in flight at a time
    bf16x8 wf[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) wf[k] = *reinterpret_cast<const bf16x8*>(W + (k0 + k) * kFragElems);
#pragma unroll
    for (int k = 0; k < 8; ++k)
      acc[k & 3] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wf[k], *reinterpret_cast<const bf16x8*>(x + (k0 + k) * 16),
                                                           acc[k & 3], 0, 0, 0);
  }
  const int64_t b = b0 + rb * 32 + lr;
  if (b < a.B) {
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const int n = nb * 32 + (e >> 2) * 8 + lh + (e & 3);
      if (n < a.nout) {
        const float o = ((acc[0][e] + acc[1][e]) + (acc[2][e] + acc[3][e])) + cb[e];
        a.ddq[n * a.ldd + b] = o * cs[e] + cm[e];
      }
    }
  }
}

template<int H, int R>
__global__ __launch_bounds__(kInfThreads) __attribute__((amdgpu_waves_per_eu(256 / R, 256 / R))) void k_mlp_infer(
    InferArgs a) {
  static_assert((H == 128 && R == 128) || ((H == 256 || H == 512) && R == 64), "fused kernel shapes");
  extern __shared__ __attribute__((aligned(16))) bf16 X[];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int64_t b0 = (int64_t)blockIdx.x * R;
  const int pitch = a.pitch;
  inferFeatures<R>(a, X, b0);
  // hidden-layer biases into LDS behind the activations (their reads inside the products
  // would otherwise wait behind the in-flight weight loads)
  float* sbias = reinterpret_cast<float*>(X + R * pitch);
  for (int i = tid; i < a.L * H; i += kInfThreads) sbias[i] = a.P[a.offB[i / H] + i % H];
  __syncthreads();

  // first layer over the padded input width, then H -> H
  const bf16* W0 = a.Wb + a.offW[0];
  switch (a.kin) {
    case 64: inferLayer<H, 64, R>(X, pitch, W0, sbias, lane, wave); break;
    case 128: inferLayer<H, 128, R>(X, pitch, W0, sbias, lane, wave); break;
    case 192: inferLayer<H, 192, R>(X, pitch, W0, sbias, lane, wave); break;
    default: inferLayer<H, 256, R>(X, pitch, W0, sbias, lane, wave); break;
  }
  for (int l = 1; l < a.L; ++l) inferLayer<H, H, R>(X, pitch, a.Wb + a.offW[l], sbias + l * H, lane, wave);

  inferOutput<H, R>(a, X, b0, lane, wave);
}

// zero rows [B, Bpad) of a bf16 matrix (padding configurations of the last tile)
__global__ void k_zero_rows(bf16* X, int64_t ld, int64_t B, int64_t Bpad) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < (Bpad - B) * ld) X[B * ld + e] = (bf16)0.f;
}

// normalised target Y[b, n] = (ddq[n, b] - mean_n) * invstd_n  (fp32, zero padding)
__global__ void k_targets(int nv, int ldy, int64_t B, int64_t Bpad, const float* ddq, int64_t ldd, const float* mean,
                          const float* invstd, float* Y) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t b = e / ldy;
  const int n = (int)(e - b * ldy);
  if (b >= Bpad) return;
  Y[e] = (n < nv && b < B) ? (ddq[n * ldd + b] - mean[n]) * invstd[n] : 0.f;
}

// dOut = 2 (out - Y) / (B nv) over real rows and outputs (bf16), loss partial sums
__global__ void k_loss_grad(int nv, int ldy, int64_t B, int64_t Bpad, const float* out, const float* Y, bf16* dOut,
                            float* lossAcc) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t b = e / ldy;
  const int n = (int)(e - b * ldy);
  float d = 0.f;
  if (b < Bpad) {
    if (n < nv && b < B) d = out[e] - Y[e];
    dOut[e] = (bf16)(d * (2.0f / ((float)B * (float)nv)));
  }
  // wave reduction of d^2, one atomic per wave (monitoring only)
  float s = d * d;
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if ((threadIdx.x & 63) == 0 && lossAcc) atomicAdd(lossAcc, s / ((float)B * (float)nv));
}

// column sums of a bf16 matrix over row splits: P[s * pstride + n] = sum_{b in split s} D[b, n]
__global__ void k_colsum(const bf16* D, int64_t ld, int N, int64_t rowsPerSplit, float* P, int64_t pstride) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  const int64_t r0 = (int64_t)blockIdx.y * rowsPerSplit;
  float s = 0.f;
  for (int64_t r = r0; r < r0 + rowsPerSplit; ++r) s += (float)D[r * ld + n];
  P[(int64_t)blockIdx.y * pstride + n] = s;
}

// Adam on fp32 master parameters with gradient = sum of split partials; refreshes the
// bf16 copy used by the products.
__global__ void k_adam(int64_t n, float* P, float* m1, float* m2, const float* G, int splits, int64_t gstride, float lr,
                       float b1, float b2, float eps, float c1, float c2, bf16* Pb) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float g = 0.f;
  for (int s = 0; s < splits; ++s) g += G[s * gstride + i];
  const float a = b1 * m1[i] + (1.f - b1) * g;
  const float c = b2 * m2[i] + (1.f - b2) * g * g;
  m1[i] = a;
  m2[i] = c;
  const float p = P[i] - lr * (a * c1) / (sqrtf(c * c2) + eps);
  P[i] = p;
  if (Pb) Pb[i] = (bf16)p;
}

// gradient = sum of split partials (exported for checks)
__global__ void k_sum_partials(int64_t n, const float* G, int splits, int64_t gstride, float* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float g = 0.f;
  for (int s = 0; s < splits; ++s) g += G[s * gstride + i];
  out[i] = g;
}

// fp32 -> bf16 copy
__global__ void k_to_bf16(int64_t n, const float* x, bf16* y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (bf16)x[i];
}

// ddq[n, b] = out[b, n] * std_n + mean_n  (batch-fastest output)
__global__ void k_outputs(int nv, int ldo, int64_t B, const float* out, const float* mean, const float* std_,
                          float* ddq, int64_t ldd) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = e / B, b = e - n * B;
  if (n >= nv) return;
  ddq[n * ldd + b] = out[b * ldo + n] * std_[n] + mean[n];
}

// Residual gate: ||tau_chk - tau|| > tol ||tau|| (2-norms) sends configuration b to the
// exact path: its index and inputs are appended to the compact lists.
__global__ void k_gate(int nq, int nv, int64_t B, const float* tauChk, int64_t ldc, const float* tau, int64_t ldt,
                       const float* q, int64_t ldq, const float* v, int64_t ldv, float tol, unsigned int* count,
                       int64_t* idx, float* cq, float* cv, float* ct, int64_t ldcmp, unsigned char* accepted,
                       float* relOut) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  float num = 0.f, den = 0.f;
  for (int k = 0; k < nv; ++k) {
    const float t = tau[k * ldt + b];
    const float d = tauChk[k * ldc + b] - t;
    num += d * d;
    den += t * t;
  }
  const float rel = sqrtf(num) / fmaxf(sqrtf(den), 1e-30f);
  if (relOut) relOut[b] = rel;
  const bool ok = rel <= tol;
  if (accepted) accepted[b] = ok ? 1 : 0;
  if (!ok) {
    const unsigned int s = atomicAdd(count, 1u);
    idx[s] = b;
    for (int k = 0; k < nq; ++k) cq[k * ldcmp + s] = q[k * ldq + b];
    for (int k = 0; k < nv; ++k) {
      cv[k * ldcmp + s] = v[k * ldv + b];
      ct[k * ldcmp + s] = tau[k * ldt + b];
    }
  }
}

__global__ void k_scatter(int nv, int64_t n, const int64_t* idx, const float* src, int64_t lds, float* dst,
                          int64_t ldd) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t k = e / n, j = e - k * n;
  if (k >= nv) return;
  dst[k * ldd + idx[j]] = src[k * lds + j];
}

// fill kernel (used instead of a memset so that call sequences stay capturable)
__global__ void k_fill_u32(unsigned int* p, int64_t n, unsigned int v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_f32(float* p, int64_t n, float v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

}  // namespace mlp
}  // namespace rbd
