// BF16 MFMA multilayer perceptron for gfx950: forward, backward and Adam, all on the device.
//
// Network: x[n_in] -> (Linear, SiLU) x n_hidden_layers -> Linear -> y[n_out].  Weights W_l [n_out_l x n_in_l]
// are held in FP32 (master copy, Adam moments) and mirrored to BF16 in two layouts: W_l [n_out x n_in] for the
// forward product and W_l^T [n_in x n_out] for the input-gradient product.  Activations are BF16, every product
// accumulates in FP32 on the matrix cores (v_mfma_f32_32x32x16_bf16).
//
// All products have the one form C[M x N] = A[M x K] * B[N x K]^T with A and B K-contiguous:
//   forward        Z_l   = X_l  * W_l^T         (A = X_l,   B = W_l)
//   input grad     dX_l  = dZ_l * W_l           (A = dZ_l,  B = W_l^T)
//   weight grad    dW_l  = dZ_l^T * X_l         (A = dZ_l^T, B = X_l^T, contraction over the batch)
// so the forward and backward epilogues also store the transposed BF16 copies the weight-gradient product reads.
// Dimensions must be multiples of 128 (M, N) and 32 (K); callers pad features and outputs with zeros.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/cloudsc_surrogate.h"
#include "capi_util.h"
#include "mlp_internal.h"
#include "synth.h"

namespace mlp {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// Block tiles: TBM x TBN per 256-lane group (2 x 2 waves, each (TBM/2) x (TBN/2) in 32x32 MFMA tiles), K-step TBK.
// 128 x 128 for small products (training batches), 256 x 128 / 256 x 256 for large ones (inference), where the
// larger wave tile reads fewer LDS bytes per MFMA (1 KB / 0.75 KB / 0.5 KB).
constexpr int BM = 128, BN = 128, BK = 32;  // smallest tile: the granularity of every dimension
constexpr int MAX_SPLIT = 4;                // stretches of a split weight-gradient contraction

enum Epi : int {
  EPI_F32 = 0,      // C (fp32)
  EPI_HIDDEN = 1,   // z = acc + bias -> Z (fp32, kept for the backward pass); a = silu(z) -> A, A^T (bf16)
  EPI_HIDDEN_INF = 2,  // inference: a = silu(acc + bias) -> A (bf16) only
  EPI_OUT = 3,      // y = acc + bias -> C (fp32)
  EPI_DHIDDEN = 4,  // g = acc * silu'(Z) -> A, A^T (bf16): the gradient w.r.t. the previous pre-activation
};

struct EpiArgs {
  const float* bias;
  float* c;        // fp32 [M x N], leading dimension ldc
  int ldc;
  bf16* act;       // bf16 [M x N], leading dimension ldc
  bf16* actT;      // bf16 [N x M], leading dimension ldt
  int ldt;
  const float* z;  // fp32 [M x N] pre-activation (EPI_DHIDDEN), leading dimension ldc
  long zstride;    // split contraction (EPI_F32, grid z > 1): elements between the partial products of the splits
};

__device__ __forceinline__ float sigmoid(float z) { return 1.0f / (1.0f + __expf(-z)); }

template <int EPI, int TBM, int TBN, int TBK>
__global__ __launch_bounds__(256) void gemm_tn(const bf16* __restrict__ A, int lda, const bf16* __restrict__ B,
                                               int ldb, int K, EpiArgs e) {
  constexpr int FM = TBM / 64, FN = TBN / 64;        // 32x32 fragments per wave in M and N
  constexpr int CPR = TBK / 8, RPP = 256 / CPR;      // 16-byte chunks per row; rows per pass of 256 lanes
  constexpr int LA = TBM / RPP, LB = TBN / RPP;      // chunks per thread per K-step, A and B
  // LDS rows of TBK bf16 (no padding); chunk c of row r sits in slot c ^ ((r >> SH) & (CPR - 1)), SH = log2(16 / CPR):
  // the 16 rows one fragment read touches (one chunk each) and the rows one store pass writes then fall in 16
  // distinct 16-byte slots of each 256-byte bank window -- no bank conflict on either side
  constexpr int LDKT = TBK, SH = CPR == 4 ? 2 : (CPR == 8 ? 1 : 0);
  static_assert(CPR == 4 || CPR == 8, "K-step 32 or 64");
  __shared__ bf16 As[2][TBM * LDKT];
  __shared__ bf16 Bs[2][TBN * LDKT];
  auto slot = [](int r, int c) { return 8 * (c ^ ((r >> SH) & (CPR - 1))); };
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w >> 1, wn = w & 1;
  const int m0 = blockIdx.y * TBM, n0 = blockIdx.x * TBN;
  // split contraction: group z multiplies the z-th stretch of K columns into its own partial product
  A += (size_t)blockIdx.z * K;
  B += (size_t)blockIdx.z * K;
  if (EPI == EPI_F32) e.c += (size_t)blockIdx.z * e.zstride;
  // global -> LDS in chunks of 8 bf16; thread t moves rows r0, r0 + RPP, ...
  const int r0 = tid / CPR, kq = (tid % CPR) * 8;
  const bf16* ag = A + (size_t)(m0 + r0) * lda + kq;
  const bf16* bg = B + (size_t)(n0 + r0) * ldb + kq;
  bf16x8 ra[LA], rb[LB];
  auto gload = [&](int k0) {
#pragma unroll
    for (int i = 0; i < LA; ++i) ra[i] = *reinterpret_cast<const bf16x8*>(ag + (size_t)(RPP * i) * lda + k0);
#pragma unroll
    for (int i = 0; i < LB; ++i) rb[i] = *reinterpret_cast<const bf16x8*>(bg + (size_t)(RPP * i) * ldb + k0);
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int i = 0; i < LA; ++i) {
      const int r = r0 + RPP * i;
      *reinterpret_cast<bf16x8*>(&As[buf][r * LDKT + slot(r, kq / 8)]) = ra[i];
    }
#pragma unroll
    for (int i = 0; i < LB; ++i) {
      const int r = r0 + RPP * i;
      *reinterpret_cast<bf16x8*>(&Bs[buf][r * LDKT + slot(r, kq / 8)]) = rb[i];
    }
  };
  f32x16 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

  const int nk = K / TBK;
  gload(0);
  sstore(0);
  __syncthreads();
  // MFMA 32x32x16 operand layout: lane l supplies row (l & 31) of its 32-row block and k = 8*(l >> 5) .. +7
  const int fr = lane & 31, fk = 8 * (lane >> 5);
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < nk) gload((kt + 1) * TBK);
#pragma unroll
    for (int ks = 0; ks < TBK; ks += 16) {
      bf16x8 a[FM], b[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) {
        const int r = wm * (TBM / 2) + i * 32 + fr;
        a[i] = *reinterpret_cast<const bf16x8*>(&As[cur][r * LDKT + slot(r, (ks + fk) / 8)]);
      }
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        const int r = wn * (TBN / 2) + j * 32 + fr;
        b[j] = *reinterpret_cast<const bf16x8*>(&Bs[cur][r * LDKT + slot(r, (ks + fk) / 8)]);
      }
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    if (kt + 1 < nk) sstore(cur ^ 1);
    __syncthreads();
  }

  // Accumulator layout: register r of lane l holds row 8*(r>>2) + 4*(l>>5) + (r&3), column l & 31.
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int col = n0 + wn * (TBN / 2) + j * 32 + (lane & 31);
      const float bias = (EPI == EPI_HIDDEN || EPI == EPI_HIDDEN_INF || EPI == EPI_OUT) ? e.bias[col] : 0.0f;
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int row = m0 + wm * (TBM / 2) + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const size_t o = (size_t)row * e.ldc + col;
        const float v = acc[i][j][r];
        if (EPI == EPI_F32) {
          e.c[o] = v;
        } else if (EPI == EPI_OUT) {
          e.c[o] = v + bias;
        } else if (EPI == EPI_HIDDEN || EPI == EPI_HIDDEN_INF) {
          const float z = v + bias;
          const bf16 a = (bf16)(z * sigmoid(z));
          if (EPI == EPI_HIDDEN) {
            e.c[o] = z;
            e.actT[(size_t)col * e.ldt + row] = a;
          }
          e.act[o] = a;
        } else {  // EPI_DHIDDEN
          const float z = e.z[o];
          const float s = sigmoid(z);
          const bf16 g = (bf16)(v * (s * (1.0f + z * (1.0f - s))));
          e.act[o] = g;
          e.actT[(size_t)col * e.ldt + row] = g;
        }
      }
    }
}

// The large-product kernel: 256 x 256 block tile, K-step 64, NW waves as 2 x (NW/2), each computing 128 rows x
// (512/NW) columns in 32x32 fragments whose accumulators live in the AGPRs.  The two LDS buffers are filled straight
// from global memory (global_load_lds_dwordx4: no staging registers), the next tile while the current one
// multiplies.  One such load writes 1 KB of LDS contiguously, lane l at 16*l: 8 rows of 128 B.  Chunk c (16 B) of
// row r is stored in slot c ^ ((r >> 1) & 7) of its row: a fragment read (16 lanes = 16 consecutive rows, one
// chunk, 256 B across the 64 banks) then finds the 8 even rows in 8 distinct slots of one 128-byte half of the banks
// and the 8 odd rows likewise in the other half -- no bank conflict.  Each lane of a load fetches the global chunk
// that belongs in its LDS slot.
template <int EPI, int NW>
__global__ __launch_bounds__(64 * NW) void gemm_tn_dl(const bf16* __restrict__ A, int lda, const bf16* __restrict__ B,
                                                       int ldb, int K, EpiArgs e) {
  constexpr int TBM = 256, TBN = 256, TBK = 64, WN = NW / 2, WCOL = TBN / WN, FM = 4, FN = WCOL / 32;
  constexpr int TILE = TBM * TBK;                    // bf16 per operand tile (32 KB)
  constexpr int NLD = 32 / NW;                       // 1 KB loads per operand, thread and K-step
  __shared__ __attribute__((aligned(1024))) bf16 lds[2][2][TILE];  // [buffer][A|B]
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w / WN, wn = w % WN;
  // Tile mapping: the dispatcher hands consecutive workgroups to the 8 XCDs in turn, each with its own L2.  With the
  // M-tile count a multiple of 8, workgroup L goes to XCD L % 8 and takes, in that XCD's order, all N-tiles of one
  // M-tile before the next, so the A tile they share is fetched into one L2 only.
  int mt = blockIdx.y, nt = blockIdx.x;
  if ((gridDim.y & 7) == 0) {
    const int L = blockIdx.y * gridDim.x + blockIdx.x, slot = L >> 3;
    nt = slot % gridDim.x;
    mt = (slot / gridDim.x) * 8 + (L & 7);
  }
  const int m0 = mt * TBM, n0 = nt * TBN;
  // loads: wave w, instruction i fills row group g = NW i + w (rows 8g .. 8g+7) of each operand; g has the parity of
  // w, so (row >> 1) & 7 = (4 (w & 1) + (lrow >> 1)) & 7
  const int lrow = lane >> 3, lchunk = (lane & 7) ^ ((4 * (w & 1) + (lrow >> 1)) & 7);
  const bf16* ag = A + (size_t)(m0 + 8 * w + lrow) * lda + 8 * lchunk;
  const bf16* bg = B + (size_t)(n0 + 8 * w + lrow) * ldb + 8 * lchunk;
  const size_t astep = (size_t)(8 * NW) * lda, bstep = (size_t)(8 * NW) * ldb;
  auto gload = [&](int buf, int k0) {
    typedef bf16 __attribute__((address_space(3))) lbf16;
#pragma unroll
    for (int i = 0; i < NLD; ++i) {
      __builtin_amdgcn_global_load_lds((const void*)(ag + astep * i + k0),
                                       (lbf16*)&lds[buf][0][(NW * i + w) * 512], 16, 0, 0);
      __builtin_amdgcn_global_load_lds((const void*)(bg + bstep * i + k0),
                                       (lbf16*)&lds[buf][1][(NW * i + w) * 512], 16, 0, 0);
    }
  };
  f32x16 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

  const int nk = K / TBK;
  gload(0, 0);
  // fragment reads: lane supplies row (lane & 31) of a 32-row block and k = 8*(lane >> 5) .. +7 of a 16-wide step
  const int fr = lane & 31, fh = lane >> 5;
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    __builtin_amdgcn_s_waitcnt(0);  // this thread's loads of the current tile have landed
    __syncthreads();                // everyone's; and everyone is done reading the other buffer
    if (kt + 1 < nk) gload(cur ^ 1, (kt + 1) * TBK);
    const bf16* as = lds[cur][0];
    const bf16* bs = lds[cur][1];
#pragma unroll
    for (int ks = 0; ks < TBK / 16; ++ks) {
      const int c = 2 * ks + fh;  // logical 16-byte chunk of the row
      bf16x8 a[FM], b[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) {
        const int r = wm * 128 + i * 32 + fr;
        a[i] = *reinterpret_cast<const bf16x8*>(as + r * TBK + 8 * (c ^ ((r >> 1) & 7)));
      }
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        const int r = wn * WCOL + j * 32 + fr;
        b[j] = *reinterpret_cast<const bf16x8*>(bs + r * TBK + 8 * (c ^ ((r >> 1) & 7)));
      }
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(b[j], a[i], acc[i][j], 0, 0, 0);
    }
  }

  // The product is formed as C^T = B A^T, so register r of lane l holds C[m][n] with m = lane & 31 of the fragment's
  // rows and n = 8*(r>>2) + 4*(l>>5) + (r&3) of its columns: each lane owns 4 runs of 4 consecutive columns of one
  // row, stored as one 16-byte (fp32) or 8-byte (bf16) write each.
  typedef float f32x4 __attribute__((ext_vector_type(4)));
  typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int row = m0 + wm * 128 + i * 32 + (lane & 31);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int col = n0 + wn * WCOL + j * 32 + 8 * q + 4 * (lane >> 5);
        const size_t o = (size_t)row * e.ldc + col;
        f32x4 v = {acc[i][j][4 * q], acc[i][j][4 * q + 1], acc[i][j][4 * q + 2], acc[i][j][4 * q + 3]};
        if (EPI == EPI_HIDDEN || EPI == EPI_HIDDEN_INF || EPI == EPI_OUT)
          v = v + *reinterpret_cast<const f32x4*>(e.bias + col);
        if (EPI == EPI_F32 || EPI == EPI_OUT) {
          *reinterpret_cast<f32x4*>(e.c + o) = v;
        } else if (EPI == EPI_HIDDEN || EPI == EPI_HIDDEN_INF) {
          bf16x4 a;
#pragma unroll
          for (int t = 0; t < 4; ++t) a[t] = (bf16)(v[t] * sigmoid(v[t]));
          if (EPI == EPI_HIDDEN) {
            *reinterpret_cast<f32x4*>(e.c + o) = v;
#pragma unroll
            for (int t = 0; t < 4; ++t) e.actT[(size_t)(col + t) * e.ldt + row] = a[t];
          }
          *reinterpret_cast<bf16x4*>(e.act + o) = a;
        } else {  // EPI_DHIDDEN
          const f32x4 z = *reinterpret_cast<const f32x4*>(e.z + o);
          bf16x4 g;
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            const float sg = sigmoid(z[t]);
            g[t] = (bf16)(v[t] * (sg * (1.0f + z[t] * (1.0f - sg))));
          }
          *reinterpret_cast<bf16x4*>(e.act + o) = g;
#pragma unroll
          for (int t = 0; t < 4; ++t) e.actT[(size_t)(col + t) * e.ldt + row] = g[t];
        }
      }
    }
}

// Output-layer loss: L = sum_{m,n} wgt[n] (y - t)^2 / M.  dY = 2 wgt[n] (y - t) / M (bf16, both layouts);
// loss partial sums are accumulated into *loss (fp32 atomics).  A 64 x 64 tile per 256-lane group: rows of dY read
// and written along n, the transposed copy written along m through LDS, so that every store is coalesced.
__global__ __launch_bounds__(256) void mse_grad(const float* __restrict__ y, const float* __restrict__ t,
                                                const float* __restrict__ wgt, int M, int N, bf16* __restrict__ dy,
                                                bf16* __restrict__ dyT, float* __restrict__ loss) {
  __shared__ bf16 tile[64][66];
  const int n0 = blockIdx.x * 64, m0 = blockIdx.y * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  float l = 0.0f;
  for (int r = ty; r < 64; r += 4) {  // lanes along n
    const int m = m0 + r, n = n0 + tx;
    bf16 g = (bf16)0.0f;
    if (m < M && n < N) {
      const size_t idx = (size_t)m * N + n;
      const float d = y[idx] - t[idx];
      const float wd = wgt[n] * d;
      l += wd * d;
      g = (bf16)(2.0f * wd / (float)M);
      dy[idx] = g;
    }
    tile[r][tx] = g;
  }
  __syncthreads();
  for (int r = ty; r < 64; r += 4) {  // lanes along m
    const int n = n0 + r, m = m0 + tx;
    if (m < M && n < N) dyT[(size_t)n * M + m] = tile[tx][r];
  }
  // wave reduction, one atomic per wave
  for (int off = 32; off > 0; off >>= 1) l += __shfl_down(l, off, 64);
  if ((threadIdx.x & 63) == 0 && l != 0.0f) atomicAdd(loss, l / (float)M);
}

// Sum of the S partial products of a split contraction, in split order: out[i] = sum_z part[z * n + i].
__global__ void sum_splits(const float* __restrict__ part, size_t n, int S, float* __restrict__ out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float v = part[i];
  for (int z = 1; z < S; ++z) v += part[(size_t)z * n + i];
  out[i] = v;
}

// Bias gradient: db[n] = sum_m G^T[n][m] (fp32 sum of the bf16 gradient the weight product also sees).
__global__ __launch_bounds__(256) void row_sums(const bf16* __restrict__ gT, int M, float* __restrict__ db) {
  const int n = blockIdx.x;
  float s = 0.0f;
  for (int m = threadIdx.x; m < M; m += 256) s += (float)gT[(size_t)n * M + m];
  __shared__ float red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[n] = red[0];
}

// Adam on FP32 master weights; refreshes the BF16 copies W [rows x cols] and W^T [cols x rows] (if wT != null).
__global__ void adam_update(float* __restrict__ w, float* __restrict__ m1, float* __restrict__ m2,
                            const float* __restrict__ g, size_t n, int cols, float lr, float b1, float b2, float eps,
                            float bc1, float bc2, bf16* __restrict__ wb, bf16* __restrict__ wT, int rows) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.0f - b1) * gi;
  const float b = b2 * m2[i] + (1.0f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  const float wi = w[i] - lr * (a / bc1) / (sqrtf(b / bc2) + eps);
  w[i] = wi;
  if (wb) {
    const bf16 h = (bf16)wi;
    wb[i] = h;
    if (wT) {
      const size_t r = i / cols, c = i - r * cols;
      wT[c * rows + r] = h;
    }
  }
}

// FP32 -> BF16 copies of a weight matrix in both layouts.
__global__ void to_bf16(const float* __restrict__ w, size_t n, int cols, int rows, bf16* __restrict__ wb,
                        bf16* __restrict__ wT) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bf16 h = (bf16)w[i];
  wb[i] = h;
  const size_t r = i / cols, c = i - r * cols;
  wT[c * rows + r] = h;
}

__global__ void fill_f32(float* p, size_t n, float v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

}  // namespace mlp

// ---------------------------------------------------------------------------------------------------------
// Model object and C-ABI
// ---------------------------------------------------------------------------------------------------------
using mlp::bf16;
using cloudsc_capi::fail;


namespace {

int dev_alloc(cloudsc_mlp* m, size_t bytes, void** p) {
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) return fail(CLOUDSC_ENOMEM, "cloudsc_mlp: hipMalloc", e);
  m->allocs.push_back(*p);
  return CLOUDSC_OK;
}

// Tile choice: the largest tile that divides M and N and still gives at least one group per CU (256 CUs).
template <int EPI>
hipError_t gemm(const bf16* A, int lda, const bf16* B, int ldb, int M, int N, int K, const mlp::EpiArgs& e,
                hipStream_t s) {
  const long groups256 = (M % 256 == 0 && N % 256 == 0) ? (long)(M / 256) * (N / 256) : 0;
  const long groups256x128 = (M % 256 == 0) ? (long)(M / 256) * (N / 128) : 0;
  const bool k64 = K % 64 == 0;
  if (groups256 >= 256 && k64) {
    hipLaunchKernelGGL((mlp::gemm_tn_dl<EPI, 8>), dim3(N / 256, M / 256), dim3(512), 0, s, A, lda,
                       B, ldb, K, e);
  } else if (groups256x128 >= 256 && k64) {
    hipLaunchKernelGGL((mlp::gemm_tn<EPI, 256, 128, 64>), dim3(N / 128, M / 256), dim3(256), 0, s, A, lda, B, ldb, K,
                       e);
  } else {
    hipLaunchKernelGGL((mlp::gemm_tn<EPI, 128, 128, 32>), dim3(N / 128, M / 128), dim3(256), 0, s, A, lda, B, ldb, K,
                       e);
  }
  return hipGetLastError();
}

inline unsigned blocks(size_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

int check_batch(const cloudsc_mlp* m, int M) {
  if (M <= 0 || M > m->max_batch || M % mlp::BM)
    return fail(CLOUDSC_EDIMS, "cloudsc_mlp: batch must be a positive multiple of 128 and <= max_batch");
  return CLOUDSC_OK;
}

}  // namespace

namespace mlp {
// Forward pass; train = keep Z and the transposed activations for the backward pass.
hipError_t forward(cloudsc_mlp* m, const bf16* x, int M, float* y, bool train, hipStream_t s) {
  hipError_t e = hipSuccess;
  const bf16* in = x;
  int ldin = m->din[0];
  for (int l = 0; l < m->nl && e == hipSuccess; ++l) {
    mlp::EpiArgs ea{};
    ea.bias = m->b[l];
    if (l < m->nl - 1) {
      ea.c = m->z[l];
      ea.ldc = m->dout[l];
      ea.act = m->a[l];
      ea.actT = m->aT[l];
      ea.ldt = M;
      e = train ? gemm<mlp::EPI_HIDDEN>(in, ldin, m->wb[l], m->din[l], M, m->dout[l], m->din[l], ea, s)
                : gemm<mlp::EPI_HIDDEN_INF>(in, ldin, m->wb[l], m->din[l], M, m->dout[l], m->din[l], ea, s);
      in = m->a[l];
      ldin = m->dout[l];
    } else {
      ea.c = y;
      ea.ldc = m->dout[l];
      e = gemm<mlp::EPI_OUT>(in, ldin, m->wb[l], m->din[l], M, m->dout[l], m->din[l], ea, s);
    }
  }
  return e;
}
}  // namespace mlp

namespace {

// dW = G^T X over the batch (C[dout x din], contraction over M columns).  With few output tiles (one 128x128 tile
// per CU or less) the contraction is split in S stretches over the grid's z dimension, each into its own partial
// product in the workspace, which sum_splits then adds in split order (deterministic).
hipError_t weight_gradient(cloudsc_mlp* m, const bf16* gT, const bf16* xT, int dout, int din, int M,
                           const mlp::EpiArgs& ew, hipStream_t s) {
  const long tiles = (long)(dout / mlp::BM) * (din / mlp::BN);
  int S = 1;
  while (S < mlp::MAX_SPLIT && tiles * S < 256 && (M / (2 * S)) % mlp::BK == 0 && M / (2 * S) >= 512) S *= 2;
  if (S == 1 || dout % mlp::BM || din % mlp::BN || (size_t)S * dout * din > m->ws_elems)
    return gemm<mlp::EPI_F32>(gT, M, xT, M, dout, din, M, ew, s);
  mlp::EpiArgs ep{};
  ep.c = m->ws;
  ep.ldc = din;
  ep.zstride = (long)dout * din;
  hipLaunchKernelGGL((mlp::gemm_tn<mlp::EPI_F32, 128, 128, 32>), dim3(din / 128, dout / 128, S), dim3(256), 0, s, gT,
                     M, xT, M, M / S, ep);
  const size_t n = (size_t)dout * din;
  hipLaunchKernelGGL(mlp::sum_splits, dim3(blocks(n, 256)), dim3(256), 0, s, m->ws, n, S, ew.c);
  return hipGetLastError();
}

// Backward pass from the output gradient (g0, g0T hold dY and dY^T); fills gw, gb.
hipError_t backward(cloudsc_mlp* m, const bf16* xT, int M, hipStream_t s) {
  hipError_t e = hipSuccess;
  bf16 *g = m->g0, *gT = m->g0T, *gn = m->g1, *gnT = m->g1T;
  for (int l = m->nl - 1; l >= 0 && e == hipSuccess; --l) {
    // dW_l = G^T * X_l : C[dout x din] = gT[dout x M] * (X_l^T [din x M])^T
    const bf16* xlT = (l == 0) ? xT : m->aT[l - 1];
    mlp::EpiArgs ew{};
    ew.c = m->gw[l];
    ew.ldc = m->din[l];
    e = weight_gradient(m, gT, xlT, m->dout[l], m->din[l], M, ew, s);
    if (e != hipSuccess) break;
    hipLaunchKernelGGL(mlp::row_sums, dim3(m->dout[l]), dim3(256), 0, s, gT, M, m->gb[l]);
    if (l == 0) break;
    // dX_l = G * W_l, then times silu'(Z_{l-1}) -> gradient of the previous layer
    mlp::EpiArgs ed{};
    ed.ldc = m->din[l];
    ed.act = gn;
    ed.actT = gnT;
    ed.ldt = M;
    ed.z = m->z[l - 1];
    e = gemm<mlp::EPI_DHIDDEN>(g, m->dout[l], m->wT[l], m->dout[l], M, m->din[l], m->dout[l], ed, s);
    std::swap(g, gn);
    std::swap(gT, gnT);
  }
  if (e == hipSuccess) e = hipGetLastError();
  return e;
}

}  // namespace

extern "C" {

int cloudsc_mlp_create(cloudsc_mlp_t** out, int n_in, int n_hidden, int n_hidden_layers, int n_out, int max_batch,
                       uint64_t seed) {
  if (!out) return fail(CLOUDSC_ENULL, "cloudsc_mlp_create: NULL output");
  *out = nullptr;
  if (n_in <= 0 || n_hidden <= 0 || n_out <= 0 || n_hidden_layers < 1 || max_batch <= 0 || n_in % 128 ||
      n_hidden % 128 || n_out % 128 || max_batch % 128)
    return fail(CLOUDSC_EDIMS, "cloudsc_mlp_create: widths and max_batch must be positive multiples of 128");
  int rc = cloudsc_capi::check_device();
  if (rc) return rc;
  cloudsc_mlp* m = new cloudsc_mlp();
  m->nl = n_hidden_layers + 1;
  m->max_batch = max_batch;
  for (int l = 0; l < m->nl; ++l) {
    m->din.push_back(l == 0 ? n_in : n_hidden);
    m->dout.push_back(l == m->nl - 1 ? n_out : n_hidden);
  }
  auto F = [&](size_t n, float** p) { return dev_alloc(m, n * sizeof(float), (void**)p); };
  auto H = [&](size_t n, bf16** p) { return dev_alloc(m, n * sizeof(bf16), (void**)p); };
  m->w.resize(m->nl); m->b.resize(m->nl); m->mw.resize(m->nl); m->vw.resize(m->nl); m->mb.resize(m->nl);
  m->vb.resize(m->nl); m->gw.resize(m->nl); m->gb.resize(m->nl); m->wb.resize(m->nl); m->wT.resize(m->nl);
  for (int l = 0; l < m->nl && !rc; ++l) {
    const size_t nw = (size_t)m->din[l] * m->dout[l], nb = m->dout[l];
    rc = F(nw, &m->w[l]) || F(nb, &m->b[l]) || F(nw, &m->mw[l]) || F(nw, &m->vw[l]) || F(nb, &m->mb[l]) ||
         F(nb, &m->vb[l]) || F(nw, &m->gw[l]) || F(nb, &m->gb[l]) || H(nw, &m->wb[l]) || H(nw, &m->wT[l]);
  }
  m->z.resize(m->nl - 1); m->a.resize(m->nl - 1); m->aT.resize(m->nl - 1);
  const size_t mh = (size_t)max_batch * n_hidden;
  for (int l = 0; l < m->nl - 1 && !rc; ++l) rc = F(mh, &m->z[l]) || H(mh, &m->a[l]) || H(mh, &m->aT[l]);
  const size_t mg = (size_t)max_batch * std::max(n_hidden, n_out);
  if (!rc) rc = H(mg, &m->g0) || H(mg, &m->g0T) || H(mg, &m->g1) || H(mg, &m->g1T);
  if (!rc) rc = F((size_t)max_batch * n_out, &m->y) || F(1, &m->loss);
  size_t wmax = 0;
  for (int l = 0; l < m->nl; ++l) wmax = std::max(wmax, (size_t)m->din[l] * m->dout[l]);
  if (!rc) {
    m->ws_elems = (size_t)mlp::MAX_SPLIT * wmax;
    rc = F(m->ws_elems, &m->ws);
  }
  if (rc) {
    rc = CLOUDSC_ENOMEM;
    cloudsc_mlp_destroy(m);
    return rc;
  }
  // He-style initialisation: W ~ N(0, 1/n_in) from Philox (seed, layer, row, col); zero biases and moments.
  std::vector<float> hw;
  hipError_t e = hipSuccess;
  for (int l = 0; l < m->nl && e == hipSuccess; ++l) {
    const size_t nw = (size_t)m->din[l] * m->dout[l];
    hw.resize(nw);
    const double sd = std::sqrt(1.0 / m->din[l]);
    for (int r = 0; r < m->dout[l]; ++r)
      for (int c = 0; c < m->din[l]; ++c)
        hw[(size_t)r * m->din[l] + c] = (float)(sd * synth::draw(seed, 100u + (uint32_t)l, (uint32_t)c, (uint32_t)r).n);
    e = hipMemcpy(m->w[l], hw.data(), nw * sizeof(float), hipMemcpyHostToDevice);
    if (e != hipSuccess) break;
    e = hipMemset(m->b[l], 0, m->dout[l] * sizeof(float));
    if (e == hipSuccess) e = hipMemset(m->mw[l], 0, nw * sizeof(float));
    if (e == hipSuccess) e = hipMemset(m->vw[l], 0, nw * sizeof(float));
    if (e == hipSuccess) e = hipMemset(m->mb[l], 0, m->dout[l] * sizeof(float));
    if (e == hipSuccess) e = hipMemset(m->vb[l], 0, m->dout[l] * sizeof(float));
    if (e == hipSuccess) {
      hipLaunchKernelGGL(mlp::to_bf16, dim3(blocks(nw, 256)), dim3(256), 0, 0, m->w[l], nw, m->din[l], m->dout[l],
                         m->wb[l], m->wT[l]);
      e = hipGetLastError();
    }
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    cloudsc_mlp_destroy(m);
    return fail(CLOUDSC_EHIP, "cloudsc_mlp_create: initialisation", e);
  }
  *out = m;
  return CLOUDSC_OK;
}

int cloudsc_mlp_destroy(cloudsc_mlp_t* m) {
  if (!m) return CLOUDSC_OK;
  for (void* p : m->allocs) (void)hipFree(p);
  delete m;
  return CLOUDSC_OK;
}

int cloudsc_mlp_dims(const cloudsc_mlp_t* m, int* n_layers, int* din, int* dout) {
  if (!m || !n_layers) return fail(CLOUDSC_ENULL, "cloudsc_mlp_dims: NULL argument");
  *n_layers = m->nl;
  for (int l = 0; l < m->nl; ++l) {
    if (din) din[l] = m->din[l];
    if (dout) dout[l] = m->dout[l];
  }
  return CLOUDSC_OK;
}

int cloudsc_mlp_forward(cloudsc_mlp_t* m, const uint16_t* x, int M, float* y, void* stream) {
  if (!m || !x || !y) return fail(CLOUDSC_ENULL, "cloudsc_mlp_forward: NULL argument");
  int rc = check_batch(m, M);
  if (rc) return rc;
  hipError_t e = mlp::forward(m, (const bf16*)x, M, y, false, (hipStream_t)stream);
  if (e != hipSuccess) return fail(CLOUDSC_EHIP, "cloudsc_mlp_forward", e);
  return CLOUDSC_OK;
}

int cloudsc_mlp_gradients(cloudsc_mlp_t* m, const uint16_t* x, const uint16_t* xT, const float* t, const float* wgt,
                          int M, void* stream) {
  if (!m || !x || !xT || !t || !wgt) return fail(CLOUDSC_ENULL, "cloudsc_mlp_gradients: NULL argument");
  int rc = check_batch(m, M);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = mlp::forward(m, (const bf16*)x, M, m->y, true, s);
  if (e == hipSuccess) {
    const int n = m->dout[m->nl - 1];
    hipLaunchKernelGGL(mlp::fill_f32, dim3(1), dim3(64), 0, s, m->loss, (size_t)1, 0.0f);
    hipLaunchKernelGGL(mlp::mse_grad, dim3((n + 63) / 64, (M + 63) / 64), dim3(256), 0, s, m->y, t, wgt, M, n, m->g0,
                       m->g0T, m->loss);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = backward(m, (const bf16*)xT, M, s);
  if (e != hipSuccess) return fail(CLOUDSC_EHIP, "cloudsc_mlp_gradients", e);
  return CLOUDSC_OK;
}

int cloudsc_mlp_adam(cloudsc_mlp_t* m, float lr, float beta1, float beta2, float eps, void* stream) {
  if (!m) return fail(CLOUDSC_ENULL, "cloudsc_mlp_adam: NULL model");
  hipStream_t s = (hipStream_t)stream;
  m->step += 1;
  const float bc1 = 1.0f - std::pow(beta1, (float)m->step), bc2 = 1.0f - std::pow(beta2, (float)m->step);
  for (int l = 0; l < m->nl; ++l) {
    const size_t nw = (size_t)m->din[l] * m->dout[l];
    hipLaunchKernelGGL(mlp::adam_update, dim3(blocks(nw, 256)), dim3(256), 0, s, m->w[l], m->mw[l], m->vw[l],
                       m->gw[l], nw, m->din[l], lr, beta1, beta2, eps, bc1, bc2, m->wb[l], m->wT[l], m->dout[l]);
    hipLaunchKernelGGL(mlp::adam_update, dim3(blocks(m->dout[l], 256)), dim3(256), 0, s, m->b[l], m->mb[l],
                       m->vb[l], m->gb[l], (size_t)m->dout[l], 1, lr, beta1, beta2, eps, bc1, bc2, (bf16*)nullptr,
                       (bf16*)nullptr, 1);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(CLOUDSC_EHIP, "cloudsc_mlp_adam", e);
  return CLOUDSC_OK;
}

int cloudsc_mlp_train_step(cloudsc_mlp_t* m, const uint16_t* x, const uint16_t* xT, const float* t, const float* wgt,
                           int M, float lr, void* stream) {
  int rc = cloudsc_mlp_gradients(m, x, xT, t, wgt, M, stream);
  if (rc) return rc;
  return cloudsc_mlp_adam(m, lr, 0.9f, 0.999f, 1e-8f, stream);
}

// Host copies of one layer's parameters (which: 0 = W [dout x din] fp32, 1 = b, 2 = dW, 3 = db, 4 = last output Y
// [M x n_out] of a training forward, 5 = loss scalar).  set: 0 = device -> host, 1 = host -> device (W, b only;
// refreshes the BF16 copies).
int cloudsc_mlp_param(cloudsc_mlp_t* m, int layer, int which, float* host, size_t n, int set) {
  if (!m || !host) return fail(CLOUDSC_ENULL, "cloudsc_mlp_param: NULL argument");
  if (layer < 0 || layer >= m->nl || which < 0 || which > 5)
    return fail(CLOUDSC_EDIMS, "cloudsc_mlp_param: layer or field out of range");
  const size_t nw = (size_t)m->din[layer] * m->dout[layer], nb = m->dout[layer];
  float* p = nullptr;
  size_t cap = 0;
  switch (which) {
    case 0: p = m->w[layer]; cap = nw; break;
    case 1: p = m->b[layer]; cap = nb; break;
    case 2: p = m->gw[layer]; cap = nw; break;
    case 3: p = m->gb[layer]; cap = nb; break;
    case 4: p = m->y; cap = (size_t)m->max_batch * m->dout[m->nl - 1]; break;
    default: p = m->loss; cap = 1; break;
  }
  if (n > cap) return fail(CLOUDSC_EDIMS, "cloudsc_mlp_param: count exceeds the field size");
  if (set && which > 1) return fail(CLOUDSC_EDIMS, "cloudsc_mlp_param: only W and b can be set");
  hipError_t e = hipDeviceSynchronize();
  if (e == hipSuccess)
    e = set ? hipMemcpy(p, host, n * sizeof(float), hipMemcpyHostToDevice)
            : hipMemcpy(host, p, n * sizeof(float), hipMemcpyDeviceToHost);
  if (e == hipSuccess && set && which == 0) {
    hipLaunchKernelGGL(mlp::to_bf16, dim3(blocks(nw, 256)), dim3(256), 0, 0, m->w[layer], nw, m->din[layer],
                       m->dout[layer], m->wb[layer], m->wT[layer]);
    e = hipDeviceSynchronize();
  }
  if (e != hipSuccess) return fail(CLOUDSC_EHIP, "cloudsc_mlp_param", e);
  return CLOUDSC_OK;
}

// Plain product for tests: C[M x N] (fp32) = A[M x K] * B[N x K]^T, bf16 operands, device pointers.
int cloudsc_mlp_gemm_tn(const uint16_t* A, int lda, const uint16_t* B, int ldb, float* C, int ldc, int M, int N,
                        int K, void* stream) {
  if (!A || !B || !C) return fail(CLOUDSC_ENULL, "cloudsc_mlp_gemm_tn: NULL argument");
  if (M <= 0 || N <= 0 || K <= 0 || M % mlp::BM || N % mlp::BN || K % mlp::BK || lda < K || ldb < K || ldc < N ||
      lda % 8 || ldb % 8)
    return fail(CLOUDSC_EDIMS, "cloudsc_mlp_gemm_tn: need M, N multiples of 128, K of 32, lda/ldb >= K and of 8");
  int rc = cloudsc_capi::check_device();
  if (rc) return rc;
  mlp::EpiArgs e{};
  e.c = C;
  e.ldc = ldc;
  hipError_t err = gemm<mlp::EPI_F32>((const bf16*)A, lda, (const bf16*)B, ldb, M, N, K, e, (hipStream_t)stream);
  if (err != hipSuccess) return fail(CLOUDSC_EHIP, "cloudsc_mlp_gemm_tn", err);
  return CLOUDSC_OK;
}

}  // extern "C"
