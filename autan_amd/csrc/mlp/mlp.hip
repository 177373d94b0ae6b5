// The warm-start network on gfx950: a multilayer perceptron with SiLU activations, BF16 weights and
// activations on the matrix cores (v_mfma_f32_16x16x32_bf16, FP32 accumulation), FP32 master
// weights with Adam. No counterpart in the reference: the learned initial state of the bump sweep.
//
// Layout: samples are rows. Layer l maps K_l inputs to N_l outputs with weights W_l stored N_l x K_l
// (row n holds the weights into output n), so every operand of the forward product is read along
// K. The input width is padded to 32 and the output width to 16 with zero weights.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/mxmlp.h"

namespace mxm {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
using bf16 = __bf16;

constexpr int kThreads = 256;
constexpr int kTileM = 64, kTileN = 64, kTileK = 32;
constexpr int kPad = 8;  // LDS row padding (bf16) against bank conflicts

inline void check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string("E1101: ") + what + ": " + hipGetErrorString(e));
}

// SiLU and its derivative through the logistic s(z) = 1 / (1 + e^-z), with the hardware exponential
// and reciprocal (about 1 ulp each; the activations are rounded to BF16 right after).
__device__ __forceinline__ float logistic(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * logistic(z); }
__device__ __forceinline__ float silu_grad(float z) {
  const float s = logistic(z);
  return s * (1.0f + z * (1.0f - s));
}

// Epilogues of the product C = op(A) op(B).
enum Epi : int {
  kStoreF32 = 0,     // Cf = C
  kAtomicF32 = 1,    // Cf += C (split over K)
  kBiasSiLU = 2,     // Zf = C + bias, Hb = silu(Z)
  kBias = 3,         // Cf = C + bias
  kSiLUGrad = 4,     // Hb = C * silu'(Z[m, n])  (Z read from Cf)
};

// C[M x N] = op(A) op(B), op(A)[m][k] = TA ? A[k lda + m] : A[m lda + k], op(B)[k][n] = TB ? B[n ldb + k]
// : B[k ldb + n]. One 64 x 64 tile per workgroup of four waves (32 x 32 each, 2 x 2 matrix-core
// tiles of 16 x 16 x 32), K in steps of 32 staged through LDS with K contiguous; blockIdx.z splits
// K into chunks of kchunk. Elements outside M, N, K read as zero.
template <bool TA, bool TB, int EPI>
__global__ __launch_bounds__(kThreads) void k_gemm(const bf16* __restrict__ A, const bf16* __restrict__ B, int M, int N,
                                                   int K, int lda, int ldb, int kchunk, float* __restrict__ Cf,
                                                   bf16* __restrict__ Hb, const float* __restrict__ bias, int ldc) {
  __shared__ __align__(16) bf16 As[kTileM][kTileK + kPad];
  __shared__ __align__(16) bf16 Bs[kTileN][kTileK + kPad];
  const int m0 = blockIdx.y * kTileM, n0 = blockIdx.x * kTileN;
  const int kb = blockIdx.z * kchunk, ke = min(K, kb + kchunk);
  const int t = threadIdx.x, lane = t % 64, wave = t / 64;
  const int wm = (wave / 2) * 32, wn = (wave % 2) * 32;
  f32x4 acc[2][2];
  for (int i = 0; i < 2; ++i) {
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  }
  const bf16 zero = static_cast<bf16>(0.0f);
  for (int k0 = kb; k0 < ke; k0 += kTileK) {
    // Stage A: 64 rows x 32 k, 8 elements per thread.
    if constexpr (!TA) {
      const int r = t / 4, kk = (t % 4) * 8;
      const int m = m0 + r, k = k0 + kk;
      bf16x8 v;
      if (m < M && k + 8 <= ke) {
        v = *reinterpret_cast<const bf16x8*>(A + static_cast<long long>(m) * lda + k);
      } else {
        for (int q = 0; q < 8; ++q) v[q] = (m < M && k + q < ke) ? A[static_cast<long long>(m) * lda + k + q] : zero;
      }
      *reinterpret_cast<bf16x8*>(&As[r][kk]) = v;
    } else {
      const int kk = t / 8, r = (t % 8) * 8;
      const int k = k0 + kk, m = m0 + r;
      bf16x8 v;
      if (k < ke && m + 8 <= M) {
        v = *reinterpret_cast<const bf16x8*>(A + static_cast<long long>(k) * lda + m);
      } else {
        for (int q = 0; q < 8; ++q) v[q] = (k < ke && m + q < M) ? A[static_cast<long long>(k) * lda + m + q] : zero;
      }
      for (int q = 0; q < 8; ++q) As[r + q][kk] = v[q];
    }
    // Stage B as Bs[n][k].
    if constexpr (TB) {
      const int r = t / 4, kk = (t % 4) * 8;
      const int n = n0 + r, k = k0 + kk;
      bf16x8 v;
      if (n < N && k + 8 <= ke) {
        v = *reinterpret_cast<const bf16x8*>(B + static_cast<long long>(n) * ldb + k);
      } else {
        for (int q = 0; q < 8; ++q) v[q] = (n < N && k + q < ke) ? B[static_cast<long long>(n) * ldb + k + q] : zero;
      }
      *reinterpret_cast<bf16x8*>(&Bs[r][kk]) = v;
    } else {
      const int kk = t / 8, r = (t % 8) * 8;
      const int k = k0 + kk, n = n0 + r;
      bf16x8 v;
      if (k < ke && n + 8 <= N) {
        v = *reinterpret_cast<const bf16x8*>(B + static_cast<long long>(k) * ldb + n);
      } else {
        for (int q = 0; q < 8; ++q) v[q] = (k < ke && n + q < N) ? B[static_cast<long long>(k) * ldb + n + q] : zero;
      }
      for (int q = 0; q < 8; ++q) Bs[r + q][kk] = v[q];
    }
    __syncthreads();
    // Lane l holds row (l mod 16) of A and column (l mod 16) of B at k = 8 (l / 16) .. + 7.
    const int row = lane % 16, kq = 8 * (lane / 16);
    bf16x8 a[2], b[2];
    for (int i = 0; i < 2; ++i) a[i] = *reinterpret_cast<const bf16x8*>(&As[wm + 16 * i + row][kq]);
    for (int j = 0; j < 2; ++j) b[j] = *reinterpret_cast<const bf16x8*>(&Bs[wn + 16 * j + row][kq]);
    for (int i = 0; i < 2; ++i) {
      for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
  // Lane l holds rows 4 (l / 16) .. + 3 of column (l mod 16) of each 16 x 16 tile.
  for (int i = 0; i < 2; ++i) {
    for (int j = 0; j < 2; ++j) {
      const int n = n0 + wn + 16 * j + lane % 16;
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm + 16 * i + 4 * (lane / 16) + r;
        if (m >= M || n >= N) continue;
        const float c = acc[i][j][r];
        const long long at = static_cast<long long>(m) * ldc + n;
        if constexpr (EPI == kStoreF32) {
          Cf[at] = c;
        } else if constexpr (EPI == kAtomicF32) {
          atomicAdd(Cf + at, c);
        } else if constexpr (EPI == kBiasSiLU) {
          const float z = c + bias[n];
          Cf[at] = z;
          Hb[at] = static_cast<bf16>(silu(z));
        } else if constexpr (EPI == kBias) {
          Cf[at] = c + bias[n];
        } else {
          Hb[at] = static_cast<bf16>(c * silu_grad(Cf[at]));
        }
      }
    }
  }
}

// Inputs normalised and rounded to BF16, padded to `width` columns with zeros.
__global__ void k_inputs(const float* __restrict__ x, int n, int d, const float* __restrict__ mean,
                         const float* __restrict__ scale, bf16* __restrict__ X, int width) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= static_cast<long long>(n) * width) return;
  const long long r = i / width;
  const int c = static_cast<int>(i % width);
  X[i] = static_cast<bf16>(c < d ? (x[r * d + c] - mean[c]) / scale[c] : 0.0f);
}

// Mean squared error of the first d of `width` outputs against normalised targets, and its
// gradient 2 (y - t) / (n d) in BF16 (zero in the padding columns); the squared errors summed into
// *sum (each block's partial added once).
__global__ void k_loss(const float* __restrict__ y, const float* __restrict__ t, int n, int d, int width,
                       const float* __restrict__ mean, const float* __restrict__ scale, bf16* __restrict__ dy,
                       double* __restrict__ sum) {
  __shared__ float red[256];
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  float e2 = 0.0f;
  if (i < static_cast<long long>(n) * width) {
    const long long r = i / width;
    const int c = static_cast<int>(i % width);
    float g = 0.0f;
    if (c < d) {
      const float e = y[i] - (t[r * d + c] - mean[c]) / scale[c];
      e2 = e * e;
      g = 2.0f * e / (static_cast<float>(n) * d);
    }
    dy[i] = static_cast<bf16>(g);
  }
  red[threadIdx.x] = e2;
  __syncthreads();
  for (int s = 128; s > 0; s /= 2) {
    if (static_cast<int>(threadIdx.x) < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(sum, static_cast<double>(red[0]));
}

// Column sums of a BF16 n x width matrix (the bias gradients) added into g: thread (x, y) sums
// column x over the rows y, y + gridDim.y blockDim.y, ...
__global__ void k_colsum(const bf16* __restrict__ dz, int n, int width, float* __restrict__ g) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= width) return;
  float s = 0.0f;
  for (int r = blockIdx.y; r < n; r += gridDim.y) s += static_cast<float>(dz[static_cast<long long>(r) * width + c]);
  atomicAdd(g + c, s);
}

// A training step's counters, kept on the device so that a recorded step replays unchanged: the
// batch of the epoch, Adam's step and its bias corrections 1 - b^step.
struct StepState {
  long long batch;
  long long step;
  float c1, c2;
};

// Adam's next step (one thread): the step count and its bias corrections; the next batch.
__global__ void k_next_step(StepState* st, float b1, float b2) {
  st->batch += 1;
  st->step += 1;
  const float t = static_cast<float>(st->step);
  st->c1 = 1.0f - powf(b1, t);
  st->c2 = 1.0f - powf(b2, t);
}

__global__ void k_set_batch(StepState* st, long long b) { st->batch = b; }

// Batch `st->batch` of the shuffled order (n samples from idx[batch n]) into bx, bt.
__global__ void k_gather(const long long* __restrict__ idx, const StepState* __restrict__ st, int n,
                         const float* __restrict__ x, int dx, const float* __restrict__ t, int dt, float* __restrict__ bx,
                         float* __restrict__ bt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long long r = idx[st->batch * n + i];
  for (int c = 0; c < dx; ++c) bx[static_cast<long long>(i) * dx + c] = x[r * dx + c];
  for (int c = 0; c < dt; ++c) bt[static_cast<long long>(i) * dt + c] = t[r * dt + c];
}

__global__ void k_fill(float* __restrict__ p, long long n, float v) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Adam (Kingma and Ba 2015) on the FP32 master weights, then their BF16 copy. `mask` marks the
// entries that exist (padding stays zero); the bias corrections are the step state's.
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m1, float* __restrict__ m2,
                       const unsigned char* __restrict__ mask, long long n, float lr, float b1, float b2, float eps,
                       const StepState* __restrict__ st, bf16* __restrict__ wb) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float c1 = st->c1, c2 = st->c2;
  if (mask[i]) {
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.0f - b1) * gi;
    const float b = b2 * m2[i] + (1.0f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    w[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
  }
  if (wb != nullptr) wb[i] = static_cast<bf16>(w[i]);
}

__global__ void k_to_bf16(const float* __restrict__ w, long long n, bf16* __restrict__ wb) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) wb[i] = static_cast<bf16>(w[i]);
}

// Network outputs back in physical units: y * scale + mean for the first d columns.
__global__ void k_out(const float* __restrict__ y, int n, int d, int width, const float* __restrict__ mean,
                      const float* __restrict__ scale, float* __restrict__ out) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= static_cast<long long>(n) * d) return;
  const long long r = i / d;
  const int c = static_cast<int>(i % d);
  out[i] = y[r * width + c] * scale[c] + mean[c];
}

// The layers of a network as k_infer reads them: weight and bias offsets into the parameter arrays.
constexpr int kMaxLayers = 18;
struct LayerTable {
  long long w[kMaxLayers], b[kMaxLayers];
  int nl;
};

// The whole forward of 16 RT samples per workgroup (inference): the normalised inputs, then every
// layer's activations, stay in one 16 RT x (hidden + 8) BF16 LDS tile (dynamic shared memory); the
// weights stream from L2, each weight column read once per workgroup and used for all 16 RT rows,
// so the L2 bytes per product fall as RT grows. Each of the WAVES waves computes 16 CT output
// columns of the tile's rows as RT x CT matrix-core tiles, K in steps of 32 in the order of
// k_gemm, and the epilogues are k_gemm's (bias + SiLU, then bias and the output scaling of
// k_out): the outputs are the layer-by-layer path's bits whatever RT, without its HBM round trip
// of every layer's activations. The hidden layers' tiles are computed transposed (W H^T), so that
// a lane's four results are four consecutive outputs of one sample: one 8-byte LDS store of two
// packed BF16 conversions each, instead of four 2-byte stores.
template <int CT, int RT, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_infer(const float* __restrict__ x, int n, int d_in, int d_out,
                                                      const float* __restrict__ norm, const bf16* __restrict__ wb,
                                                      const float* __restrict__ params, LayerTable L,
                                                      float* __restrict__ y) {
  constexpr int kThreads = 64 * WAVES;
  constexpr int H = 16 * CT * WAVES, ld = H + kPad, kIn = 32, M = 16 * RT;
  extern __shared__ __align__(16) unsigned char infer_smem[];
  bf16* act = reinterpret_cast<bf16*>(infer_smem);
  const long long m0 = static_cast<long long>(blockIdx.x) * M;
  const int t = threadIdx.x, lane = t % 64, wave = t / 64;
  const int row = lane % 16, kq = 8 * (lane / 16);
  for (int i = t; i < M * kIn; i += kThreads) {
    const int r = i / kIn, c = i % kIn;
    const long long m = m0 + r;
    act[r * ld + c] = static_cast<bf16>(m < n && c < d_in ? (x[m * d_in + c] - norm[c]) / norm[32 + c] : 0.0f);
  }
  for (int l = 0; l < L.nl; ++l) {
    const int K = l == 0 ? kIn : H;
    const bf16* W = wb + L.w[l];
    const float* bias = params + L.b[l];
    __syncthreads();  // the layer's input is complete
    if (l < L.nl - 1) {
      f32x4 acc[RT][CT];
#pragma unroll
      for (int i = 0; i < RT; ++i) {
#pragma unroll
        for (int j = 0; j < CT; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
      }
      const int nb = wave * 16 * CT;
      const bf16* wrow[CT];
#pragma unroll
      for (int j = 0; j < CT; ++j) wrow[j] = W + static_cast<long long>(nb + 16 * j + row) * K + kq;
      // The next K step's weights are in flight while this step's products run.
      bf16x8 b_next[CT];
#pragma unroll
      for (int j = 0; j < CT; ++j) b_next[j] = *reinterpret_cast<const bf16x8*>(wrow[j]);
      for (int k0 = 0; k0 < K; k0 += kTileK) {
        bf16x8 b[CT];
#pragma unroll
        for (int j = 0; j < CT; ++j) b[j] = b_next[j];
        if (k0 + kTileK < K) {
#pragma unroll
          for (int j = 0; j < CT; ++j) b_next[j] = *reinterpret_cast<const bf16x8*>(wrow[j] + k0 + kTileK);
        }
#pragma unroll
        for (int i = 0; i < RT; ++i) {
          const bf16x8 a = *reinterpret_cast<const bf16x8*>(&act[(16 * i + row) * ld + k0 + kq]);
          // The weights as the first operand: the tile is W H^T, so each lane ends with four
          // consecutive outputs of one sample. Every output is the same sum of the same exact
          // BF16 products as in k_gemm's H W^T.
#pragma unroll
          for (int j = 0; j < CT; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b[j], a, acc[i][j], 0, 0, 0);
        }
      }
      __syncthreads();  // every wave has read the layer's input: the tile takes its output
#pragma unroll
      for (int j = 0; j < CT; ++j) {
        // Lane l holds outputs f .. f + 3 of sample 16 i + l mod 16 of each row tile i.
        const int f = nb + 16 * j + 4 * (lane / 16);
        float bn[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) bn[r] = bias[f + r];
#pragma unroll
        for (int i = 0; i < RT; ++i) {
          f32x4 h;
#pragma unroll
          for (int r = 0; r < 4; ++r) h[r] = silu(acc[i][j][r] + bn[r]);
          *reinterpret_cast<bf16x4*>(&act[(16 * i + lane % 16) * ld + f]) = __builtin_convertvector(h, bf16x4);
        }
      }
    } else {
      // The padded output layer (16 columns): wave w takes the row tiles w, w + 4, ...
      for (int i = wave; i < RT; i += kThreads / 64) {
        f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
        for (int k0 = 0; k0 < K; k0 += kTileK) {
          const bf16x8 a = *reinterpret_cast<const bf16x8*>(&act[(16 * i + row) * ld + k0 + kq]);
          const bf16x8 b = *reinterpret_cast<const bf16x8*>(W + static_cast<long long>(row) * K + k0 + kq);
          acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
        }
        const int nn = lane % 16;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const long long m = m0 + 16 * i + 4 * (lane / 16) + r;
          if (m < n && nn < d_out) {
            const float z = acc[r] + bias[nn];
            y[m * d_out + nn] = z * norm[96 + nn] + norm[64 + nn];
          }
        }
      }
    }
  }
}

// LDS bytes of the tile of 16 rt samples of width h.
constexpr std::size_t infer_lds(int h, int rt) { return std::size_t(16 * rt) * std::size_t(h + kPad) * 2; }

// Launches k_infer<CT, RT, WAVES> for n samples, raising its dynamic LDS limit once per process.
template <int CT, int RT, int WAVES>
void launch_infer(const float* x, int n, int d_in, int d_out, const float* norm, const bf16* wb, const float* params,
                  const LayerTable& L, float* y, hipStream_t s) {
  constexpr int h = 16 * CT * WAVES;
  static const hipError_t raised = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_infer<CT, RT, WAVES>),
                                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                                       static_cast<int>(infer_lds(h, RT)));
  check(raised, "hipFuncSetAttribute(k_infer)");
  const int g = static_cast<int>((static_cast<long long>(n) + 16 * RT - 1) / (16 * RT));
  k_infer<CT, RT, WAVES><<<g, 64 * WAVES, infer_lds(h, RT), s>>>(x, n, d_in, d_out, norm, wb, params, L, y);
}

// The fused forward for hidden width 64 q (q = 1 .. 4) and a tile of `rows` samples: four waves of
// 16 q columns each (64, 96 or 128 samples), or, for 128 samples and an even q, eight waves of 8 q
// columns each (half the accumulators per wave, so twice the waves per SIMD).
template <int Q>
void launch_infer_rows(int rows, const float* x, int n, int d_in, int d_out, const float* norm, const bf16* wb,
                       const float* params, const LayerTable& L, float* y, hipStream_t s) {
  switch (rows) {
    case 64: launch_infer<Q, 4, 4>(x, n, d_in, d_out, norm, wb, params, L, y, s); break;
    case 96: launch_infer<Q, 6, 4>(x, n, d_in, d_out, norm, wb, params, L, y, s); break;
    case 128:
      if constexpr (Q % 2 == 0) {
        launch_infer<Q / 2, 8, 8>(x, n, d_in, d_out, norm, wb, params, L, y, s);
      } else {
        launch_infer<Q, 8, 4>(x, n, d_in, d_out, norm, wb, params, L, y, s);
      }
      break;
    default: throw std::invalid_argument("inference tile: 64, 96 or 128 samples per workgroup");
  }
}

template <class T>
class Buf {
 public:
  Buf() = default;
  explicit Buf(std::size_t n) : n_(n) {
    if (n) check(hipMalloc(reinterpret_cast<void**>(&p_), n * sizeof(T)), "hipMalloc");
  }
  Buf(const Buf&) = delete;
  Buf& operator=(const Buf&) = delete;
  Buf(Buf&& o) noexcept : p_(o.p_), n_(o.n_) {
    o.p_ = nullptr;
    o.n_ = 0;
  }
  Buf& operator=(Buf&& o) noexcept {
    std::swap(p_, o.p_);
    std::swap(n_, o.n_);
    return *this;
  }
  ~Buf() {
    if (p_) (void)hipFree(p_);
  }
  T* get() const { return p_; }
  std::size_t size() const { return n_; }

 private:
  T* p_ = nullptr;
  std::size_t n_ = 0;
};

inline int cdiv(long long a, long long b) { return static_cast<int>((a + b - 1) / b); }

class Net {
 public:
  Net(int inputs, int hidden, int layers, int outputs, uint64_t seed) : d_in_(inputs), d_out_(outputs) {
    if (inputs < 1 || inputs > 32 || outputs < 1 || outputs > 16 || hidden < 16 || hidden % 64 != 0 || layers < 1 ||
        layers > 16) {
      throw std::invalid_argument("mlp: 1-32 inputs, 1-16 outputs, hidden a multiple of 64, 1-16 hidden layers");
    }
    check(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking), "hipStreamCreate");
    // Widths: 32 (padded inputs), hidden x layers, 16 (padded outputs).
    w_.push_back(kInPad);
    for (int l = 0; l < layers; ++l) w_.push_back(hidden);
    w_.push_back(kOutPad);
    const int nl = static_cast<int>(w_.size()) - 1;
    std::mt19937_64 rng(seed);
    std::vector<float> host;
    std::vector<unsigned char> mask;
    for (int l = 0; l < nl; ++l) {
      const int K = w_[l], N = w_[l + 1];
      const int k_real = l == 0 ? inputs : K, n_real = l == nl - 1 ? outputs : N;
      woff_.push_back(static_cast<long long>(host.size()));
      // He-style uniform initialisation over the real fan-in; biases zero.
      const float bound = std::sqrt(6.0f / static_cast<float>(k_real));
      std::uniform_real_distribution<float> u(-bound, bound);
      for (int n = 0; n < N; ++n) {
        for (int k = 0; k < K; ++k) {
          const bool real = n < n_real && k < k_real;
          host.push_back(real ? u(rng) : 0.0f);
          mask.push_back(real ? 1 : 0);
        }
      }
      boff_.push_back(static_cast<long long>(host.size()));
      for (int n = 0; n < N; ++n) {
        host.push_back(0.0f);
        mask.push_back(n < n_real ? 1 : 0);
      }
    }
    np_ = static_cast<long long>(host.size());
    params_ = Buf<float>(std::size_t(np_));
    grads_ = Buf<float>(std::size_t(np_));
    m1_ = Buf<float>(std::size_t(np_));
    m2_ = Buf<float>(std::size_t(np_));
    wb_ = Buf<bf16>(std::size_t(np_));
    mask_ = Buf<unsigned char>(std::size_t(np_));
    check(hipMemcpy(params_.get(), host.data(), np_ * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
    check(hipMemcpy(mask_.get(), mask.data(), np_, hipMemcpyHostToDevice), "hipMemcpy");
    check(hipMemset(m1_.get(), 0, np_ * sizeof(float)), "hipMemset");
    check(hipMemset(m2_.get(), 0, np_ * sizeof(float)), "hipMemset");
    state_ = Buf<StepState>(1);
    check(hipMemset(state_.get(), 0, sizeof(StepState)), "hipMemset");
    refresh_bf16();
    std::vector<float> ones(32, 1.0f), zeros(32, 0.0f);
    norm_ = Buf<float>(4 * 32);
    set_normalization(zeros.data(), ones.data(), zeros.data(), ones.data());
  }
  ~Net() {
    (void)hipStreamSynchronize(s_);
    (void)hipStreamDestroy(s_);
  }

  long long param_count() const { return np_; }
  int layers() const { return static_cast<int>(w_.size()) - 1; }

  void set_normalization(const float* in_mean, const float* in_scale, const float* out_mean, const float* out_scale) {
    std::vector<float> h(4 * 32, 0.0f);
    for (int i = 0; i < 32; ++i) h[32 + i] = h[96 + i] = 1.0f;
    for (int i = 0; i < d_in_; ++i) {
      h[i] = in_mean[i];
      h[32 + i] = in_scale[i];
      if (!(in_scale[i] > 0.0f)) throw std::invalid_argument("mlp: input scales must be positive");
    }
    for (int i = 0; i < d_out_; ++i) {
      h[64 + i] = out_mean[i];
      h[96 + i] = out_scale[i];
      if (!(out_scale[i] > 0.0f)) throw std::invalid_argument("mlp: output scales must be positive");
    }
    check(hipMemcpy(norm_.get(), h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
  }

  void get_params(float* out) {
    check(hipMemcpy(out, params_.get(), np_ * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy");
  }
  void set_params(const float* in) {
    check(hipMemcpy(params_.get(), in, np_ * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
    refresh_bf16();
  }

  // Forward of n samples (host arrays, n x inputs -> n x outputs, physical units).
  void predict(const float* x, long long n, float* y) {
    const long long chunk = 1 << 16;
    Buf<float> xd(std::size_t(std::min(n, chunk) * d_in_)), yd(std::size_t(std::min(n, chunk) * d_out_));
    for (long long at = 0; at < n; at += chunk) {
      const int m = static_cast<int>(std::min(chunk, n - at));
      check(hipMemcpyAsync(xd.get(), x + at * d_in_, std::size_t(m) * d_in_ * sizeof(float), hipMemcpyHostToDevice, s_),
            "hipMemcpyAsync");
      infer(xd.get(), m, yd.get());
      check(hipMemcpyAsync(y + at * d_out_, yd.get(), std::size_t(m) * d_out_ * sizeof(float), hipMemcpyDeviceToHost, s_),
            "hipMemcpyAsync");
      check(hipStreamSynchronize(s_), "predict");
    }
  }

  // Loss and gradients (FP32, in parameter order) of one batch; grads may be null.
  float gradients(const float* x, const float* t, int n, float* grads) {
    Buf<float> xd(std::size_t(n) * d_in_), td(std::size_t(n) * d_out_);
    check(hipMemcpyAsync(xd.get(), x, std::size_t(n) * d_in_ * sizeof(float), hipMemcpyHostToDevice, s_), "hipMemcpy");
    check(hipMemcpyAsync(td.get(), t, std::size_t(n) * d_out_ * sizeof(float), hipMemcpyHostToDevice, s_), "hipMemcpy");
    ensure(n);
    check(hipMemsetAsync(loss_sum_.get(), 0, sizeof(double), s_), "hipMemsetAsync");
    backward(xd.get(), td.get(), n);
    double sum = 0.0;
    check(hipMemcpyAsync(&sum, loss_sum_.get(), sizeof(double), hipMemcpyDeviceToHost, s_), "hipMemcpyAsync");
    check(hipStreamSynchronize(s_), "gradients");
    const float loss = static_cast<float>(sum / (static_cast<double>(n) * d_out_));
    if (grads != nullptr) check(hipMemcpy(grads, grads_.get(), np_ * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy");
    return loss;
  }

  // One batch's forward and backward, then one of its tensors to the host as FP32, row-major
  // n x width: what 0 the input activations of layer `layer` (BF16: the normalised inputs for
  // layer 0), 1 its pre-activations Z (FP32), 2 the gradient dZ of its pre-activations (BF16).
  void trace(const float* x, const float* t, int n, int what, int layer, float* out) {
    const int nl = layers();
    if (layer < 0 || layer >= nl || what < 0 || what > 2) throw std::invalid_argument("trace: no such tensor");
    Buf<float> xd(std::size_t(n) * d_in_), td(std::size_t(n) * d_out_);
    check(hipMemcpyAsync(xd.get(), x, std::size_t(n) * d_in_ * sizeof(float), hipMemcpyHostToDevice, s_), "hipMemcpy");
    check(hipMemcpyAsync(td.get(), t, std::size_t(n) * d_out_ * sizeof(float), hipMemcpyHostToDevice, s_), "hipMemcpy");
    ensure(n);
    check(hipMemsetAsync(loss_sum_.get(), 0, sizeof(double), s_), "hipMemsetAsync");
    backward(xd.get(), td.get(), n);
    const std::size_t l = static_cast<std::size_t>(layer);
    const std::size_t count = std::size_t(n) * std::size_t(what == 0 ? w_[l] : w_[l + 1]);
    if (what == 1) {
      check(hipMemcpyAsync(out, z_[l].get(), count * sizeof(float), hipMemcpyDeviceToHost, s_), "hipMemcpy");
      check(hipStreamSynchronize(s_), "trace");
      return;
    }
    std::vector<bf16> h(count);
    const bf16* src = what == 0 ? acts_[l].get() : dz_[l].get();
    check(hipMemcpyAsync(h.data(), src, count * sizeof(bf16), hipMemcpyDeviceToHost, s_), "hipMemcpy");
    check(hipStreamSynchronize(s_), "trace");
    for (std::size_t i = 0; i < count; ++i) out[i] = static_cast<float>(h[i]);
  }

  // Training data resident on the device; epochs of shuffled batches with Adam.
  void set_data(const float* x, const float* t, long long n) {
    data_x_ = Buf<float>(std::size_t(n) * d_in_);
    data_t_ = Buf<float>(std::size_t(n) * d_out_);
    check(hipMemcpy(data_x_.get(), x, std::size_t(n) * d_in_ * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
    check(hipMemcpy(data_t_.get(), t, std::size_t(n) * d_out_ * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
    n_data_ = n;
  }

  // `epochs` passes over the data in batches of `batch` (the order shuffled per epoch with
  // `seed`), one Adam step each. Returns the last epoch's mean batch loss.
  float train(int epochs, int batch, float lr, uint64_t seed, double* seconds) {
    if (n_data_ < batch || batch < 64) throw std::invalid_argument("mlp: batch >= 64 and <= the data");
    const long long nb = n_data_ / batch;
    Buf<float> bx(std::size_t(batch) * d_in_), bt(std::size_t(batch) * d_out_);
    Buf<long long> idx(static_cast<std::size_t>(n_data_));
    std::vector<long long> order(static_cast<std::size_t>(n_data_));
    std::mt19937_64 rng(seed);
    float last = 0.0f;
    hipGraphExec_t step_graph = nullptr;
    struct GraphGuard {
      hipGraphExec_t* g;
      ~GraphGuard() {
        if (*g != nullptr) (void)hipGraphExecDestroy(*g);
      }
    } guard{&step_graph};
    hipEvent_t e0, e1;
    check(hipEventCreate(&e0), "hipEventCreate");
    check(hipEventCreate(&e1), "hipEventCreate");
    check(hipEventRecord(e0, s_), "hipEventRecord");
    for (int ep = 0; ep < epochs; ++ep) {
      for (long long i = 0; i < n_data_; ++i) order[std::size_t(i)] = i;
      std::shuffle(order.begin(), order.end(), rng);
      check(hipMemcpyAsync(idx.get(), order.data(), order.size() * sizeof(long long), hipMemcpyHostToDevice, s_),
            "hipMemcpyAsync");
      ensure(batch);
      check(hipMemsetAsync(loss_sum_.get(), 0, sizeof(double), s_), "hipMemsetAsync");
      k_set_batch<<<1, 1, 0, s_>>>(state_.get(), 0);
      check(hipGetLastError(), "k_set_batch");
      if (step_graph == nullptr) {
        // One step (gather of the state's batch, forward, loss, backward, Adam, next batch)
        // recorded once and replayed for every batch: its launches cost one graph launch.
        check(hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        hipGraph_t g = nullptr;
        try {
          k_gather<<<cdiv(batch, 256), 256, 0, s_>>>(idx.get(), state_.get(), batch, data_x_.get(), d_in_,
                                                     data_t_.get(), d_out_, bx.get(), bt.get());
          check(hipGetLastError(), "k_gather");
          backward(bx.get(), bt.get(), batch);
          adam(lr);
        } catch (...) {
          (void)hipStreamEndCapture(s_, &g);
          if (g != nullptr) (void)hipGraphDestroy(g);
          throw;
        }
        check(hipStreamEndCapture(s_, &g), "hipStreamEndCapture");
        const hipError_t e = hipGraphInstantiate(&step_graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        check(e, "hipGraphInstantiate");
      }
      for (long long b = 0; b < nb; ++b) check(hipGraphLaunch(step_graph, s_), "hipGraphLaunch");
      double sum = 0.0;
      check(hipMemcpyAsync(&sum, loss_sum_.get(), sizeof(double), hipMemcpyDeviceToHost, s_), "hipMemcpyAsync");
      check(hipStreamSynchronize(s_), "train");
      last = static_cast<float>(sum / (static_cast<double>(nb) * batch * d_out_));
    }
    check(hipEventRecord(e1, s_), "hipEventRecord");
    check(hipEventSynchronize(e1), "hipEventSynchronize");
    float ms = 0.0f;
    check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    if (seconds) *seconds = ms * 1e-3;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return last;
  }

  void adam_step(float lr) {
    adam(lr);
    check(hipStreamSynchronize(s_), "adam");
  }

  // Inference: n device-resident samples to n x outputs in physical units, by the fused forward
  // (hidden widths up to 256) or, with fused_ off or wider layers, the training path's
  // layer-by-layer products; the same bits either way.
  void infer(const float* x, int n, float* y) {
    const int h = w_[1];
    if (fused_ && h <= 256 && w_.size() - 1 <= std::size_t(kMaxLayers)) {
      LayerTable L{};
      L.nl = layers();
      for (int l = 0; l < L.nl; ++l) {
        L.w[l] = woff_[std::size_t(l)];
        L.b[l] = boff_[std::size_t(l)];
      }
      const float* nm = norm_.get();
      const float* pr = params_.get();
      switch (h / 64) {
        case 1: launch_infer_rows<1>(infer_rows_, x, n, d_in_, d_out_, nm, wb_.get(), pr, L, y, s_); break;
        case 2: launch_infer_rows<2>(infer_rows_, x, n, d_in_, d_out_, nm, wb_.get(), pr, L, y, s_); break;
        case 3: launch_infer_rows<3>(infer_rows_, x, n, d_in_, d_out_, nm, wb_.get(), pr, L, y, s_); break;
        default: launch_infer_rows<4>(infer_rows_, x, n, d_in_, d_out_, nm, wb_.get(), pr, L, y, s_); break;
      }
      check(hipGetLastError(), "k_infer");
      return;
    }
    forward(x, n);
    k_out<<<cdiv(static_cast<long long>(n) * d_out_, 256), 256, 0, s_>>>(z_.back().get(), n, d_out_, kOutPad,
                                                                         norm_.get() + 64, norm_.get() + 96, y);
    check(hipGetLastError(), "k_out");
  }

  void set_fused(bool on) { fused_ = on; }
  void set_infer_rows(int rows) {
    if (rows != 64 && rows != 96 && rows != 128)
      throw std::invalid_argument("inference tile: 64, 96 or 128 samples per workgroup");
    infer_rows_ = rows;
  }

  // Milliseconds of one inference of n samples (device-resident inputs and outputs, HIP events).
  double time_forward(int n, int repeats) {
    Buf<float> xd(std::size_t(n) * d_in_), yd(std::size_t(n) * d_out_);
    check(hipMemsetAsync(xd.get(), 0, std::size_t(n) * d_in_ * sizeof(float), s_), "hipMemset");
    infer(xd.get(), n, yd.get());
    hipEvent_t e0, e1;
    check(hipEventCreate(&e0), "hipEventCreate");
    check(hipEventCreate(&e1), "hipEventCreate");
    check(hipEventRecord(e0, s_), "hipEventRecord");
    for (int r = 0; r < repeats; ++r) infer(xd.get(), n, yd.get());
    check(hipEventRecord(e1, s_), "hipEventRecord");
    check(hipEventSynchronize(e1), "hipEventSynchronize");
    float ms = 0.0f;
    check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ms / repeats;
  }

 private:
  static constexpr int kInPad = 32, kOutPad = 16;

  void refresh_bf16() {
    k_to_bf16<<<cdiv(np_, 256), 256, 0, s_>>>(params_.get(), np_, wb_.get());
    check(hipGetLastError(), "k_to_bf16");
    check(hipStreamSynchronize(s_), "refresh");
  }

  void ensure(int n) {
    if (n <= cap_) return;
    cap_ = n;
    acts_.clear();
    z_.clear();
    dz_.clear();
    acts_.emplace_back(std::size_t(n) * kInPad);
    for (std::size_t l = 1; l < w_.size(); ++l) {
      acts_.emplace_back(std::size_t(n) * w_[l]);
      z_.emplace_back(std::size_t(n) * w_[l]);
      dz_.emplace_back(std::size_t(n) * w_[l]);
    }
    if (loss_sum_.size() == 0) loss_sum_ = Buf<double>(1);
  }

  template <bool TA, bool TB, int EPI>
  void gemm(const bf16* A, const bf16* B, int M, int N, int K, int lda, int ldb, float* Cf, bf16* Hb, const float* bias,
            int ldc, int splits = 1) {
    const int kchunk = ((cdiv(K, splits) + kTileK - 1) / kTileK) * kTileK;
    dim3 grid(cdiv(N, kTileN), cdiv(M, kTileM), cdiv(K, kchunk));
    k_gemm<TA, TB, EPI><<<grid, kThreads, 0, s_>>>(A, B, M, N, K, lda, ldb, kchunk, Cf, Hb, bias, ldc);
    check(hipGetLastError(), "k_gemm");
  }

  const bf16* W(int l) const { return wb_.get() + woff_[std::size_t(l)]; }
  const float* bias(int l) const { return params_.get() + boff_[std::size_t(l)]; }

  void forward(const float* x, int n) {
    ensure(n);
    k_inputs<<<cdiv(static_cast<long long>(n) * kInPad, 256), 256, 0, s_>>>(x, n, d_in_, norm_.get(), norm_.get() + 32,
                                                                             acts_[0].get(), kInPad);
    check(hipGetLastError(), "k_inputs");
    const int nl = layers();
    for (int l = 0; l < nl; ++l) {
      const int K = w_[std::size_t(l)], N = w_[std::size_t(l) + 1];
      if (l < nl - 1) {
        gemm<false, true, kBiasSiLU>(acts_[std::size_t(l)].get(), W(l), n, N, K, K, K, z_[std::size_t(l)].get(),
                                      acts_[std::size_t(l) + 1].get(), bias(l), N);
      } else {
        gemm<false, true, kBias>(acts_[std::size_t(l)].get(), W(l), n, N, K, K, K, z_[std::size_t(l)].get(), nullptr,
                                 bias(l), N);
      }
    }
  }


  // Forward and backward of one batch: gradients into grads_, the squared errors added to
  // loss_sum_ (cleared by the caller). No host synchronisation.
  void backward(const float* x, const float* t, int n) {
    forward(x, n);
    const int nl = layers();
    const int blocks = cdiv(static_cast<long long>(n) * kOutPad, 256);
    k_loss<<<blocks, 256, 0, s_>>>(z_.back().get(), t, n, d_out_, kOutPad, norm_.get() + 64, norm_.get() + 96,
                                  dz_.back().get(), loss_sum_.get());
    check(hipGetLastError(), "k_loss");
    k_fill<<<cdiv(np_, 256), 256, 0, s_>>>(grads_.get(), np_, 0.0f);  // a kernel: the step may be recorded
    check(hipGetLastError(), "k_fill");
    const int splits = std::max(1, std::min(32, n / 512));
    for (int l = nl - 1; l >= 0; --l) {
      const int K = w_[std::size_t(l)], N = w_[std::size_t(l) + 1];
      // dW_l = dZ_l^T H_{l-1} (N x K, over the samples), db_l = column sums of dZ_l.
      gemm<true, false, kAtomicF32>(dz_[std::size_t(l)].get(), acts_[std::size_t(l)].get(), N, K, n, N, K,
                                    grads_.get() + woff_[std::size_t(l)], nullptr, nullptr, K, splits);
      k_colsum<<<dim3(cdiv(N, 64), std::min(n, 256)), 64, 0, s_>>>(dz_[std::size_t(l)].get(), n, N,
                                                                    grads_.get() + boff_[std::size_t(l)]);
      check(hipGetLastError(), "k_colsum");
      if (l > 0) {
        // dZ_{l-1} = (dZ_l W_l) * silu'(Z_{l-1}).
        gemm<false, false, kSiLUGrad>(dz_[std::size_t(l)].get(), W(l), n, K, N, N, K, z_[std::size_t(l) - 1].get(),
                                      dz_[std::size_t(l) - 1].get(), nullptr, K);
      }
    }
  }

  void adam(float lr) {
    const float b1 = 0.9f, b2 = 0.999f;
    k_next_step<<<1, 1, 0, s_>>>(state_.get(), b1, b2);
    check(hipGetLastError(), "k_next_step");
    k_adam<<<cdiv(np_, 256), 256, 0, s_>>>(params_.get(), grads_.get(), m1_.get(), m2_.get(), mask_.get(), np_, lr, b1,
                                            b2, 1e-8f, state_.get(), wb_.get());
    check(hipGetLastError(), "k_adam");
  }

  int d_in_, d_out_;
  hipStream_t s_ = nullptr;
  std::vector<int> w_;
  std::vector<long long> woff_, boff_;
  long long np_ = 0;
  Buf<float> params_, grads_, m1_, m2_, norm_;
  Buf<StepState> state_;
  Buf<double> loss_sum_;
  Buf<bf16> wb_;
  Buf<unsigned char> mask_;
  std::vector<Buf<bf16>> acts_, dz_;
  std::vector<Buf<float>> z_;
  int cap_ = 0;
  bool fused_ = true;
  int infer_rows_ = 128;  // the fastest tile on an MI355X (eight waves, 4 waves per SIMD)
  Buf<float> data_x_, data_t_;
  long long n_data_ = 0;
};

int fail(const std::exception& e, char* err, size_t len) {
  if (err != nullptr && len > 0) {
    std::strncpy(err, e.what(), len - 1);
    err[len - 1] = '\0';
  }
  return 1;
}

template <class F>
int guarded(char* err, size_t len, F&& f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, len);
  }
}

}  // namespace mxm

struct mxmlp {
  std::unique_ptr<mxm::Net> net;
};

extern "C" {

int mxmlp_create(int32_t inputs, int32_t hidden, int32_t layers, int32_t outputs, uint64_t seed, mxmlp** out, char* err,
                 size_t err_len) {
  return mxm::guarded(err, err_len, [&] {
    if (!out) throw std::invalid_argument("null argument");
    auto h = std::make_unique<mxmlp>();
    h->net = std::make_unique<mxm::Net>(inputs, hidden, layers, outputs, seed);
    *out = h.release();
  });
}

void mxmlp_destroy(mxmlp* h) { delete h; }

int64_t mxmlp_param_count(const mxmlp* h) { return h ? h->net->param_count() : 0; }

int mxmlp_set_normalization(mxmlp* h, const float* in_mean, const float* in_scale, const float* out_mean,
                            const float* out_scale, char* err, size_t err_len) {
  return mxm::guarded(err, err_len, [&] { h->net->set_normalization(in_mean, in_scale, out_mean, out_scale); });
}

int mxmlp_get_params(mxmlp* h, float* params, char* err, size_t err_len) {
  return mxm::guarded(err, err_len, [&] { h->net->get_params(params); });
}

int mxmlp_set_params(mxmlp* h, const float* params, char* err, size_t err_len) {
  return mxm::guarded(err, err_len, [&] { h->net->set_params(params); });
}

int mxmlp_predict(mxmlp* h, const float* x, int64_t n, float* y, char* err, size_t err_len) {
  return mxm::guarded(err, err_len, [&] {
    if (n < 0) throw std::invalid_argument("n < 0");
    h->net->predict(x, n, y);
  });
}

int mxmlp_gradients(mxmlp* h, const float* x, const float* t, int32_t n, float* grads, float* loss, char* err,
                    size_t err_len) {
  return mxm::guarded(err, err_len, [&] {
    if (n < 1) throw std::invalid_argument("n >= 1");
    *loss = h->net->gradients(x, t, n, grads);
  });
}

int mxmlp_trace(mxmlp* h, const float* x, const float* t, int32_t n, int32_t what, int32_t layer, float* out,
                char* err, size_t err_len) {
  return mxm::guarded(err, err_len, [&] {
    if (n < 1 || !x || !t || !out) throw std::invalid_argument("trace: n >= 1 and arrays");
    h->net->trace(x, t, n, what, layer, out);
  });
}

int mxmlp_adam_step(mxmlp* h, float lr, char* err, size_t err_len) {
  return mxm::guarded(err, err_len, [&] { h->net->adam_step(lr); });
}

int mxmlp_set_data(mxmlp* h, const float* x, const float* t, int64_t n, char* err, size_t err_len) {
  return mxm::guarded(err, err_len, [&] { h->net->set_data(x, t, n); });
}

int mxmlp_train(mxmlp* h, int32_t epochs, int32_t batch, float lr, uint64_t seed, float* loss, double* seconds,
                char* err, size_t err_len) {
  return mxm::guarded(err, err_len, [&] { *loss = h->net->train(epochs, batch, lr, seed, seconds); });
}

int mxmlp_time_forward(mxmlp* h, int32_t n, int32_t repeats, double* ms, char* err, size_t err_len) {
  return mxm::guarded(err, err_len, [&] {
    if (n < 1 || repeats < 1) throw std::invalid_argument("n >= 1 and repeats >= 1");
    *ms = h->net->time_forward(n, repeats);
  });
}

int mxmlp_set_fused(mxmlp* h, int32_t on, char* err, size_t err_len) {
  return mxm::guarded(err, err_len, [&] { h->net->set_fused(on != 0); });
}

int mxmlp_set_infer_rows(mxmlp* h, int32_t rows, char* err, size_t err_len) {
  return mxm::guarded(err, err_len, [&] { h->net->set_infer_rows(rows); });
}

}  // extern "C"
