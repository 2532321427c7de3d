// Per-element MLPs in plain FP32 (VALU), layer by layer.
//
// Restates the energy and charge branches of energy_inference /
// dipole_inference (TensorMol/TFNetworks/TFMolInstanceDirect.py:5164-5212,
// 5219-5272): h_k = act(h_{k-1} W_k + b_k) for the hidden layers, y = h W_o + b_o,
// act(z) = log(1 + exp(alpha z)) / alpha (Util.py:200-201; alpha = 1 is
// tf.nn.softplus), and the input gradient dy/dG by the hand-written adjoint
//   dZ_last = w_o * act'(Z_last),  dH_{k-1} = dZ_k W_k^T,  dZ_{k-1} = dH_{k-1} * act'(Z_{k-1})
// with act'(z) = sigmoid(alpha z).  This is the plain FP32 path the BF16/FP16
// MFMA kernel is measured against; it stores every intermediate in HBM.
//
// Rows of one launch all belong to one (net, element) segment: row r of the
// segment is atom perm[row0 + r]; layer inputs/outputs are dense [rows, width]
// scratch matrices.
#include "tm_common.h"

namespace {

constexpr int TB = 64;     // tile rows and cols
constexpr int TK = 16;     // k step

__device__ __forceinline__ float softplus_a(float z, float alpha) {
  float x = alpha * z;
  // log(1 + e^x) evaluated stably
  return (fmaxf(x, 0.f) + log1pf(__expf(-fabsf(x)))) / alpha;
}
__device__ __forceinline__ float sigmoid_a(float z, float alpha) {
  return 1.0f / (1.0f + __expf(-alpha * z));
}

// C[M,N] = A[M,K] * op(B) with op(B) = B[K,N] (trans=0) or B[N,K]^T (trans=1).
// A rows optionally gathered through arow (A row r = arow[r]) with leading dim lda.
// epi 0: C = AB + bias, H = act(C) -> out0, S = act'(C) -> out1
// epi 1: out0 = (AB) * S (S = aux [M,N])
// epi 2: out0 = AB (scattered to rows orow[r] when orow != null)
__global__ void __launch_bounds__(256)
gemm_f32_kernel(int M, int N, int K, const float* __restrict__ A, int lda,
                const int* __restrict__ arow, const float* __restrict__ B, int trans,
                const float* __restrict__ bias, int epi, float alpha,
                float* __restrict__ out0, float* __restrict__ out1, int ldo,
                const float* __restrict__ aux, const int* __restrict__ orow) {
  __shared__ float As[TK][TB + 4];
  __shared__ float Bs[TK][TB + 4];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int row0 = blockIdx.y * TB, col0 = blockIdx.x * TB;
  float acc[4][4] = {};
  for (int k0 = 0; k0 < K; k0 += TK) {
    // load A tile: 64 rows x 16 k, 256 threads -> 4 each
    for (int e = threadIdx.x; e < TB * TK; e += 256) {
      int r = e / TK, k = e % TK;
      int gr = row0 + r, gk = k0 + k;
      float v = 0.f;
      if (gr < M && gk < K) {
        long ar = arow ? arow[gr] : gr;
        v = A[ar * lda + gk];
      }
      As[k][r] = v;
    }
    for (int e = threadIdx.x; e < TB * TK; e += 256) {
      int k = e / TB, c = e % TB;
      int gk = k0 + k, gc = col0 + c;
      float v = 0.f;
      if (gk < K && gc < N) v = trans ? B[(long)gc * K + gk] : B[(long)gk * N + gc];
      Bs[k][c] = v;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < TK; ++k) {
      float a[4], b[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) a[u] = As[k][ty * 4 + u];
#pragma unroll
      for (int v = 0; v < 4; ++v) b[v] = Bs[k][tx * 4 + v];
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] += a[u] * b[v];
    }
    __syncthreads();
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    int r = row0 + ty * 4 + u;
    if (r >= M) continue;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      int c = col0 + tx * 4 + v;
      if (c >= N) continue;
      float z = acc[u][v];
      if (epi == 0) {
        z += bias[c];
        out0[(long)r * ldo + c] = softplus_a(z, alpha);
        out1[(long)r * ldo + c] = sigmoid_a(z, alpha);
      } else if (epi == 1) {
        out0[(long)r * ldo + c] = z * aux[(long)r * ldo + c];
      } else {
        long orr = orow ? orow[r] : r;
        out0[orr * ldo + c] = z;
      }
    }
  }
}

// y[r] = h[r,:] . w + b ; dZ[r,:] = w * S[r,:] ; y scattered to atom orow[r]
__global__ void head_kernel(int M, int H, const float* __restrict__ h, const float* __restrict__ S,
                            const float* __restrict__ w, const float* __restrict__ b,
                            const int* __restrict__ orow,
                            float* __restrict__ y, float* __restrict__ dZ) {
  const int lane = threadIdx.x & 63;
  const int r = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (r >= M) return;
  float acc = 0.f;
  for (int c = lane; c < H; c += 64) {
    float wc = w[c];
    acc += h[(long)r * H + c] * wc;
    dZ[(long)r * H + c] = wc * S[(long)r * H + c];
  }
  acc = tm_wave_sum(acc);
  if (lane == 0) y[orow[r]] = acc + b[0];
}

}  // namespace

extern "C" {

// One (net, element) segment.  W: packed float32 [W1 (D x H1), b1, W2 (H1 x H2), b2,
// W3 (H2 x H3), b3, wo (H3), bo] for nh = 3 hidden layers (general nh <= 4 with
// widths hid[0..nh-1]).  rows = atoms perm[0..M).  Outputs y[atom], dG[atom*ldg + f].
// ws: scratch of 2*nh*M*Hmax + 2*M*Hmax floats.
int tm_mlp_f32(int M, int D, int nh, const int* hid, const float* G, int ldg, const int* perm,
               const float* W, float alpha, float* y, float* dG, float* ws, void* stream) {
  TM_REQUIRE(M >= 0 && D > 0 && nh >= 1 && nh <= 4, "tm_mlp_f32: bad sizes");
  if (M == 0) return TM_OK;
  hipStream_t s = (hipStream_t)stream;
  int Hmax = 0;
  for (int k = 0; k < nh; ++k) Hmax = hid[k] > Hmax ? hid[k] : Hmax;
  // weight offsets
  const float* Wk[5];
  const float* bk[5];
  long o = 0;
  int in = D;
  for (int k = 0; k < nh; ++k) {
    Wk[k] = W + o; o += (long)in * hid[k];
    bk[k] = W + o; o += hid[k];
    in = hid[k];
  }
  const float* wo = W + o;
  o += in;
  const float* bo = W + o;
  float* Hs[4];
  float* Ss[4];
  for (int k = 0; k < nh; ++k) {
    Hs[k] = ws + (long)(2 * k) * M * Hmax;
    Ss[k] = ws + (long)(2 * k + 1) * M * Hmax;
  }
  float* dA = ws + (long)(2 * nh) * M * Hmax;
  float* dB = dA + (long)M * Hmax;
  dim3 blk(256);
  // forward
  in = D;
  for (int k = 0; k < nh; ++k) {
    dim3 grid(tm_div_up(hid[k], TB), tm_div_up(M, TB));
    const float* A = k == 0 ? G : Hs[k - 1];
    int la = k == 0 ? ldg : hid[k - 1];
    gemm_f32_kernel<<<grid, blk, 0, s>>>(M, hid[k], in, A, la, k == 0 ? perm : nullptr, Wk[k], 0,
                                         bk[k], 0, alpha, Hs[k], Ss[k], hid[k], nullptr, nullptr);
    TM_LAUNCH_CHECK();
    in = hid[k];
  }
  head_kernel<<<tm_div_up(M, 4), 256, 0, s>>>(M, in, Hs[nh - 1], Ss[nh - 1], wo, bo, perm, y, dA);
  TM_LAUNCH_CHECK();
  // backward: dZ_k in dA -> dH_{k-1} = dZ_k W_k^T -> dZ_{k-1} = dH * S_{k-1}
  float* cur = dA;
  float* nxt = dB;
  for (int k = nh - 1; k >= 1; --k) {
    dim3 grid(tm_div_up(hid[k - 1], TB), tm_div_up(M, TB));
    gemm_f32_kernel<<<grid, blk, 0, s>>>(M, hid[k - 1], hid[k], cur, hid[k], nullptr, Wk[k], 1,
                                         nullptr, 1, alpha, nxt, nullptr, hid[k - 1], Ss[k - 1],
                                         nullptr);
    TM_LAUNCH_CHECK();
    float* t = cur; cur = nxt; nxt = t;
  }
  dim3 grid(tm_div_up(D, TB), tm_div_up(M, TB));
  gemm_f32_kernel<<<grid, blk, 0, s>>>(M, D, hid[0], cur, hid[0], nullptr, Wk[0], 1, nullptr, 2,
                                       alpha, dG, nullptr, ldg, nullptr, perm);
  TM_LAUNCH_CHECK();
  return TM_OK;
}

}  // extern "C"
