// Kernels of the force-matching training step's MLP part: a BF16-operand MFMA
// GEMM with the epilogues the step needs, the elementwise steps between its
// GEMMs, and row gathers / scatters.
//
// The step's second-order graph (loss_op_EandG and train_step_EandG,
// TensorMol/TFNetworks/TFMolInstanceDirect.py:4889-4903, 5288-5348; the force
// is tf.gradients(E, xyz), :5761) is taken apart by hand.  Per element, rows
// r of its atoms, layers z_l = h_{l-1} W_l + b_l, h_l = act(z_l) (l = 1..3),
// y = h_3 w_4 + b_4, act(z) = log(1 + e^{a z}) / a, s_l = act'(z_l) =
// sigmoid(a z_l), act'' = a s (1 - s):
//   J = dy/dG:      zd_3 = s_3 * w_4, hd_{l-1} = zd_l W_l^T, zd_{l-1} = s_{l-1} * hd_{l-1},
//                   J = zd_1 W_1^T;
//   tangent along V (= dLoss/dJ): zt_l = ht_{l-1} W_l, ht_l = s_l * zt_l, ht_0 = V;
//   reverse of Phi = ybar y + J . V over both (ybar = dLoss/dy):
//                   zb_3 = s_3 * (ybar w_4) + a s_3 (1 - s_3) * zt_3 * w_4,
//                   hb_{l-1} = zb_l W_l^T,
//                   zb_{l-1} = s_{l-1} * hb_{l-1} + a s (1 - s) * zt_{l-1} * hd_{l-1};
//   weight gradients: dW_l = h_{l-1}^T zb_l + ht_{l-1}^T zd_l, db_l = sum_r zb_l,
//                   dw_4 = h_3^T ybar + ht_3^T 1, db_4 = sum_r ybar.
// Each weight gradient is one GEMM over the stacked rows [h; ht]^T [zb; zd],
// with a column of ones appended to h (zeros to ht) so that the bias gradient
// comes out as the last row.
#include "tm_common.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int GT = 64;        // C tile edge per workgroup
constexpr int GK = 32;        // K step (one v_mfma_f32_16x16x32_bf16)
constexpr int LROW = GK + 8;  // LDS row of a staged operand, in bf16 (padded)

enum { EPI_STORE = 0, EPI_SOFTPLUS = 1, EPI_TANGENT = 2 };

struct GemmArgs {
  int M, N, K;
  const float* A;
  int lda, ta;        // A(m, k) = ta ? A[k * lda + m] : A[m * lda + k]
  const float* B;
  int ldb, tb;        // B(k, n) = tb ? B[n * ldb + k] : B[k * ldb + n]
  float* C;
  int ldc;
  int epi;
  const float* bias;  // [N] (EPI_STORE: may be NULL; EPI_SOFTPLUS: required)
  const float* s;     // EPI_TANGENT: aux = acc * s (s[m * ldc + n])
  float* aux;         // EPI_SOFTPLUS: sigmoid(a z); EPI_TANGENT: acc * s
  float alpha;
  int kchunk;         // K range of one split (blockIdx.z)
  long cstride;       // element offset between the splits' partial outputs
};

__device__ __forceinline__ unsigned short to_bf16(float x) {
  return __builtin_bit_cast(unsigned short, (__bf16)x);
}

// C tile (GT x GT) of one workgroup, four waves of 32 x 32 (2 x 2 MFMA tiles).
// Operands are staged through LDS as BF16, A as [m][k] and B as [n][k], so
// every MFMA fragment is one 16-byte read: lane l takes row l % 16 and k
// elements 8 (l / 16) .. 8 (l / 16) + 7; the accumulator holds rows
// 4 (l / 16) .. + 3 of column l % 16.
__global__ void __launch_bounds__(256) gemm_bf16_kernel(GemmArgs g) {
  __shared__ unsigned short As[GT][LROW];
  __shared__ unsigned short Bs[GT][LROW];
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
  const int k_begin = blockIdx.z * g.kchunk;
  const int k_end = min(g.K, k_begin + g.kchunk);
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  floatx4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = k_begin; k0 < k_end; k0 += GK) {
    // stage A (GT x GK) and B (GK x GT): 8 elements per thread each
    if (!g.ta) {   // contiguous along k
      const int r = tid >> 2, kk = (tid & 3) * 8;
      const int m = m0 + r;
      unsigned short v[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int k = k0 + kk + i;
        v[i] = to_bf16((m < g.M && k < k_end) ? g.A[(long)m * g.lda + k] : 0.f);
      }
#pragma unroll
      for (int i = 0; i < 8; ++i) As[r][kk + i] = v[i];
    } else {       // contiguous along m
      const int kk = tid >> 3, r = (tid & 7) * 8;
      const int k = k0 + kk;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int m = m0 + r + i;
        As[r + i][kk] = to_bf16((m < g.M && k < k_end) ? g.A[(long)k * g.lda + m] : 0.f);
      }
    }
    if (g.tb) {    // contiguous along k
      const int c = tid >> 2, kk = (tid & 3) * 8;
      const int n = n0 + c;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int k = k0 + kk + i;
        Bs[c][kk + i] = to_bf16((n < g.N && k < k_end) ? g.B[(long)n * g.ldb + k] : 0.f);
      }
    } else {       // contiguous along n
      const int kk = tid >> 3, c = (tid & 7) * 8;
      const int k = k0 + kk;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int n = n0 + c + i;
        Bs[c + i][kk] = to_bf16((n < g.N && k < k_end) ? g.B[(long)k * g.ldb + n] : 0.f);
      }
    }
    __syncthreads();
    const int fr = lane & 15, fk = (lane >> 4) * 8;
    uint4 a[2], b[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) a[i] = *reinterpret_cast<const uint4*>(&As[wm + 16 * i + fr][fk]);
#pragma unroll
    for (int j = 0; j < 2; ++j) b[j] = *reinterpret_cast<const uint4*>(&Bs[wn + 16 * j + fr][fk]);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a[i]),
                                                             __builtin_bit_cast(bf16x8, b[j]),
                                                             acc[i][j], 0, 0, 0);
    __syncthreads();
  }
  float* C = g.C + (long)blockIdx.z * g.cstride;
  const float a = g.alpha;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int n = n0 + wn + 16 * j + (lane & 15);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int m = m0 + wm + 16 * i + (lane >> 4) * 4 + q;
        if (m >= g.M || n >= g.N) continue;
        const long o = (long)m * g.ldc + n;
        float v = acc[i][j][q];
        if (g.epi == EPI_SOFTPLUS) {
          const float z = a * (v + g.bias[n]);
          // log(1 + e^z) / a, and its derivative sigmoid(z), without overflow
          C[o] = (fmaxf(z, 0.f) + log1pf(__expf(-fabsf(z)))) / a;
          g.aux[o] = 1.0f / (1.0f + __expf(-z));
        } else if (g.epi == EPI_TANGENT) {
          C[o] = v;
          g.aux[o] = v * g.s[o];
        } else {
          C[o] = g.bias != nullptr ? v + g.bias[n] : v;
        }
      }
    }
}

// C = sum over the splits of their partial outputs, in split order.
__global__ void split_sum_kernel(int M, int N, int ldc, int nsplit, long cstride,
                                 const float* __restrict__ part, float* __restrict__ C) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)M * N) return;
  const long o = (t / N) * ldc + t % N;
  float v = 0.f;
  for (int z = 0; z < nsplit; ++z) v += part[z * cstride + o];
  C[o] = v;
}

// Elementwise steps between the GEMMs (row r < M, column n < H; ld = row stride
// of every [M][H] operand):
//  op 0  zd = s * w4[n]                               (start of the J chain)
//  op 1  zd = s * hd                                   (J chain)
//  op 2  zb = s * ybar[r] * w4[n] + a s (1 - s) zt w4[n] (reverse, last layer)
//  op 3  zb = s * hb + a s (1 - s) zt hd               (reverse)
__global__ void train_elem_kernel(int op, int M, int H, int ld, float alpha,
                                  const float* __restrict__ s, const float* __restrict__ x,
                                  const float* __restrict__ zt, const float* __restrict__ hd,
                                  const float* __restrict__ w4, const float* __restrict__ ybar,
                                  float* __restrict__ out) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)M * H) return;
  const int r = (int)(t / H), n = (int)(t % H);
  const long o = (long)r * ld + n;
  const float sv = s[o];
  float v;
  if (op == 0) {
    v = sv * w4[n];
  } else if (op == 1) {
    v = sv * x[o];
  } else if (op == 2) {
    v = sv * ybar[r] * w4[n] + alpha * sv * (1.0f - sv) * zt[o] * w4[n];
  } else {
    v = sv * x[o] + alpha * sv * (1.0f - sv) * zt[o] * hd[o];
  }
  out[o] = v;
}

// out[r * ldo + c] = in[idx[r] * ldi + c] (gather), or in[r * ldi + c] to
// out[idx[r] * ldo + c] (scatter); c < cols.
__global__ void rows_kernel(int scatter, int M, int cols, const int* __restrict__ idx,
                            const float* __restrict__ in, int ldi, float* __restrict__ out,
                            int ldo) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)M * cols) return;
  const int r = (int)(t / cols), c = (int)(t % cols);
  if (scatter) out[(long)idx[r] * ldo + c] = in[(long)r * ldi + c];
  else out[(long)r * ldo + c] = in[(long)idx[r] * ldi + c];
}

}  // namespace

extern "C" {

int tm_gemm_bf16(int M, int N, int K, const float* A, int lda, int ta, const float* B, int ldb,
                 int tb, float* C, int ldc, int epi, const float* bias, const float* s,
                 float* aux, float alpha, int nsplit, float* work, void* stream) {
  TM_REQUIRE(M >= 0 && N >= 0 && K >= 0, "tm_gemm_bf16: negative size");
  TM_REQUIRE(epi >= EPI_STORE && epi <= EPI_TANGENT, "tm_gemm_bf16: bad epilogue %d", epi);
  TM_REQUIRE(epi != EPI_SOFTPLUS || (bias != nullptr && aux != nullptr && alpha > 0.f),
             "tm_gemm_bf16: the softplus epilogue needs bias, aux and alpha > 0");
  TM_REQUIRE(epi != EPI_TANGENT || (s != nullptr && aux != nullptr),
             "tm_gemm_bf16: the tangent epilogue needs s and aux");
  TM_REQUIRE(nsplit >= 1 && (nsplit == 1 || (epi == EPI_STORE && bias == nullptr &&
                                             work != nullptr)),
             "tm_gemm_bf16: split K needs the plain epilogue and a workspace");
  if (M == 0 || N == 0) return TM_OK;
  hipStream_t st = (hipStream_t)stream;
  GemmArgs g;
  g.M = M; g.N = N; g.K = K;
  g.A = A; g.lda = lda; g.ta = ta;
  g.B = B; g.ldb = ldb; g.tb = tb;
  g.C = nsplit > 1 ? work : C;
  g.ldc = ldc;
  g.epi = epi; g.bias = bias; g.s = s; g.aux = aux; g.alpha = alpha;
  // whole K steps per split
  const int steps = (K + GK - 1) / GK;
  g.kchunk = ((steps + nsplit - 1) / nsplit) * GK;
  const int nz = K == 0 ? 1 : (K + g.kchunk - 1) / g.kchunk;
  g.cstride = (long)M * ldc;
  dim3 grid(tm_div_up(N, GT), tm_div_up(M, GT), nz);
  gemm_bf16_kernel<<<grid, 256, 0, st>>>(g);
  TM_LAUNCH_CHECK();
  if (nsplit > 1) {
    split_sum_kernel<<<tm_div_up((long)M * N, 256), 256, 0, st>>>(M, N, ldc, nz, g.cstride,
                                                                  work, C);
    TM_LAUNCH_CHECK();
  }
  return TM_OK;
}

int tm_train_elem(int op, int M, int H, int ld, float alpha, const float* s, const float* x,
                  const float* zt, const float* hd, const float* w4, const float* ybar, float* out,
                  void* stream) {
  TM_REQUIRE(op >= 0 && op <= 3, "tm_train_elem: bad op %d", op);
  if ((long)M * H == 0) return TM_OK;
  train_elem_kernel<<<tm_div_up((long)M * H, 256), 256, 0, (hipStream_t)stream>>>(
      op, M, H, ld, alpha, s, x, zt, hd, w4, ybar, out);
  TM_LAUNCH_CHECK();
  return TM_OK;
}

int tm_rows(int scatter, int M, int cols, const int* idx, const float* in, int ldi, float* out,
            int ldo, void* stream) {
  if ((long)M * cols == 0) return TM_OK;
  rows_kernel<<<tm_div_up((long)M * cols, 256), 256, 0, (hipStream_t)stream>>>(
      scatter, M, cols, idx, in, ldi, out, ldo);
  TM_LAUNCH_CHECK();
  return TM_OK;
}

}  // extern "C"
