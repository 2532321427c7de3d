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
#include <algorithm>
#include <type_traits>

#include "tensormol_amd.h"
#include "tm_common.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int GT = 128;       // C tile edge per workgroup (four waves of 64 x 64)
constexpr int GK = 64;        // K per staged step (two v_mfma_f32_16x16x32_bf16 deep)
constexpr int LROW = GK + 8;  // LDS row of a staged operand, in bf16 (144 B, padded)
constexpr int MAXG = 16;      // GEMMs per grouped launch
constexpr int NXCD = 8;       // XCDs; workgroups are dealt to them round-robin

enum { EPI_STORE = 0, EPI_SOFTPLUS = 1, EPI_TANGENT = 2, EPI_REVERSE = 3 };

// One GEMM of a grouped launch: the caller's tm_gemm_group and its tiling.
struct GemmDev {
  const float* A;
  const float* B;
  float* C;            // where the GEMM kernel writes: out, or the split partials
  float* out;
  const float* bias;
  const float* s;
  float* aux;
  const float* x1;
  const float* x2;
  float* out2;
  const int* crow;     // output row of GEMM row m (plain epilogue), or NULL
  long cstride;        // floats between the splits' partial outputs
  int M, N, K, lda, ldb, ldc;
  int kchunk;          // K range of one split (a multiple of GK)
  int vec_out;         // every row-indexed epilogue array takes 16-byte accesses
  int tiles_n;         // column tiles
  int tile0;           // first linear tile of this GEMM
};

struct GemmBatch {
  GemmDev g[MAXG];
  int ngroups, ntiles, per_xcd, epi, io;
  float alpha;
};

__device__ __forceinline__ unsigned bf16_bits(float x) {
  return __builtin_bit_cast(unsigned short, (__bf16)x);
}

__device__ __forceinline__ uint4 pack8(const float* v) {
  return make_uint4(bf16_bits(v[0]) | (bf16_bits(v[1]) << 16),
                    bf16_bits(v[2]) | (bf16_bits(v[3]) << 16),
                    bf16_bits(v[4]) | (bf16_bits(v[5]) << 16),
                    bf16_bits(v[6]) | (bf16_bits(v[7]) << 16));
}

// One operand's share of a K step (GT rows x GK), 32 values per thread, held
// in registers from its global load to its LDS store.  Both forms leave the
// operand in LDS as S[row][k] in BF16, so an MFMA fragment (8 consecutive k
// of one row) is one 16-byte read, and each thread's LDS writes are 16 bytes.
// KCONT, element (r, k) at X[r * ld + k]: thread t takes k = 8 (t % 8) .. + 7
// of rows t / 8 + 32 i (VEC: two 16-byte loads per row, which needs ld % 4 == 0
// and a 16-byte aligned X).
// Otherwise element (r, k) at X[k * ld + r]: thread t takes row t % 128 at
// k = 8 (t / 128 + 2 i) .. + 7, one 4-byte load per k, so a wave's load reads
// 256 contiguous bytes.
// Loads are branch-free: a row past the operand's last is read as its last
// row (it only feeds output rows or columns that are never stored), and in
// the last, partial K step a k past kend is read at kend - 1 and replaced by
// zero.  Every address stays inside the operand.
template <bool KCONT, bool VEC>
struct Stage {
  float v[32];

  template <bool TAIL>
  __device__ __forceinline__ void load(const float* __restrict__ X, long ld, int rows, int r0,
                                       int k0, int kend, int tid) {
    if (KCONT) {
      const int k = k0 + (tid & 7) * 8;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = min(r0 + (tid >> 3) + 32 * i, rows - 1);
        const float* p = X + (long)r * ld;
        float* d = v + 8 * i;
        if (VEC && !TAIL) {
          const float4 f0 = *reinterpret_cast<const float4*>(p + k);
          const float4 f1 = *reinterpret_cast<const float4*>(p + k + 4);
          d[0] = f0.x; d[1] = f0.y; d[2] = f0.z; d[3] = f0.w;
          d[4] = f1.x; d[5] = f1.y; d[6] = f1.z; d[7] = f1.w;
        } else {
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            const float x = p[TAIL ? min(k + q, kend - 1) : k + q];
            d[q] = (!TAIL || k + q < kend) ? x : 0.f;
          }
        }
      }
    } else {
      const float* p = X + min(r0 + (tid & 127), rows - 1);
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const int k = k0 + 8 * ((tid >> 7) + 2 * i) + q;
          const float x = p[(long)(TAIL ? min(k, kend - 1) : k) * ld];
          v[8 * i + q] = (!TAIL || k < kend) ? x : 0.f;
        }
    }
  }

  __device__ __forceinline__ void store(unsigned short (*S)[LROW], int tid) const {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int row = KCONT ? (tid >> 3) + 32 * i : (tid & 127);
      const int col = KCONT ? (tid & 7) * 8 : 8 * ((tid >> 7) + 2 * i);
      *reinterpret_cast<uint4*>(&S[row][col]) = pack8(v + 8 * i);
    }
  }
};

// The same staging for operands stored as BF16 (the bits the FP32 path
// would stage): KCONT with VEC is one 16-byte load per row, stored as is.
template <bool KCONT, bool VEC>
struct StageBF {
  uint4 v[4];

  template <bool TAIL>
  __device__ __forceinline__ void load(const unsigned short* __restrict__ X, long ld, int rows,
                                       int r0, int k0, int kend, int tid) {
    if (KCONT) {
      const int k = k0 + (tid & 7) * 8;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = min(r0 + (tid >> 3) + 32 * i, rows - 1);
        const unsigned short* p = X + (long)r * ld;
        if (VEC && !TAIL) {
          v[i] = *reinterpret_cast<const uint4*>(p + k);
        } else {
          unsigned h[8];
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            const unsigned x = p[TAIL ? min(k + q, kend - 1) : k + q];
            h[q] = (!TAIL || k + q < kend) ? x : 0u;
          }
          v[i] = make_uint4(h[0] | (h[1] << 16), h[2] | (h[3] << 16), h[4] | (h[5] << 16),
                            h[6] | (h[7] << 16));
        }
      }
    } else {
      const unsigned short* p = X + min(r0 + (tid & 127), rows - 1);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        unsigned h[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const int k = k0 + 8 * ((tid >> 7) + 2 * i) + q;
          const unsigned x = p[(long)(TAIL ? min(k, kend - 1) : k) * ld];
          h[q] = (!TAIL || k < kend) ? x : 0u;
        }
        v[i] = make_uint4(h[0] | (h[1] << 16), h[2] | (h[3] << 16), h[4] | (h[5] << 16),
                          h[6] | (h[7] << 16));
      }
    }
  }

  __device__ __forceinline__ void store(unsigned short (*S)[LROW], int tid) const {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int row = KCONT ? (tid >> 3) + 32 * i : (tid & 127);
      const int col = KCONT ? (tid & 7) * 8 : 8 * ((tid >> 7) + 2 * i);
      *reinterpret_cast<uint4*>(&S[row][col]) = v[i];
    }
  }
};

constexpr int CLD = GT + 4;   // LDS row of the staged C tile, in floats

// Four consecutive values of row o .. o + 3 (nv of them valid).
__device__ __forceinline__ void ld4(const float* __restrict__ p, long o, int nv, bool vec,
                                    float (&v)[4]) {
  if (vec && nv == 4) {
    const float4 f = *reinterpret_cast<const float4*>(p + o);
    v[0] = f.x; v[1] = f.y; v[2] = f.z; v[3] = f.w;
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = k < nv ? p[o + k] : 0.f;
  }
}

__device__ __forceinline__ void st4(float* __restrict__ p, long o, int nv, bool vec,
                                    const float (&v)[4]) {
  if (vec && nv == 4) {
    *reinterpret_cast<float4*>(p + o) = make_float4(v[0], v[1], v[2], v[3]);
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (k < nv) p[o + k] = v[k];
  }
}

// st4 for a BF16 array (round to nearest even, as the staging does).
__device__ __forceinline__ void st4h(unsigned short* __restrict__ p, long o, int nv, bool vec,
                                     const float (&v)[4]) {
  if (vec && nv == 4) {
    *reinterpret_cast<uint2*>(p + o) = make_uint2(bf16_bits(v[0]) | (bf16_bits(v[1]) << 16),
                                                  bf16_bits(v[2]) | (bf16_bits(v[3]) << 16));
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (k < nv) p[o + k] = (unsigned short)bf16_bits(v[k]);
  }
}

// Epilogue of one C tile staged in LDS (Cs[row * CLD + col]): thread t takes
// columns 4 (t % 32) .. + 3 of rows t / 32 + 8 r.  The outputs (C, aux, out2)
// never overlap the inputs (bias, s, x1, x2), which restrict states, so a
// row's loads need not wait for the previous row's stores.
//  EPI_STORE    C = acc + bias[n]
//  EPI_SOFTPLUS z = a (acc + bias[n]), C = log(1 + e^z) / a, aux = sigmoid(z);
//               out2 (if given) = aux * x1[n]
//  EPI_TANGENT  C (if given) = acc, aux = acc * s; out2 (if given) =
//               s * x2[m] * x1[n] + a s (1 - s) * acc * x1[n]
//  EPI_REVERSE  C = s * acc + a s (1 - s) * x1 * x2  (s, x1, x2 indexed as C)
__device__ __forceinline__ void epilogue_tile(
    const float* Cs, int m0, int n0, int M, int N, long ldc, int epi, float a, bool vec, int io,
    const int* __restrict__ crow,
    float* __restrict__ C, float* __restrict__ aux, float* __restrict__ out2,
    const float* __restrict__ bias, const float* __restrict__ s,
    const float* __restrict__ x1, const float* __restrict__ x2, int tid) {
  const int c = (tid & 31) * 4, n = n0 + c;
  if (n >= N) return;
  const int nv = min(4, N - n);
  float bv[4], wv[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    bv[k] = (bias != nullptr && k < nv) ? bias[n + k] : 0.f;
    wv[k] = (out2 != nullptr && k < nv) ? x1[n + k] : 0.f;
  }
  const float ia = 1.0f / a;
#pragma unroll 2
  for (int r = tid >> 5; r < GT; r += 8) {
    const int m = m0 + r;
    if (m >= M) break;
    const long o = (long)(crow != nullptr ? crow[m] : m) * ldc + n;
    float v[4], sv[4] = {0.f, 0.f, 0.f, 0.f}, u1[4] = {0.f, 0.f, 0.f, 0.f},
                u2[4] = {0.f, 0.f, 0.f, 0.f}, r0[4], r1[4], r2[4];
    const float4 cv = *reinterpret_cast<const float4*>(Cs + r * CLD + c);
    v[0] = cv.x; v[1] = cv.y; v[2] = cv.z; v[3] = cv.w;
    if (epi == EPI_TANGENT || epi == EPI_REVERSE) ld4(s, o, nv, vec, sv);
    if (epi == EPI_REVERSE) {
      ld4(x1, o, nv, vec, u1);
      ld4(x2, o, nv, vec, u2);
    }
    const float ym = (epi == EPI_TANGENT && out2 != nullptr) ? x2[m] : 0.f;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (epi == EPI_SOFTPLUS) {
        const float z = a * (v[k] + bv[k]);
        // log(1 + e^z) / a and its derivative sigmoid(z), without overflow
        const float e = __expf(-fabsf(z));
        r0[k] = (fmaxf(z, 0.f) + __logf(1.0f + e)) * ia;
        r1[k] = z >= 0.f ? __fdividef(1.0f, 1.0f + e) : __fdividef(e, 1.0f + e);
        r2[k] = r1[k] * wv[k];
      } else if (epi == EPI_TANGENT) {
        const float sg = sv[k];
        r0[k] = v[k];
        r1[k] = v[k] * sg;
        r2[k] = sg * ym * wv[k] + a * sg * (1.0f - sg) * v[k] * wv[k];
      } else if (epi == EPI_REVERSE) {
        const float sg = sv[k];
        r0[k] = sg * v[k] + a * sg * (1.0f - sg) * u1[k] * u2[k];
      } else {
        r0[k] = v[k] + bv[k];
      }
    }
    // io bits 1, 2, 3: C, aux, out2 stored as BF16
    if (C != nullptr) {
      if (io & 2) st4h(reinterpret_cast<unsigned short*>(C), o, nv, vec, r0);
      else st4(C, o, nv, vec, r0);
    }
    if (epi == EPI_SOFTPLUS || epi == EPI_TANGENT) {
      if (io & 4) st4h(reinterpret_cast<unsigned short*>(aux), o, nv, vec, r1);
      else st4(aux, o, nv, vec, r1);
    }
    if (out2 != nullptr) {
      if (io & 8) st4h(reinterpret_cast<unsigned short*>(out2), o, nv, vec, r2);
      else st4(out2, o, nv, vec, r2);
    }
  }
}

// C tile (GT x GT) of one GEMM of the batch per workgroup, four waves of
// 64 x 64 (4 x 4 MFMA tiles); blockIdx.y is the K split.  Operands are staged
// through two LDS buffers as BF16 [row][k]; MFMA fragment: lane l takes row
// l % 16 and k elements 8 (l / 16) .. + 7; the accumulator holds rows
// 4 (l / 16) .. + 3 of column l % 16.  The next K step's operands are loaded
// into registers while the current one is multiplied.
// Workgroups are dealt to the XCDs round-robin, so workgroup w takes tile
// (w % 8) * per_xcd + w / 8: each XCD runs a contiguous run of tiles, and the
// column tiles of a row block (which share their A rows) meet in one L2.
template <bool TA, bool TB, bool VEC, bool BF>
__global__ void __launch_bounds__(256) gemm_bf16_kernel(const GemmBatch gb) {
  // two staged K steps of A and B, and after the K loop the C tile
  __shared__ __align__(16) unsigned char lds[2 * 2 * GT * LROW * 2];
  auto As = reinterpret_cast<unsigned short (*)[GT][LROW]>(lds);
  auto Bs = reinterpret_cast<unsigned short (*)[GT][LROW]>(lds + 2 * GT * LROW * 2);
  const int t = (blockIdx.x % NXCD) * gb.per_xcd + blockIdx.x / NXCD;
  if (t >= gb.ntiles) return;
  int gi = 0;
  while (gi + 1 < gb.ngroups && t >= gb.g[gi + 1].tile0) ++gi;
  const GemmDev& g = gb.g[gi];
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int local = t - g.tile0;
  const int m0 = (local / g.tiles_n) * GT, n0 = (local % g.tiles_n) * GT;
  const int k_begin = blockIdx.y * g.kchunk;
  const int k_end = min(g.K, k_begin + g.kchunk);
  const int nsteps = k_end > k_begin ? (k_end - k_begin + GK - 1) / GK : 0;
  const int wm = (w >> 1) * 64, wn = (w & 1) * 64;
  floatx4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
  // A(m, k): contiguous along k unless transposed; B(k, n) as rows n:
  // contiguous along k when transposed.  BF: both stored as BF16.
  typedef typename std::conditional<BF, StageBF<!TA, VEC>, Stage<!TA, VEC>>::type SA;
  typedef typename std::conditional<BF, StageBF<TB, VEC>, Stage<TB, VEC>>::type SB;
  typedef typename std::conditional<BF, unsigned short, float>::type Elem;
  SA sa;
  SB sb;
  const Elem* Ap = reinterpret_cast<const Elem*>(g.A);
  const Elem* Bp = reinterpret_cast<const Elem*>(g.B);
  auto load = [&](int k0) {
    if (k0 + GK <= k_end) {
      sa.template load<false>(Ap, g.lda, g.M, m0, k0, k_end, tid);
      sb.template load<false>(Bp, g.ldb, g.N, n0, k0, k_end, tid);
    } else {
      sa.template load<true>(Ap, g.lda, g.M, m0, k0, k_end, tid);
      sb.template load<true>(Bp, g.ldb, g.N, n0, k0, k_end, tid);
    }
  };
  if (nsteps > 0) load(k_begin);
  const int fr = lane & 15, fk = (lane >> 4) * 8;
  for (int step = 0; step < nsteps; ++step) {
    const int buf = step & 1;
    sa.store(As[buf], tid);
    sb.store(Bs[buf], tid);
    __syncthreads();
    if (step + 1 < nsteps) load(k_begin + (step + 1) * GK);
#pragma unroll
    for (int kk = 0; kk < GK; kk += 32) {
      uint4 a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i)
        a[i] = *reinterpret_cast<const uint4*>(&As[buf][wm + 16 * i + fr][kk + fk]);
#pragma unroll
      for (int j = 0; j < 4; ++j)
        b[j] = *reinterpret_cast<const uint4*>(&Bs[buf][wn + 16 * j + fr][kk + fk]);
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
              __builtin_bit_cast(bf16x8, a[i]), __builtin_bit_cast(bf16x8, b[j]), acc[i][j], 0,
              0, 0);
    }
  }
  // Epilogue: the C tile goes through LDS, then epilogue_tile() applies the
  // epilogue in a rolled loop with coalesced rows (a fully unrolled per-lane
  // epilogue is tens of KB of straight-line code that every workgroup
  // fetches once, which costs more than the tile's arithmetic).
  __syncthreads();   // every wave is done with the staged operands
  float* Cs = reinterpret_cast<float*>(lds);
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q)
        Cs[(wm + 16 * i + (lane >> 4) * 4 + q) * CLD + wn + 16 * j + (lane & 15)] = acc[i][j][q];
  __syncthreads();
  epilogue_tile(Cs, m0, n0, g.M, g.N, g.ldc, gb.epi, gb.alpha, g.vec_out != 0, gb.io, g.crow,
                g.C != nullptr ? g.C + (long)blockIdx.y * g.cstride : nullptr, g.aux, g.out2,
                g.bias, g.s, g.x1, g.x2, tid);
}

// out = sum over the splits of their partial outputs, in split order; one
// GEMM of the batch per blockIdx.y.
__global__ void split_sum_kernel(const GemmBatch gb, int nsplit) {
  const GemmDev& g = gb.g[blockIdx.y];
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)g.M * g.N) return;
  const long o = (t / g.N) * g.ldc + t % g.N;
  float v = 0.f;
  for (int z = 0; z < nsplit; ++z) v += g.C[z * g.cstride + o];
  g.out[o] = v;
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

// rows_kernel over several (idx, in, out) sets at once, one per blockIdx.y.
struct RowsBatch {
  tm_rows_group r[MAXG];
  int scatter;
};

// Four rows per 256-thread block, 64 lanes per row walking its columns.
__global__ void rows_grouped_kernel(const RowsBatch rb) {
  const tm_rows_group& d = rb.r[blockIdx.y];
  const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= d.M) return;
  const long src = rb.scatter ? (long)r * d.ldi : (long)d.idx[r] * d.ldi;
  const long dst = rb.scatter ? (long)d.idx[r] * d.ldo : (long)r * d.ldo;
  if (d.out_bf16) {
    unsigned short* out = static_cast<unsigned short*>(d.out);
    for (int c = threadIdx.x & 63; c < d.cols; c += 64)
      out[dst + c] = (unsigned short)bf16_bits(d.in[src + c]);
  } else {
    float* out = static_cast<float*>(d.out);
    for (int c = threadIdx.x & 63; c < d.cols; c += 64) out[dst + c] = d.in[src + c];
  }
}

// out = BF16 of in (round to nearest even) for several arrays, one per blockIdx.y.
struct CastBatch {
  tm_cast_group c[MAXG];
};

__global__ void cast_bf16_kernel(const CastBatch cb) {
  const tm_cast_group& d = cb.c[blockIdx.y];
  unsigned short* out = static_cast<unsigned short*>(d.out);
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < d.n;
       t += (long)gridDim.x * blockDim.x)
    out[t] = (unsigned short)bf16_bits(d.in[t]);
}

// Force-matching loss terms, one wave per molecule m (atoms s .. e - 1):
//   E[m] = sum of eps over its atoms (FP64, lane sums then a fixed butterfly),
//   dE = E[m] - E_ref[m], ybar[i] = 2 w_e[m] dE,
//   F = -grad, gg[i] = -2 w_f[m] (F_i - F_ref_i) (= dLoss/dgrad),
//   part[m] = w_e[m] dE^2 + sum_i w_f[m] |F_i - F_ref_i|^2.
__global__ void fm_loss_kernel(int B, const int* __restrict__ mol_start,
                               const float* __restrict__ eps, const double* __restrict__ E_ref,
                               const double* __restrict__ w_e, const double* __restrict__ grad,
                               const double* __restrict__ F_ref, const double* __restrict__ w_f,
                               double* __restrict__ E, float* __restrict__ ybar,
                               double* __restrict__ gg, double* __restrict__ part) {
  const int m = blockIdx.x, lane = threadIdx.x;
  if (m >= B) return;
  const int s = mol_start[m], e = mol_start[m + 1];
  double esum = 0.0;
  for (int i = s + lane; i < e; i += 64) esum += (double)eps[i];
  esum = tm_wave_sum_d(esum);
  const double dE = esum - E_ref[m], we = w_e[m], wf = w_f[m];
  const float yb = (float)(2.0 * we * dE);
  double fsum = 0.0;
  for (int i = s + lane; i < e; i += 64) {
    ybar[i] = yb;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double d = -grad[3 * (long)i + c] - F_ref[3 * (long)i + c];
      gg[3 * (long)i + c] = -2.0 * wf * d;
      fsum += wf * d * d;
    }
  }
  fsum = tm_wave_sum_d(fsum);
  if (lane == 0) {
    E[m] = esum;
    part[m] = we * dE * dE + fsum;
  }
}

// L = sum of part[0 .. B) in a fixed order (one workgroup).
__global__ void fm_loss_sum_kernel(int B, const double* __restrict__ part, double* __restrict__ L) {
  __shared__ double red[256];
  double v = 0.0;
  for (int m = threadIdx.x; m < B; m += 256) v += part[m];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) L[0] = red[0];
}

}  // namespace

extern "C" {

int tm_gemm_bf16_grouped(int ngroups, const tm_gemm_group* groups, int ta, int tb, int epi,
                         int io, float alpha, int nsplit, void* stream) {
  TM_REQUIRE(ngroups >= 0 && (ngroups == 0 || groups != nullptr),
             "tm_gemm_bf16_grouped: bad group list");
  TM_REQUIRE(epi >= EPI_STORE && epi <= EPI_REVERSE, "tm_gemm_bf16: bad epilogue %d", epi);
  TM_REQUIRE((io & ~15) == 0, "tm_gemm_bf16: bad io flags %d", io);
  TM_REQUIRE(nsplit >= 1 && nsplit <= 64, "tm_gemm_bf16: nsplit %d outside 1..64", nsplit);
  TM_REQUIRE(nsplit == 1 || (epi == EPI_STORE && (io & 2) == 0),
             "tm_gemm_bf16: split K needs the plain epilogue and an FP32 C");
  TM_REQUIRE(epi != EPI_SOFTPLUS || alpha > 0.f, "tm_gemm_bf16: softplus needs alpha > 0");
  for (int i = 0; i < ngroups; ++i) {
    const tm_gemm_group& q = groups[i];
    TM_REQUIRE(q.M >= 0 && q.N >= 0 && q.K >= 0, "tm_gemm_bf16: negative size in GEMM %d", i);
    if (q.M == 0 || q.N == 0) continue;   // nothing to write, no operand read
    TM_REQUIRE(epi != EPI_SOFTPLUS || (q.bias != nullptr && q.aux != nullptr),
               "tm_gemm_bf16: the softplus epilogue needs bias and aux (GEMM %d)", i);
    TM_REQUIRE(epi != EPI_TANGENT || (q.s != nullptr && q.aux != nullptr),
               "tm_gemm_bf16: the tangent epilogue needs s and aux (GEMM %d)", i);
    TM_REQUIRE(epi != EPI_REVERSE || (q.s != nullptr && q.x1 != nullptr && q.x2 != nullptr &&
                                      q.C != nullptr),
               "tm_gemm_bf16: the reverse epilogue needs s, x1, x2 and C (GEMM %d)", i);
    TM_REQUIRE(q.out2 == nullptr ||
                   (epi == EPI_SOFTPLUS && q.x1 != nullptr) ||
                   (epi == EPI_TANGENT && q.x1 != nullptr && q.x2 != nullptr),
               "tm_gemm_bf16: out2 needs the softplus (x1) or tangent (x1, x2) epilogue "
               "(GEMM %d)", i);
    TM_REQUIRE(q.C != nullptr || epi == EPI_TANGENT,
               "tm_gemm_bf16: C may be NULL only with the tangent epilogue (GEMM %d)", i);
    TM_REQUIRE(nsplit == 1 || (q.bias == nullptr && q.work != nullptr),
               "tm_gemm_bf16: split K needs a workspace and no bias (GEMM %d)", i);
    TM_REQUIRE(q.crow == nullptr || (epi == EPI_STORE && nsplit == 1),
               "tm_gemm_bf16: a row map needs the plain epilogue without split K (GEMM %d)", i);
  }
  hipStream_t st = (hipStream_t)stream;
  const bool bf = (io & 1) != 0;
  auto aligned = [](const void* p, int bytes) {
    return (reinterpret_cast<uintptr_t>(p) & (uintptr_t)(bytes - 1)) == 0;
  };
  // 16-byte loads of the operands stored along k when every one allows them
  bool vec = true;
  const int per16 = bf ? 8 : 4;   // elements per 16 bytes
  for (int i = 0; i < ngroups; ++i) {
    const tm_gemm_group& q = groups[i];
    if (q.M == 0 || q.N == 0) continue;
    if (!ta) vec = vec && q.lda % per16 == 0 && aligned(q.A, 16);
    if (tb) vec = vec && q.ldb % per16 == 0 && aligned(q.B, 16);
  }
  using Kern = void (*)(const GemmBatch);
#define TM_GEMM_K(a, b)                                                                  \
  {gemm_bf16_kernel<a, b, false, false>, gemm_bf16_kernel<a, b, false, true>,            \
   gemm_bf16_kernel<a, b, true, false>, gemm_bf16_kernel<a, b, true, true>}
  const Kern kerns[2][2][4] = {{TM_GEMM_K(false, false), TM_GEMM_K(false, true)},
                               {TM_GEMM_K(true, false), TM_GEMM_K(true, true)}};
#undef TM_GEMM_K
  const Kern kern = kerns[ta ? 1 : 0][tb ? 1 : 0][(vec ? 2 : 0) + (bf ? 1 : 0)];
  for (int g0 = 0; g0 < ngroups;) {
    GemmBatch gb;
    gb.ngroups = 0;
    gb.ntiles = 0;
    gb.epi = epi;
    gb.io = io;
    gb.alpha = alpha;
    long max_mn = 0;
    for (; g0 < ngroups && gb.ngroups < MAXG; ++g0) {
      const tm_gemm_group& q = groups[g0];
      if (q.M == 0 || q.N == 0) continue;
      GemmDev& d = gb.g[gb.ngroups++];
      d.A = static_cast<const float*>(q.A);
      d.B = static_cast<const float*>(q.B);
      d.out = static_cast<float*>(q.C);
      d.C = nsplit > 1 ? q.work : static_cast<float*>(q.C);
      d.bias = q.bias; d.s = q.s; d.aux = static_cast<float*>(q.aux);
      d.x1 = q.x1; d.x2 = q.x2; d.out2 = static_cast<float*>(q.out2);
      d.crow = q.crow;
      {
        // 16-byte epilogue accesses: FP32 arrays 16-byte aligned, BF16 outputs 8-byte
        const bool rev = epi == EPI_REVERSE;
        d.vec_out = q.ldc % 4 == 0 && aligned(d.C, (io & 2) && nsplit == 1 ? 8 : 16) &&
                    aligned(q.aux, (io & 4) ? 8 : 16) && aligned(q.out2, (io & 8) ? 8 : 16) &&
                    aligned(q.s, 16) && (!rev || (aligned(q.x1, 16) && aligned(q.x2, 16)));
      }
      d.M = q.M; d.N = q.N; d.K = q.K;
      d.lda = q.lda; d.ldb = q.ldb; d.ldc = q.ldc;
      d.cstride = (long)q.M * q.ldc;
      const int steps = (q.K + GK - 1) / GK;
      d.kchunk = steps == 0 ? GK : ((steps + nsplit - 1) / nsplit) * GK;
      d.tiles_n = tm_div_up(q.N, GT);
      d.tile0 = gb.ntiles;
      gb.ntiles += d.tiles_n * tm_div_up(q.M, GT);
      max_mn = std::max(max_mn, (long)q.M * q.N);
    }
    if (gb.ngroups == 0) continue;
    gb.per_xcd = tm_div_up(gb.ntiles, NXCD);
    kern<<<dim3(NXCD * gb.per_xcd, nsplit), 256, 0, st>>>(gb);
    TM_LAUNCH_CHECK();
    if (nsplit > 1) {
      split_sum_kernel<<<dim3(tm_div_up(max_mn, 256), gb.ngroups), 256, 0, st>>>(gb, nsplit);
      TM_LAUNCH_CHECK();
    }
  }
  return TM_OK;
}

int tm_gemm_bf16(int M, int N, int K, const float* A, int lda, int ta, const float* B, int ldb,
                 int tb, float* C, int ldc, int epi, const float* bias, const float* s,
                 float* aux, float alpha, int nsplit, float* work, void* stream) {
  tm_gemm_group q;
  q.M = M; q.N = N; q.K = K;
  q.A = A; q.lda = lda; q.B = B; q.ldb = ldb; q.C = C; q.ldc = ldc;
  q.bias = bias; q.s = s; q.aux = aux; q.work = work;
  q.x1 = nullptr; q.x2 = nullptr; q.out2 = nullptr; q.crow = nullptr;
  return tm_gemm_bf16_grouped(1, &q, ta, tb, epi, 0, alpha, nsplit, stream);
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

int tm_rows_grouped(int ngroups, const tm_rows_group* groups, int scatter, void* stream) {
  TM_REQUIRE(ngroups >= 0 && (ngroups == 0 || groups != nullptr),
             "tm_rows_grouped: bad group list");
  for (int g0 = 0; g0 < ngroups;) {
    RowsBatch rb;
    rb.scatter = scatter;
    int n = 0;
    int most = 0;
    for (; g0 < ngroups && n < MAXG; ++g0) {
      const tm_rows_group& q = groups[g0];
      TM_REQUIRE(q.M >= 0 && q.cols >= 0, "tm_rows_grouped: negative size (set %d)", g0);
      if ((long)q.M * q.cols == 0) continue;
      rb.r[n++] = q;
      most = std::max(most, q.M);
    }
    if (n == 0) continue;
    rows_grouped_kernel<<<dim3(tm_div_up(most, 4), n), 256, 0, (hipStream_t)stream>>>(rb);
    TM_LAUNCH_CHECK();
  }
  return TM_OK;
}

int tm_fm_loss(int B, const int* mol_start, const float* eps, const double* E_ref,
               const double* w_e, const double* grad, const double* F_ref, const double* w_f,
               double* E, float* ybar, double* gg, double* part, double* L, void* stream) {
  TM_REQUIRE(B >= 0, "tm_fm_loss: negative molecule count");
  hipStream_t st = (hipStream_t)stream;
  if (B > 0) {
    fm_loss_kernel<<<B, 64, 0, st>>>(B, mol_start, eps, E_ref, w_e, grad, F_ref, w_f, E, ybar,
                                     gg, part);
    TM_LAUNCH_CHECK();
  }
  fm_loss_sum_kernel<<<1, 256, 0, st>>>(B, part, L);
  TM_LAUNCH_CHECK();
  return TM_OK;
}

int tm_cast_bf16_grouped(int ngroups, const tm_cast_group* groups, void* stream) {
  TM_REQUIRE(ngroups >= 0 && (ngroups == 0 || groups != nullptr),
             "tm_cast_bf16_grouped: bad group list");
  for (int g0 = 0; g0 < ngroups;) {
    CastBatch cb;
    int n = 0;
    long most = 0;
    for (; g0 < ngroups && n < MAXG; ++g0) {
      TM_REQUIRE(groups[g0].n >= 0, "tm_cast_bf16_grouped: negative size (array %d)", g0);
      if (groups[g0].n == 0) continue;
      cb.c[n++] = groups[g0];
      most = std::max(most, groups[g0].n);
    }
    if (n == 0) continue;
    const int blocks = (int)std::min<long>(tm_div_up(most, 256), 1024);
    cast_bf16_kernel<<<dim3(blocks, n), 256, 0, (hipStream_t)stream>>>(cb);
    TM_LAUNCH_CHECK();
  }
  return TM_OK;
}

}  // extern "C"
