// Fused per-element MLPs on MFMA (gfx950): forward and input gradient of
// both the energy and the charge network in one launch.
//
// Math restates energy_inference / dipole_inference
// (TensorMol/TFNetworks/TFMolInstanceDirect.py:5164-5212, 5219-5272), three
// hidden layers h_k = act(h_{k-1} W_k + b_k), act(z) = log(1 + e^{a z})/a
// (Util.py:200-201), output y = h_3 w_o + b_o; and the adjoint that
// tf.gradients builds (:5761) for dy/dG:
//   dZ3 = w_o * act'(Z3), dH_{k-1} = dZ_k W_k^T, dZ_{k-1} = dH_{k-1} * act'(Z_{k-1}),
//   dG = dZ1 W_1^T, act'(z) = sigmoid(a z).
//
// Work decomposition: one 512-thread workgroup per tile of MT atoms of one
// (net, element) segment (atoms are element-sorted so a tile never straddles
// an element).  Every GEMM is computed transposed, out^T = W^T in^T, with
// v_mfma_f32_16x16x32_{bf16,f16}: A = weights (pre-packed in fragment order,
// one 1 KiB coalesced load per fragment), B = activations from LDS
// ([atom][feature], 16-byte reads), C = 16 features x 16 atoms.  Wave w owns
// output features [64w, 64w+64) of every hidden layer, so the activation
// derivative of its features stays in its registers from the forward to the
// backward pass; only activations / upstream gradients cross waves, through
// two LDS ping-pong buffers.  The descriptor block streams through LDS in
// 64-feature chunks, converted from FP32 on the fly.
//
// Modes: 0 bf16, 1 fp16 (operands rounded once, FP32 accumulation),
// 2 bf16 hi/lo split: x = hi + lo for weights and activations and
// acc += Whi*xhi + Wlo*xhi + Whi*xlo (near-FP32 products),
// 3 bf16 with split forward weights: the three forward GEMMs take
// acc += Whi*x + Wlo*x (weights carried to 16 significant bits, activations
// and descriptors rounded once to BF16), the input-gradient GEMMs plain BF16;
// 4 the same with FP16 operands (weights to 22 bits, activations to 11);
// 5 bf16 with the forward GEMMs split in weights and activations (three
// MFMAs, as mode 2) and plain BF16 input-gradient GEMMs.
// Rounding the weights shifts every atom's energy the same way, so that error
// adds up over a system; rounding activations gives an unbiased error that is
// decorrelated across atoms, so it averages out over large systems but not
// over a small molecule: modes 3 and 4 remove the first, mode 5 both.
#include "tm_common.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int NW = 8;       // waves per workgroup
constexpr int FT = 4;       // 16-feature tiles per wave (64 features)
constexpr int KC = 64;      // the descriptor length must be a multiple of KC
constexpr int HMAX = 512;   // widest hidden layer

struct MlpArgs {
  int N, D, H, ldg;
  int ntiles;
  const int* tiles;        // [ntiles][2] = (segment, first row in segment)
  const int* seg;          // [nseg][4] = (net, blob index, first row in perm, rows)
  const float* G;
  const int* perm;
  const unsigned short* const* blobs;   // device array of per-(net,element) weight blobs
  float alpha;
  float* y;                // [2][N]
  float* dG;               // [2][N][ldg]
  int* flags;
};

// ---- packed blob layout (elements of 2 bytes), identical on the host ----
__host__ __device__ inline long frag_elems(int rows16, int cols32) {
  return (long)rows16 * cols32 * 512;
}
struct BlobLayout {
  long f1, f2, f3, b3, b2, b1, fend;   // fragment planes (offsets in 2-byte elements)
  long bias_bytes;                     // byte offset of the float section
};
// planes of the forward (f*) and backward (b*) fragment arrays: 1 (one rounding)
// or 2 (hi plane, then lo plane)
__host__ __device__ constexpr int fwd_planes(int mode) { return mode >= 2 ? 2 : 1; }
__host__ __device__ constexpr int bwd_planes(int mode) { return mode == 2 ? 2 : 1; }
// activation (B operand) lo planes in the forward / backward GEMMs
__host__ __device__ constexpr bool act_split_fwd(int mode) { return mode == 2 || mode == 5; }
__host__ __device__ constexpr bool act_split_bwd(int mode) { return mode == 2; }
__host__ __device__ inline BlobLayout blob_layout(int D, int H, int fp, int bp) {
  BlobLayout L;
  long o = 0;
  L.f1 = o; o += fp * frag_elems(H / 16, D / 32);
  L.f2 = o; o += fp * frag_elems(H / 16, H / 32);
  L.f3 = o; o += fp * frag_elems(H / 16, H / 32);
  L.b3 = o; o += bp * frag_elems(H / 16, H / 32);
  L.b2 = o; o += bp * frag_elems(H / 16, H / 32);
  L.b1 = o; o += bp * frag_elems(D / 16, H / 32);
  L.fend = o;
  L.bias_bytes = ((o * 2 + 255) / 256) * 256;
  return L;
}
// float section: b1[H], b2[H], b3[H], wo[H], bo

template <int MODE>
struct Ty;
template <>
struct Ty<0> {
  typedef bf16x8 vec;
  static __device__ __forceinline__ unsigned short cvt(float x) {
    return __builtin_bit_cast(unsigned short, (__bf16)x);
  }
  static __device__ __forceinline__ float back(unsigned short b) {
    return (float)__builtin_bit_cast(__bf16, b);
  }
  static __device__ __forceinline__ floatx4 mma(const uint4& a, const uint4& b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                   __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
  }
};
template <>
struct Ty<1> {
  static __device__ __forceinline__ unsigned short cvt(float x) {
    return __builtin_bit_cast(unsigned short, (_Float16)x);
  }
  static __device__ __forceinline__ float back(unsigned short b) {
    return (float)__builtin_bit_cast(_Float16, b);
  }
  static __device__ __forceinline__ floatx4 mma(const uint4& a, const uint4& b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a),
                                                  __builtin_bit_cast(f16x8, b), c, 0, 0, 0);
  }
};
template <>
struct Ty<2> : Ty<0> {};
template <>
struct Ty<3> : Ty<0> {};
template <>
struct Ty<4> : Ty<1> {};
template <>
struct Ty<5> : Ty<0> {};

// act(z) = log(1 + e^{a z}) / a and act'(z) = sigmoid(a z), sharing e = e^{-|a z|}:
// act = (max(a z, 0) + log(1 + e)) / a, sigmoid = 1/(1+e) for a z >= 0, e/(1+e) below.
// The hardware exp2 / log2 are used directly: the exp2 argument is <= 0 (a
// result below the normal range is flushed to 0, harmless here) and the log2
// argument lies in [1, 2], so the library's range fix-ups are not needed.
struct ActConst {
  float alpha, inv_alpha, k_exp, k_log;   // k_exp = -alpha log2(e), k_log = ln(2) / alpha
};
__device__ __forceinline__ ActConst act_const(float alpha) {
  ActConst c;
  c.alpha = alpha;
  c.inv_alpha = 1.0f / alpha;
  c.k_exp = -alpha * 1.4426950408889634f;
  c.k_log = 0.6931471805599453f / alpha;
  return c;
}
__device__ __forceinline__ void act_pair(float z, const ActConst& c, float& act, float& dact) {
  const float e = __builtin_amdgcn_exp2f(fabsf(z) * c.k_exp);
  const float r = __builtin_amdgcn_rcpf(1.0f + e);
  act = fmaxf(z, 0.f) + __builtin_amdgcn_logf(1.0f + e) * c.k_log;
  dact = z >= 0.f ? r : e * r;
}

// Activation derivatives of a wave's features, kept from the forward to the
// backward pass: packed FP16 pairs in the BF16 / FP16 modes (the backward
// operands are rounded to 8 / 11 bits anyway), FP32 in the split mode.
typedef __fp16 half2v __attribute__((ext_vector_type(2)));
template <int MODE, int NFT, int AT>
struct DerivRegs {
  half2v h[NFT][AT][2];   // two derivatives per register
  __device__ __forceinline__ void set4(int ft, int at, const float d[4]) {
    h[ft][at][0] = __builtin_amdgcn_cvt_pkrtz(d[0], d[1]);
    h[ft][at][1] = __builtin_amdgcn_cvt_pkrtz(d[2], d[3]);
  }
  __device__ __forceinline__ float get(int ft, int at, int i) const {
    return (float)h[ft][at][i >> 1][i & 1];
  }
};
template <int NFT, int AT>
struct DerivRegs<2, NFT, AT> {
  float h[NFT][AT][4];
  __device__ __forceinline__ void set4(int ft, int at, const float d[4]) {
#pragma unroll
    for (int i = 0; i < 4; ++i) h[ft][at][i] = d[i];
  }
  __device__ __forceinline__ float get(int ft, int at, int i) const { return h[ft][at][i]; }
};

// Activation buffers.  Two ping-pong buffers X, Y (the descriptor tile staged
// across both), or, where two would not fit (split activations in 64-row
// tiles), one buffer used in place by every layer: the descriptor tile
// streams through its first columns, and each layer's GEMM reads the whole
// buffer before (a barrier) its epilogue overwrites it with the layer's output.
template <int MODE, int MT>
constexpr bool in_place() { return act_split_fwd(MODE) && MT > 32; }

template <int MODE, int MT, bool INPLACE = in_place<MODE, MT>()>
struct Smem {
  static constexpr int PL = act_split_fwd(MODE) ? 2 : 1;
  static constexpr int SH = HMAX + 8;                // activation row stride (elements)
  union {
    struct {
      unsigned short X[PL][MT][SH];
      unsigned short Y[PL][MT][SH];
    };
    // layer 1: the whole descriptor tile, row stride D + 8 <= 2 SH
    unsigned short Gt[2 * PL * MT * SH];
  };
  float epart[NW][MT];
  int atoms[MT];
};
template <int MODE, int MT>
struct Smem<MODE, MT, true> {
  static constexpr int PL = act_split_fwd(MODE) ? 2 : 1;
  static constexpr int SH = HMAX + 8;
  unsigned short X[PL][MT][SH];
  float epart[NW][MT];
  int atoms[MT];
};

// write 4 consecutive features (values v) of one atom into an activation
// buffer; LO: also the lo plane (the residual v - hi, rounded)
template <int MODE, int MT, bool LO>
__device__ __forceinline__ void put4(unsigned short (*buf)[MT][HMAX + 8], int atom, int f,
                                     const float v[4]) {
  unsigned short h[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) h[i] = Ty<MODE>::cvt(v[i]);
  uint2 w;
  w.x = (unsigned)h[0] | ((unsigned)h[1] << 16);
  w.y = (unsigned)h[2] | ((unsigned)h[3] << 16);
  *reinterpret_cast<uint2*>(&buf[0][atom][f]) = w;
  if constexpr (LO) {
    unsigned short l[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) l[i] = Ty<MODE>::cvt(v[i] - Ty<MODE>::back(h[i]));
    uint2 u;
    u.x = (unsigned)l[0] | ((unsigned)l[1] << 16);
    u.y = (unsigned)l[2] | ((unsigned)l[3] << 16);
    *reinterpret_cast<uint2*>(&buf[1][atom][f]) = u;
  }
}

// acc[ft][at] += sum over K of W-fragments (rows tiles rt0 + ft) x B from LDS buffer.
// The weight fragments of k-step ks+1 are in flight while the MFMAs of ks run:
// two register sets alternate (the loop is unrolled by two; nks is even), and
// scheduling barriers keep the loads ahead of the MFMAs that do not need them.
// First weight fragments of a GEMM (k-step 0).  Weights do not depend on LDS,
// so a wave loads them before the epilogue and barrier that precede the GEMM.
template <int NFT>
__device__ __forceinline__ void gemm_first(uint4 (&a)[NFT], const unsigned short* __restrict__ wf,
                                           int nks, const int* rtile, int lane) {
#pragma unroll
  for (int ft = 0; ft < NFT; ++ft)
    a[ft] = *reinterpret_cast<const uint4*>(wf + ((long)(rtile[ft] * nks) * 64 + lane) * 8);
}

template <int MODE, int MT, int NFT, bool SPLIT>
__device__ __forceinline__ void gemm_tiles(floatx4 (&acc)[NFT][MT / 16],
                                           const unsigned short* __restrict__ wf, long plane_elems,
                                           int nks, int kstride, const int* rtile,
                                           const unsigned short* bbase, long bplane, int lane,
                                           int STRIDE, const uint4 (&a_first)[NFT]) {
  constexpr int AT = MT / 16;
  const int brow = lane & 15;
  const int bk = (lane >> 4) * 8;
  auto afrag = [&](int ft, int ks, long plane) {
    return *reinterpret_cast<const uint4*>(wf + plane + ((long)(rtile[ft] * kstride + ks) * 64 + lane) * 8);
  };
  auto step = [&](const uint4 (&a)[NFT], int ks) {
    uint4 b[AT];
#pragma unroll
    for (int at = 0; at < AT; ++at)
      b[at] = *reinterpret_cast<const uint4*>(bbase + (long)(at * 16 + brow) * STRIDE + ks * 32 + bk);
#pragma unroll
    for (int ft = 0; ft < NFT; ++ft)
#pragma unroll
      for (int at = 0; at < AT; ++at) acc[ft][at] = Ty<MODE>::mma(a[ft], b[at], acc[ft][at]);
    if constexpr (SPLIT) {
      uint4 al[NFT], bl[AT];
#pragma unroll
      for (int ft = 0; ft < NFT; ++ft) al[ft] = afrag(ft, ks, plane_elems);
#pragma unroll
      for (int at = 0; at < AT; ++at)
        bl[at] = *reinterpret_cast<const uint4*>(bbase + bplane + (long)(at * 16 + brow) * STRIDE +
                                                 ks * 32 + bk);
#pragma unroll
      for (int ft = 0; ft < NFT; ++ft)
#pragma unroll
        for (int at = 0; at < AT; ++at) {
          acc[ft][at] = Ty<MODE>::mma(al[ft], b[at], acc[ft][at]);
          acc[ft][at] = Ty<MODE>::mma(a[ft], bl[at], acc[ft][at]);
        }
    }
  };
  uint4 a0[NFT], a1[NFT];
#pragma unroll
  for (int ft = 0; ft < NFT; ++ft) a0[ft] = a_first[ft];
  for (int ks = 0; ks < nks; ks += 2) {
#pragma unroll
    for (int ft = 0; ft < NFT; ++ft) a1[ft] = afrag(ft, ks + 1, 0);
    __builtin_amdgcn_sched_barrier(0);
    step(a0, ks);
    __builtin_amdgcn_sched_barrier(0);
    if (ks + 2 < nks) {
#pragma unroll
      for (int ft = 0; ft < NFT; ++ft) a0[ft] = afrag(ft, ks + 2, 0);
    }
    __builtin_amdgcn_sched_barrier(0);
    step(a1, ks + 1);
    __builtin_amdgcn_sched_barrier(0);
  }
}

// Split-weight GEMM (mode 3 forward): acc += (Whi + Wlo) x B with one B plane.
// Each k-step is two MFMA passes over the same B fragments, the hi fragments
// (a0) then the lo fragments (a1); the lo fragments load while the hi MFMAs
// run and the next hi fragments while the lo MFMAs run, so the register
// footprint and the load-ahead distance are those of gemm_tiles.
template <int MODE, int MT, int NFT>
__device__ __forceinline__ void gemm_tiles_ws(floatx4 (&acc)[NFT][MT / 16],
                                              const unsigned short* __restrict__ wf,
                                              long plane_elems, int nks, int kstride,
                                              const int* rtile, const unsigned short* bbase,
                                              int lane, int STRIDE, const uint4 (&a_first)[NFT]) {
  constexpr int AT = MT / 16;
  const int brow = lane & 15;
  const int bk = (lane >> 4) * 8;
  auto afrag = [&](int ft, int ks, long plane) {
    return *reinterpret_cast<const uint4*>(wf + plane + ((long)(rtile[ft] * kstride + ks) * 64 + lane) * 8);
  };
  uint4 a0[NFT], a1[NFT];
#pragma unroll
  for (int ft = 0; ft < NFT; ++ft) a0[ft] = a_first[ft];
  for (int ks = 0; ks < nks; ++ks) {
#pragma unroll
    for (int ft = 0; ft < NFT; ++ft) a1[ft] = afrag(ft, ks, plane_elems);
    __builtin_amdgcn_sched_barrier(0);
    uint4 b[AT];
#pragma unroll
    for (int at = 0; at < AT; ++at)
      b[at] = *reinterpret_cast<const uint4*>(bbase + (long)(at * 16 + brow) * STRIDE + ks * 32 + bk);
#pragma unroll
    for (int ft = 0; ft < NFT; ++ft)
#pragma unroll
      for (int at = 0; at < AT; ++at) acc[ft][at] = Ty<MODE>::mma(a0[ft], b[at], acc[ft][at]);
    __builtin_amdgcn_sched_barrier(0);
    if (ks + 1 < nks) {
#pragma unroll
      for (int ft = 0; ft < NFT; ++ft) a0[ft] = afrag(ft, ks + 1, 0);
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int ft = 0; ft < NFT; ++ft)
#pragma unroll
      for (int at = 0; at < AT; ++at) acc[ft][at] = Ty<MODE>::mma(a1[ft], b[at], acc[ft][at]);
    __builtin_amdgcn_sched_barrier(0);
  }
}

// Forward GEMM of a layer: split weights only (modes 3, 4), split weights and
// activations (modes 2, 5), or one rounding of each (modes 0, 1).
template <int MODE, int MT, int NFT>
__device__ __forceinline__ void gemm_fwd(floatx4 (&acc)[NFT][MT / 16],
                                         const unsigned short* __restrict__ wf, long plane_elems,
                                         int nks, int kstride, const int* rtile,
                                         const unsigned short* bbase, long bplane, int lane,
                                         int STRIDE, const uint4 (&a_first)[NFT]) {
  if constexpr (MODE == 3 || MODE == 4)
    gemm_tiles_ws<MODE, MT, NFT>(acc, wf, plane_elems, nks, kstride, rtile, bbase, lane, STRIDE,
                                 a_first);
  else
    gemm_tiles<MODE, MT, NFT, act_split_fwd(MODE)>(acc, wf, plane_elems, nks, kstride, rtile,
                                                   bbase, bplane, lane, STRIDE, a_first);
}

// 64-row tiles: one workgroup per CU (LDS); 32-row tiles in the BF16 / FP16
// modes fit two workgroups per CU when the waves stay within 128 VGPRs.
template <int MODE, int MT>
__global__ void __launch_bounds__(NW * 64)
__attribute__((amdgpu_waves_per_eu(MT == 32 && !act_split_fwd(MODE) ? 4 : 2)))
mlp_mfma_kernel(MlpArgs a) {
  constexpr int AT = MT / 16;
  constexpr bool LOF = act_split_fwd(MODE);   // forward operands carry lo planes
  constexpr bool LOB = act_split_bwd(MODE);   // backward operands carry lo planes
  constexpr bool INPLACE = in_place<MODE, MT>();
  constexpr int SH = HMAX + 8;
  typedef Smem<MODE, MT> S;
  __shared__ S sm;
  // output buffer of layers 2, bwd 3 (read by layer 3, bwd 2): Y, or X in place
  unsigned short (*const Yb)[MT][SH] = [&]() {
    if constexpr (INPLACE) return sm.X; else return sm.Y;
  }();
  const int tid = threadIdx.x;
  const int w = tid >> 6, lane = tid & 63;
  const int tile = blockIdx.x;
  const int s = a.tiles[2 * tile], row0 = a.tiles[2 * tile + 1];
  if (s < 0) return;   // padding of the XCD-ordered tile table (whole workgroup)
  const int net = a.seg[4 * s], blob = a.seg[4 * s + 1], first = a.seg[4 * s + 2];
  const int rows = a.seg[4 * s + 3];
  const int D = a.D, H = a.H;
  const BlobLayout BL = blob_layout(D, H, fwd_planes(MODE), bwd_planes(MODE));
  const unsigned short* wb = a.blobs[blob];
  const float* fsec = reinterpret_cast<const float*>(reinterpret_cast<const char*>(wb) + BL.bias_bytes);
  const float* b1 = fsec;
  const float* b2 = fsec + H;
  const float* b3 = fsec + 2 * H;
  const float* wo = fsec + 3 * H;
  const float bo = fsec[4 * H];
  if (tid < MT) {
    int r = row0 + tid;
    sm.atoms[tid] = r < rows ? a.perm[first + r] : -1;
  }
  const bool wact = (w * 64) < H;   // this wave owns hidden features
  int rt[FT];
#pragma unroll
  for (int ft = 0; ft < FT; ++ft) rt[ft] = 4 * w + ft;
  const int fq = (lane >> 4) * 4;   // first of the 4 features this lane holds in a C tile
  const int ca = lane & 15;         // atom column of this lane in a C tile
  // activation derivatives of layers 1 and 2 for this wave's features
  DerivRegs<MODE, FT, AT> s1, s2;
  const ActConst ac = act_const(a.alpha);
  floatx4 acc[FT][AT];
  const long phh = frag_elems(H / 16, H / 32);
  uint4 pf[FT];   // first weight fragments of the next GEMM, loaded early
  if (wact) gemm_first<FT>(pf, wb + BL.f1, D / 32, rt, lane);

  // ------------------------------------------------------------ layer 1
#pragma unroll
  for (int ft = 0; ft < FT; ++ft)
#pragma unroll
    for (int at = 0; at < AT; ++at) acc[ft][at] = floatx4{0.f, 0.f, 0.f, 0.f};
  const long p1 = frag_elems(H / 16, D / 32);
  const long bpl = (long)MT * SH;   // element offset of the lo plane
  if constexpr (INPLACE) {
    // the descriptor tile G[atom][0 .. D) streams through the buffer in
    // column blocks of up to HMAX features (one block when D <= HMAX)
    for (int c0 = 0; c0 < D; c0 += HMAX) {
      const int cw = min(HMAX, D - c0);
      __syncthreads();   // atoms known; every wave done with the previous block
      for (int e = tid; e < MT * cw / 4; e += NW * 64) {
        int r = e / (cw / 4), c4 = (e % (cw / 4)) * 4;
        int atom = sm.atoms[r];
        float4 g = atom >= 0 ? *reinterpret_cast<const float4*>(a.G + (long)atom * a.ldg + c0 + c4)
                             : make_float4(0.f, 0.f, 0.f, 0.f);
        float v[4] = {g.x, g.y, g.z, g.w};
        put4<MODE, MT, LOF>(sm.X, r, c4, v);
      }
      __syncthreads();
      if (wact) {
        if (c0 > 0) gemm_first<FT>(pf, wb + BL.f1 + (long)(c0 / 32) * 512, D / 32, rt, lane);
        gemm_fwd<MODE, MT, FT>(acc, wb + BL.f1 + (long)(c0 / 32) * 512, p1, cw / 32, D / 32, rt,
                               &sm.X[0][0][0], bpl, lane, SH, pf);
      }
    }
  } else {
    // the whole descriptor tile G[atom][0 .. D) (hi / lo), row stride D + 8,
    // across both buffers
    const int SG = D + 8;
    __syncthreads();
    for (int e = tid; e < MT * D / 4; e += NW * 64) {
      int r = e / (D / 4), c4 = (e % (D / 4)) * 4;
      int atom = sm.atoms[r];
      float4 g = atom >= 0 ? *reinterpret_cast<const float4*>(a.G + (long)atom * a.ldg + c4)
                           : make_float4(0.f, 0.f, 0.f, 0.f);
      float v[4] = {g.x, g.y, g.z, g.w};
      unsigned short* dst = sm.Gt + (long)r * SG + c4;
      unsigned short h[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) h[i] = Ty<MODE>::cvt(v[i]);
      uint2 ww;
      ww.x = (unsigned)h[0] | ((unsigned)h[1] << 16);
      ww.y = (unsigned)h[2] | ((unsigned)h[3] << 16);
      *reinterpret_cast<uint2*>(dst) = ww;
      if constexpr (LOF) {
        unsigned short l[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) l[i] = Ty<MODE>::cvt(v[i] - Ty<MODE>::back(h[i]));
        uint2 u;
        u.x = (unsigned)l[0] | ((unsigned)l[1] << 16);
        u.y = (unsigned)l[2] | ((unsigned)l[3] << 16);
        *reinterpret_cast<uint2*>(dst + (long)MT * SG) = u;
      }
    }
    __syncthreads();
    if (wact)
      gemm_fwd<MODE, MT, FT>(acc, wb + BL.f1, p1, D / 32, D / 32, rt, sm.Gt, (long)MT * SG, lane,
                             SG, pf);
  }
  if (wact) gemm_first<FT>(pf, wb + BL.f2, H / 32, rt, lane);
  __syncthreads();   // every wave has read the tile before it is overwritten
  // epilogue 1 -> X, s1
  if (wact) {
#pragma unroll
    for (int ft = 0; ft < FT; ++ft)
#pragma unroll
      for (int at = 0; at < AT; ++at) {
        int f = 64 * w + 16 * ft + fq;
        float v[4], d[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) act_pair(acc[ft][at][i] + b1[f + i], ac, v[i], d[i]);
        s1.set4(ft, at, d);
        put4<MODE, MT, LOF>(sm.X, at * 16 + ca, f, v);
      }
  }
  __syncthreads();
  // ------------------------------------------------------------ layer 2
  if (wact) {
#pragma unroll
    for (int ft = 0; ft < FT; ++ft)
#pragma unroll
      for (int at = 0; at < AT; ++at) acc[ft][at] = floatx4{0.f, 0.f, 0.f, 0.f};
    gemm_fwd<MODE, MT, FT>(acc, wb + BL.f2, phh, H / 32, H / 32, rt, &sm.X[0][0][0], bpl, lane, SH, pf);
    gemm_first<FT>(pf, wb + BL.f3, H / 32, rt, lane);
  }
  if constexpr (INPLACE) __syncthreads();
  if (wact) {
#pragma unroll
    for (int ft = 0; ft < FT; ++ft)
#pragma unroll
      for (int at = 0; at < AT; ++at) {
        int f = 64 * w + 16 * ft + fq;
        float v[4], d[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) act_pair(acc[ft][at][i] + b2[f + i], ac, v[i], d[i]);
        s2.set4(ft, at, d);
        put4<MODE, MT, LOF>(Yb, at * 16 + ca, f, v);
      }
  }
  __syncthreads();
  // ------------------------------------------------------------ layer 3 + head
  if (wact) {
#pragma unroll
    for (int ft = 0; ft < FT; ++ft)
#pragma unroll
      for (int at = 0; at < AT; ++at) acc[ft][at] = floatx4{0.f, 0.f, 0.f, 0.f};
    gemm_fwd<MODE, MT, FT>(acc, wb + BL.f3, phh, H / 32, H / 32, rt, &Yb[0][0][0], bpl, lane, SH, pf);
    gemm_first<FT>(pf, wb + BL.b3, H / 32, rt, lane);
  }
  if constexpr (INPLACE) __syncthreads();
  if (wact) {
    float ep[AT];
#pragma unroll
    for (int at = 0; at < AT; ++at) ep[at] = 0.f;
#pragma unroll
    for (int ft = 0; ft < FT; ++ft)
#pragma unroll
      for (int at = 0; at < AT; ++at) {
        int f = 64 * w + 16 * ft + fq;
        float v[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          float wv = wo[f + i];
          float act, d;
          act_pair(acc[ft][at][i] + b3[f + i], ac, act, d);
          ep[at] += act * wv;
          v[i] = wv * d;   // dZ3 with upstream 1
        }
        put4<MODE, MT, LOB>(sm.X, at * 16 + ca, f, v);
      }
#pragma unroll
    for (int at = 0; at < AT; ++at) {
      float p = ep[at];
      p += __shfl_xor(p, 16, 64);
      p += __shfl_xor(p, 32, 64);
      if (lane < 16) sm.epart[w][at * 16 + lane] = p;
    }
  } else {
    for (int t = lane; t < MT; t += 64) sm.epart[w][t] = 0.f;
  }
  __syncthreads();
  if (tid < MT) {
    int atom = sm.atoms[tid];
    if (atom >= 0) {
      float yv = bo;
#pragma unroll
      for (int q = 0; q < NW; ++q) yv += sm.epart[q][tid];
      a.y[(long)net * a.N + atom] = yv;
    }
  }
  // ------------------------------------------------------------ backward 3: dZ3 -> dZ2
  if (wact) {
#pragma unroll
    for (int ft = 0; ft < FT; ++ft)
#pragma unroll
      for (int at = 0; at < AT; ++at) acc[ft][at] = floatx4{0.f, 0.f, 0.f, 0.f};
    gemm_tiles<MODE, MT, FT, LOB>(acc, wb + BL.b3, phh, H / 32, H / 32, rt, &sm.X[0][0][0], bpl, lane,
                                  SH, pf);
    gemm_first<FT>(pf, wb + BL.b2, H / 32, rt, lane);
  }
  if constexpr (INPLACE) __syncthreads();
  if (wact) {
#pragma unroll
    for (int ft = 0; ft < FT; ++ft)
#pragma unroll
      for (int at = 0; at < AT; ++at) {
        int f = 64 * w + 16 * ft + fq;
        float v[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) v[i] = acc[ft][at][i] * s2.get(ft, at, i);
        put4<MODE, MT, LOB>(Yb, at * 16 + ca, f, v);
      }
  }
  __syncthreads();
  // ------------------------------------------------------------ backward 2: dZ2 -> dZ1
  if (wact) {
#pragma unroll
    for (int ft = 0; ft < FT; ++ft)
#pragma unroll
      for (int at = 0; at < AT; ++at) acc[ft][at] = floatx4{0.f, 0.f, 0.f, 0.f};
    gemm_tiles<MODE, MT, FT, LOB>(acc, wb + BL.b2, phh, H / 32, H / 32, rt, &Yb[0][0][0], bpl, lane,
                                  SH, pf);
  }
  if constexpr (INPLACE) __syncthreads();
  if (wact) {
#pragma unroll
    for (int ft = 0; ft < FT; ++ft)
#pragma unroll
      for (int at = 0; at < AT; ++at) {
        int f = 64 * w + 16 * ft + fq;
        float v[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) v[i] = acc[ft][at][i] * s1.get(ft, at, i);
        put4<MODE, MT, LOB>(sm.X, at * 16 + ca, f, v);
      }
  }
  uint4 pf2[2];   // first fragments of backward 1 (rows of dG tiles w, w + NW)
  if (w < D / 16) {
    int r0[2] = {w, w + NW < D / 16 ? w + NW : w};
    gemm_first<2>(pf2, wb + BL.b1, H / 32, r0, lane);
  }
  __syncthreads();
  // ------------------------------------------------------------ backward 1: dZ1 -> dG
  {
    const long pd = frag_elems(D / 16, H / 32);
    float* dGn = a.dG + (long)net * a.N * a.ldg;
    const int ntile = D / 16;
    for (int t0 = w; t0 < ntile; t0 += 2 * NW) {
      int rt2[2] = {t0, t0 + NW < ntile ? t0 + NW : t0};
      floatx4 acc2[2][AT];
#pragma unroll
      for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int at = 0; at < AT; ++at) acc2[u][at] = floatx4{0.f, 0.f, 0.f, 0.f};
      gemm_tiles<MODE, MT, 2, LOB>(acc2, wb + BL.b1, pd, H / 32, H / 32, rt2, &sm.X[0][0][0], bpl, lane,
                                   SH, pf2);
      const int tn = t0 + 2 * NW;   // the next iteration's first fragments
      if (tn < ntile) {
        int rn[2] = {tn, tn + NW < ntile ? tn + NW : tn};
        gemm_first<2>(pf2, wb + BL.b1, H / 32, rn, lane);
      }
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        if (u == 1 && rt2[1] == rt2[0]) continue;
#pragma unroll
        for (int at = 0; at < AT; ++at) {
          int atom = sm.atoms[at * 16 + ca];
          if (atom < 0) continue;
          int f = 16 * rt2[u] + fq;
          *reinterpret_cast<float4*>(dGn + (long)atom * a.ldg + f) =
              make_float4(acc2[u][at][0], acc2[u][at][1], acc2[u][at][2], acc2[u][at][3]);
        }
      }
    }
  }
}

}  // namespace

extern "C" {

// out[0] = blob bytes, out[1] = byte offset of the float section,
// out[2..7] = element offsets of f1, f2, f3, b3, b2, b1 fragment arrays,
// out[8] = planes of the forward arrays, out[9] = planes of the backward arrays
int tm_mlp_mfma_pack_info(int D, int H, int nh, int mode, long* out) {
  TM_REQUIRE(nh == 3, "tm_mlp_mfma: exactly 3 hidden layers supported (got %d)", nh);
  TM_REQUIRE(H % 64 == 0 && H <= HMAX && H >= 64, "tm_mlp_mfma: hidden width %d not a multiple of 64 in 64..512", H);
  TM_REQUIRE(D % KC == 0 && D > 0 && D + 8 <= 2 * (HMAX + 8),
             "tm_mlp_mfma: descriptor length %d not a multiple of %d in %d..%d", D, KC, KC,
             2 * HMAX + 8);
  TM_REQUIRE(mode >= 0 && mode <= 5, "tm_mlp_mfma: bad mode %d", mode);
  BlobLayout L = blob_layout(D, H, fwd_planes(mode), bwd_planes(mode));
  out[0] = L.bias_bytes + (4l * H + 4) * 4;
  out[1] = L.bias_bytes;
  out[2] = L.f1; out[3] = L.f2; out[4] = L.f3; out[5] = L.b3; out[6] = L.b2; out[7] = L.b1;
  out[8] = fwd_planes(mode);
  out[9] = bwd_planes(mode);
  return TM_OK;
}

// tiles: device int32 [ntiles][2] = (segment, first row), one workgroup per
// entry (entries with segment -1 are skipped); seg: device int32
// [nseg][4] = (net, blob index, first row in perm, rows); blobs: device array
// of per-(net, element) weight blob pointers (layout of tm_mlp_mfma_pack_info).
int tm_mlp_mfma(int N, int D, int H, int ntiles, const int* tiles, const int* seg, int ldg,
                const float* G, const int* perm, const void* blobs, float alpha, int mode,
                int tile_rows, float* y, float* dG, void* stream) {
  long info[10];
  int st = tm_mlp_mfma_pack_info(D, H, 3, mode, info);
  if (st) return st;
  TM_REQUIRE(ldg % 4 == 0 && ldg >= D, "tm_mlp_mfma: ldg %d", ldg);
  TM_REQUIRE(alpha > 0.f, "tm_mlp_mfma: activation alpha %g must be positive", (double)alpha);
  TM_REQUIRE(ntiles >= 0, "tm_mlp_mfma: bad tile count");
  TM_REQUIRE(tile_rows == 32 || (tile_rows == 64 && mode != 2),
             "tm_mlp_mfma: tile_rows %d (64 or 32; 32 in mode 2)", tile_rows);
  if (ntiles == 0) return TM_OK;
  MlpArgs a;
  a.N = N; a.D = D; a.H = H; a.ldg = ldg;
  a.ntiles = ntiles;
  a.tiles = tiles;
  a.seg = seg;
  a.G = G;
  a.perm = perm;
  a.blobs = (const unsigned short* const*)blobs;
  a.alpha = alpha;
  a.y = y;
  a.dG = dG;
  a.flags = nullptr;
  hipStream_t s = (hipStream_t)stream;
  // 32-row tiles fit several workgroups per CU.  With no more tiles than CUs
  // that would let the dispatcher stack two on one CU while others idle, so
  // the launch then asks for extra LDS: one workgroup per CU (over half of
  // the CU's 160 KiB).
  size_t pad = 0;
  if (tile_rows == 32 && mode != 2) {
    int dev = 0, ncu = 0;
    TM_HIP_CHECK(hipGetDevice(&dev));
    TM_HIP_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const size_t lds = mode == 5 ? sizeof(Smem<5, 32>) : sizeof(Smem<0, 32>);
    if (ntiles <= ncu && lds < 81920) pad = 81920 - lds + 256;
  }
  if (mode == 0 && tile_rows == 64) mlp_mfma_kernel<0, 64><<<ntiles, NW * 64, 0, s>>>(a);
  else if (mode == 0) mlp_mfma_kernel<0, 32><<<ntiles, NW * 64, pad, s>>>(a);
  else if (mode == 1 && tile_rows == 64) mlp_mfma_kernel<1, 64><<<ntiles, NW * 64, 0, s>>>(a);
  else if (mode == 1) mlp_mfma_kernel<1, 32><<<ntiles, NW * 64, pad, s>>>(a);
  else if (mode == 3 && tile_rows == 64) mlp_mfma_kernel<3, 64><<<ntiles, NW * 64, 0, s>>>(a);
  else if (mode == 3) mlp_mfma_kernel<3, 32><<<ntiles, NW * 64, pad, s>>>(a);
  else if (mode == 4 && tile_rows == 64) mlp_mfma_kernel<4, 64><<<ntiles, NW * 64, 0, s>>>(a);
  else if (mode == 4) mlp_mfma_kernel<4, 32><<<ntiles, NW * 64, pad, s>>>(a);
  else if (mode == 5 && tile_rows == 64) mlp_mfma_kernel<5, 64><<<ntiles, NW * 64, 0, s>>>(a);
  else if (mode == 5) mlp_mfma_kernel<5, 32><<<ntiles, NW * 64, pad, s>>>(a);
  else mlp_mfma_kernel<2, 32><<<ntiles, NW * 64, 0, s>>>(a);
  TM_LAUNCH_CHECK();
  return TM_OK;
}

}  // extern "C"
