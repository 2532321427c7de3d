// Charges, DSF-elu electrostatics, vdW and per-molecule reductions.
//
// Charges: q_i = qraw_i - (sum_{j in m} qraw_j) / natom_m, the neutrality step of
// dipole_inference (TensorMol/TFNetworks/TFMolInstanceDirect.py:5274-5277);
// dipole mu_m = sum q_i x_i * BOHRPERA (:5278-5279).
//
// Pair energy (TFCoulombEluSRDSFLR, TFDescriptors/RawSymFunc.py:1307-1359),
// r in bohr, alpha' = DSFAlpha / BOHRPERA:
//   r > R_lr:          K = 0
//   R_sr < r <= R_lr:  K = erfc(alpha' r)/r - ZZ + (r - R_lr)(ZZ/R_lr + YY)
//   r <= R_sr:         K = elu_a (exp(r - R_sr) - 1) + elu_shift
// with ZZ = erfc(alpha' R_lr)/R_lr, YY = (2/sqrt(pi)) alpha' exp(-(alpha' R_lr)^2)/R_lr.
// vdW (TFVdwPolyLR, RawSymFunc.py:1361-1414) with the reference's double
// A->bohr scaling: r_v = r_A * BOHRPERA^2, t = r_v / (Poly_Width * BOHRPERA),
//   V = -S(t) sqrt(C6_i C6_j) / r_v^6 / (1 + 6 (r_v / (R0_i + R0_j))^-12),
//   S = -t^2 (2t - 3) on (0,1], 1 beyond.
// Both run over every pair inside the EE cutoff (the 15 A list of
// Neighbors.py:262-282), evaluated on the fly from the cell grid; no pair
// list is stored.  One wavefront per centre gathers its own terms, so every
// per-atom sum is formed in a fixed order (no atomics) and accumulated in FP64.
#include <mutex>

#include "tm_cells.h"

namespace {

constexpr int C_WAVES = 4;
constexpr double BOHRPERA_D = 1.889725989;
constexpr int C_RANGES = 480;    // 15 A stencil over ~2.3 A cells: <= ~230 rows (+ wraps)

struct CoulArgs {
  float alpha, r_sr, r_lr, zz, yy, shift, ea;     // bohr units
  float lin;                                     // ZZ / R_lr + YY (slope of the shift)
  float rc, rc2;                                 // EE cutoff (A) and its square
  float inv_width;                               // 1 / vdW poly width (bohr after scaling)
  int add_ecc, has_vdw;
  float c6[TM_MAX_ELE], r0[TM_MAX_ELE];
};

// erfc(z) for z >= 0 given e2 = exp(-z^2): the Chebyshev fit "erfcc" of
// Numerical Recipes (Press, Teukolsky, Vetterling, Flannery, 2nd ed., sec. 6.2),
// erfc(z) = t exp(-z^2 + P(t)), t = 1/(1 + z/2), fractional error < 1.2e-7.
__device__ __forceinline__ float erfc_cheb(float z, float e2) {
  const float t = __builtin_amdgcn_rcpf(1.0f + 0.5f * z);
  float P = 0.17087277f;
  P = fmaf(P, t, -0.82215223f);
  P = fmaf(P, t, 1.48851587f);
  P = fmaf(P, t, -1.13520398f);
  P = fmaf(P, t, 0.27886807f);
  P = fmaf(P, t, -0.18628806f);
  P = fmaf(P, t, 0.09678418f);
  P = fmaf(P, t, 0.37409196f);
  P = fmaf(P, t, 1.00002368f);
  P = fmaf(P, t, -1.26551223f);
  return t * e2 * __expf(P);
}

// minimum image with the reciprocal edge (inv_L = 0 for open boundaries)
__device__ __forceinline__ float min_image_f(float d, float L, float inv_L) {
  return d - L * rintf(d * inv_L);
}

struct PairAcc {
  double phi, gx, gy, gz, ev;
};

// energy terms of one pair (i, j) seen from centre i; d = x_j - x_i (A).
// vc = sqrt(C6_i C6_j), vq = 6 (R0_i + R0_j)^12 (per element pair).
__device__ __forceinline__ void pair_terms(const CoulArgs& a, float qi, float qj, float vc, float vq,
                                           float dx, float dy, float dz, PairAcc& acc) {
  const float B1 = (float)BOHRPERA_D, B2 = (float)(BOHRPERA_D * BOHRPERA_D);
  const float iB1 = (float)(1.0 / BOHRPERA_D), iB2 = (float)(1.0 / (BOHRPERA_D * BOHRPERA_D));
  // r_A = sqrt(s + 1e-27), the reference's regulariser (RawSymFunc.py:1330)
  const float s = fmaf(dx, dx, fmaf(dy, dy, dz * dz)) + 1e-27f;
  const float inv_rA = __builtin_amdgcn_rsqf(s);
  const float rA = s * inv_rA;
  float dEdr = 0.f;               // dE_pair/dr_A
  if (a.add_ecc) {
    const float r = rA * B1, inv_r = inv_rA * iB1;
    // one exponential serves both regions: exp(-z^2) beyond R_sr, exp(r - R_sr)
    // inside it (a wave usually holds pairs of both, so no branch)
    const bool lr = r > a.r_sr;
    const float z = a.alpha * r;
    const float ex = __expf(lr ? -z * z : r - a.r_sr);
    float K, dK;
    if (lr) {
      const float ec = erfc_cheb(z, ex);
      K = ec * inv_r - a.zz + (r - a.r_lr) * a.lin;
      dK = -ec * inv_r * inv_r - 1.1283791671f * a.alpha * ex * inv_r + a.lin;
      if (r > a.r_lr) { K = 0.f; dK = 0.f; }
    } else {
      K = a.ea * (ex - 1.0f) + a.shift;
      dK = a.ea * ex;
    }
    acc.phi += (double)(K * qj);
    dEdr += qi * qj * dK * B1;
  }
  if (a.has_vdw) {
    const float r = rA * B2, inv_r = inv_rA * iB2;
    // t clamped to [0, 1] gives S = 1, dS = 0 beyond the width exactly, and
    // S = dS = 0 at t = 0, without branches
    const float t = fminf(fmaxf(r * a.inv_width, 0.f), 1.f);
    const float S = -t * t * (2.f * t - 3.f);
    const float dS = 6.f * t * (1.f - t) * a.inv_width;
    // C r^-6 / (1 + 6 (r/R0)^-12) written as C r^6 / (r^12 + 6 R0^12): the same
    // function, finite for every r >= 0 in FP32 (r <= 15 A BOHRPERA^2 here)
    const float r2 = r * r, r6 = r2 * r2 * r2, r12 = r6 * r6;
    const float iden = __builtin_amdgcn_rcpf(r12 + vq);
    const float f = vc * r6 * iden;
    const float df = 6.0f * vc * (r6 * inv_r) * iden * ((vq - r12) * iden);
    const float V = -S * f;
    const float dV = -(dS * f + S * df);
    acc.ev += 0.5 * (double)V;
    dEdr += dV * B2;
  }
  // dE/dx_i = dE/dr_A * (x_i - x_j)/r_A = -dEdr * d / r_A
  const float sc = -dEdr * inv_rA;
  acc.gx += (double)(sc * dx);
  acc.gy += (double)(sc * dy);
  acc.gz += (double)(sc * dz);
}

constexpr int QCAP = 128;   // compacted pair queue per wave

// One wavefront per centre.  Candidates (contiguous in cell order, or the
// molecule's atoms) are tested 64 at a time; those inside the cutoff are
// compacted into an LDS queue (minimum-image separation and charge) and
// evaluated 64 at a time with every lane busy.  Queue order = candidate
// order, so the per-lane sums and the final wave reduction are formed in a
// fixed order.
// MOLC: the cell-ordered candidates carry molecule ids (several systems share
// the grid); a template parameter so the candidate loads are branch-free.
template <bool MOLC>
__global__ void __launch_bounds__(C_WAVES * 64)
coulomb_kernel(int N, const float4* __restrict__ xq, const int* __restrict__ elem,
               const float4* __restrict__ xqc, const int* __restrict__ elc,
               const int* __restrict__ mol, const int* __restrict__ mol_start,
               const double* __restrict__ box, CellGrid grid, CoulArgs a, double* __restrict__ phi,
               double* __restrict__ grad, double* __restrict__ eatom, int* __restrict__ flags) {
  __shared__ float4 qx[C_WAVES][QCAP];
  __shared__ int qe[C_WAVES][QCAP];
  __shared__ RangeLds<C_RANGES> rng[C_WAVES];
  __shared__ float vdw_c[TM_MAX_ELE * TM_MAX_ELE], vdw_q[TM_MAX_ELE * TM_MAX_ELE];
  if (threadIdx.x < TM_MAX_ELE * TM_MAX_ELE) {
    const int ea = threadIdx.x / TM_MAX_ELE, eb = threadIdx.x % TM_MAX_ELE;
    const float R0 = a.r0[ea] + a.r0[eb];
    const float q2 = R0 * R0, q6 = q2 * q2 * q2;
    vdw_c[threadIdx.x] = sqrtf(a.c6[ea] * a.c6[eb]);
    vdw_q[threadIdx.x] = 6.0f * q6 * q6;
  }
  __syncthreads();
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wi = blockIdx.x * C_WAVES + w;
  if (wi >= N) return;   // wave-level synchronisation only below
  // centres are visited in cell order so neighbouring waves share candidates in L2
  const int i = grid.cell_atoms != nullptr ? grid.cell_atoms[wi] : wi;
  const int cpi = grid.cell_atoms != nullptr ? wi : -1;
  const int mi = mol[i];
  float Lx = 0.f, Ly = 0.f, Lz = 0.f, iLx = 0.f, iLy = 0.f, iLz = 0.f;
  if (box != nullptr) {
    Lx = (float)box[3 * mi]; Ly = (float)box[3 * mi + 1]; Lz = (float)box[3 * mi + 2];
    iLx = 1.0f / Lx; iLy = 1.0f / Ly; iLz = 1.0f / Lz;
  }
  const float4 pi = cpi >= 0 ? xqc[cpi] : xq[i];
  const int ei = cpi >= 0 ? elc[cpi] : elem[i];
  const float* vci = vdw_c + ei * TM_MAX_ELE;
  const float* vqi = vdw_q + ei * TM_MAX_ELE;
  PairAcc acc = {0.0, 0.0, 0.0, 0.0, 0.0};
  int qn = 0;   // wave-uniform queue length
  auto drain = [&](int n) {   // evaluate the first n (<= 64) queued pairs
    if (lane < n) {
      const float4 d = qx[w][lane];
      const int ej = qe[w][lane];
      pair_terms(a, pi.w, d.w, vci[ej], vqi[ej], d.x, d.y, d.z, acc);
    }
  };
  // candidate payload, fetched one chunk ahead of its test
  struct Cand {
    float4 x;
    int e, m;
  };
  // Branch-free, so the same number of loads is in flight on every path and
  // the wait for a chunk's data does not also wait for the next chunk's.
  // Invalid lanes read entry 0 (N >= 1) and are masked by `valid` in visit.
  auto load = [&](int p, bool valid, int mapped) {
    const int q = valid ? p : 0;
    Cand c;
    c.x = mapped ? xqc[q] : xq[q];
    c.e = mapped ? elc[q] : elem[q];
    c.m = (MOLC && mapped) ? grid.molc[q] : mi;
    return c;
  };
  auto visit = [&](int p, bool valid, int mapped, const Cand& c) {
    // every lane holds a loaded entry (invalid lanes entry 0), so the
    // separation is formed branch-free and validity enters the test
    float4 d;
    d.x = min_image_f(c.x.x - pi.x, Lx, iLx);
    d.y = min_image_f(c.x.y - pi.y, Ly, iLy);
    d.z = min_image_f(c.x.z - pi.z, Lz, iLz);
    d.w = c.x.w;
    const bool ok = valid && (mapped ? p != cpi : p != i) && c.m == mi &&
                    d.x * d.x + d.y * d.y + d.z * d.z < a.rc2;
    unsigned long long bal = __ballot(ok);
    int before = tm_lanes_below(bal);
    if (ok) {
      qx[w][qn + before] = d;
      qe[w][qn + before] = c.e;
    }
    qn += __popcll(bal);
    __syncwarp();
    if (qn >= 64) {
      drain(64);
      __syncwarp();
      // move the remainder (< 64 entries) to the front
      float4 t4;
      int te = 0;
      bool mv = lane < qn - 64;
      if (mv) { t4 = qx[w][64 + lane]; te = qe[w][64 + lane]; }
      __syncwarp();
      if (mv) { qx[w][lane] = t4; qe[w][lane] = te; }
      qn -= 64;
      __syncwarp();
    }
  };
  tm_sweep_pf(i, mol, mol_start, grid, a.rc, rng[w], lane, flags, (double)pi.x, (double)pi.y,
              (double)pi.z, load, visit);
  drain(qn);
  double sphi = tm_wave_sum_d(acc.phi);
  double gx = tm_wave_sum_d(acc.gx);
  double gy = tm_wave_sum_d(acc.gy);
  double gz = tm_wave_sum_d(acc.gz);
  double ev = tm_wave_sum_d(acc.ev);
  if (lane == 0) {
    phi[i] = sphi;
    grad[3 * i] = gx; grad[3 * i + 1] = gy; grad[3 * i + 2] = gz;
    eatom[2 * i] = 0.5 * (double)pi.w * sphi;
    eatom[2 * i + 1] = ev;
  }
}

// positions and charges in cell order for the candidate sweeps
__global__ void cell_gather_kernel(int N, const int* __restrict__ cell_atoms,
                                   const float4* __restrict__ xq, const int* __restrict__ elem,
                                   float4* __restrict__ xqc, int* __restrict__ elc) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  int i = cell_atoms[p];
  xqc[p] = xq[i];
  elc[p] = elem[i];
}

// Per-molecule column sums of an [N, C] array (float or double input) into
// out[B, C] (double); one block per molecule, fixed reduction tree.
template <class T, int NT>
__global__ void __launch_bounds__(NT)
mol_sum_kernel(const int* __restrict__ mol_start, int C, const T* __restrict__ in, int ldin,
               double* __restrict__ out, int scale_by_count) {
  constexpr int CB = 4;   // columns accumulated per pass over the atoms
  __shared__ double part[CB][NT / 64];
  const int m = blockIdx.x;
  const int lo = mol_start[m], hi = mol_start[m + 1];
  for (int c0 = 0; c0 < C; c0 += CB) {
    const int nc = min(CB, C - c0);
    double acc[CB] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
    for (int a = lo + threadIdx.x; a < hi; a += NT) {
      const T* row = in + (long)a * ldin + c0;
#pragma unroll
      for (int c = 0; c < CB; ++c)
        if (c < nc) acc[c] += (double)row[c];
    }
    // fixed-order reduction: butterfly inside each wave, then the wave sums in order
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < CB; ++c) {
      double v = acc[c];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
      if (lane == 0) part[c][wv] = v;
    }
    __syncthreads();
    if (threadIdx.x < nc) {
      double v = 0.0;
      for (int q = 0; q < NT / 64; ++q) v += part[threadIdx.x][q];
      if (scale_by_count) v = hi > lo ? v / (double)(hi - lo) : 0.0;
      out[(long)m * C + c0 + threadIdx.x] = v;
    }
    __syncthreads();
  }
}

// One large molecule in one launch: block k sums atoms [1024 k, 1024 (k + 1))
// for every column into partial[k][c]; the last block to finish (completion
// counter in work, agent-scope acquire/release) sums the partials in block
// order and resets the counter to zero.  Deterministic whatever the block
// that finishes last.
template <class T>
__global__ void __launch_bounds__(256)
sum_last_block_kernel(int n, int C, const T* __restrict__ in, int ldin, double* work,
                      double* __restrict__ out, int scale_by_count) {
  constexpr int CB = 4;                    // columns per pass
  __shared__ double part[CB][256 / 64];
  __shared__ double stage[256];
  __shared__ int last;
  unsigned int* counter = reinterpret_cast<unsigned int*>(work);
  double* partial = work + 1;
  const int k = blockIdx.x;
  const int lo = k * 1024, hi = min(n, lo + 1024);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int c0 = 0; c0 < C; c0 += CB) {
    const int nc = min(CB, C - c0);
    // all loads issued before any is used (rows past the end read row n - 1)
    T v[4][CB];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const long a = min(lo + u * 256 + (int)threadIdx.x, n - 1);
#pragma unroll
      for (int c = 0; c < CB; ++c) v[u][c] = in[a * ldin + c0 + min(c, nc - 1)];
    }
#pragma unroll
    for (int c = 0; c < CB; ++c) {
      double acc = 0.0;
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (lo + u * 256 + (int)threadIdx.x < hi) acc += (double)v[u][c];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
      if (lane == 0) part[c][wv] = acc;
    }
    __syncthreads();
    if ((int)threadIdx.x < nc)
      partial[(long)k * C + c0 + threadIdx.x] =
          ((part[threadIdx.x][0] + part[threadIdx.x][1]) + part[threadIdx.x][2]) +
          part[threadIdx.x][3];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const unsigned int done =
        __hip_atomic_fetch_add(counter, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    last = done == gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  // the other blocks' partials are visible after the acquire: stage them in
  // LDS 256 at a time, then sum each column in block order
  double tot = 0.0;   // thread c < C: column c
  const int nb = gridDim.x;
  for (int c = 0; c < C; ++c) {
    for (int b0 = 0; b0 < nb; b0 += 256) {
      const int b = b0 + threadIdx.x;
      stage[threadIdx.x] = b < nb ? partial[(long)b * C + c] : 0.0;
      __syncthreads();
      if ((int)threadIdx.x == c) {
        const int m = min(256, nb - b0);
        for (int q = 0; q < m; ++q) tot += stage[q];
      }
      __syncthreads();
    }
  }
  if ((int)threadIdx.x < C) {
    if (scale_by_count) tot = n > 0 ? tot / (double)n : 0.0;
    out[threadIdx.x] = tot;
  }
  if (threadIdx.x == 0)
    __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Several per-molecule column sums in one launch.  Source k is an [N, ld[k]]
// array (float64 when f64[k], else float32) of which the first cols[k]
// columns are summed; columns are numbered across sources in order.
constexpr int SUMS_MAX_SRC = 4, SUMS_MAX_COL = 16;
struct SumSrc {
  const void* in[SUMS_MAX_SRC];
  double* out[SUMS_MAX_SRC];
  int f64[SUMS_MAX_SRC], cols[SUMS_MAX_SRC], ld[SUMS_MAX_SRC];
  int nsrc, ctot;
};

__device__ __forceinline__ double src_val(const SumSrc& S, int k, long a, int c) {
  return S.f64[k] ? ((const double*)S.in[k])[a * S.ld[k] + c]
                  : (double)((const float*)S.in[k])[a * S.ld[k] + c];
}

// one large system: block k sums atoms [1024 k, 1024 (k + 1)) of every column;
// the last block to finish sums the partials in block order (as
// sum_last_block_kernel) and resets the counter
__global__ void __launch_bounds__(256)
sums_last_block_kernel(int n, SumSrc S, double* work) {
  __shared__ double part[SUMS_MAX_COL][4];
  __shared__ double stage[256];
  __shared__ int last;
  unsigned int* counter = reinterpret_cast<unsigned int*>(work);
  double* partial = work + 1;
  const int k = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int lo = k * 1024, hi = min(n, lo + 1024);
  int cg = 0;
  for (int q = 0; q < S.nsrc; ++q) {
    for (int c = 0; c < S.cols[q]; ++c, ++cg) {
      double v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) v[u] = src_val(S, q, min(lo + u * 256 + tid, n - 1), c);
      double acc = 0.0;
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (lo + u * 256 + tid < hi) acc += v[u];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
      if (lane == 0) part[cg][wv] = acc;
    }
  }
  __syncthreads();
  const int C = S.ctot;
  if (tid < C) partial[(long)k * C + tid] = ((part[tid][0] + part[tid][1]) + part[tid][2]) + part[tid][3];
  __syncthreads();
  if (tid == 0) {
    const unsigned int done =
        __hip_atomic_fetch_add(counter, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    last = done == gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  // partials [nb][C] staged 256 at a time; thread c < C adds column c in block order
  const int tot_n = (int)gridDim.x * C;
  double tot = 0.0;
  for (int e0 = 0; e0 < tot_n; e0 += 256) {
    stage[tid] = e0 + tid < tot_n ? partial[e0 + tid] : 0.0;
    __syncthreads();
    if (tid < C) {
      const int e1 = min(e0 + 256, tot_n);
      for (int e = e0 + ((tid - e0 % C) + C) % C; e < e1; e += C) tot += stage[e - e0];
    }
    __syncthreads();
  }
  if (tid < C) {
    int q = 0, c = tid;
    while (c >= S.cols[q]) { c -= S.cols[q]; ++q; }
    S.out[q][c] = tot;
  }
  if (tid == 0) __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// one block per molecule, every column; fixed reduction tree
template <int NT>
__global__ void __launch_bounds__(NT)
sums_per_mol_kernel(const int* __restrict__ mol_start, SumSrc S) {
  __shared__ double part[NT / 64];
  const int m = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int lo = mol_start[m], hi = mol_start[m + 1];
  for (int q = 0; q < S.nsrc; ++q) {
    for (int c = 0; c < S.cols[q]; ++c) {
      double acc = 0.0;
      for (int a = lo + tid; a < hi; a += NT) acc += src_val(S, q, a, c);
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
      if (lane == 0) part[wv] = acc;
      __syncthreads();
      if (tid == 0) {
        double v = 0.0;
        for (int w = 0; w < NT / 64; ++w) v += part[w];
        S.out[q][(long)m * S.cols[q] + c] = v;
      }
      __syncthreads();
    }
  }
}

template <class T>
int mol_sum(int B, const int* mol_start, int C, const T* in, int ldin, double* out, int mean,
            double* work, int n_total, hipStream_t s) {
  if (B == 1 && n_total > 2048 && work != nullptr) {
    sum_last_block_kernel<T><<<(n_total + 1023) / 1024, 256, 0, s>>>(n_total, C, in, ldin,
                                                                     work, out, mean);
    TM_LAUNCH_CHECK();
    return TM_OK;
  }
  if (B == 1 && n_total > 2048) {
    // one system without work: a single 1024-thread block, one launch
    mol_sum_kernel<T, 1024><<<1, 1024, 0, s>>>(mol_start, C, in, ldin, out, mean);
    TM_LAUNCH_CHECK();
    return TM_OK;
  }
  mol_sum_kernel<T, 256><<<B, 256, 0, s>>>(mol_start, C, in, ldin, out, mean);
  TM_LAUNCH_CHECK();
  return TM_OK;
}

// Charges and their cell-ordered copies.  MEANS: the molecule means of qraw
// are formed in the block itself: every block sums each molecule its atoms
// belong to over all of that molecule's atoms, in one fixed order (thread t
// takes atoms lo + t + 256 j into partial sum j mod 8, the eight partials in a
// fixed tree, a butterfly per wave, then the four wave totals in order), so
// all blocks agree bit for bit and no launch or cross-block wait precedes the
// charges.  The block that holds a molecule's
// first atom writes its mean to qmean (when given).  Otherwise qmean is input.
// The wave totals alternate between two LDS slots, so one barrier per
// molecule suffices: a slot is rewritten two molecules later, after the
// barrier that every reader of its previous contents has passed.
template <bool MEANS>
__global__ void __launch_bounds__(256)
charges_kernel(int N, const int* __restrict__ mol, const int* __restrict__ mol_start,
               double* __restrict__ qmean, const float* __restrict__ qraw,
               const double* __restrict__ pos, float4* __restrict__ xq, double* __restrict__ q,
               double* __restrict__ qx, const int* __restrict__ cell_pos,
               const int* __restrict__ elem, float4* __restrict__ xqc, int* __restrict__ elc) {
  __shared__ double wpart[2][4];   // wave totals, alternating between molecules
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double mine = 0.0;   // the mean of this thread's molecule
  if (MEANS) {
    const int i0 = blockIdx.x * blockDim.x;
    const int m0 = mol[i0];
    const int m1 = mol[min(i0 + (int)blockDim.x, N) - 1];
    const int my_m = i < N ? mol[i] : -1;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int mm = m0; mm <= m1; ++mm) {   // molecules without atoms cost one pass
      const int lo = mol_start[mm], hi = mol_start[mm + 1];
      // eight independent partial sums (loads in flight), combined in a fixed tree
      double p8[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
      int a = lo + (int)threadIdx.x;
      for (; a + 7 * 256 < hi; a += 8 * 256) {
#pragma unroll
        for (int k = 0; k < 8; ++k) p8[k] += (double)qraw[a + k * 256];
      }
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (a + k * 256 < hi) p8[k] += (double)qraw[a + k * 256];
      double acc = ((p8[0] + p8[1]) + (p8[2] + p8[3])) + ((p8[4] + p8[5]) + (p8[6] + p8[7]));
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
      double* part = wpart[(mm - m0) & 1];
      if (lane == 0) part[w] = acc;
      __syncthreads();   // part of the molecule before this one is no longer read
      const double tot = ((part[0] + part[1]) + part[2]) + part[3];
      const double mean = hi > lo ? tot / (double)(hi - lo) : 0.0;
      if (my_m == mm) mine = mean;
      if (threadIdx.x == 0 && qmean != nullptr && lo >= i0) qmean[mm] = mean;
    }
  }
  if (i >= N) return;
  const double mi = MEANS ? mine : qmean[mol[i]];
  double qi = (double)qraw[i] - mi;
  q[i] = qi;
  double x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
  const float4 v = make_float4((float)x, (float)y, (float)z, (float)qi);
  if (xq != nullptr) xq[i] = v;
  if (cell_pos != nullptr) {   // the cell-ordered copies the Coulomb sweep reads
    const int p = cell_pos[i];
    xqc[p] = v;
    elc[p] = elem[i];
  }
  qx[3 * i] = qi * x * BOHRPERA_D;
  qx[3 * i + 1] = qi * y * BOHRPERA_D;
  qx[3 * i + 2] = qi * z * BOHRPERA_D;
}

__global__ void charge_upstream_kernel(int N, const int* __restrict__ mol,
                                       const double* __restrict__ phi,
                                       const double* __restrict__ phimean, float* __restrict__ u) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  u[i] = (float)(phi[i] - phimean[mol[i]]);
}

}  // namespace

extern "C" {

// One large molecule (B == 1, n_total = N > 2048) is summed by ceil(N/1024)
// blocks in one launch when work is given (1 + ceil(N/1024)*C doubles, its
// first word a completion counter that must be zero before the first call and
// is zero again after each), else by a single 1024-thread block.
int tm_mol_sum_f32(int B, const int* mol_start, int C, const float* in, int ldin, double* out,
                   int mean, double* work, int n_total, void* stream) {
  TM_REQUIRE(B >= 0 && C > 0 && C <= 64, "tm_mol_sum_f32: bad sizes");
  if (B == 0) return TM_OK;
  return mol_sum<float>(B, mol_start, C, in, ldin, out, mean, work, n_total, (hipStream_t)stream);
}

int tm_mol_sum_f64(int B, const int* mol_start, int C, const double* in, int ldin, double* out,
                   int mean, double* work, int n_total, void* stream) {
  TM_REQUIRE(B >= 0 && C > 0 && C <= 64, "tm_mol_sum_f64: bad sizes");
  if (B == 0) return TM_OK;
  return mol_sum<double>(B, mol_start, C, in, ldin, out, mean, work, n_total,
                         (hipStream_t)stream);
}

}  // extern "C"

namespace {

// The per-molecule sums of tm_mol_sums; `work` (may be null) is the block
// partials and completion counter of the many-block path for one system.
int mol_sums_impl(const char* name, int B, const int* mol_start, int n_total, int nsrc,
                  const void* const* in, const int* f64, const int* cols, const int* ld,
                  double* const* out, double* work, void* stream) {
  TM_REQUIRE(B >= 0 && n_total >= 0 && nsrc > 0 && nsrc <= SUMS_MAX_SRC, "%s: bad sizes", name);
  SumSrc S;
  S.nsrc = nsrc;
  S.ctot = 0;
  for (int k = 0; k < SUMS_MAX_SRC; ++k) {
    const bool on = k < nsrc;
    S.in[k] = on ? in[k] : nullptr;
    S.out[k] = on ? out[k] : nullptr;
    S.f64[k] = on ? f64[k] : 0;
    S.cols[k] = on ? cols[k] : 0;
    S.ld[k] = on ? ld[k] : 0;
    if (on) {
      TM_REQUIRE(cols[k] > 0 && ld[k] >= cols[k] && in[k] != nullptr && out[k] != nullptr,
                 "%s: bad source %d", name, k);
      S.ctot += cols[k];
    }
  }
  TM_REQUIRE(S.ctot <= SUMS_MAX_COL, "%s: %d columns above %d", name, S.ctot, SUMS_MAX_COL);
  if (B == 0) return TM_OK;
  hipStream_t s = (hipStream_t)stream;
  if (B == 1 && n_total > 2048 && work != nullptr) {
    sums_last_block_kernel<<<(n_total + 1023) / 1024, 256, 0, s>>>(n_total, S, work);
  } else if (B == 1 && n_total > 2048) {
    sums_per_mol_kernel<1024><<<1, 1024, 0, s>>>(mol_start, S);
  } else {
    sums_per_mol_kernel<256><<<B, 256, 0, s>>>(mol_start, S);
  }
  TM_LAUNCH_CHECK();
  return TM_OK;
}

// ---- the entry point that allocates nothing --------------------------------
// The many-block path's partials and counter live in this file-scope array
// instead of a caller's workspace: room for one system of up to
// STATIC_SUM_MAX_ATOMS atoms and SUMS_MAX_COL columns (512 KiB per GPU, held
// until the process ends).  Word 0 is the completion counter, cleared once per
// device at set-up and left at zero by every call (the last block resets it).
constexpr int STATIC_SUM_MAX_ATOMS = 1 << 22;
constexpr int STATIC_SUM_WORDS = 1 + (STATIC_SUM_MAX_ATOMS / 1024) * SUMS_MAX_COL;
static __device__ __align__(256) double scratch[STATIC_SUM_WORDS];

constexpr int STATIC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[STATIC_MAX_DEVICES] = {};   // device address of `scratch`, per device

__global__ void clear_counter_kernel(double* p) { p[0] = 0.0; }

// Device address of this GPU's `scratch` (set up on the first call on a device:
// hipGetSymbolAddress with that device current, then the counter cleared by a
// kernel on the caller's stream).  The array's name is a host-side stand-in and
// is never passed to a kernel.
int static_scratch(hipStream_t s, double** out) {
  int dev = 0;
  TM_HIP_CHECK(hipGetDevice(&dev));
  TM_REQUIRE(dev >= 0 && dev < STATIC_MAX_DEVICES, "tm_mol_sums_static: device %d", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    TM_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    clear_counter_kernel<<<1, 1, 0, s>>>(static_cast<double*>(p));
    TM_LAUNCH_CHECK();
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return TM_OK;
}

}  // namespace

extern "C" {

int tm_mol_sums(int B, const int* mol_start, int n_total, int nsrc, const void* const* in,
                const int* f64, const int* cols, const int* ld, double* const* out,
                double* work, void* stream) {
  return mol_sums_impl("tm_mol_sums", B, mol_start, n_total, nsrc, in, f64, cols, ld, out, work,
                       stream);
}

int tm_mol_sums_static(int B, const int* mol_start, int n_total, int nsrc, const void* const* in,
                       const int* f64, const int* cols, const int* ld, double* const* out,
                       void* stream) {
  if (n_total > STATIC_SUM_MAX_ATOMS) {
    tm_set_error("tm_mol_sums_static: %d atoms above the %d its static scratch holds", n_total,
                 STATIC_SUM_MAX_ATOMS);
    return TM_ERR_CAPACITY;
  }
  double* work = nullptr;
  int st = static_scratch((hipStream_t)stream, &work);
  if (st) return st;
  return mol_sums_impl("tm_mol_sums_static", B, mol_start, n_total, nsrc, in, f64, cols, ld, out,
                       work, stream);
}

int tm_charges(int N, const int* mol, const int* mol_start, double* qmean, const float* qraw,
               const double* pos, float* xq, double* q, double* qx, const int* cell_pos,
               const int* elem, float* xqc, int* elc, void* stream) {
  TM_REQUIRE(cell_pos == nullptr || (elem != nullptr && xqc != nullptr && elc != nullptr),
             "tm_charges: cell_pos needs elem, xqc and elc");
  TM_REQUIRE(xq != nullptr || cell_pos != nullptr, "tm_charges: no position output");
  TM_REQUIRE(mol_start != nullptr || qmean != nullptr,
             "tm_charges: needs qmean, or mol_start to form the means");
  if (N == 0) return TM_OK;
  hipStream_t s = (hipStream_t)stream;
  auto k = mol_start != nullptr ? charges_kernel<true> : charges_kernel<false>;
  k<<<tm_div_up(N, 256), 256, 0, s>>>(N, mol, mol_start, qmean, qraw, pos, (float4*)xq, q, qx,
                                      cell_pos, elem, (float4*)xqc, elc);
  TM_LAUNCH_CHECK();
  return TM_OK;
}

int tm_charge_upstream(int N, const int* mol, const double* phi, const double* phimean, float* u,
                       void* stream) {
  if (N == 0) return TM_OK;
  charge_upstream_kernel<<<tm_div_up(N, 256), 256, 0, (hipStream_t)stream>>>(N, mol, phi, phimean,
                                                                             u);
  TM_LAUNCH_CHECK();
  return TM_OK;
}

// consts: [alpha', R_sr, R_lr, ZZ, YY, elu_shift, elu_alpha] (bohr), rc_ee (A),
// poly width (bohr*BOHRPERA), c6[nele], r0[nele] (host arrays, atomic units).
// With a cell grid, xqc (float4 [N]) and elc (int [N]) receive positions,
// charges and elements in cell order.
int tm_coulomb(int N, const float* xq, const int* elem, const int* mol, const int* mol_start,
               const double* box, const int* cell_start, const int* cell_atoms, const int* cell_of,
               const int* cell_pos, const int* molc, int nx, int ny, int nz, double csx, double csy,
               double csz, double ox, double oy, double oz, int periodic, const double* consts,
               double rc_ee, double width,
               int add_ecc, int has_vdw, int nele, const double* c6, const double* r0, float* xqc,
               int* elc, double* phi, double* grad, double* eatom, int* flags, void* stream) {
  TM_REQUIRE(nele > 0 && nele <= TM_MAX_ELE, "tm_coulomb: bad nele");
  TM_REQUIRE(xq != nullptr || cell_atoms != nullptr, "tm_coulomb: xq is needed without a cell grid");
  TM_REQUIRE(cell_atoms == nullptr || (xqc != nullptr && elc != nullptr && cell_pos != nullptr),
             "tm_coulomb: cell sweeps need cell_pos and the xqc/elc buffers");
  if (cell_atoms != nullptr) {
    const double cs[3] = {csx, csy, csz};
    const int nc[3] = {nx, ny, nz};
    TM_REQUIRE(tm_range_bound(rc_ee, cs, nc, periodic) <= C_RANGES,
               "tm_coulomb: cells of %.2f A are too small for a %.1f A cutoff (a sweep could "
               "need %d candidate ranges, room for %d)", fmin(csx, fmin(csy, csz)), rc_ee,
               tm_range_bound(rc_ee, cs, nc, periodic), C_RANGES);
  }
  if (N == 0) return TM_OK;
  CoulArgs a;
  a.alpha = (float)consts[0]; a.r_sr = (float)consts[1]; a.r_lr = (float)consts[2];
  a.zz = (float)consts[3]; a.yy = (float)consts[4]; a.shift = (float)consts[5];
  a.ea = (float)consts[6];
  a.lin = (float)(consts[3] / consts[2] + consts[4]);
  a.rc = (float)rc_ee;
  a.rc2 = (float)(rc_ee * rc_ee);
  a.inv_width = (float)(1.0 / width);
  a.add_ecc = add_ecc;
  a.has_vdw = has_vdw;
  for (int e = 0; e < TM_MAX_ELE; ++e) {
    a.c6[e] = e < nele ? (float)c6[e] : 0.f;
    a.r0[e] = e < nele ? (float)r0[e] : 0.f;
  }
  CellGrid g;
  g.cell_start = cell_start; g.cell_atoms = cell_atoms; g.cell_of = cell_of;
  g.cell_pos = cell_pos; g.posc = nullptr; g.molc = molc;
  g.nc[0] = nx; g.nc[1] = ny; g.nc[2] = nz;
  g.cs[0] = (float)csx; g.cs[1] = (float)csy; g.cs[2] = (float)csz;
  g.org[0] = (float)ox; g.org[1] = (float)oy; g.org[2] = (float)oz;
  g.periodic = periodic;
  hipStream_t s = (hipStream_t)stream;
  if (cell_atoms != nullptr && xq != nullptr) {
    cell_gather_kernel<<<tm_div_up(N, 256), 256, 0, s>>>(N, cell_atoms, (const float4*)xq, elem,
                                                         (float4*)xqc, elc);
    TM_LAUNCH_CHECK();
  }
  auto k = (cell_atoms != nullptr && molc != nullptr) ? coulomb_kernel<true>
                                                      : coulomb_kernel<false>;
  k<<<tm_div_up(N, C_WAVES), C_WAVES * 64, 0, s>>>(N, (const float4*)xq, elem,
                                                   (const float4*)xqc, elc, mol, mol_start, box, g,
                                                   a, phi, grad, eatom, flags);
  TM_LAUNCH_CHECK();
  return TM_OK;
}

}  // extern "C"
