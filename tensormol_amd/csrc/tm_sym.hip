// ANI-style radial and angular symmetry functions and their adjoint.
//
// Forward restates TFSymRSet_Linear_WithEle (TensorMol/TFDescriptors/
// RawSymFunc.py:1696-1755), TFSymASet_Linear_WithEle (:868-962) and the
// concatenation of TFSymSet_Scattered_Linear_WithEle (:2223-2265):
//   G[i, l*NR + s]          = sum_j exp(-eta (r_ij - Rs_s)^2) fc_r(r_ij)
//   G[i, nele*NR + p*NA*NT + t*NA + s]
//        = sum_{j<k} 2^(1-zeta) (1 + cos(theta_ijk - theta_t))^zeta
//                     exp(-eta ((r_ij + r_ik)/2 - Rs_s)^2) fc_a(r_ij) fc_a(r_ik)
// with fc(r) = 0.5 (cos(3.14159265359 r / Rc) + 1) and the neighbour list as
// the only cutoff mask.  cos(theta - theta_t) is evaluated as
// c cos(theta_t) + s sin(theta_t) with c = cos(theta) and s = sin(theta) >= 0
// (taken from the cross product of the two bonds), which is the same function
// as the reference's acos-then-cos.
//
// One wavefront per centre atom.  Neighbour geometry is staged in LDS; the
// angular triples of a centre are enumerated on the fly from its short
// sub-list, grouped by element pair, so nothing per-triple is stored in HBM.
// Every output element is written exactly once; no atomics.
//
// The adjoint takes dE/dG per atom and returns, for every neighbour slot
// (i -> j), the vector w_ij = dE/d(x_j - x_i) of centre i's descriptor; a
// gather over mirror slots turns those into per-atom gradients without
// atomics, so forces are bitwise reproducible.
#include "tm_common.h"

namespace {

constexpr int SYM_WAVES = 4;
constexpr int MAX_ROW = 256;      // widest radial row (row stride) the kernels stage in LDS
constexpr int NR_MAX = 64;        // radial shells
constexpr int NA = 8;             // angular shells (fixed by the lane mapping)
constexpr int NT = 8;             // angular theta points

struct SymArgs {
  int nele, nelep, nr;            // elements, element pairs, radial shells
  int D;                          // descriptor length
  float eta, zeta, rr, ra;        // AN1 parameters and cutoffs (A)
  float two_1mz;                  // 2^(1-zeta)
  int zeta_int;                   // zeta when it is a small positive integer, else 0
  float kr, ka;                    // pi / Rc_r, pi / Rc_a (pi as the reference's literal)
  float rs_r[NR_MAX];
  float rs_a[NA];
  float cos_t[NT], sin_t[NT];
};

constexpr float PI_LIT = 3.14159265359f;

// b^zeta: ZI > 0 is zeta known at compile time (a few multiplies); ZI = 0 reads
// zeta from the arguments (square-and-multiply when it is a small integer).
template <int ZI>
__device__ __forceinline__ float powz(float b, const SymArgs& a) {
  const int zi = ZI > 0 ? ZI : a.zeta_int;
  if (zi > 0) {   // square-and-multiply over the 7 bits of zeta <= 64
    float r = 1.f, x = b;
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      r = ((zi >> k) & 1) ? r * x : r;
      x *= x;
    }
    return r;
  }
  return b > 0.f ? __expf(a.zeta * __logf(b)) : 0.f;
}

// (1 + v)^zeta and zeta (1 + v)^(zeta - 1) for b = 1 + v >= 0, without a division
// when zeta is a small integer: r = b^(zeta - 1) by square-and-multiply.
template <int ZI>
__device__ __forceinline__ void powz_pair(float b, const SymArgs& a, float& Az, float& dAz) {
  const int zi = ZI > 0 ? ZI : a.zeta_int;
  if (zi > 0) {
    const int e = zi - 1;
    float r = 1.f, x = b;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      r = ((e >> k) & 1) ? r * x : r;
      x *= x;
    }
    Az = r * b;
    dAz = (ZI > 0 ? (float)ZI : a.zeta) * r;
    return;
  }
  Az = b > 0.f ? __expf(a.zeta * __logf(b)) : 0.f;
  dAz = b > 0.f ? a.zeta * Az * __builtin_amdgcn_rcpf(b) : 0.f;
}

template <int RC>
struct RowLds {
  static constexpr bool has_fc = true, has_fa = true, has_el = true;
  float4 d[SYM_WAVES][RC];      // (x_j - x_i, r)
  float fc[SYM_WAVES][RC];      // radial cutoff function
  float fa[SYM_WAVES][RC];      // angular cutoff function
  int el[SYM_WAVES][RC];
};
// The forward's row: the angular cutoff is formed once per triple instead,
// which keeps its LDS small enough for 7 waves per SIMD.
template <int RC>
struct RowLdsFwd {
  static constexpr bool has_fc = true, has_fa = false, has_el = false;
  float4 d[SYM_WAVES][RC];
  float fc[SYM_WAVES][RC];
};
// The adjoint's row: its radial pass forms fc and reads the element itself
// (lane = neighbour), which keeps its LDS small enough for 6 waves per SIMD.
template <int RC>
struct RowLdsAdj {
  static constexpr bool has_fc = false, has_fa = true, has_el = false;
  float4 d[SYM_WAVES][RC];
  float fa[SYM_WAVES][RC];
};

__device__ __forceinline__ float cutoff_fc(float k, float r) { return 0.5f * (__cosf(k * r) + 1.0f); }

// sin of the angle between the bond vectors dj and dk (inv = 1 / (|dj| |dk|)),
// from the cross product: its relative error grows as eps / sin(theta) for
// nearly collinear bonds, where sqrt(1 - cos^2) loses every digit (eps / sin^2).
__device__ __forceinline__ float sin_between(const float4& dj, const float4& dk, float inv) {
  const float cx = dj.y * dk.z - dj.z * dk.y;
  const float cy = dj.z * dk.x - dj.x * dk.z;
  const float cz = dj.x * dk.y - dj.y * dk.x;
  return fminf(sqrtf(cx * cx + cy * cy + cz * cz) * inv, 1.0f);
}
__device__ __forceinline__ float cutoff_dfc(float k, float r) { return -0.5f * k * __sinf(k * r); }

template <class LDS>
__device__ __forceinline__ void load_row(int i, int w, int lane, const double* __restrict__ pos,
                                         const double* __restrict__ box, const int* __restrict__ mol,
                                         const int* __restrict__ elem, const int* __restrict__ nbr,
                                         int n, int ldr, const SymArgs& a, LDS& L) {
  // minimum image with the reciprocal edge (iL = 0 for open boundaries): it
  // differs from d - L rint(d / L) only at |d| = L/2, far beyond the cutoff
  double Lx = 0, Ly = 0, Lz = 0, iLx = 0, iLy = 0, iLz = 0;
  if (box != nullptr) {
    int m = mol[i];
    Lx = box[3 * m]; Ly = box[3 * m + 1]; Lz = box[3 * m + 2];
    iLx = 1.0 / Lx; iLy = 1.0 / Ly; iLz = 1.0 / Lz;
  }
  const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
  for (int q = lane; q < n; q += 64) {
    int j = nbr[(long)i * ldr + q];
    double ex = pos[3 * j] - xi, ey = pos[3 * j + 1] - yi, ez = pos[3 * j + 2] - zi;
    float dx = (float)(ex - Lx * rint(ex * iLx));
    float dy = (float)(ey - Ly * rint(ey * iLy));
    float dz = (float)(ez - Lz * rint(ez * iLz));
    float r = sqrtf(dx * dx + dy * dy + dz * dz + 1e-27f);
    L.d[w][q] = make_float4(dx, dy, dz, r);
    if constexpr (LDS::has_fa) L.fa[w][q] = cutoff_fc(a.ka, r);
    if constexpr (LDS::has_fc) L.fc[w][q] = cutoff_fc(a.kr, r);
    if constexpr (LDS::has_el) L.el[w][q] = elem[j];
  }
}

// decode triple index t (in the centre's enumeration) into sub-list
// positions (ja, kb) and the element pair p; blocks are ordered by
// (ea <= eb) and inside a block by (k, j)
__device__ __forceinline__ void decode_triple(int t, const int* seg, int nele, int& ja, int& kb,
                                              int& p) {
  p = 0;
  for (int ea = 0; ea < nele; ++ea) {
    int la = seg[ea + 1] - seg[ea];
    for (int eb = ea; eb < nele; ++eb, ++p) {
      int lb = seg[eb + 1] - seg[eb];
      int nb = (ea == eb) ? la * (la - 1) / 2 : la * lb;
      if (t < nb) {
        if (ea != eb) {
          kb = seg[eb] + t / la;
          ja = seg[ea] + t % la;
        } else {
          // k-major triangle: k = 1.., j < k; t = k(k-1)/2 + j
          int k = (int)((1.0f + sqrtf(1.0f + 8.0f * (float)t)) * 0.5f);
          while (k * (k - 1) / 2 > t) --k;
          while ((k + 1) * k / 2 <= t) ++k;
          kb = seg[ea] + k;
          ja = seg[ea] + (t - k * (k - 1) / 2);
        }
        return;
      }
      t -= nb;
    }
  }
  ja = kb = 0;
  p = -1;
}

template <int ZI, int RC>
__global__ void __launch_bounds__(SYM_WAVES * 64)
sym_fwd_kernel(int N, const double* __restrict__ pos, const double* __restrict__ box,
               const int* __restrict__ mol, const int* __restrict__ elem,
               const int* __restrict__ nbr, const int* __restrict__ ncnt,
               const int* __restrict__ rseg, const int* __restrict__ ang,
               const int* __restrict__ acnt, const int* __restrict__ aseg, int ldr, int lda,
               SymArgs a, float* __restrict__ G, int ldg, int* __restrict__ rev,
               int* __restrict__ flags) {
  __shared__ RowLdsFwd<RC> L;
  __shared__ float4 tq[SYM_WAVES][64];        // per triple: (cos theta, sin theta, fc_j fc_k, -)
  __shared__ float4 tR[SYM_WAVES][64][NA / 4];   // per triple: exp(-eta (ravg - Rs_s)^2), s = 0..7
  __shared__ int segs[SYM_WAVES][2][TM_MAX_ELE + 1];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = blockIdx.x * SYM_WAVES + w;
  if (i >= N) return;   // only wave-level synchronisation below
  const int n = min(ncnt[i], RC);
  if (lane <= a.nele) {
    segs[w][0][lane] = min(rseg[(long)i * (a.nele + 1) + lane], n);
    segs[w][1][lane] = aseg[(long)i * (a.nele + 1) + lane];
  }
  load_row(i, w, lane, pos, box, mol, elem, nbr, n, ldr, a, L);
  __syncwarp();
  float* Gi = G + (long)i * ldg;
  // ---------------- radial: lane = (half h, shell s)
  {
    const int h = lane >> 5, s = lane & 31;
    for (int e = 0; e < a.nele; ++e) {
      int lo = segs[w][0][e], hi = segs[w][0][e + 1];
      for (int s0 = 0; s0 < a.nr; s0 += 32) {
        int ss = s0 + s;
        float rs = ss < a.nr ? a.rs_r[ss] : 0.f;
        float acc = 0.f;
        for (int q = lo + h; q < hi; q += 2) {
          float r = L.d[w][q].w;
          float x = r - rs;
          acc += __expf(-a.eta * x * x) * L.fc[w][q];
        }
        acc += __shfl_xor(acc, 32, 64);
        if (h == 0 && ss < a.nr) Gi[e * a.nr + ss] = acc;
      }
    }
  }
  // ---------------- angular
  // Phase A: lane = triple, stages its geometry in LDS.  Phase B: lane =
  // (theta point t, triple group g); a lane takes every 8th triple of an
  // element-pair segment, forms (1 + cos(theta - theta_t))^zeta fc_j fc_k once
  // and adds it times R_s to one accumulator per shell s.  When a segment
  // ends, a butterfly over g sums the 8 groups and lane (t, g) writes shell
  // s = g.  Triples are ordered by element pair (blocks ea <= eb), so each
  // segment is a contiguous range of the enumeration.
  const int m = min(acnt[i], lda);
  const int* arow = ang + (long)i * lda;
  const int ntrip = m * (m - 1) / 2;
  const int off = a.nele * a.nr;
  const int ot = lane >> 3, g = lane & 7;
  const float ct = a.cos_t[ot], st = a.sin_t[ot];
  const int* sg = segs[w][1];
  float acc[NA];
#pragma unroll
  for (int s = 0; s < NA; ++s) acc[s] = 0.f;
  int pdone = 0;   // segments [0, pdone) are written
  auto flush = [&](int p) {
    float v[NA];
#pragma unroll
    for (int s = 0; s < NA; ++s) {
      float x = acc[s];
      x += __shfl_xor(x, 1, 64);
      x += __shfl_xor(x, 2, 64);
      x += __shfl_xor(x, 4, 64);
      v[s] = x;
      acc[s] = 0.f;
    }
    float out = v[0];
#pragma unroll
    for (int s = 1; s < NA; ++s) out = g == s ? v[s] : out;
    Gi[off + p * 64 + ot * NA + g] = a.two_1mz * out;
  };
  for (int base = 0; base < ntrip; base += 64) {
    int t = base + lane;
    if (t < ntrip) {
      int ja, kb, p;
      decode_triple(t, sg, a.nele, ja, kb, p);
      int qj = arow[ja], qk = arow[kb];
      float4 dj = L.d[w][qj], dk = L.d[w][qk];
      const float inv = __builtin_amdgcn_rcpf(dj.w * dk.w);
      float c = (dj.x * dk.x + dj.y * dk.y + dj.z * dk.z) * inv;
      c = fminf(fmaxf(c, -1.0f), 1.0f);
      float sn = sin_between(dj, dk, inv);
      float ravg = 0.5f * (dj.w + dk.w);
      tq[w][lane] = make_float4(c, sn, cutoff_fc(a.ka, dj.w) * cutoff_fc(a.ka, dk.w), 0.f);
      float R[NA];
#pragma unroll
      for (int s = 0; s < NA; ++s) {
        float x = ravg - a.rs_a[s];
        R[s] = __expf(-a.eta * x * x);
      }
      tR[w][lane][0] = make_float4(R[0], R[1], R[2], R[3]);
      tR[w][lane][1] = make_float4(R[4], R[5], R[6], R[7]);
    }
    __syncwarp();
    const int cnt = min(64, ntrip - base);
    int T = 0, p = 0;   // first triple of segment p
    for (int ea = 0; ea < a.nele; ++ea) {
      const int la = sg[ea + 1] - sg[ea];
      for (int eb = ea; eb < a.nele; ++eb, ++p) {
        const int lb = sg[eb + 1] - sg[eb];
        const int Te = T + (ea == eb ? la * (la - 1) / 2 : la * lb);
        const int lo = max(T - base, 0), hi = min(Te - base, cnt);   // chunk-local part
        for (int u = lo + ((g - lo) & 7); u < hi; u += 8) {
          const float4 q = tq[w][u];
          const float4 r0 = tR[w][u][0], r1 = tR[w][u][1];
          const float AF = powz<ZI>(1.0f + (q.x * ct + q.y * st), a) * q.z;
          acc[0] += AF * r0.x; acc[1] += AF * r0.y; acc[2] += AF * r0.z; acc[3] += AF * r0.w;
          acc[4] += AF * r1.x; acc[5] += AF * r1.y; acc[6] += AF * r1.z; acc[7] += AF * r1.w;
        }
        if (p >= pdone && Te <= base + cnt) {   // segment complete (wave-uniform)
          flush(p);
          pdone = p + 1;
        }
        T = Te;
      }
    }
    __syncwarp();
  }
  for (int p = pdone; p < a.nelep; ++p) Gi[off + p * 64 + lane] = 0.f;
  if (rev == nullptr) return;
  // Mirror slots (the rev output of the neighbour build): rev[i*ldr+q] =
  // j*ldr + position of i in row j, found by binary search in the element
  // segment of row j that holds i (rows are sorted by (element, index)).
  const int ei = elem[i];
  for (int q = lane; q < n; q += 64) {
    const int j = nbr[(long)i * ldr + q];
    int lo = rseg[(long)j * (a.nele + 1) + ei];
    int hi = rseg[(long)j * (a.nele + 1) + ei + 1];
    const int* row = nbr + (long)j * ldr;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (row[mid] < i) lo = mid + 1; else hi = mid;
    }
    if (lo < ncnt[j] && row[lo] == i) {
      rev[(long)i * ldr + q] = (int)((long)j * ldr + lo);
    } else {
      rev[(long)i * ldr + q] = (int)((long)i * ldr + q);   // unmatched (truncated row)
      atomicOr(&flags[TM_FLAG_NB_OVERFLOW], 2);
    }
  }
}

// dE/dG (two contributions combined: gE + u_i * gQ) -> w slots
template <int ZI, int RC>
__global__ void __launch_bounds__(SYM_WAVES * 64) __attribute__((amdgpu_waves_per_eu(4)))
sym_adj_kernel(int N, const double* __restrict__ pos, const double* __restrict__ box,
               const int* __restrict__ mol, const int* __restrict__ elem,
               const int* __restrict__ nbr, const int* __restrict__ ncnt,
               const int* __restrict__ rseg, const int* __restrict__ ang,
               const int* __restrict__ acnt, const int* __restrict__ aseg, int ldr, int lda,
               SymArgs a, const float* __restrict__ gE, const float* __restrict__ gQ,
               const float* __restrict__ u, const double* __restrict__ phi,
               const double* __restrict__ phimean, int ldg, float4* __restrict__ wslot) {
  __shared__ RowLdsAdj<RC> L;
  extern __shared__ float gdyn[];   // SYM_WAVES * D floats: dE/dG of each wave's centre
  float* gw = gdyn + (threadIdx.x >> 6) * a.D;
  __shared__ float wacc[3][SYM_WAVES][RC];    // w per neighbour slot (x, y, z planes)
  __shared__ float cab[2][3][SYM_WAVES][64];       // per triple: contributions to j (0) and k (1)
  __shared__ int segs[SYM_WAVES][2][TM_MAX_ELE + 1];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = blockIdx.x * SYM_WAVES + w;
  if (i >= N) return;
  const int n = min(ncnt[i], RC);
  if (lane <= a.nele) {
    segs[w][0][lane] = min(rseg[(long)i * (a.nele + 1) + lane], n);
    segs[w][1][lane] = aseg[(long)i * (a.nele + 1) + lane];
  }
  // charge upstream u_i = phi_i - mean_m(phi) (as tm_charge_upstream) unless given
  const float ui = gQ == nullptr ? 0.f
                   : (u != nullptr ? u[i] : (float)(phi[i] - phimean[mol[i]]));
  for (int f = lane; f < a.D; f += 64) {
    float v = gE[(long)i * ldg + f];
    if (gQ != nullptr) v += ui * gQ[(long)i * ldg + f];
    gw[f] = v;
  }
  load_row(i, w, lane, pos, box, mol, elem, nbr, n, ldr, a, L);
  __syncwarp();
  // ---------------- radial: lane = neighbour
  const float kr = a.kr;
  for (int q = lane; q < n; q += 64) {
    float4 d = L.d[w][q];
    float r = d.w;
    int e = elem[nbr[(long)i * ldr + q]];
    float fc = cutoff_fc(kr, r);
    float dfc = cutoff_dfc(kr, r);
    float dEdr = 0.f;
    for (int s = 0; s < a.nr; ++s) {
      float x = r - a.rs_r[s];
      float ex = __expf(-a.eta * x * x);
      dEdr += gw[e * a.nr + s] * ex * (dfc - 2.0f * a.eta * x * fc);
    }
    float sc = dEdr * __builtin_amdgcn_rcpf(r);
    wacc[0][w][q] = sc * d.x;
    wacc[1][w][q] = sc * d.y;
    wacc[2][w][q] = sc * d.z;
  }
  __syncwarp();
  // ---------------- angular: lane = triple, contributions staged per chunk
  const int m = min(acnt[i], lda);
  const int* arow = ang + (long)i * lda;
  const int ntrip = m * (m - 1) / 2;
  const int off = a.nele * a.nr;
  float4 wang = make_float4(0.f, 0.f, 0.f, 0.f);   // lane = sub-list member
  const int* sg = segs[w][1];
  int eq = -1, lq = 0;   // element and in-element index of member `lane`
  if (lane < m)
    for (int e = 0; e < a.nele; ++e)
      if (lane >= sg[e] && lane < sg[e + 1]) { eq = e; lq = lane - sg[e]; }
  for (int base = 0; base < ntrip; base += 64) {
    int t = base + lane;
    if (t < ntrip) {
      int ja, kb, p;
      decode_triple(t, segs[w][1], a.nele, ja, kb, p);
      int qj = arow[ja], qk = arow[kb];
      float4 dj = L.d[w][qj], dk = L.d[w][qk];
      float rj = dj.w, rk = dk.w;
      const float irj = __builtin_amdgcn_rcpf(rj), irk = __builtin_amdgcn_rcpf(rk);
      float inv = irj * irk;
      float c = (dj.x * dk.x + dj.y * dk.y + dj.z * dk.z) * inv;
      c = fminf(fmaxf(c, -1.0f), 1.0f);
      float sn = sin_between(dj, dk, inv);
      float cos_over_sin = sn > 1e-6f ? c * __builtin_amdgcn_rcpf(sn) : 0.f;
      float ravg = 0.5f * (rj + rk);
      float fj = L.fa[w][qj], fk = L.fa[w][qk];
      float dfj = cutoff_dfc(a.ka, rj), dfk = cutoff_dfc(a.ka, rk);
      float F = fj * fk;
      float R[NA], Rp[NA];
#pragma unroll
      for (int s = 0; s < NA; ++s) {
        float x = ravg - a.rs_a[s];
        R[s] = __expf(-a.eta * x * x);
        Rp[s] = -2.0f * a.eta * x * R[s];
      }
      const float* gp = gw + off + p * 64;
      float dEdc = 0.f, dEdrav = 0.f, dEdF = 0.f;
#pragma unroll
      for (int tt = 0; tt < NT; ++tt) {
        float S = 0.f, Sp = 0.f;
#pragma unroll
        for (int s = 0; s < NA; ++s) {
          float gv = gp[tt * NA + s];
          S += gv * R[s];
          Sp += gv * Rp[s];
        }
        float v = c * a.cos_t[tt] + sn * a.sin_t[tt];
        float b = 1.0f + v;
        float Az, dAz;                             // (1+v)^zeta, zeta (1+v)^(zeta-1)
        powz_pair<ZI>(b, a, Az, dAz);
        float dv = a.cos_t[tt] - cos_over_sin * a.sin_t[tt];
        dEdc += dAz * dv * S;
        dEdrav += Az * Sp;
        dEdF += Az * S;
      }
      // the common factor 2^(1 - zeta) of every term, applied once
      dEdc *= a.two_1mz * F;
      dEdrav *= a.two_1mz * F;
      dEdF *= a.two_1mz;
      // chain rule into d_j, d_k
      float cj = (dEdrav * 0.5f + dEdF * dfj * fk - dEdc * c * irj) * irj;
      float ck = (dEdrav * 0.5f + dEdF * fj * dfk - dEdc * c * irk) * irk;
      float x = dEdc * inv;
      cab[0][0][w][lane] = cj * dj.x + x * dk.x;
      cab[0][1][w][lane] = cj * dj.y + x * dk.y;
      cab[0][2][w][lane] = cj * dj.z + x * dk.z;
      cab[1][0][w][lane] = ck * dk.x + x * dj.x;
      cab[1][1][w][lane] = ck * dk.y + x * dj.y;
      cab[1][2][w][lane] = ck * dk.z + x * dj.z;
    }
    __syncwarp();
    // Member `lane` gathers its contributions from the triples [base, base + cnt)
    // in ascending triple order, walking only its own partners: in a same-element
    // block it is k of (j < k) at t = k(k-1)/2 + j, then j at t = k'(k'-1)/2 + j
    // for k' > j; in a cross block (ea < eb) t = kb_local * la + ja_local.
    if (eq >= 0) {
      const int cnt = min(64, ntrip - base);
      // contribution of triple slot u to its j (h = 0) or k (h = 1) member
      auto add = [&](int h, int u) {
        wang.x += cab[h][0][w][u]; wang.y += cab[h][1][w][u]; wang.z += cab[h][2][w][u];
      };
      int bo = 0;   // first triple of the block
      for (int ea = 0; ea < a.nele; ++ea) {
        const int la = sg[ea + 1] - sg[ea];
        for (int eb = ea; eb < a.nele; ++eb) {
          const int lb = sg[eb + 1] - sg[eb];
          const int nb = ea == eb ? la * (la - 1) / 2 : la * lb;
          const int lo = base - bo, hi = lo + cnt;   // chunk window, block-local
          if (hi > 0 && lo < nb) {
            if (ea == eb) {
              if (eq == ea) {
                const int t0 = lq * (lq - 1) / 2;
                for (int r = max(0, lo - t0); r < min(lq, hi - t0); ++r) add(1, t0 + r - lo);
                for (int r = lq + 1; r < la; ++r) {
                  const int t = r * (r - 1) / 2 + lq;
                  if (t >= hi) break;
                  if (t >= lo) add(0, t - lo);
                }
              }
            } else if (eq == ea) {
              for (int kl = 0; kl < lb; ++kl) {
                const int t = kl * la + lq;
                if (t >= hi) break;
                if (t >= lo) add(0, t - lo);
              }
            } else if (eq == eb) {
              const int t0 = lq * la;
              for (int r = max(0, lo - t0); r < min(la, hi - t0); ++r) add(1, t0 + r - lo);
            }
          }
          bo += nb;
        }
      }
    }
    __syncwarp();
  }
  if (lane < m) {
    int q = arow[lane];
    wacc[0][w][q] += wang.x;
    wacc[1][w][q] += wang.y;
    wacc[2][w][q] += wang.z;
  }
  __syncwarp();
  for (int q = lane; q < n; q += 64)
    wslot[(long)i * ldr + q] = make_float4(wacc[0][w][q], wacc[1][w][q], wacc[2][w][q], 0.f);
}

// Velocity Verlet state advanced by the force gather (VV = true): the kick of
// the step with the gathered force, then the drift of the next step, with the
// operations of vv_kick_kernel and vv_drift_kernel.
struct VVState {
  double dt;
  const double* m;
  double *v, *a, *ke, *x;
};

// grad_a = sum_n (w[rev(a,n)] - w[(a,n)]) + extra_a ; optional force output
// force = -scale * grad (J/mol/A when scale = JOULEPERHARTREE).  16 lanes per
// atom, each summing every 16th slot, then a fixed butterfly over the 16.
template <bool VV>
__global__ void force_gather_kernel(int N, int ldr, const int* __restrict__ ncnt,
                                    const int* __restrict__ rev, const float4* __restrict__ wslot,
                                    const double* __restrict__ extra, double scale,
                                    double* __restrict__ grad, double* __restrict__ force,
                                    VVState vv) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int a = (int)(t >> 4);
  const int sub = threadIdx.x & 15;
  double gx = 0, gy = 0, gz = 0;
  if (a < N) {
    const int n = ncnt[a];
    for (int q = sub; q < n; q += 16) {
      long sl = (long)a * ldr + q;
      float4 wo = wslot[sl];
      float4 wi = wslot[rev[sl]];
      gx += (double)wi.x - (double)wo.x;
      gy += (double)wi.y - (double)wo.y;
      gz += (double)wi.z - (double)wo.z;
    }
  }
#pragma unroll
  for (int o = 8; o > 0; o >>= 1) {
    gx += __shfl_xor(gx, o, 16);
    gy += __shfl_xor(gy, o, 16);
    gz += __shfl_xor(gz, o, 16);
  }
  if (a >= N || sub != 0) return;
  if (extra != nullptr) {
    gx += extra[3 * a]; gy += extra[3 * a + 1]; gz += extra[3 * a + 2];
  }
  if (grad != nullptr) {
    grad[3 * a] = gx; grad[3 * a + 1] = gy; grad[3 * a + 2] = gz;
  }
  const double f[3] = {-scale * gx, -scale * gy, -scale * gz};
  if (force != nullptr) {
    force[3 * a] = f[0]; force[3 * a + 1] = f[1]; force[3 * a + 2] = f[2];
  }
  if (VV) {
    const double mi = vv.m[a], dt = vv.dt;
    double s = 0.0;
    for (int d = 0; d < 3; ++d) {
      double an = 1e-10 * f[d] / mi;
      double vn = vv.v[3 * a + d] + 0.5 * (vv.a[3 * a + d] + an) * dt;
      vv.v[3 * a + d] = vn;
      vv.a[3 * a + d] = an;
      s += vn * vn;
      vv.x[3 * a + d] = vv.x[3 * a + d] + vn * dt + 0.5 * an * dt * dt;
    }
    if (vv.ke) vv.ke[a] = 0.5 * mi * s;
  }
}

// Tangent (forward-mode derivative) of the descriptor along per-atom
// displacements du: V[i, f] = sum_j dG[i, f]/d(x_j - x_i) . (du_j - du_i).
// It is the transpose of the adjoint above (same partial derivatives, the
// same guard where sin(theta) vanishes), contracted over neighbours instead
// of features; the lane layout and the write-once order are the forward's.
// Used for the weight gradient of a force loss: with u = dLoss/dF,
// dLoss/dW = -d/dW sum_i (dE_i/dG_i)(W) . V_i.
template <int RC>
struct JvpLds {
  float4 t[SYM_WAVES][RC];     // (du_j - du_i, dr_j = d_j . (du_j - du_i) / r_j)
  float dfc[SYM_WAVES][RC];    // d fc_r / dr
};

template <int ZI, int RC>
__global__ void __launch_bounds__(SYM_WAVES * 64)
sym_jvp_kernel(int N, const double* __restrict__ pos, const double* __restrict__ box,
               const int* __restrict__ mol, const int* __restrict__ elem,
               const int* __restrict__ nbr, const int* __restrict__ ncnt,
               const int* __restrict__ rseg, const int* __restrict__ ang,
               const int* __restrict__ acnt, const int* __restrict__ aseg, int ldr, int lda,
               SymArgs a, const double* __restrict__ du, float* __restrict__ V, int ldg) {
  // the forward's row (the angular cutoff and its derivative are formed per
  // triple), which with the per-triple staging below fits 4 waves per SIMD
  __shared__ RowLdsFwd<RC> L;
  __shared__ JvpLds<RC> T;
  // per triple: (cos theta, sin theta, cos/sin guarded, fc_j fc_k), (d(fc_j fc_k),
  // d cos theta, d ravg, ravg), R_s for s = 0..7
  __shared__ float4 tqa[SYM_WAVES][64], tqb[SYM_WAVES][64];
  __shared__ float4 tR[SYM_WAVES][64][NA / 4];
  __shared__ int segs[SYM_WAVES][2][TM_MAX_ELE + 1];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = blockIdx.x * SYM_WAVES + w;
  if (i >= N) return;   // only wave-level synchronisation below
  const int n = min(ncnt[i], RC);
  if (lane <= a.nele) {
    segs[w][0][lane] = min(rseg[(long)i * (a.nele + 1) + lane], n);
    segs[w][1][lane] = aseg[(long)i * (a.nele + 1) + lane];
  }
  load_row(i, w, lane, pos, box, mol, elem, nbr, n, ldr, a, L);
  const float kr = a.kr, ka = a.ka;
  {
    const double uxi = du[3 * i], uyi = du[3 * i + 1], uzi = du[3 * i + 2];
    __syncwarp();
    for (int q = lane; q < n; q += 64) {
      const int j = nbr[(long)i * ldr + q];
      const float tx = (float)(du[3 * j] - uxi), ty = (float)(du[3 * j + 1] - uyi);
      const float tz = (float)(du[3 * j + 2] - uzi);
      const float4 d = L.d[w][q];
      T.t[w][q] = make_float4(tx, ty, tz, (d.x * tx + d.y * ty + d.z * tz) * __builtin_amdgcn_rcpf(d.w));
      T.dfc[w][q] = -0.5f * kr * __sinf(kr * d.w);
    }
  }
  __syncwarp();
  float* Vi = V + (long)i * ldg;
  // ---------------- radial: lane = (half h, shell s)
  {
    const int h = lane >> 5, s = lane & 31;
    for (int e = 0; e < a.nele; ++e) {
      int lo = segs[w][0][e], hi = segs[w][0][e + 1];
      for (int s0 = 0; s0 < a.nr; s0 += 32) {
        int ss = s0 + s;
        float rs = ss < a.nr ? a.rs_r[ss] : 0.f;
        float acc = 0.f;
        for (int q = lo + h; q < hi; q += 2) {
          float r = L.d[w][q].w;
          float x = r - rs;
          acc += __expf(-a.eta * x * x) * (T.dfc[w][q] - 2.0f * a.eta * x * L.fc[w][q]) *
                 T.t[w][q].w;
        }
        acc += __shfl_xor(acc, 32, 64);
        if (h == 0 && ss < a.nr) Vi[e * a.nr + ss] = acc;
      }
    }
  }
  // ---------------- angular
  // The forward's layout (sym_fwd_kernel): phase A lane = triple stages the
  // triple's geometry and its tangents in LDS; phase B lane = (theta point t,
  // triple group g) takes every 8th triple of an element-pair segment, forms
  // (1 + cos(theta - theta_t))^zeta and its derivative once, and adds
  //   X R_s + Y R'_s,  X = A' v' c' F + A F',  Y = A r' F
  // to one accumulator per shell s (A = (1 + v)^zeta, v = cos(theta - theta_t),
  // c', r', F' the tangents of cos theta, ravg and fc_j fc_k); a butterfly
  // over g at each segment end, lane (t, g) writes shell s = g.
  const int m = min(acnt[i], lda);
  const int* arow = ang + (long)i * lda;
  const int ntrip = m * (m - 1) / 2;
  const int off = a.nele * a.nr;
  const int ot = lane >> 3, g = lane & 7;
  const float ct = a.cos_t[ot], st = a.sin_t[ot];
  const int* sg = segs[w][1];
  float acc[NA];
#pragma unroll
  for (int s = 0; s < NA; ++s) acc[s] = 0.f;
  int pdone = 0;   // segments [0, pdone) are written
  auto flush = [&](int p) {
    float v[NA];
#pragma unroll
    for (int s = 0; s < NA; ++s) {
      float x = acc[s];
      x += __shfl_xor(x, 1, 64);
      x += __shfl_xor(x, 2, 64);
      x += __shfl_xor(x, 4, 64);
      v[s] = x;
      acc[s] = 0.f;
    }
    float out = v[0];
#pragma unroll
    for (int s = 1; s < NA; ++s) out = g == s ? v[s] : out;
    Vi[off + p * 64 + ot * NA + g] = a.two_1mz * out;
  };
  for (int base = 0; base < ntrip; base += 64) {
    int t = base + lane;
    if (t < ntrip) {
      int ja, kb, p;
      decode_triple(t, sg, a.nele, ja, kb, p);
      int qj = arow[ja], qk = arow[kb];
      float4 dj = L.d[w][qj], dk = L.d[w][qk];
      float4 uj = T.t[w][qj], uk = T.t[w][qk];
      float rj = dj.w, rk = dk.w;
      const float irj = __builtin_amdgcn_rcpf(rj), irk = __builtin_amdgcn_rcpf(rk);
      float inv = irj * irk;
      float c = (dj.x * dk.x + dj.y * dk.y + dj.z * dk.z) * inv;
      c = fminf(fmaxf(c, -1.0f), 1.0f);
      float sn = sin_between(dj, dk, inv);
      float dc = (uj.x * dk.x + uj.y * dk.y + uj.z * dk.z + dj.x * uk.x + dj.y * uk.y +
                  dj.z * uk.z) * inv - c * (uj.w * irj + uk.w * irk);
      const float fj = cutoff_fc(ka, rj), fk = cutoff_fc(ka, rk);
      const float dfj = cutoff_dfc(ka, rj), dfk = cutoff_dfc(ka, rk);
      float ravg = 0.5f * (rj + rk);
      tqa[w][lane] = make_float4(c, sn, sn > 1e-6f ? c * __builtin_amdgcn_rcpf(sn) : 0.f,
                                 fj * fk);
      tqb[w][lane] = make_float4(dfj * uj.w * fk + fj * dfk * uk.w, dc, 0.5f * (uj.w + uk.w),
                                 ravg);
      float R[NA];
#pragma unroll
      for (int s = 0; s < NA; ++s) {
        float x = ravg - a.rs_a[s];
        R[s] = __expf(-a.eta * x * x);
      }
      tR[w][lane][0] = make_float4(R[0], R[1], R[2], R[3]);
      tR[w][lane][1] = make_float4(R[4], R[5], R[6], R[7]);
    }
    __syncwarp();
    const int cnt = min(64, ntrip - base);
    int T0 = 0, p = 0;   // first triple of segment p
    for (int ea = 0; ea < a.nele; ++ea) {
      const int la = sg[ea + 1] - sg[ea];
      for (int eb = ea; eb < a.nele; ++eb, ++p) {
        const int lb = sg[eb + 1] - sg[eb];
        const int Te = T0 + (ea == eb ? la * (la - 1) / 2 : la * lb);
        const int lo = max(T0 - base, 0), hi = min(Te - base, cnt);   // chunk-local part
        for (int u = lo + ((g - lo) & 7); u < hi; u += 8) {
          const float4 qa = tqa[w][u], qb = tqb[w][u];
          float Az, dAz;
          powz_pair<ZI>(1.0f + (qa.x * ct + qa.y * st), a, Az, dAz);
          const float dv = ct - qa.z * st;
          const float X = dAz * dv * qb.y * qa.w + Az * qb.x;
          const float Y2 = -2.0f * a.eta * Az * qb.z * qa.w;   // times (ravg - R_s) R_s
          const float4 r0 = tR[w][u][0], r1 = tR[w][u][1];
          const float R[NA] = {r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w};
#pragma unroll
          for (int s = 0; s < NA; ++s) acc[s] += R[s] * (X + Y2 * (qb.w - a.rs_a[s]));
        }
        if (p >= pdone && Te <= base + cnt) {   // segment complete (wave-uniform)
          flush(p);
          pdone = p + 1;
        }
        T0 = Te;
      }
    }
    __syncwarp();
  }
  for (int p = pdone; p < a.nelep; ++p) Vi[off + p * 64 + lane] = 0.f;
}

int fill_args(SymArgs& a, int nele, int nr, float eta, float zeta, float rr, float ra,
              const float* rs_r, const float* rs_a, const float* thetas) {
  TM_REQUIRE(nele > 0 && nele <= TM_MAX_ELE, "sym: bad nele %d", nele);
  TM_REQUIRE(nr > 0 && nr <= NR_MAX, "sym: radial shells %d outside 1..%d", nr, NR_MAX);
  a.nele = nele;
  a.nelep = nele * (nele + 1) / 2;
  a.nr = nr;
  a.D = nele * nr + a.nelep * NA * NT;
  TM_REQUIRE(a.D <= 1024, "sym: descriptor length %d above 1024", a.D);
  a.eta = eta; a.zeta = zeta; a.rr = rr; a.ra = ra;
  a.kr = PI_LIT / rr;
  a.ka = PI_LIT / ra;
  a.two_1mz = powf(2.0f, 1.0f - zeta);
  a.zeta_int = (zeta == floorf(zeta) && zeta >= 1.f && zeta <= 64.f) ? (int)zeta : 0;
  for (int s = 0; s < nr; ++s) a.rs_r[s] = rs_r[s];
  for (int s = 0; s < NA; ++s) a.rs_a[s] = rs_a[s];
  for (int t = 0; t < NT; ++t) {
    a.cos_t[t] = cosf(thetas[t]);
    a.sin_t[t] = sinf(thetas[t]);
  }
  return TM_OK;
}

}  // namespace

extern "C" {

// Host-side parameter arrays (rs_r[nr], rs_a[8], thetas[8]) are read on the
// host and passed to the kernels by value.
int tm_sym_forward(int N, const double* pos, const double* box, const int* mol, const int* elem,
                   const int* nbr, const int* ncnt, const int* rseg, const int* ang,
                   const int* acnt, const int* aseg, int ldr, int lda, int nele, int nr,
                   float eta, float zeta, float rr, float ra, const float* rs_r,
                   const float* rs_a, const float* thetas, float* G, int ldg, int* rev,
                   int* flags, void* stream) {
  SymArgs a;
  int st = fill_args(a, nele, nr, eta, zeta, rr, ra, rs_r, rs_a, thetas);
  if (st) return st;
  TM_REQUIRE(ldg >= a.D, "tm_sym_forward: ldg %d < D %d", ldg, a.D);
  TM_REQUIRE(rev == nullptr || flags != nullptr, "tm_sym_forward: rev needs flags");
  TM_REQUIRE(ldr > 0 && ldr <= MAX_ROW && lda > 0 && lda <= ldr,
             "tm_sym_forward: strides %d/%d (need 0 < lda <= ldr <= %d)", ldr, lda, MAX_ROW);
  if (N == 0) return TM_OK;
  hipStream_t s = (hipStream_t)stream;
  // rows up to 128 entries stage in the smaller LDS layout (more waves per CU)
  auto k = ldr <= 128 ? (a.zeta_int == 8 ? sym_fwd_kernel<8, 128> : sym_fwd_kernel<0, 128>)
                      : (a.zeta_int == 8 ? sym_fwd_kernel<8, 256> : sym_fwd_kernel<0, 256>);
  k<<<tm_div_up(N, SYM_WAVES), SYM_WAVES * 64, 0, s>>>(N, pos, box, mol, elem, nbr, ncnt, rseg,
                                                       ang, acnt, aseg, ldr, lda, a, G, ldg, rev,
                                                       flags);
  TM_LAUNCH_CHECK();
  return TM_OK;
}

int tm_sym_adjoint(int N, const double* pos, const double* box, const int* mol, const int* elem,
                   const int* nbr, const int* ncnt, const int* rseg, const int* ang,
                   const int* acnt, const int* aseg, int ldr, int lda, int nele, int nr,
                   float eta, float zeta, float rr, float ra, const float* rs_r,
                   const float* rs_a, const float* thetas, const float* gE, const float* gQ,
                   const float* u, const double* phi, const double* phimean, int ldg,
                   float* wslot, void* stream) {
  SymArgs a;
  int st = fill_args(a, nele, nr, eta, zeta, rr, ra, rs_r, rs_a, thetas);
  if (st) return st;
  TM_REQUIRE(ldg >= a.D, "tm_sym_adjoint: ldg %d < D %d", ldg, a.D);
  TM_REQUIRE(gQ == nullptr || u != nullptr || (phi != nullptr && phimean != nullptr),
             "tm_sym_adjoint: gQ needs u or phi and phimean");
  TM_REQUIRE(ldr > 0 && ldr <= MAX_ROW && lda > 0 && lda <= ldr,
             "tm_sym_adjoint: strides %d/%d (need 0 < lda <= ldr <= %d)", ldr, lda, MAX_ROW);
  if (N == 0) return TM_OK;
  hipStream_t s = (hipStream_t)stream;
  auto k = ldr <= 128 ? (a.zeta_int == 8 ? sym_adj_kernel<8, 128> : sym_adj_kernel<0, 128>)
                      : (a.zeta_int == 8 ? sym_adj_kernel<8, 256> : sym_adj_kernel<0, 256>);
  k<<<tm_div_up(N, SYM_WAVES), SYM_WAVES * 64, sizeof(float) * SYM_WAVES * a.D, s>>>(
      N, pos, box, mol, elem, nbr, ncnt, rseg, ang, acnt, aseg, ldr, lda, a, gE, gQ, u, phi,
      phimean, ldg, (float4*)wslot);
  TM_LAUNCH_CHECK();
  return TM_OK;
}

int tm_sym_jvp(int N, const double* pos, const double* box, const int* mol, const int* elem,
               const int* nbr, const int* ncnt, const int* rseg, const int* ang, const int* acnt,
               const int* aseg, int ldr, int lda, int nele, int nr, float eta, float zeta,
               float rr, float ra, const float* rs_r, const float* rs_a, const float* thetas,
               const double* du, float* V, int ldg, void* stream) {
  SymArgs a;
  int st = fill_args(a, nele, nr, eta, zeta, rr, ra, rs_r, rs_a, thetas);
  if (st) return st;
  TM_REQUIRE(ldg >= a.D, "tm_sym_jvp: ldg %d < D %d", ldg, a.D);
  TM_REQUIRE(ldr > 0 && ldr <= MAX_ROW && lda > 0 && lda <= ldr,
             "tm_sym_jvp: strides %d/%d (need 0 < lda <= ldr <= %d)", ldr, lda, MAX_ROW);
  if (N == 0) return TM_OK;
  hipStream_t s = (hipStream_t)stream;
  auto k = ldr <= 128 ? (a.zeta_int == 8 ? sym_jvp_kernel<8, 128> : sym_jvp_kernel<0, 128>)
                      : (a.zeta_int == 8 ? sym_jvp_kernel<8, 256> : sym_jvp_kernel<0, 256>);
  k<<<tm_div_up(N, SYM_WAVES), SYM_WAVES * 64, 0, s>>>(N, pos, box, mol, elem, nbr, ncnt, rseg,
                                                       ang, acnt, aseg, ldr, lda, a, du, V, ldg);
  TM_LAUNCH_CHECK();
  return TM_OK;
}

int tm_force_gather(int N, int ldr, const int* ncnt, const int* rev, const float* wslot,
                    const double* extra, double scale, double* grad, double* force,
                    void* stream) {
  if (N == 0) return TM_OK;
  hipStream_t s = (hipStream_t)stream;
  force_gather_kernel<false><<<tm_div_up(16l * N, 256), 256, 0, s>>>(
      N, ldr, ncnt, rev, (const float4*)wslot, extra, scale, grad, force, VVState{});
  TM_LAUNCH_CHECK();
  return TM_OK;
}

int tm_force_gather_vv(int N, int ldr, const int* ncnt, const int* rev, const float* wslot,
                       const double* extra, double scale, double* grad, double* force, double dt,
                       const double* m, double* v, double* a, double* ke, double* x,
                       void* stream) {
  TM_REQUIRE(m != nullptr && v != nullptr && a != nullptr && x != nullptr,
             "tm_force_gather_vv: m, v, a and x are required");
  if (N == 0) return TM_OK;
  hipStream_t s = (hipStream_t)stream;
  force_gather_kernel<true><<<tm_div_up(16l * N, 256), 256, 0, s>>>(
      N, ldr, ncnt, rev, (const float4*)wslot, extra, scale, grad, force,
      VVState{dt, m, v, a, ke, x});
  TM_LAUNCH_CHECK();
  return TM_OK;
}

}  // extern "C"
