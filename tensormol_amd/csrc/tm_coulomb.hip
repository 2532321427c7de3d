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
#include "tm_common.h"

namespace {

constexpr int C_WAVES = 4;
constexpr double BOHRPERA_D = 1.889725989;

struct CoulArgs {
  float alpha, r_sr, r_lr, zz, yy, shift, ea;     // bohr units
  float rc2;                                     // EE cutoff squared (A^2)
  float width;                                   // vdW poly width (bohr after scaling)
  int add_ecc, has_vdw;
  float c6[TM_MAX_ELE], r0[TM_MAX_ELE];
};

struct Grid {
  const int* cell_start;
  const int* cell_atoms;
  const int* cell_of;
  int nc[3];
  int periodic;
  int reach;
};

template <class F>
__device__ __forceinline__ void candidate_ranges(int i, const int* __restrict__ mol,
                                                 const int* __restrict__ mol_start, const Grid& g,
                                                 F&& f) {
  if (g.cell_atoms == nullptr) {
    int m = mol[i];
    f(mol_start[m], mol_start[m + 1], 0);
    return;
  }
  int c = g.cell_of[i];
  int cx = c % g.nc[0], cy = (c / g.nc[0]) % g.nc[1], cz = c / (g.nc[0] * g.nc[1]);
  int k = g.reach;
  bool full[3];
  for (int d = 0; d < 3; ++d) full[d] = g.periodic ? (2 * k + 1 >= g.nc[d]) : false;
  int zlo = full[2] ? 0 : cz - k, zhi = full[2] ? g.nc[2] - 1 : cz + k;
  int ylo = full[1] ? 0 : cy - k, yhi = full[1] ? g.nc[1] - 1 : cy + k;
  if (!g.periodic) {
    zlo = max(zlo, 0); zhi = min(zhi, g.nc[2] - 1);
    ylo = max(ylo, 0); yhi = min(yhi, g.nc[1] - 1);
  }
  for (int zz = zlo; zz <= zhi; ++zz) {
    int z = g.periodic ? ((zz % g.nc[2]) + g.nc[2]) % g.nc[2] : zz;
    for (int yy = ylo; yy <= yhi; ++yy) {
      int y = g.periodic ? ((yy % g.nc[1]) + g.nc[1]) % g.nc[1] : yy;
      int base = g.nc[0] * (y + g.nc[1] * z);
      if (full[0]) {
        f(g.cell_start[base], g.cell_start[base + g.nc[0]], 1);
      } else if (!g.periodic) {
        int xlo = max(cx - k, 0), xhi = min(cx + k, g.nc[0] - 1);
        f(g.cell_start[base + xlo], g.cell_start[base + xhi + 1], 1);
      } else {
        int xlo = cx - k, xhi = cx + k;
        if (xlo < 0) {
          f(g.cell_start[base + xlo + g.nc[0]], g.cell_start[base + g.nc[0]], 1);
          f(g.cell_start[base], g.cell_start[base + xhi + 1], 1);
        } else if (xhi >= g.nc[0]) {
          f(g.cell_start[base + xlo], g.cell_start[base + g.nc[0]], 1);
          f(g.cell_start[base], g.cell_start[base + xhi - g.nc[0] + 1], 1);
        } else {
          f(g.cell_start[base + xlo], g.cell_start[base + xhi + 1], 1);
        }
      }
    }
  }
}

__device__ __forceinline__ float min_image_f(float d, float L) {
  return L > 0.f ? d - L * rintf(d / L) : d;
}

__global__ void __launch_bounds__(C_WAVES * 64)
coulomb_kernel(int N, const float4* __restrict__ xq, const int* __restrict__ elem,
               const int* __restrict__ mol, const int* __restrict__ mol_start,
               const double* __restrict__ box, Grid grid, CoulArgs a, double* __restrict__ phi,
               double* __restrict__ grad, double* __restrict__ eatom) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = blockIdx.x * C_WAVES + w;
  if (i >= N) return;
  const int mi = mol[i];
  float Lx = 0.f, Ly = 0.f, Lz = 0.f;
  if (box != nullptr) {
    Lx = (float)box[3 * mi]; Ly = (float)box[3 * mi + 1]; Lz = (float)box[3 * mi + 2];
  }
  const float4 pi = xq[i];
  const int ei = elem[i];
  const float c6i = a.c6[ei], r0i = a.r0[ei];
  const float B1 = (float)BOHRPERA_D, B2 = (float)(BOHRPERA_D * BOHRPERA_D);
  double sphi = 0.0, gx = 0.0, gy = 0.0, gz = 0.0, ev = 0.0;
  auto visit = [&](int lo, int hi, int mapped) {
    for (int base = lo; base < hi; base += 64) {
      int p = base + lane;
      if (p >= hi) continue;
      int j = mapped ? grid.cell_atoms[p] : p;
      if (j == i || (mapped && mol[j] != mi)) continue;
      float4 pj = xq[j];
      float dx = min_image_f(pj.x - pi.x, Lx);
      float dy = min_image_f(pj.y - pi.y, Ly);
      float dz = min_image_f(pj.z - pi.z, Lz);
      float s = dx * dx + dy * dy + dz * dz;
      if (!(s < a.rc2)) continue;
      float rA = sqrtf(s);
      float inv_rA = 1.0f / rA;
      float dEdr = 0.f;   // dE_pair/dr_A
      if (a.add_ecc) {
        float r = rA * B1;
        float K, dK;
        if (r > a.r_sr) {
          if (r > a.r_lr) {
            K = 0.f; dK = 0.f;
          } else {
            float ar = a.alpha * r;
            float ec = erfcf(ar);
            float inv = 1.0f / r;
            K = ec * inv - a.zz + (r - a.r_lr) * (a.zz / a.r_lr + a.yy);
            dK = -ec * inv * inv - 1.1283791671f * a.alpha * __expf(-ar * ar) * inv +
                 a.zz / a.r_lr + a.yy;
          }
        } else {
          float ex = __expf(r - a.r_sr);
          K = a.ea * (ex - 1.0f) + a.shift;
          dK = a.ea * ex;
        }
        sphi += (double)(K * pj.w);
        dEdr += pi.w * pj.w * dK * B1;
      }
      if (a.has_vdw) {
        int ej = elem[j];
        float r = rA * B2;
        float t = r / a.width;
        float S, dS;
        if (t > 1.f) { S = 1.f; dS = 0.f; }
        else if (t > 0.f) { S = -t * t * (2.f * t - 3.f); dS = 6.f * t * (1.f - t) / a.width; }
        else { S = 0.f; dS = 0.f; }
        float C = sqrtf(c6i * a.c6[ej]);
        float R0 = r0i + a.r0[ej];
        float ur = R0 / r;
        float u2 = ur * ur, u4 = u2 * u2, u12 = u4 * u4 * u4;   // (r/R0)^-12
        float den = 1.0f + 6.0f * u12;
        float ir = 1.0f / r;
        float ir2 = ir * ir;
        float ir6 = ir2 * ir2 * ir2;
        float f = C * ir6 / den;
        float df = C * ir6 * ir * (-6.0f + 36.0f * u12) / (den * den);
        float V = -S * f;
        float dV = -(dS * f + S * df);
        ev += 0.5 * (double)V;
        dEdr += dV * B2;
      }
      // dE/dx_i = dE/dr_A * (x_i - x_j)/r_A = -dEdr * d / r_A
      float sc = -dEdr * inv_rA;
      gx += (double)(sc * dx);
      gy += (double)(sc * dy);
      gz += (double)(sc * dz);
    }
  };
  candidate_ranges(i, mol, mol_start, grid, visit);
  sphi = tm_wave_sum_d(sphi);
  gx = tm_wave_sum_d(gx);
  gy = tm_wave_sum_d(gy);
  gz = tm_wave_sum_d(gz);
  ev = tm_wave_sum_d(ev);
  if (lane == 0) {
    phi[i] = sphi;
    grad[3 * i] = gx; grad[3 * i + 1] = gy; grad[3 * i + 2] = gz;
    eatom[2 * i] = 0.5 * (double)pi.w * sphi;
    eatom[2 * i + 1] = ev;
  }
}

// Per-molecule column sums of an [N, C] array (float or double input) into
// out[B, C] (double); one block per molecule, fixed reduction tree.
template <class T>
__global__ void __launch_bounds__(256)
mol_sum_kernel(const int* __restrict__ mol_start, int C, const T* __restrict__ in, int ldin,
               double* __restrict__ out, int scale_by_count) {
  __shared__ double part[256];
  const int m = blockIdx.x;
  const int lo = mol_start[m], hi = mol_start[m + 1];
  for (int c = 0; c < C; ++c) {
    double acc = 0.0;
    for (int a = lo + threadIdx.x; a < hi; a += 256) acc += (double)in[(long)a * ldin + c];
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if (threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      double v = part[0];
      if (scale_by_count) v = hi > lo ? v / (double)(hi - lo) : 0.0;
      out[(long)m * C + c] = v;
    }
    __syncthreads();
  }
}

__global__ void charges_kernel(int N, const int* __restrict__ mol, const double* __restrict__ qmean,
                               const float* __restrict__ qraw, const double* __restrict__ pos,
                               float4* __restrict__ xq, double* __restrict__ q,
                               double* __restrict__ qx) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double qi = (double)qraw[i] - qmean[mol[i]];
  q[i] = qi;
  double x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
  xq[i] = make_float4((float)x, (float)y, (float)z, (float)qi);
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

int tm_mol_sum_f32(int B, const int* mol_start, int C, const float* in, int ldin, double* out,
                   int mean, void* stream) {
  TM_REQUIRE(B >= 0 && C > 0, "tm_mol_sum_f32: bad sizes");
  if (B == 0) return TM_OK;
  mol_sum_kernel<float><<<B, 256, 0, (hipStream_t)stream>>>(mol_start, C, in, ldin, out, mean);
  TM_LAUNCH_CHECK();
  return TM_OK;
}

int tm_mol_sum_f64(int B, const int* mol_start, int C, const double* in, int ldin, double* out,
                   int mean, void* stream) {
  TM_REQUIRE(B >= 0 && C > 0, "tm_mol_sum_f64: bad sizes");
  if (B == 0) return TM_OK;
  mol_sum_kernel<double><<<B, 256, 0, (hipStream_t)stream>>>(mol_start, C, in, ldin, out, mean);
  TM_LAUNCH_CHECK();
  return TM_OK;
}

int tm_charges(int N, const int* mol, const double* qmean, const float* qraw, const double* pos,
               float* xq, double* q, double* qx, void* stream) {
  if (N == 0) return TM_OK;
  charges_kernel<<<tm_div_up(N, 256), 256, 0, (hipStream_t)stream>>>(N, mol, qmean, qraw, pos,
                                                                     (float4*)xq, q, qx);
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
int tm_coulomb(int N, const float* xq, const int* elem, const int* mol, const int* mol_start,
               const double* box, const int* cell_start, const int* cell_atoms, const int* cell_of,
               int nx, int ny, int nz, int periodic, int reach, const double* consts,
               double rc_ee, double width, int add_ecc, int has_vdw, int nele, const double* c6,
               const double* r0, double* phi, double* grad, double* eatom, void* stream) {
  TM_REQUIRE(nele > 0 && nele <= TM_MAX_ELE, "tm_coulomb: bad nele");
  if (N == 0) return TM_OK;
  CoulArgs a;
  a.alpha = (float)consts[0]; a.r_sr = (float)consts[1]; a.r_lr = (float)consts[2];
  a.zz = (float)consts[3]; a.yy = (float)consts[4]; a.shift = (float)consts[5];
  a.ea = (float)consts[6];
  a.rc2 = (float)(rc_ee * rc_ee);
  a.width = (float)width;
  a.add_ecc = add_ecc;
  a.has_vdw = has_vdw;
  for (int e = 0; e < TM_MAX_ELE; ++e) {
    a.c6[e] = e < nele ? (float)c6[e] : 0.f;
    a.r0[e] = e < nele ? (float)r0[e] : 0.f;
  }
  Grid g;
  g.cell_start = cell_start; g.cell_atoms = cell_atoms; g.cell_of = cell_of;
  g.nc[0] = nx; g.nc[1] = ny; g.nc[2] = nz;
  g.periodic = periodic;
  g.reach = reach;
  coulomb_kernel<<<tm_div_up(N, C_WAVES), C_WAVES * 64, 0, (hipStream_t)stream>>>(
      N, (const float4*)xq, elem, mol, mol_start, box, g, a, phi, grad, eatom);
  TM_LAUNCH_CHECK();
  return TM_OK;
}

}  // extern "C"
