// Coulomb / vdW sweep of one system on a cell grid with the pair functions
// read from cubic tables.
//
// The energy terms are those of tm_coulomb.hip: DSF-elu electrostatics
// (TFCoulombEluSRDSFLR, TensorMol/TFDescriptors/RawSymFunc.py:1307-1359) and
// the polynomial-switched vdW (TFVdwPolyLR, RawSymFunc.py:1361-1414) over
// every pair within the EE cutoff (the 15 A list of Neighbors.py:262-282),
// swept as tm_coulomb sweeps them: one wavefront per centre in cell order,
// candidates from the trimmed cell stencil, the pairs inside the cutoff
// compacted into an LDS queue and evaluated 64 at a time.  What differs:
//  - Tables.  Over [0, rc] the Coulomb kernel K(r) (bohr units, per unit
//    charge product) and each element pair's vdW energy V(r) are cubic pieces
//    over `nint` equal intervals of r (A), built on the host in float64 from
//    the values and r-derivatives at the interval ends (cubic Hermite), stored
//    as float4 (c0, c1, c2, c3) of c0 + c1 t + c2 t^2 + c3 t^3, t in [0, 1).
//    The grid is shifted so that a node falls on the elu matching radius,
//    where K's second derivative jumps.  A pair costs two table reads and ten
//    FMAs for both energies and both derivatives instead of the erfc / exp /
//    reciprocal chains.  The tables are read through the L1 cache (16 KiB for
//    two elements), so the sweep keeps tm_coulomb's LDS footprint and
//    occupancy.
//  - Sums.  Each lane sums its pairs in float32 (a few tens per centre) and
//    the lane sums are added in float64 in a fixed butterfly, instead of a
//    float64 add per pair.  Queue order is candidate order, so the result
//    depends only on the inputs.
#include <math.h>

#include "tm_cells.h"

namespace {

constexpr double CL_BOHRPERA = 1.889725989;
constexpr int CT_WAVES = 4;
// Candidate ranges per wave: CT_RANGES covers every grid of cells >= 2.3 A at
// a 15 A cutoff (at most 386 ranges); CT_RANGES_WIDE (tm_coulomb's bound) wider
// stencils.  The smaller LDS footprint fits one more workgroup per CU.
constexpr int CT_RANGES = 392;
constexpr int CT_RANGES_WIDE = 480;
constexpr int CT_QCAP = 128;     // compacted pair queue per wave
constexpr int CL_TABLE_MAX = 3072;   // float4 table entries (48 KiB)

struct ClArgs {
  float rc, rc2;   // cutoff (A) and its square
  float inv_h;     // intervals per A
  float u0;        // table coordinate of r = 0 (in [0, 1))
  int nint, nele;
};

__device__ __forceinline__ float cl_min_image(float d, float L, float inv_L) {
  return d - L * rintf(d * inv_L);
}

struct ClAcc {
  float phi, gx, gy, gz, ev;
};

// Energies and gradient of one pair from the tables; d = (x_j - x_i, q_j),
// vo = the offset of the (e_i, e_j) vdW table.
__device__ __forceinline__ void ct_pair(const ClArgs& a, const float4* __restrict__ tab, int vo,
                                        float qi, float4 d, ClAcc& acc) {
  // r_A = sqrt(s + 1e-27), the reference's regulariser (RawSymFunc.py:1330)
  const float s = fmaf(d.x, d.x, fmaf(d.y, d.y, d.z * d.z)) + 1e-27f;
  const float ir = __builtin_amdgcn_rsqf(s);
  const float u = fmaf(s * ir, a.inv_h, a.u0);
  const int k = max(0, min((int)u, a.nint - 1));
  const float t = u - (float)k;
  const float4 C = tab[k];
  const float4 V = tab[vo + k];
  // value and t-derivative of c0 + c1 t + c2 t^2 + c3 t^3 (Horner)
  const float pc = fmaf(t, C.w, C.z), qc = fmaf(t, pc, C.y);
  const float K = fmaf(t, qc, C.x), dK = fmaf(t, fmaf(t, C.w, pc), qc);
  const float pv = fmaf(t, V.w, V.z), qv = fmaf(t, pv, V.y);
  const float E = fmaf(t, qv, V.x), dV = fmaf(t, fmaf(t, V.w, pv), qv);
  acc.phi = fmaf(K, d.w, acc.phi);
  acc.ev += E;
  // dE_pair/dr_A = (q_i q_j dK/dt + dV/dt) / h; dE/dx_i = -dE/dr_A (x_j - x_i) / r_A
  const float sc = -fmaf(qi * d.w, dK, dV) * (a.inv_h * ir);
  acc.gx = fmaf(sc, d.x, acc.gx);
  acc.gy = fmaf(sc, d.y, acc.gy);
  acc.gz = fmaf(sc, d.z, acc.gz);
}

template <int RCAP>
__global__ void __launch_bounds__(CT_WAVES * 64)
coulomb_tab_kernel(int N, const float4* __restrict__ xqc, const int* __restrict__ elc,
                   const double* __restrict__ box, CellGrid grid, ClArgs a,
                   const float4* __restrict__ tab, double* __restrict__ phi,
                   double* __restrict__ grad, double* __restrict__ eatom,
                   int* __restrict__ flags) {
  __shared__ float4 qx[CT_WAVES][CT_QCAP];
  __shared__ int qe[CT_WAVES][CT_QCAP];
  __shared__ RangeLds<RCAP> rng[CT_WAVES];
  __shared__ int voff[TM_MAX_ELE * TM_MAX_ELE];
  if (threadIdx.x < TM_MAX_ELE * TM_MAX_ELE) {
    const int ea = threadIdx.x / TM_MAX_ELE, eb = threadIdx.x % TM_MAX_ELE;
    const int lo = min(ea, eb), hi = max(ea, eb);
    voff[threadIdx.x] = (ea < a.nele && eb < a.nele)
                            ? (1 + lo * a.nele - lo * (lo - 1) / 2 + (hi - lo)) * a.nint
                            : 0;
  }
  __syncthreads();
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int cpi = blockIdx.x * CT_WAVES + w;   // centre, in cell order
  if (cpi >= N) return;   // wave-level synchronisation only below
  const int i = grid.cell_atoms[cpi];
  float Lx = 0.f, Ly = 0.f, Lz = 0.f, iLx = 0.f, iLy = 0.f, iLz = 0.f;
  if (grid.periodic) {
    Lx = (float)box[0]; Ly = (float)box[1]; Lz = (float)box[2];
    iLx = 1.0f / Lx; iLy = 1.0f / Ly; iLz = 1.0f / Lz;
  }
  const float4 pi = xqc[cpi];
  const int* vrow = voff + (elc[cpi] & (TM_MAX_ELE - 1)) * TM_MAX_ELE;
  ClAcc acc = {0.f, 0.f, 0.f, 0.f, 0.f};
  int qn = 0;   // wave-uniform queue length
  auto drain = [&](int n) {   // evaluate the first n (<= 64) queued pairs
    if (lane < n) ct_pair(a, tab, vrow[qe[w][lane] & (TM_MAX_ELE - 1)], pi.w, qx[w][lane], acc);
  };
  struct Cand {
    float4 x;
    int e;
  };
  // branch-free loads (invalid lanes read entry 0), masked by `valid` below
  auto load = [&](int p, bool valid, int) {
    const int q = valid ? p : 0;
    return Cand{xqc[q], elc[q]};
  };
  auto visit = [&](int p, bool valid, int, const Cand& c) {
    float4 d;
    d.x = cl_min_image(c.x.x - pi.x, Lx, iLx);
    d.y = cl_min_image(c.x.y - pi.y, Ly, iLy);
    d.z = cl_min_image(c.x.z - pi.z, Lz, iLz);
    d.w = c.x.w;
    const bool ok = valid && p != cpi && d.x * d.x + d.y * d.y + d.z * d.z < a.rc2;
    const unsigned long long bal = __ballot(ok);
    const int before = tm_lanes_below(bal);
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
      const bool mv = lane < qn - 64;
      if (mv) { t4 = qx[w][64 + lane]; te = qe[w][64 + lane]; }
      __syncwarp();
      if (mv) { qx[w][lane] = t4; qe[w][lane] = te; }
      qn -= 64;
      __syncwarp();
    }
  };
  tm_sweep_pf(i, nullptr, nullptr, grid, a.rc, rng[w], lane, flags, (double)pi.x, (double)pi.y,
              (double)pi.z, load, visit);
  drain(qn);
  const double sphi = tm_wave_sum_d((double)acc.phi);
  const double gx = tm_wave_sum_d((double)acc.gx);
  const double gy = tm_wave_sum_d((double)acc.gy);
  const double gz = tm_wave_sum_d((double)acc.gz);
  const double ev = tm_wave_sum_d((double)acc.ev);
  if (lane == 0) {
    phi[i] = sphi;
    grad[3 * i] = gx; grad[3 * i + 1] = gy; grad[3 * i + 2] = gz;
    eatom[2 * i] = 0.5 * (double)pi.w * sphi;
    eatom[2 * i + 1] = 0.5 * ev;
  }
}

// ---- host: the tables ---------------------------------------------------------
// Pair functions of r_A in float64, with their r_A-derivatives.
struct PairFns {
  double alpha, r_sr, r_lr, zz, yy, shift, ea, lin;   // DSF-elu, bohr units
  double inv_width;                                    // vdW switch, r in A * BOHRPERA^2
  int add_ecc, has_vdw;

  void coulomb(double rA, double* K, double* dK) const {
    *K = 0.0; *dK = 0.0;
    if (!add_ecc) return;
    const double r = rA * CL_BOHRPERA;
    if (r > r_lr) return;
    if (r > r_sr) {
      const double ec = erfc(alpha * r);
      *K = ec / r - zz + (r - r_lr) * lin;
      *dK = (-ec / (r * r) - 1.1283791670955126 * alpha * exp(-alpha * r * alpha * r) / r + lin) *
            CL_BOHRPERA;
    } else {
      const double ex = exp(r - r_sr);
      *K = ea * (ex - 1.0) + shift;
      *dK = ea * ex * CL_BOHRPERA;
    }
  }
  // vc = sqrt(C6_a C6_b), vq = 6 (R0_a + R0_b)^12
  void vdw(double rA, double vc, double vq, double* V, double* dV) const {
    *V = 0.0; *dV = 0.0;
    if (!has_vdw) return;
    const double B2 = CL_BOHRPERA * CL_BOHRPERA;
    const double r = rA * B2;
    const double tt = fmin(fmax(r * inv_width, 0.0), 1.0);
    const double S = -tt * tt * (2.0 * tt - 3.0);
    const double dS = (r * inv_width > 0.0 && r * inv_width < 1.0) ? 6.0 * tt * (1.0 - tt) * inv_width
                                                                  : 0.0;
    const double r5 = r * r * r * r * r, r6 = r5 * r, r12 = r6 * r6;
    const double den = r12 + vq;
    const double f = vc * r6 / den;
    const double df = 6.0 * vc * r5 * (vq - r12) / (den * den);   // d/dr of r^6 / (r^12 + vq)
    *V = -S * f;
    *dV = -(dS * f + S * df) * B2;
  }
};

// Cubic pieces of fn over the nodes r_k = k h - off, k = 0..nint.
template <class F>
void fill_table(F fn, double h, double off, int nint, float* out) {
  double f0 = 0.0, d0 = 0.0;
  fn(-off, &f0, &d0);
  for (int k = 0; k < nint; ++k) {
    double f1 = 0.0, d1 = 0.0;
    fn((double)(k + 1) * h - off, &f1, &d1);
    const double c1 = h * d0, e1 = h * d1;
    out[4 * k] = (float)f0;
    out[4 * k + 1] = (float)c1;
    out[4 * k + 2] = (float)(3.0 * (f1 - f0) - 2.0 * c1 - e1);
    out[4 * k + 3] = (float)(2.0 * (f0 - f1) + c1 + e1);
    f0 = f1;
    d0 = d1;
  }
}

}  // namespace

extern "C" {

int tm_coulomb_table(const double* consts, double rc_ee, double width, int add_ecc, int has_vdw,
                     int nele, const double* c6, const double* r0, int nint, float* out,
                     double* scale) {
  TM_REQUIRE(nele > 0 && nele <= TM_MAX_ELE, "tm_coulomb_table: bad nele %d", nele);
  TM_REQUIRE(nint >= 16, "tm_coulomb_table: nint %d below 16", nint);
  TM_REQUIRE(rc_ee > 0.0 && width > 0.0, "tm_coulomb_table: bad cutoff / width");
  TM_REQUIRE(consts != nullptr && c6 != nullptr && r0 != nullptr && out != nullptr &&
                 scale != nullptr,
             "tm_coulomb_table: null argument");
  PairFns P;
  P.alpha = consts[0]; P.r_sr = consts[1]; P.r_lr = consts[2]; P.zz = consts[3];
  P.yy = consts[4]; P.shift = consts[5]; P.ea = consts[6];
  P.lin = consts[3] / consts[2] + consts[4];
  P.inv_width = 1.0 / width;
  P.add_ecc = add_ecc;
  P.has_vdw = has_vdw;
  // nint - 1 intervals span the cutoff; the grid is shifted by off in [0, h)
  // so that the node m h - off lands on the matching radius R_sr (in A)
  const double h = rc_ee / (nint - 1);
  const double rsr = P.r_sr / CL_BOHRPERA;
  const double off = (rsr > 0.0 && rsr < rc_ee) ? ceil(rsr / h) * h - rsr : 0.0;
  scale[0] = 1.0 / h;
  scale[1] = off / h;
  fill_table([&](double r, double* v, double* d) { P.coulomb(fmax(r, 0.0), v, d); }, h, off, nint,
             out);
  int t = 1;
  for (int ea = 0; ea < nele; ++ea)
    for (int eb = ea; eb < nele; ++eb, ++t) {
      const double vc = sqrt(c6[ea] * c6[eb]);
      const double R0 = r0[ea] + r0[eb];
      const double q6 = R0 * R0 * R0 * R0 * R0 * R0;
      const double vq = 6.0 * q6 * q6;
      fill_table([&](double r, double* v, double* d) { P.vdw(fmax(r, 0.0), vc, vq, v, d); }, h,
                 off, nint, out + 4L * t * nint);
    }
  return TM_OK;
}

int tm_coulomb_tab(int N, const double* box, const int* cell_start, const int* cell_atoms,
                       const int* cell_of, int nx, int ny, int nz, double csx, double csy,
                       double csz, double ox, double oy, double oz, int periodic, double rc_ee,
                       int nele, const float* table, int nint, double inv_h, double u0,
                       const float* xqc, const int* elc, double* phi, double* grad, double* eatom,
                       int* flags, void* stream) {
  TM_REQUIRE(nele > 0 && nele <= TM_MAX_ELE, "tm_coulomb_tab: bad nele %d", nele);
  const int ntab = 1 + nele * (nele + 1) / 2;
  TM_REQUIRE(nint >= 16 && (long)ntab * nint <= CL_TABLE_MAX,
             "tm_coulomb_tab: %d tables of %d intervals exceed %d entries", ntab, nint,
             CL_TABLE_MAX);
  TM_REQUIRE(cell_start != nullptr && cell_atoms != nullptr && cell_of != nullptr &&
                 table != nullptr && xqc != nullptr && elc != nullptr,
             "tm_coulomb_tab: null argument");
  TM_REQUIRE(!periodic || box != nullptr, "tm_coulomb_tab: periodic without box");
  TM_REQUIRE(nx > 0 && ny > 0 && nz > 0 && csx > 0 && csy > 0 && csz > 0,
             "tm_coulomb_tab: bad grid");
  TM_REQUIRE(inv_h > 0.0 && u0 >= 0.0 && u0 < 1.0 && (rc_ee * inv_h + u0) < nint,
             "tm_coulomb_tab: table scale (%g, %g) does not cover the cutoff", inv_h, u0);
  const double cs[3] = {csx, csy, csz};
  const int nc[3] = {nx, ny, nz};
  const int need = tm_range_bound(rc_ee, cs, nc, periodic);
  TM_REQUIRE(need <= CT_RANGES_WIDE,
             "tm_coulomb_tab: cells of %.2f A are too small for a %.1f A cutoff (a sweep "
             "could need %d candidate ranges, room for %d)", fmin(csx, fmin(csy, csz)), rc_ee,
             need, CT_RANGES_WIDE);
  if (N == 0) return TM_OK;
  ClArgs a;
  a.rc = (float)rc_ee;
  a.rc2 = (float)(rc_ee * rc_ee);
  a.inv_h = (float)inv_h;
  a.u0 = (float)u0;
  a.nint = nint;
  a.nele = nele;
  CellGrid g;
  g.cell_start = cell_start; g.cell_atoms = cell_atoms; g.cell_of = cell_of;
  g.cell_pos = nullptr; g.posc = nullptr; g.molc = nullptr;
  g.nc[0] = nx; g.nc[1] = ny; g.nc[2] = nz;
  g.cs[0] = (float)csx; g.cs[1] = (float)csy; g.cs[2] = (float)csz;
  g.org[0] = (float)ox; g.org[1] = (float)oy; g.org[2] = (float)oz;
  g.periodic = periodic;
  const dim3 grid(tm_div_up(N, CT_WAVES)), block(CT_WAVES * 64);
  if (need <= CT_RANGES)
    coulomb_tab_kernel<CT_RANGES><<<grid, block, 0, (hipStream_t)stream>>>(
        N, (const float4*)xqc, elc, box, g, a, (const float4*)table, phi, grad, eatom, flags);
  else
    coulomb_tab_kernel<CT_RANGES_WIDE><<<grid, block, 0, (hipStream_t)stream>>>(
        N, (const float4*)xqc, elc, box, g, a, (const float4*)table, phi, grad, eatom, flags);
  TM_LAUNCH_CHECK();
  return TM_OK;
}

}  // extern "C"
