// Candidate generation over a cell grid, shared by the neighbour-row,
// pair-list and Coulomb sweeps.
//
// The stencil of a centre is the set of x-rows (dy, dz) whose band of cells
// comes within the cutoff of the centre, each trimmed in x to the cells that
// can still hold a neighbour; with cells of about half the short cutoff this
// is a close cover of the cutoff sphere.  Every row is one contiguous
// range of the cell-ordered atom array (two when it wraps a periodic edge).
// The ranges of one centre are collected in LDS with a prefix sum of their
// lengths, so a wavefront sweeps the concatenated candidates 64 at a time
// whatever the cell occupancy.
//
// Minimum distances are computed from cell indices; for an open grid atoms
// outside it are clamped to edge cells, which only lowers index differences,
// so the bound stays a lower bound of the true distance.
#pragma once
#include "tm_common.h"

struct CellGrid {
  const int* cell_start;   // ncell+1, offsets into cell_atoms
  const int* cell_atoms;   // atoms sorted by cell (index order inside a cell)
  const int* cell_of;      // cell of each atom
  const int* cell_pos;     // position of each atom in cell_atoms
  const double* posc;      // positions in cell order (float64 [N,3])
  const int* molc;         // molecule of each cell-ordered atom (NULL: one system)
  int nc[3];               // cells per axis
  float cs[3];             // cell edges (A)
  float org[3];            // grid origin (A): cell k of an axis spans [org + k cs, org + (k+1) cs)
  int periodic;            // 1: wrap, 0: clamp
};

// Margin (A) added to the cutoff when rows are trimmed with the centre's own
// position: it covers the FP32 rounding of positions and cell bounds, so the
// candidates stay a superset of the atoms inside the cutoff.
constexpr float TM_TRIM_MARGIN = 2e-3f;

// Distance along one axis from a centre at offset f inside its cell to the band
// of cells d cells away (0 for its own band).
__device__ __forceinline__ float tm_band_dist(int d, float f, float cs) {
  const float v = d > 0 ? (float)d * cs - f : (d < 0 ? f - (float)(d + 1) * cs : 0.f);
  return fmaxf(v, 0.f);
}

// Lanes below this one whose bit is set in a wave-wide ballot.
__device__ __forceinline__ int tm_lanes_below(unsigned long long bal) {
  return __builtin_amdgcn_mbcnt_hi((unsigned)(bal >> 32),
                                   __builtin_amdgcn_mbcnt_lo((unsigned)bal, 0u));
}

// Ranges per centre (at most two per stencil row).  A sweep of cutoff rc over
// cells of edge cs has at most (2*ceil(rc/cs)+3)^2 rows.
template <int CAP>
struct RangeLds {
  static constexpr int cap = CAP;
  int lo[CAP];
  int pre[CAP + 1];
  int mark[64];   // range starts inside the current chunk, tagged with the chunk's base
};

// Builds the candidate ranges of atom i for cutoff rc into R (one wavefront,
// all lanes participate).  Returns the number of candidates; *nrange gets the
// number of ranges (capped at CAP, overflow flagged).  Every range is
// non-empty, so the range starts pre[] increase strictly.
// Rows are trimmed with the centre's position (px, py, pz): along y and z the
// distance from the centre to a row's band of cells, along x the cells that
// [x - s, x + s] covers with s^2 = (rc + margin)^2 - dy^2 - dz^2.  On an axis
// whose stencil wraps the whole periodic box the cell-gap bound (|d| - 1) cs is
// used instead (two images of a band can be equally near there).
template <int CAP>
__device__ inline int tm_build_ranges(int i, const CellGrid& g, float rc, RangeLds<CAP>& R, int lane,
                                      int* flags, int* nrange, double px, double py, double pz) {
  const int c = g.cell_of[i];
  const int cx = c % g.nc[0], cy = (c / g.nc[0]) % g.nc[1], cz = c / (g.nc[0] * g.nc[1]);
  int k[3];
  bool full[3];
  float f[3];
  {
    const int ci[3] = {cx, cy, cz};
    const double p[3] = {px, py, pz};
    for (int d = 0; d < 3; ++d) {
      k[d] = (int)ceilf(rc / g.cs[d]) + 1;
      full[d] = g.periodic && (2 * k[d] + 1 >= g.nc[d]);
      double u = p[d] - (double)g.org[d];
      if (g.periodic) {   // wrapped in FP64, so unwrapped coordinates keep their precision
        const double Ld = (double)g.nc[d] * (double)g.cs[d];
        u -= Ld * floor(u / Ld);
      }
      f[d] = (float)(u - (double)ci[d] * (double)g.cs[d]);   // offset inside the centre's cell
    }
  }
  R.mark[lane] = -1;   // no chunk base is negative; ordered by the wave barriers below
  const int ny = full[1] ? g.nc[1] : 2 * k[1] + 1;
  const int nz = full[2] ? g.nc[2] : 2 * k[2] + 1;
  const float rcm = rc + TM_TRIM_MARGIN;
  const float rc2 = rcm * rcm;
  const float inv_cs0 = 1.0f / g.cs[0];
  int nr = 0;   // wave-uniform
  for (int base = 0; base < ny * nz; base += 64) {
    int t = base + lane;
    int lo0 = 0, hi0 = 0, lo1 = 0, hi1 = 0, cnt = 0;
    if (t < ny * nz) {
      int oy = t % ny, oz = t / ny;
      int dy, dz, y, z;
      bool ok = true;
      if (full[1]) { y = oy; dy = y - cy; if (g.periodic) { if (dy > g.nc[1] / 2) dy -= g.nc[1]; if (dy < -g.nc[1] / 2) dy += g.nc[1]; } }
      else { dy = oy - k[1]; y = cy + dy; if (g.periodic) y = ((y % g.nc[1]) + g.nc[1]) % g.nc[1]; else ok = ok && y >= 0 && y < g.nc[1]; }
      if (full[2]) { z = oz; dz = z - cz; if (g.periodic) { if (dz > g.nc[2] / 2) dz -= g.nc[2]; if (dz < -g.nc[2] / 2) dz += g.nc[2]; } }
      else { dz = oz - k[2]; z = cz + dz; if (g.periodic) z = ((z % g.nc[2]) + g.nc[2]) % g.nc[2]; else ok = ok && z >= 0 && z < g.nc[2]; }
      if (ok) {
        float my = full[1] ? fmaxf(0.f, (float)(abs(dy) - 1)) * g.cs[1] : tm_band_dist(dy, f[1], g.cs[1]);
        float mz = full[2] ? fmaxf(0.f, (float)(abs(dz) - 1)) * g.cs[2] : tm_band_dist(dz, f[2], g.cs[2]);
        float rem = rc2 - my * my - mz * mz;
        if (rem > 0.f) {
          const float sx = sqrtf(rem);
          int xlo = cx + (int)floorf((f[0] - sx) * inv_cs0);
          int xhi = cx + (int)floorf((f[0] + sx) * inv_cs0);
          int row = g.nc[0] * (y + g.nc[1] * z);
          if (g.periodic && xhi - xlo + 1 >= g.nc[0]) {
            lo0 = g.cell_start[row]; hi0 = g.cell_start[row + g.nc[0]]; cnt = 1;
          } else if (!g.periodic) {
            xlo = max(xlo, 0);
            xhi = min(xhi, g.nc[0] - 1);
            if (xlo <= xhi) {
              lo0 = g.cell_start[row + xlo]; hi0 = g.cell_start[row + xhi + 1]; cnt = 1;
            }
          } else {
            if (xlo < 0) {
              lo0 = g.cell_start[row + xlo + g.nc[0]]; hi0 = g.cell_start[row + g.nc[0]];
              lo1 = g.cell_start[row]; hi1 = g.cell_start[row + xhi + 1]; cnt = 2;
            } else if (xhi >= g.nc[0]) {
              lo0 = g.cell_start[row + xlo]; hi0 = g.cell_start[row + g.nc[0]];
              lo1 = g.cell_start[row]; hi1 = g.cell_start[row + xhi - g.nc[0] + 1]; cnt = 2;
            } else {
              lo0 = g.cell_start[row + xlo]; hi0 = g.cell_start[row + xhi + 1]; cnt = 1;
            }
          }
          // keep only non-empty ranges (both halves of a wrapped row can be empty)
          if (cnt == 2 && hi1 <= lo1) cnt = 1;
          if (cnt >= 1 && hi0 <= lo0) { lo0 = lo1; hi0 = hi1; cnt -= 1; }
        }
      }
    }
    // compact the non-empty ranges of this round into R (lane order = row order)
    unsigned long long b1 = __ballot(cnt >= 1);
    unsigned long long b2 = __ballot(cnt >= 2);
    int before = tm_lanes_below(b1) + tm_lanes_below(b2);
    int slot = nr + before;
    if (cnt >= 1 && slot < CAP) { R.lo[slot] = lo0; R.pre[slot + 1] = hi0 - lo0; }
    if (cnt >= 2 && slot + 1 < CAP) { R.lo[slot + 1] = lo1; R.pre[slot + 2] = hi1 - lo1; }
    nr += __popcll(b1) + __popcll(b2);
  }
  if (nr > CAP) {
    if (lane == 0 && flags) atomicOr(&flags[TM_FLAG_NB_OVERFLOW], 4);
    nr = CAP;
  }
  __syncwarp();
  // inclusive scan of lengths into pre[1..nr] (sequential over 64-wide blocks)
  if (lane == 0) R.pre[0] = 0;
  int carry = 0;
  for (int base = 0; base < nr; base += 64) {
    int t = base + lane;
    int v = t < nr ? R.pre[t + 1] : 0;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      int u = __shfl_up(v, o, 64);
      if (lane >= o) v += u;
    }
    if (t < nr) R.pre[t + 1] = carry + v;
    carry += __shfl(v, 63, 64);
  }
  __syncwarp();
  *nrange = nr;
  return __builtin_amdgcn_readfirstlane(carry);   // the same in every lane
}

// Position (in cell order) of the t-th candidate: binary search in pre.
template <int CAP>
__device__ __forceinline__ int tm_candidate(const RangeLds<CAP>& R, int nr, int t) {
  int lo = 0, hi = nr;   // find r with pre[r] <= t < pre[r+1]
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (R.pre[mid] <= t) lo = mid; else hi = mid;
  }
  return R.lo[lo] + (t - R.pre[lo]);
}

// Cell-order position of candidate base + lane of a sweep over the ranges in R
// (0 for base + lane >= total), and r0 advanced from the range holding
// candidate base to the range holding base + 64.  Call with base = 0, 64,
// 128, ... in turn, r0 = 0 first (wave-uniform); needs nr >= 1.
//
// Lane k looks at range r0 + k.  Ranges r0 + 1 .. r0 + span - 1 start inside
// the chunk, at offsets pre - base in [1, 63], distinct because every range
// is non-empty.  Each such lane tags slot (pre - base) of the wave's mark array
// with the chunk's base; every lane reads its own slot, and a ballot gives the
// start positions, so lane L's range is r0 + (number of starts at or below L).
// mark[0] is never written and older tags are other chunks' bases, so only
// this chunk's starts match.
template <int CAP>
__device__ __forceinline__ int tm_locate(RangeLds<CAP>& R, int nr, int total, int lane, int base,
                                         int& r0) {
  const int rk = r0 + lane;
  const int pv = R.pre[min(rk, nr)];
  const int lo_k = R.lo[min(rk, nr - 1)];
  const int pre_k = rk < nr ? pv : 0x7fffffff;   // past the last range: never inside
  const int span = __popcll(__ballot(pre_k <= base + 63));   // >= 1: range r0 holds base
  if (lane >= 1 && lane < span) R.mark[pre_k - base] = base;
  __syncwarp();
  const unsigned long long starts = __ballot(R.mark[lane] == base);
  const int idx = __popcll(starts & ((2ull << lane) - 1ull));
  const int lo_r = __shfl(lo_k, idx, 64);
  const int pre_r = __shfl(pre_k, idx, 64);
  // start of range r0 + span, the first one past the chunk (lane span holds it)
  const int nxt = span < 64 ? __builtin_amdgcn_readlane(pre_k, span)
                            : (r0 + 64 < nr ? R.pre[r0 + 64] : 0x7fffffff);
  r0 += span - 1 + (nxt <= base + 64 ? 1 : 0);
  const int t = base + lane;
  return t < total ? lo_r + (t - pre_r) : 0;
}

// Sweep the candidates of centre i: calls f(p, valid, mapped) with all 64
// lanes in lockstep (so f may use ballots); p indexes the cell-ordered arrays
// when mapped, the atom arrays otherwise (molecule range, no grid).
template <int CAP, class F>
__device__ __forceinline__ void tm_sweep(int i, const int* __restrict__ mol,
                                         const int* __restrict__ mol_start, const CellGrid& g,
                                         float rc, RangeLds<CAP>& R, int lane, int* flags,
                                         double px, double py, double pz, F&& f) {
  if (g.cell_atoms == nullptr) {
    const int m = mol[i];
    const int lo = mol_start[m], hi = mol_start[m + 1];
    for (int base = lo; base < hi; base += 64) {
      int p = base + lane;
      f(p, p < hi, 0);
    }
    return;
  }
  int nr = 0;
  const int total = tm_build_ranges(i, g, rc, R, lane, flags, &nr, px, py, pz);
  // Walk the concatenated ranges 64 candidates at a time.
  int r0 = 0;
  for (int base = 0; base < total; base += 64) {
    const int p = tm_locate(R, nr, total, lane, base, r0);
    f(p, base + lane < total, 1);
  }
}

// tm_sweep with the candidate loads of the next chunk in flight while the
// current chunk is processed: load(p, valid, mapped) fetches what f needs for
// one candidate and returns it by value; f(p, valid, mapped, payload) consumes
// it.  Same candidates in the same order as tm_sweep.
template <int CAP, class LD, class F>
__device__ __forceinline__ void tm_sweep_pf(int i, const int* __restrict__ mol,
                                            const int* __restrict__ mol_start, const CellGrid& g,
                                            float rc, RangeLds<CAP>& R, int lane, int* flags,
                                            double px, double py, double pz, LD&& load, F&& f) {
  if (g.cell_atoms == nullptr) {
    const int m = mol[i];
    const int lo = mol_start[m], hi = mol_start[m + 1];
    auto cur = load(lo + lane, lo + lane < hi, 0);
    for (int base = lo; base < hi; base += 64) {
      const int pn = base + 64 + lane;
      auto nxt = load(pn, pn < hi, 0);
      f(base + lane, base + lane < hi, 0, cur);
      cur = nxt;
    }
    return;
  }
  int nr = 0;
  const int total = tm_build_ranges(i, g, rc, R, lane, flags, &nr, px, py, pz);
  int r0 = 0;
  auto locate = [&](int base) { return tm_locate(R, nr, total, lane, base, r0); };
  if (total <= 0) return;
  int p = locate(0);
  auto cur = load(p, lane < total, 1);
  for (int base = 0; base < total; base += 64) {
    const int pc = p;
    const bool vc = base + lane < total;
    // issued unconditionally (one chunk past the end on the last pass: r0
    // stays in [0, nr - 1] and every lane is invalid), so the loads in flight
    // are the same on every path
    p = locate(base + 64);
    auto nxt = load(p, base + 64 + lane < total, 1);
    f(pc, vc, 1, cur);
    cur = nxt;
  }
}

// Upper bound of the candidate ranges one centre's sweep can need (the
// ranges tm_build_ranges keeps): stencil rows (dy, dz) whose cell-gap distance
// max(|d| - 1, 0) cs is within the trimmed cutoff (the band distance the sweep
// uses is never smaller), two ranges per row where a periodic row can wrap.
inline int tm_range_bound(double rc, const double cs[3], const int nc[3], int periodic) {
  const double r = rc + (double)TM_TRIM_MARGIN;
  int span[2];
  int full[2];
  for (int d = 1; d < 3; ++d) {
    int k = (int)ceil(rc / cs[d]) + 1;
    full[d - 1] = periodic && 2 * k + 1 >= nc[d];
    span[d - 1] = full[d - 1] ? nc[d] : 2 * k + 1;
  }
  auto gap = [&](int axis, int o) {
    int d = full[axis] ? (o <= nc[axis + 1] / 2 ? o : o - nc[axis + 1]) : o - span[axis] / 2;
    return fmax(0.0, (double)(abs(d) - 1)) * cs[axis + 1];
  };
  int rows = 0;
  for (int oz = 0; oz < span[1]; ++oz)
    for (int oy = 0; oy < span[0]; ++oy) {
      double gy = gap(0, oy), gz = gap(1, oz);
      rows += gy * gy + gz * gz < r * r;
    }
  return periodic ? 2 * rows : rows;
}
