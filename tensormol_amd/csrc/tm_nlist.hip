// Neighbour lists for the BP+EE path.
//
// Replaces MolEmb.Make_NListNaive (C_API/MolEmb.cpp:1180-1247) and the Python
// pair/triple assembly of NeighborList.buildPairsAndTriples and
// NeighborListSet.buildPairsAndTriplesWithEleIndex
// (TensorMol/ForceModifiers/Neighbors.py:117-201, 344-423).
//
// Pair predicate: the reference keeps j when sqrt(dx*dx+dy*dy+dz*dz)+1e-13 < rc
// in float64.  With correctly rounded sqrt and add that predicate is monotone
// in s = (dx*dx+dy*dy)+dz*dz, so the host computes the first double s* that
// fails it and the kernel tests s < s* with s evaluated without contraction:
// the same set, bit for bit, and no dependence on the device sqrt.
//
// Output (ELL rows, one per centre atom, stride ldr):
//   nbr[i*ldr + n]   neighbour atom j (global packed index), sorted by
//                    (element index of j, j) = the (m,i,l,j) order of trpE
//   ncnt[i]          row length
//   rseg[i*(nele+1)+e] start of element e's segment in the row (last = ncnt)
//   ang[i*lda + a]   position (in row i) of the a-th neighbour closer than
//                    the angular cutoff; the sub-list keeps the row order
//   acnt[i], aseg[i*(nele+1)+e]  as above for the sub-list
//   rev[i*ldr + n]   slot (j*ldr + n') of the mirror entry j -> i
//
// Candidate atoms come either from the atom's molecule (batches of small
// molecules) or from a cell grid (one large open or periodic system).
#include "tm_cells.h"

namespace {

constexpr int NL_WAVES = 4;           // waves per block, one centre per wave
constexpr int NL_LDS_CAP = 256;       // candidate buffer per wave
constexpr int NL_RANGES = 192;        // stencil ranges of a 4.6 A sweep over ~2.3 A cells

__device__ __forceinline__ double sq_dist_nocontract(double xi, double yi, double zi, double xj,
                                                     double yj, double zj, double Lx, double Ly,
                                                     double Lz) {
#pragma clang fp contract(off)
  double dx = tm_min_image(xi - xj, Lx);
  double dy = tm_min_image(yi - yj, Ly);
  double dz = tm_min_image(zi - zj, Lz);
  double s = dx * dx + dy * dy;
  s = s + dz * dz;
  return s;
}

__global__ void __launch_bounds__(NL_WAVES * 64)
nl_rows_kernel(int N, const double* __restrict__ pos, const int* __restrict__ elem,
               const int* __restrict__ mol, const int* __restrict__ mol_start,
               const double* __restrict__ box, CellGrid grid, int nele,
               double s_r, double s_a, float rc, int ldr, int lda,
               int* __restrict__ nbr, int* __restrict__ ncnt, int* __restrict__ rseg,
               int* __restrict__ ang, int* __restrict__ acnt, int* __restrict__ aseg,
               int* __restrict__ flags) {
  // key = element << 28 | j << 1 | (within the angular cutoff): the order of
  // keys is the order of (element, j) for distinct atoms (j < 2^27)
  __shared__ unsigned keys[NL_WAVES][NL_LDS_CAP];
  __shared__ unsigned sorted[NL_WAVES][NL_LDS_CAP];
  __shared__ RangeLds<NL_RANGES> rng[NL_WAVES];
  const int w = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * NL_WAVES + w;
  const bool active = i < N;   // wave-uniform; inactive waves still reach every barrier
  int count = 0;               // wave-uniform
  if (active) {
    double Lx = 0, Ly = 0, Lz = 0;
    const int mi = mol[i];
    if (box != nullptr) {
      Lx = box[3 * mi]; Ly = box[3 * mi + 1]; Lz = box[3 * mi + 2];
    }
    const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
    const int cpi = grid.cell_pos != nullptr ? grid.cell_pos[i] : -1;
    auto visit = [&](int p, bool valid, int mapped) {
      // invalid lanes read entry 0, so the loads and the distance are branch-free
      const int pq = valid ? p : 0;
      const double* q = mapped ? grid.posc + 3 * (long)pq : pos + 3 * (long)pq;
      const double s = sq_dist_nocontract(xi, yi, zi, q[0], q[1], q[2], Lx, Ly, Lz);
      const bool same_mol = !mapped || grid.molc == nullptr || grid.molc[pq] == mi;
      const bool take = valid && (mapped ? p != cpi : p != i) && same_mol;
      const bool ok = take && s < s_r;
      const bool in_a = take && s < s_a;
      unsigned long long bal = __ballot(ok);
      int before = tm_lanes_below(bal);
      if (ok) {
        int slot = count + before;
        int j = mapped ? grid.cell_atoms[p] : p;
        if (slot < NL_LDS_CAP)
          keys[w][slot] = ((unsigned)elem[j] << 28) | ((unsigned)j << 1) | (in_a ? 1u : 0u);
      }
      count += __popcll(bal);
    };
    tm_sweep(i, mol, mol_start, grid, rc, rng[w], lane, flags, xi, yi, zi, visit);
    if (count > NL_LDS_CAP || count > ldr) {
      if (lane == 0) atomicOr(&flags[TM_FLAG_NB_OVERFLOW], 1);
      count = min(count, min(NL_LDS_CAP, ldr));
    }
  }
  __syncthreads();
  // rank sort by (element, atom index)
  for (int a = lane; a < count; a += 64) {
    const unsigned ka = keys[w][a];
    int rank = 0;
    for (int b = 0; b < count; ++b) rank += keys[w][b] < ka;
    sorted[w][rank] = ka;
  }
  __syncthreads();
  if (!active) return;   // no barriers below
  long rowr = (long)i * ldr;
  long rowa = (long)i * lda;
  int acount = 0;
  int cr = 0, ca = 0;   // lane e <= nele: entries (all / angular) with element < e
  for (int base = 0; base < count; base += 64) {
    int a = base + lane;
    bool have = a < count;
    unsigned k = have ? sorted[w][a] : 0u;
    if (have) nbr[rowr + a] = (int)((k >> 1) & ((1u << 27) - 1u));
    bool isang = have && (k & 1u);
    unsigned long long bal = __ballot(isang);
    int before = tm_lanes_below(bal);
    if (isang) {
      int slot = acount + before;
      if (slot < lda) ang[rowa + slot] = a;
    }
    acount += __popcll(bal);
    const int el = (int)(k >> 28);
    for (int e = 1; e <= nele; ++e) {   // element segment starts of this chunk
      const unsigned long long br = __ballot(have && el < e);
      const unsigned long long ba = __ballot(isang && el < e);
      if (lane == e) { cr += __popcll(br); ca += __popcll(ba); }
    }
  }
  if (acount > lda) {
    if (lane == 0) atomicOr(&flags[TM_FLAG_ANG_OVERFLOW], 1);
    acount = lda;
  }
  if (lane == 0) { ncnt[i] = count; acnt[i] = acount; }
  if (lane <= nele) {
    if (ca > acount) ca = acount;
    rseg[(long)i * (nele + 1) + lane] = cr;
    aseg[(long)i * (nele + 1) + lane] = ca;
  }
}

// Mirror slot of every entry: rev[i*ldr+n] = j*ldr + position of i in row j.
__global__ void nl_rev_kernel(int N, int ldr, int nele, const int* __restrict__ elem,
                              const int* __restrict__ nbr, const int* __restrict__ ncnt,
                              const int* __restrict__ rseg, int* __restrict__ rev,
                              int* __restrict__ flags) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int i = (int)(t / ldr);
  int n = (int)(t % ldr);
  if (i >= N || n >= ncnt[i]) return;
  int j = nbr[(long)i * ldr + n];
  int e = elem[i];
  int lo = rseg[(long)j * (nele + 1) + e];
  int hi = rseg[(long)j * (nele + 1) + e + 1];
  const int* row = nbr + (long)j * ldr;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (row[mid] < i) lo = mid + 1; else hi = mid;
  }
  if (lo < ncnt[j] && row[lo] == i) {
    rev[(long)i * ldr + n] = (int)((long)j * ldr + lo);
  } else {
    rev[(long)i * ldr + n] = (int)((long)i * ldr + n);   // unmatched (truncated row)
    atomicOr(&flags[TM_FLAG_NB_OVERFLOW], 2);
  }
}

// ---------------------------------------------------------------- cell grid
__global__ void cell_assign_kernel(int N, const double* __restrict__ pos, double ox, double oy,
                                   double oz, double cx, double cy, double cz, int nx, int ny,
                                   int nz, int periodic, int* __restrict__ cell_of,
                                   int* __restrict__ cell_count, int* __restrict__ flags) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double p[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
  if (!isfinite(p[0]) || !isfinite(p[1]) || !isfinite(p[2])) {
    atomicOr(&flags[TM_FLAG_NONFINITE], 1);
    p[0] = p[1] = p[2] = 0.0;
  }
  double o[3] = {ox, oy, oz}, cs[3] = {cx, cy, cz};
  int n[3] = {nx, ny, nz}, c[3];
  for (int d = 0; d < 3; ++d) {
    long q = (long)floor((p[d] - o[d]) / cs[d]);
    if (periodic) {
      q %= n[d];
      if (q < 0) q += n[d];
    } else if (q < 0 || q >= n[d]) {
      atomicOr(&flags[TM_FLAG_CELL_OVERFLOW], 1);
      q = q < 0 ? 0 : n[d] - 1;
    }
    c[d] = (int)q;
  }
  int cell = c[0] + nx * (c[1] + ny * c[2]);
  cell_of[i] = cell;
  atomicAdd(&cell_count[cell], 1);
}

// Exclusive prefix sum of n ints (out[n] = total), zeroing the input as it is
// read (the per-cell counts become the fill counters).  A tile of 16384
// entries is scanned by a 1024-thread block: 16 consecutive entries per thread
// (four 16-byte loads; in and out are 16-byte aligned), a shuffle scan of the
// thread totals inside each wave, the 16 wave totals scanned through LDS.
// One block walks the tiles with a running carry when there are few; for
// more, a first launch writes the tile totals and a second scans every tile
// from the sum of the totals before it.
constexpr int SCAN_PT = 16;
constexpr int SCAN_TILE = 1024 * SCAN_PT;

__device__ __forceinline__ int scan_tile(int n, int* __restrict__ in, int* __restrict__ out,
                                         int base, int carry, int* wtot) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int i0 = base + SCAN_PT * tid;
  int v[SCAN_PT];
  if (i0 + SCAN_PT <= n) {
#pragma unroll
    for (int q = 0; q < SCAN_PT / 4; ++q) {
      const int4 x = reinterpret_cast<const int4*>(in + i0)[q];
      v[4 * q] = x.x; v[4 * q + 1] = x.y; v[4 * q + 2] = x.z; v[4 * q + 3] = x.w;
    }
#pragma unroll
    for (int q = 0; q < SCAN_PT / 4; ++q) reinterpret_cast<int4*>(in + i0)[q] = make_int4(0, 0, 0, 0);
  } else {
#pragma unroll
    for (int k = 0; k < SCAN_PT; ++k) v[k] = i0 + k < n ? in[i0 + k] : 0;
#pragma unroll
    for (int k = 0; k < SCAN_PT; ++k)
      if (i0 + k < n) in[i0 + k] = 0;
  }
  int t = 0;
#pragma unroll
  for (int k = 0; k < SCAN_PT; ++k) t += v[k];
  int incl = t;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int u = __shfl_up(incl, o, 64);
    if (lane >= o) incl += u;
  }
  if (lane == 63) wtot[w] = incl;
  __syncthreads();
  if (w == 0) {
    int x = lane < 16 ? wtot[lane] : 0;
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) {
      int u = __shfl_up(x, o, 64);
      if (lane >= o) x += u;
    }
    if (lane < 16) wtot[lane] = x;   // inclusive over waves
  }
  __syncthreads();
  int run = carry + (w > 0 ? wtot[w - 1] : 0) + incl - t;
  int e[SCAN_PT];
#pragma unroll
  for (int k = 0; k < SCAN_PT; ++k) { e[k] = run; run += v[k]; }
  if (i0 + SCAN_PT <= n) {
#pragma unroll
    for (int q = 0; q < SCAN_PT / 4; ++q)
      reinterpret_cast<int4*>(out + i0)[q] = make_int4(e[4 * q], e[4 * q + 1], e[4 * q + 2], e[4 * q + 3]);
  } else {
#pragma unroll
    for (int k = 0; k < SCAN_PT; ++k)
      if (i0 + k < n) out[i0 + k] = e[k];
  }
  const int total = wtot[15];
  __syncthreads();   // wtot is reused by the next tile
  return total;
}

__global__ void __launch_bounds__(1024) scan_kernel(int n, int* __restrict__ in,
                                                    int* __restrict__ out) {
  __shared__ int wtot[16];
  int carry = 0;
  for (int base = 0; base < n; base += SCAN_TILE) carry += scan_tile(n, in, out, base, carry, wtot);
  if (threadIdx.x == 0) out[n] = carry;
}

// totals of the 16384-entry tiles (fixed-order block reduction)
__global__ void __launch_bounds__(1024) scan_sums_kernel(int n, const int* __restrict__ in,
                                                         int* __restrict__ sums) {
  __shared__ int wsum[16];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int i0 = blockIdx.x * SCAN_TILE + SCAN_PT * tid;
  int t = 0;
  if (i0 + SCAN_PT <= n) {
#pragma unroll
    for (int q = 0; q < SCAN_PT / 4; ++q) {
      const int4 x = reinterpret_cast<const int4*>(in + i0)[q];
      t += (x.x + x.y) + (x.z + x.w);
    }
  } else {
    for (int k = 0; k < SCAN_PT; ++k) t += i0 + k < n ? in[i0 + k] : 0;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
  if (lane == 0) wsum[w] = t;
  __syncthreads();
  if (tid == 0) {
    int x = 0;
    for (int q = 0; q < 16; ++q) x += wsum[q];
    sums[blockIdx.x] = x;
  }
}

// each block scans its tile from the sum of the totals of the tiles before it
__global__ void __launch_bounds__(1024) scan_apply_kernel(int n, int* __restrict__ in,
                                                          int* __restrict__ out,
                                                          const int* __restrict__ sums) {
  __shared__ int wtot[16];
  __shared__ int carry_s;
  const int b = blockIdx.x, tid = threadIdx.x, lane = tid & 63;
  if (tid < 64) {
    int x = 0;
    for (int q = lane; q < b; q += 64) x += sums[q];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    if (tid == 0) carry_s = x;
  }
  __syncthreads();
  const int carry = carry_s;
  const int total = scan_tile(n, in, out, b * SCAN_TILE, carry, wtot);
  if (b == (int)gridDim.x - 1 && tid == 0) out[n] = carry + total;
}

// unsorted cell order into the staging array; the counters start at zero
__global__ void cell_fill_kernel(int N, const int* __restrict__ cell_of,
                                 const int* __restrict__ cell_start, int* __restrict__ counter,
                                 int* __restrict__ staging) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  int c = cell_of[i];
  int slot = cell_start[c] + atomicAdd(&counter[c], 1);
  staging[slot] = i;
}

// order atoms inside each cell by index, so every later sweep is
// deterministic: one wavefront per cell ranks its entries (cells hold a few
// dozen atoms; larger cells are ranked in 64-entry rounds).  Each wave also
// clears its cell's counter, so the counters are zero for the next build.
__global__ void cell_sort_kernel(int ncell, const int* __restrict__ cell_start,
                                 const int* __restrict__ unsorted, int* __restrict__ cell_atoms,
                                 int* __restrict__ cell_pos, const double* __restrict__ pos,
                                 double* __restrict__ posc, const int* __restrict__ mol,
                                 int* __restrict__ molc, int* __restrict__ counter) {
  const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (c >= ncell) return;
  if (lane == 0) counter[c] = 0;
  const int lo = cell_start[c], hi = cell_start[c + 1];
  for (int a = lo + lane; a < hi; a += 64) {
    int v = unsorted[a];
    int rank = 0;
    for (int b = lo; b < hi; ++b) rank += unsorted[b] < v;
    int q = lo + rank;
    cell_atoms[q] = v;
    cell_pos[v] = q;
    posc[3 * (long)q] = pos[3 * (long)v];
    posc[3 * (long)q + 1] = pos[3 * (long)v + 1];
    posc[3 * (long)q + 2] = pos[3 * (long)v + 2];
    if (molc != nullptr) molc[q] = mol[v];
  }
}

// Materialise the sorted triple list of one row set (API parity with
// buildPairsAndTriplesWithEleIndex): for each centre, triples grouped by
// element pair p, inside a group ordered by (k, j); rank = position in group.
// trip_start[i] is the exclusive prefix of per-centre triple counts.
__global__ void triples_kernel(int N, int ldr, int lda, int nele, const int* __restrict__ nbr,
                               const int* __restrict__ ang, const int* __restrict__ acnt,
                               const int* __restrict__ aseg, const long* __restrict__ trip_start,
                               int* __restrict__ out_jkpr) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  long o = trip_start[i];
  const int* seg = aseg + (long)i * (nele + 1);
  const int* arow = ang + (long)i * lda;
  const int* row = nbr + (long)i * ldr;
  int p = 0;
  for (int ea = 0; ea < nele; ++ea) {
    for (int eb = ea; eb < nele; ++eb, ++p) {
      int rank = 0;
      // order (k, j): k runs over the second element's segment ascending,
      // j over the first's ascending with j < k when the elements agree
      for (int kb = seg[eb]; kb < seg[eb + 1]; ++kb) {
        int jhi = (ea == eb) ? kb : seg[ea + 1];
        for (int ja = seg[ea]; ja < jhi; ++ja) {
          int j = row[arow[ja]], k = row[arow[kb]];
          out_jkpr[4 * o + 0] = j;
          out_jkpr[4 * o + 1] = k;
          out_jkpr[4 * o + 2] = p;
          out_jkpr[4 * o + 3] = rank++;
          ++o;
        }
      }
    }
  }
}

// Pair list for an arbitrary cutoff (the EE list of Neighbors.py:262-282 and
// Make_NListNaive with DoPerms 0/1): pass 1 counts, pass 2 writes each row in
// candidate order (ascending j for molecule sweeps, cell order otherwise).
// Centres i with local index >= nreal[m] get empty rows; with half = 1 only
// j > i is kept (pairs stored once under min(i, j)).
template <bool FILL>
__global__ void __launch_bounds__(NL_WAVES * 64)
pairs_kernel(int N, const double* __restrict__ pos, const int* __restrict__ mol,
             const int* __restrict__ mol_start, const int* __restrict__ nreal,
             const double* __restrict__ box, CellGrid grid, float rc, double s_c, int half,
             int* __restrict__ count, const long* __restrict__ start, int* __restrict__ out,
             int* __restrict__ flags) {
  __shared__ RangeLds<1024> rng[NL_WAVES];
  const int w = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * NL_WAVES + w;
  if (i >= N) return;   // wave-level synchronisation only
  const int mi = mol[i];
  if (nreal != nullptr && i - mol_start[mi] >= nreal[mi]) {
    if (!FILL && lane == 0) count[i] = 0;
    return;
  }
  double Lx = 0, Ly = 0, Lz = 0;
  if (box != nullptr) {
    Lx = box[3 * mi]; Ly = box[3 * mi + 1]; Lz = box[3 * mi + 2];
  }
  const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
  const int cpi = grid.cell_pos != nullptr ? grid.cell_pos[i] : -1;
  long o = FILL ? start[i] : 0;
  int cnt = 0;
  auto visit = [&](int p, bool valid, int mapped) {
    bool ok = false;
    int j = -1;
    if (valid && (mapped ? p != cpi : p != i) &&
        (!mapped || grid.molc == nullptr || grid.molc[p] == mi)) {
      j = mapped ? grid.cell_atoms[p] : p;
      const double* q = mapped ? grid.posc + 3 * (long)p : pos + 3 * (long)p;
      ok = (!half || j > i) && sq_dist_nocontract(xi, yi, zi, q[0], q[1], q[2], Lx, Ly, Lz) < s_c;
    }
    unsigned long long bal = __ballot(ok);
    if (FILL && ok) out[o + cnt + tm_lanes_below(bal)] = j;
    cnt += __popcll(bal);
  };
  tm_sweep(i, mol, mol_start, grid, rc, rng[w], lane, flags, xi, yi, zi, visit);
  if (!FILL && lane == 0) count[i] = cnt;
}

}  // namespace

extern "C" {

static CellGrid make_grid(const int* cell_start, const int* cell_atoms, const int* cell_of,
                          const int* cell_pos, const double* posc, const int* molc, int nx, int ny,
                          int nz, double csx, double csy, double csz, double ox, double oy,
                          double oz, int periodic) {
  CellGrid g;
  g.cell_start = cell_start; g.cell_atoms = cell_atoms; g.cell_of = cell_of;
  g.cell_pos = cell_pos; g.posc = posc; g.molc = molc;
  g.nc[0] = nx; g.nc[1] = ny; g.nc[2] = nz;
  g.cs[0] = (float)csx; g.cs[1] = (float)csy; g.cs[2] = (float)csz;
  g.org[0] = (float)ox; g.org[1] = (float)oy; g.org[2] = (float)oz;
  g.periodic = periodic;
  return g;
}

// cell_count (ncell ints) must be zero on entry (zero-fill it once; every call
// leaves it zero); it holds the per-cell counts, then the fill counters.
// cell_fill (at least N ints) receives the unsorted staging copy of cell_atoms.
// Outputs cell_atoms (index order inside each cell), cell_pos (position of
// every atom in cell order), posc (positions in cell order) and molc
// (molecules in cell order; may be NULL for one system).
int tm_cell_build(int N, const double* pos, const int* mol, double ox, double oy, double oz,
                  double cx, double cy, double cz, int nx, int ny, int nz, int periodic,
                  int* cell_of, int* cell_count, int* cell_start, int* cell_fill, int* cell_atoms,
                  int* cell_pos, double* posc, int* molc, int* flags, void* stream) {
  TM_REQUIRE(N >= 0 && nx > 0 && ny > 0 && nz > 0, "tm_cell_build: bad sizes");
  TM_REQUIRE((long)nx * ny * nz < (1l << 26), "tm_cell_build: %d x %d x %d cells is too many", nx, ny, nz);
  hipStream_t s = (hipStream_t)stream;
  long ncell = (long)nx * ny * nz;
  if (N == 0) return TM_OK;
  // cell_count is zero here: the previous call's sort cleared it (no memset
  // nodes when the caller captures a graph)
  cell_assign_kernel<<<tm_div_up(N, 256), 256, 0, s>>>(N, pos, ox, oy, oz, cx, cy, cz, nx, ny, nz,
                                                       periodic, cell_of, cell_count, flags);
  TM_LAUNCH_CHECK();
  const int ntile = (int)((ncell + SCAN_TILE - 1) / SCAN_TILE);
  if (ntile <= 2 || ntile > N) {
    scan_kernel<<<1, 1024, 0, s>>>((int)ncell, cell_count, cell_start);
    TM_LAUNCH_CHECK();
  } else {   // tile totals in cell_fill (not yet in use: the fill comes next)
    scan_sums_kernel<<<ntile, 1024, 0, s>>>((int)ncell, cell_count, cell_fill);
    TM_LAUNCH_CHECK();
    scan_apply_kernel<<<ntile, 1024, 0, s>>>((int)ncell, cell_count, cell_start, cell_fill);
    TM_LAUNCH_CHECK();
  }
  cell_fill_kernel<<<tm_div_up(N, 256), 256, 0, s>>>(N, cell_of, cell_start, cell_count,
                                                     cell_fill);
  TM_LAUNCH_CHECK();
  cell_sort_kernel<<<tm_div_up(ncell, 4), 256, 0, s>>>((int)ncell, cell_start, cell_fill,
                                                       cell_atoms, cell_pos, pos, posc, mol, molc,
                                                       cell_count);
  TM_LAUNCH_CHECK();
  return TM_OK;
}

int tm_nl_build(int N, const double* pos, const int* elem, const int* mol, const int* mol_start,
                const double* box, const int* cell_start, const int* cell_atoms,
                const int* cell_of, const int* cell_pos, const double* posc, const int* molc,
                int nx, int ny, int nz, double csx, double csy, double csz, double ox, double oy, double oz, int periodic, int nele,
                double rc, double s_r, double s_a, int ldr, int lda, int* nbr, int* ncnt,
                int* rseg, int* ang, int* acnt, int* aseg, int* rev, int* flags, void* stream) {
  TM_REQUIRE(N >= 0 && nele > 0 && nele <= TM_MAX_ELE, "tm_nl_build: bad nele %d", nele);
  TM_REQUIRE(ldr > 0 && ldr <= 256 && lda > 0 && lda <= ldr, "tm_nl_build: bad strides %d %d", ldr,
             lda);
  TM_REQUIRE((long)N * ldr < (1l << 31), "tm_nl_build: N*ldr overflows int32 slots");
  TM_REQUIRE(N <= (1 << 27), "tm_nl_build: %d atoms above 2^27", N);
  if (cell_atoms != nullptr) {
    TM_REQUIRE(cell_start && cell_of && cell_pos && posc, "tm_nl_build: incomplete cell grid");
    int k = (int)ceil(rc / fmin(csx, fmin(csy, csz))) + 1;
    TM_REQUIRE((2 * k + 3) * (2 * k + 3) * 2 <= NL_RANGES,
               "tm_nl_build: cells of %.2f A are too small for rc %.2f", fmin(csx, fmin(csy, csz)), rc);
  }
  if (N == 0) return TM_OK;
  hipStream_t s = (hipStream_t)stream;
  CellGrid g = make_grid(cell_start, cell_atoms, cell_of, cell_pos, posc, molc, nx, ny, nz, csx, csy,
                         csz, ox, oy, oz, periodic);
  nl_rows_kernel<<<tm_div_up(N, NL_WAVES), NL_WAVES * 64, 0, s>>>(
      N, pos, elem, mol, mol_start, box, g, nele, s_r, s_a, (float)rc, ldr, lda, nbr, ncnt, rseg,
      ang, acnt, aseg, flags);
  TM_LAUNCH_CHECK();
  if (rev != nullptr) {   // else tm_sym_forward fills rev
    long total = (long)N * ldr;
    nl_rev_kernel<<<tm_div_up(total, 256), 256, 0, s>>>(N, ldr, nele, elem, nbr, ncnt, rseg, rev,
                                                        flags);
    TM_LAUNCH_CHECK();
  }
  return TM_OK;
}

int tm_triples(int N, int ldr, int lda, int nele, const int* nbr, const int* ang, const int* acnt,
               const int* aseg, const long* trip_start, int* out_jkpr, void* stream) {
  TM_REQUIRE(N >= 0 && nele > 0 && nele <= TM_MAX_ELE, "tm_triples: bad sizes");
  if (N == 0) return TM_OK;
  hipStream_t s = (hipStream_t)stream;
  triples_kernel<<<tm_div_up(N, 128), 128, 0, s>>>(N, ldr, lda, nele, nbr, ang, acnt, aseg,
                                                   trip_start, out_jkpr);
  TM_LAUNCH_CHECK();
  return TM_OK;
}


int tm_pairs(int N, const double* pos, const int* mol, const int* mol_start, const int* nreal,
             const double* box, const int* cell_start, const int* cell_atoms, const int* cell_of,
             const int* cell_pos, const double* posc, const int* molc, int nx, int ny, int nz,
             double csx, double csy, double csz, double ox, double oy, double oz, int periodic, double rc, double s_c, int half,
             int* count, const long* start, int* out, int* flags, void* stream) {
  TM_REQUIRE(N >= 0, "tm_pairs: bad sizes");
  TM_REQUIRE(start == nullptr || out != nullptr, "tm_pairs: fill pass needs an output array");
  if (cell_atoms != nullptr) {
    int k = (int)ceil(rc / fmin(csx, fmin(csy, csz))) + 1;
    TM_REQUIRE((2 * k + 3) * (2 * k + 3) * 2 <= 1024,
               "tm_pairs: cells of %.2f A are too small for rc %.2f", fmin(csx, fmin(csy, csz)), rc);
  }
  if (N == 0) return TM_OK;
  hipStream_t s = (hipStream_t)stream;
  CellGrid g = make_grid(cell_start, cell_atoms, cell_of, cell_pos, posc, molc, nx, ny, nz, csx, csy,
                         csz, ox, oy, oz, periodic);
  if (start == nullptr)
    pairs_kernel<false><<<tm_div_up(N, NL_WAVES), NL_WAVES * 64, 0, s>>>(
        N, pos, mol, mol_start, nreal, box, g, (float)rc, s_c, half, count, nullptr, nullptr, flags);
  else
    pairs_kernel<true><<<tm_div_up(N, NL_WAVES), NL_WAVES * 64, 0, s>>>(
        N, pos, mol, mol_start, nreal, box, g, (float)rc, s_c, half, count, start, out, flags);
  TM_LAUNCH_CHECK();
  return TM_OK;
}

}  // extern "C"
