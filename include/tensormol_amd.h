/* tensormol_amd.h -- C ABI of libtmamd.so, the MI355X (gfx950) BP+EE engine.
 *
 * Conventions
 *   - Every pointer argument is a device pointer unless documented as a host
 *     array; sizes are element counts; `stream` is a hipStream_t.
 *   - Calls only enqueue kernels on `stream` (no host synchronisation, no
 *     memset or copy nodes), so a sequence of them can be captured into a HIP
 *     graph and replayed.
 *   - Return value: 0 on success, negative on error; tm_last_error() then
 *     holds a message (thread-local).
 *   - Packed atoms: molecule m owns atoms [mol_start[m], mol_start[m+1]);
 *     mol[i] is the molecule of atom i; elem[i] is the index of atom i's
 *     element in the model's ascending element list; pos is float64 [N,3] in A;
 *     box is float64 [B,3] orthorhombic edges (NULL = open boundaries).
 *   - Neighbour rows use an ELL layout with stride ldr (4.6 A list) and lda
 *     (3.1 A sub-list), see tm_nl_build.
 *
 * Reference paths below are relative to the TensorMol repository.
 */
#ifndef TENSORMOL_AMD_H
#define TENSORMOL_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

int tm_version(void);
const char* tm_last_error(void);
int tm_zero_i32(int n, int* p, void* stream);

/* ---- neighbour lists ---------------------------------------------------
 * Replaces MolEmb.Make_NListNaive (C_API/MolEmb.cpp:1180-1247) and the pair /
 * triple assembly of NeighborList.buildPairsAndTriples and
 * NeighborListSet.buildPairsAndTriplesWithEleIndex
 * (TensorMol/ForceModifiers/Neighbors.py:117-201, 344-423).
 *
 * Cell grid (one large open or periodic system, or a batch): tm_cell_build
 * bins atoms into an nx*ny*nz grid of cells with origin o and edges c
 * (periodic: positions wrapped; open: clamped, flag 2 raised) and sorts each
 * cell's atoms by index.  Buffers: cell_of[N], cell_count[ncell] (zero-filled
 * before the first call; every call leaves it zero, so calls sharing it must
 * not overlap), cell_start[ncell+1], cell_fill[N] (staging), cell_atoms[N]
 * (atoms in cell order), cell_pos[N] (cell-order position of each atom),
 * posc[3N] (positions in cell order), molc[N] (molecules in cell order; NULL
 * for one system).
 * Sweeps over the grid cover the cutoff sphere with pruned stencil rows; cells
 * of about half the short cutoff are the intended size.
 */
int tm_cell_build(int N, const double* pos, const int* mol, double ox, double oy, double oz,
                  double cx, double cy, double cz, int nx, int ny, int nz, int periodic,
                  int* cell_of, int* cell_count, int* cell_start, int* cell_fill, int* cell_atoms,
                  int* cell_pos, double* posc, int* molc, int* flags, void* stream);

/* Builds, for every atom i, the full (DoPerms=1) list of neighbours j with
 * squared distance s < s_r and marks those with s < s_a.  s_r/s_a are the
 * first doubles failing sqrt(s) + 1e-13 < rc (MolEmb.cpp:1214-1218), so the
 * sets are bit-exact; rc (>= the pair cutoff) sizes the cell stencil.  Rows
 * are sorted by (element of j, j).  Candidates come from the molecule
 * (cell_atoms == NULL) or from the grid of tm_cell_build (nx..nz cells of edges
 * csx..csz from the origin ox..oz, the arguments given to tm_cell_build; each
 * centre's stencil rows are trimmed with its own position).
 * Outputs: nbr[i*ldr+n], ncnt[i], rseg[i*(nele+1)+e], ang[i*lda+a] (position
 * in row i), acnt[i], aseg[i*(nele+1)+e], rev[i*ldr+n] (mirror slot; may be
 * NULL, and tm_sym_forward can fill it instead), flags (0: row overflow, 1:
 * sub-list overflow, 3: non-finite input). */
int tm_nl_build(int N, const double* pos, const int* elem, const int* mol, const int* mol_start,
                const double* box, const int* cell_start, const int* cell_atoms,
                const int* cell_of, const int* cell_pos, const double* posc, const int* molc,
                int nx, int ny, int nz, double csx, double csy, double csz, double ox, double oy,
                double oz, int periodic, int nele, double rc, double s_r, double s_a, int ldr,
                int lda, int* nbr, int* ncnt,
                int* rseg, int* ang, int* acnt, int* aseg, int* rev, int* flags, void* stream);

/* Materialises the sorted triples of every centre (Neighbors.py:369-420):
 * out_jkpr[4*t] = (j, k, element-pair index, rank in its (i, pair) group),
 * ordered by (i, pair, k, j); trip_start[i] = exclusive prefix of
 * acnt[i]*(acnt[i]-1)/2 (int64). */
int tm_triples(int N, int ldr, int lda, int nele, const int* nbr, const int* ang, const int* acnt,
               const int* aseg, const long* trip_start, int* out_jkpr, void* stream);

/* Pair list for any cutoff (replaces Make_NListNaive with DoPerms 0/1 and the
 * EE list NeighborListSet(..., DoPerms=False).buildPairs, Neighbors.py:262-282):
 * call with start == NULL to get count[i], then with start = exclusive prefix
 * (int64) to fill out[start[i] ..].  nreal (per molecule, may be NULL) limits
 * the centres; half = 1 keeps j > i only. */
int tm_pairs(int N, const double* pos, const int* mol, const int* mol_start, const int* nreal,
             const double* box, const int* cell_start, const int* cell_atoms, const int* cell_of,
             const int* cell_pos, const double* posc, const int* molc, int nx, int ny, int nz,
             double csx, double csy, double csz, double ox, double oy, double oz, int periodic,
             double rc, double s_c, int half,
             int* count, const long* start, int* out, int* flags, void* stream);

/* ---- descriptors -------------------------------------------------------
 * TFSymRSet_Linear_WithEle + TFSymASet_Linear_WithEle concatenated as in
 * TFSymSet_Scattered_Linear_WithEle (TensorMol/TFDescriptors/RawSymFunc.py:
 * 1696-1755, 868-962, 2223-2265).  G float32 [N, ldg], column l*nr + s
 * (radial) then nele*nr + p*64 + t*8 + s (angular).  Host arrays: rs_r[nr],
 * rs_a[8], thetas[8].  rev (may be NULL) receives the mirror slots of the
 * rows, as tm_nl_build's rev output (unmatched entries raise bit 2 of flags[0]),
 * so a caller may pass rev = NULL to tm_nl_build and have it filled here. */
int tm_sym_forward(int N, const double* pos, const double* box, const int* mol, const int* elem,
                   const int* nbr, const int* ncnt, const int* rseg, const int* ang,
                   const int* acnt, const int* aseg, int ldr, int lda, int nele, int nr,
                   float eta, float zeta, float rr, float ra, const float* rs_r,
                   const float* rs_a, const float* thetas, float* G, int ldg, int* rev,
                   int* flags, void* stream);

/* Descriptor adjoint (replaces tf.gradients through the descriptor graph,
 * TFMolInstanceDirect.py:5761): with g_i = gE_i + u_i * gQ_i (gQ may be NULL),
 * writes wslot[(i*ldr+n)*4 + 0..2] = d(sum_f g_if G_if)/d(x_j - x_i).  u (float
 * [N]) may be NULL when phi (double [N]) and phimean (double [B]) are given:
 * u_i = phi_i - phimean[mol_i] is then formed in the kernel, as
 * tm_charge_upstream forms it. */
int tm_sym_adjoint(int N, const double* pos, const double* box, const int* mol, const int* elem,
                   const int* nbr, const int* ncnt, const int* rseg, const int* ang,
                   const int* acnt, const int* aseg, int ldr, int lda, int nele, int nr,
                   float eta, float zeta, float rr, float ra, const float* rs_r,
                   const float* rs_a, const float* thetas, const float* gE, const float* gQ,
                   const float* u, const double* phi, const double* phimean, int ldg,
                   float* wslot, void* stream);

/* grad_a = sum_n (w[rev(a,n)] - w[(a,n)]) + extra_a (Ha/A);
 * force = -scale * grad (scale = JOULEPERHARTREE gives J/mol/A as returned by
 * TFMolManage.EvalBPDirectEEUpdateSingle, TFMolManage.py:1320). */
/* Descriptor tangent V[i, f] = sum_j dG[i, f]/d(x_j - x_i) . (du_j - du_i) for
 * per-atom displacements du (device double [N,3]); V float [N, ldg].  The
 * transpose of tm_sym_adjoint.  Replaces the second-order part of the
 * tf.gradients graph that the force term of the training loss differentiates
 * through (loss_op_EandG, TFMolInstanceDirect.py:4889-4903, with the energy
 * gradient of :5761): dLoss/dW = -d/dW sum_i (dE_i/dG_i)(W) . V_i for
 * du = dLoss/dF. */
int tm_sym_jvp(int N, const double* pos, const double* box, const int* mol, const int* elem,
               const int* nbr, const int* ncnt, const int* rseg, const int* ang, const int* acnt,
               const int* aseg, int ldr, int lda, int nele, int nr, float eta, float zeta,
               float rr, float ra, const float* rs_r, const float* rs_a, const float* thetas,
               const double* du, float* V, int ldg, void* stream);

/* ---- training-step MLP kernels ------------------------------------------
 * The MLP half of the force-matching step's second-order graph
 * (loss_op_EandG / train_step_EandG, TFMolInstanceDirect.py:4889-4903,
 * 5288-5348, differentiating tf.gradients(E, xyz), :5761), written out as
 * GEMMs: forward, the input gradient J = dy/dG, the tangent along
 * V = dLoss/dJ, the reverse pass and the weight gradients
 * (tensormol_amd/train.py lists the sequence).
 *
 * C = op(A) op(B) on v_mfma_f32_16x16x32_bf16: FP32 operands in memory,
 * rounded to BF16 when staged, FP32 accumulation.  A(m,k) = ta ? A[k*lda+m] :
 * A[m*lda+k]; B(k,n) = tb ? B[n*ldb+k] : B[k*ldb+n].  epi 0: C = acc (+ bias[n]
 * when bias is given); 1: z = acc + bias[n], C = log(1 + e^{alpha z})/alpha,
 * aux = sigmoid(alpha z); 2: C = acc, aux = acc * s (s, aux indexed as C).
 * nsplit > 1 (epi 0, no bias): K is split over nsplit workgroup layers whose
 * partial products go to work (nsplit * M * ldc floats) and are added in
 * split order.  C, aux and work must not overlap A, B, bias or s. */
int tm_gemm_bf16(int M, int N, int K, const float* A, int lda, int ta, const float* B, int ldb,
                 int tb, float* C, int ldc, int epi, const float* bias, const float* s,
                 float* aux, float alpha, int nsplit, float* work, void* stream);

/* One GEMM of a grouped call; fields as the arguments of tm_gemm_bf16, plus
 * the fused elementwise steps of the double backward (x1, x2, out2; NULL when
 * unused):
 *   epi 1 with out2: out2 = aux * x1[n]                  (zd_3 = s_3 w_4)
 *   epi 2 with out2: out2 = s * x2[m] * x1[n] + alpha s (1 - s) * acc * x1[n]
 *                    (zb_3 from the last tangent layer, x2 = ybar, x1 = w_4);
 *                    C may be NULL with epi 2
 *   epi 3:           C = s * acc + alpha s (1 - s) * x1 * x2
 *                    (zb_{l-1} = s hb + a s (1 - s) zt hd; s, x1, x2 indexed
 *                    as C)
 * The formulas are those of tm_train_elem ops 0, 2 and 3 (op 1 is epi 2's
 * aux). */
typedef struct tm_gemm_group {
  int M, N, K;
  const void* A;       /* FP32, or BF16 with io bit 0 */
  int lda;
  const void* B;       /* as A */
  int ldb;
  void* C;             /* FP32, or BF16 with io bit 1 */
  int ldc;
  const float* bias;
  const float* s;
  void* aux;           /* FP32, or BF16 with io bit 2 */
  float* work;
  const float* x1;
  const float* x2;
  void* out2;          /* FP32, or BF16 with io bit 3 */
  const int* crow;     /* epi 0 without split K: row m of the product goes to row
                          crow[m] of C (a scatter), or NULL */
} tm_gemm_group;

/* ngroups independent GEMMs (groups: host array) with the same operand forms,
 * epilogue and split count, in one launch per 16 of them: the per-element
 * networks of one layer (and the four weight-gradient GEMMs of every element)
 * run side by side instead of one after another.  Each result equals the
 * corresponding tm_gemm_bf16 call bit for bit.  io selects storage types for
 * the whole call: bit 0 A and B are BF16 (the staged bits: a BF16 operand
 * gives the result of its FP32 value exactly), bits 1-3 C, aux, out2 are
 * written as BF16 (rounded to nearest even; split K needs an FP32 C); bias, s,
 * x1, x2 and work are FP32.  The outputs of a GEMM (C,
 * aux, out2, work) must not overlap any input of the call (A, B, bias, s, x1,
 * x2 of any GEMM) or each other. */
int tm_gemm_bf16_grouped(int ngroups, const tm_gemm_group* groups, int ta, int tb, int epi,
                         int io, float alpha, int nsplit, void* stream);

/* Elementwise steps between those GEMMs over rows r < M, columns n < H (every
 * [M][H] operand with row stride ld): op 0 out = s * w4[n]; 1 out = s * x;
 * 2 out = s * ybar[r] * w4[n] + alpha s (1 - s) zt w4[n];
 * 3 out = s * x + alpha s (1 - s) zt hd. */
int tm_train_elem(int op, int M, int H, int ld, float alpha, const float* s, const float* x,
                  const float* zt, const float* hd, const float* w4, const float* ybar, float* out,
                  void* stream);

/* Row gather (scatter = 0: out[r*ldo + c] = in[idx[r]*ldi + c]) or scatter
 * (scatter = 1: out[idx[r]*ldo + c] = in[r*ldi + c]) of M rows, c < cols. */
int tm_rows(int scatter, int M, int cols, const int* idx, const float* in, int ldi, float* out,
            int ldo, void* stream);

/* One row gather / scatter of a grouped call; fields as tm_rows' arguments;
 * out_bf16: out is BF16 (the FP32 input rounded to nearest even). */
typedef struct tm_rows_group {
  int M, cols;
  const int* idx;
  const float* in;
  int ldi;
  void* out;
  int ldo;
  int out_bf16;
} tm_rows_group;

/* ngroups tm_rows calls (groups: host array; the per-element row sets of one
 * step) in one launch per 16 of them. */
int tm_rows_grouped(int ngroups, const tm_rows_group* groups, int scatter, void* stream);

/* out[0 .. n) = BF16 of in (round to nearest even) for ngroups arrays (host
 * array of records), one launch per 16: the weights of one training step,
 * cast once for every GEMM that reads them. */
typedef struct tm_cast_group {
  const float* in;
  void* out;
  long n;
} tm_cast_group;
int tm_cast_bf16_grouped(int ngroups, const tm_cast_group* groups, void* stream);

/* ---- serving ------------------------------------------------------------
 * Result staging of one batch step: copies nparts device arrays (host array
 * of records; bytes and offset in whole 4-byte words) into dst at their
 * offsets in one launch per 16, so the host reads every result of the step
 * (status flags, energies, forces, positions, row counts) through one
 * device-to-host copy.  No counterpart in the reference (its loop returns one
 * molecule's results per TF session run). */
typedef struct tm_pack_part {
  const void* src;
  long bytes;
  long offset;
} tm_pack_part;
int tm_pack_copy(int nparts, const tm_pack_part* parts, void* dst, void* stream);

/* The loss side of the force-matching step (loss_op_EandG,
 * TFMolInstanceDirect.py:4889-4903): per molecule m of atoms
 * [mol_start[m], mol_start[m+1]), E[m] = sum of the atom energies eps,
 * ybar[i] = 2 w_e[m] (E[m] - E_ref[m]) (dLoss/dy of each atom), F = -grad,
 * gg[i] = -2 w_f[m] (F_i - F_ref_i) (dLoss/dgrad, [N][3]), and the loss
 * L[0] = sum_m w_e[m] (E[m] - E_ref[m])^2 + sum_i w_f[m] |F_i - F_ref_i|^2 in
 * FP64 (part: B doubles of workspace).  Deterministic: fixed summation order. */
int tm_fm_loss(int B, const int* mol_start, const float* eps, const double* E_ref,
               const double* w_e, const double* grad, const double* F_ref, const double* w_f,
               double* E, float* ybar, double* gg, double* part, double* L, void* stream);

int tm_force_gather(int N, int ldr, const int* ncnt, const int* rev, const float* wslot,
                    const double* extra, double scale, double* grad, double* force,
                    void* stream);
/* tm_force_gather followed by tm_vv_kick_drift with the gathered force (scale
 * = JOULEPERHARTREE gives the force tm_vv_kick takes), in one launch and
 * bit-identical to the two calls: one MD step of VelocityVerletStep
 * (SimpleMD.py:14-38) ends with the force evaluation.  ke may be NULL. */
int tm_force_gather_vv(int N, int ldr, const int* ncnt, const int* rev, const float* wslot,
                       const double* extra, double scale, double* grad, double* force, double dt,
                       const double* m, double* v, double* a, double* ke, double* x,
                       void* stream);

/* ---- atomic networks ---------------------------------------------------
 * Energy / charge branches of energy_inference and dipole_inference
 * (TFMolInstanceDirect.py:5164-5212, 5219-5272) with act = log(1+e^{a z})/a
 * (Util.py:200-201) and their input gradient.
 *
 * tm_mlp_f32: plain FP32 path, one (net, element) segment of M atoms
 * perm[0..M); W = float32 [W1, b1, ..., W_nh, b_nh, wo, bo]; writes y[atom]
 * and dG[atom*ldg + f]; ws holds (2*nh+2)*M*max(hid) floats. */
int tm_mlp_f32(int M, int D, int nh, const int* hid, const float* G, int ldg, const int* perm,
               const float* W, float alpha, float* y, float* dG, float* ws, void* stream);

/* tm_mlp_mfma: fused BF16/FP16 MFMA path, FP32 accumulation, both nets and all
 * elements in one launch, one workgroup per entry of tiles (device int32
 * [ntiles][2]) = (segment, first row of the tile inside the segment); entries
 * with segment -1 are skipped (padding of an XCD-ordered table, so each XCD's
 * L2 serves few weight blobs); seg (device int32 [nseg][4]) = (net,
 * blob index, first row in perm, rows); blobs: device array of weight-blob
 * pointers packed as tm_mlp_mfma_pack_info describes; mode 0 = bf16, 1 = fp16,
 * 2 = bf16 hi/lo split of weights and activations, 3 = bf16 with hi/lo split
 * weights in the three forward GEMMs only (activations and the input-gradient
 * GEMMs plain bf16); tile_rows = atoms per tile, 64 or 32 (32 in mode 2).
 * Writes y[net*N + atom] and dG[(net*N + atom)*ldg + f]. */
int tm_mlp_mfma(int N, int D, int H, int ntiles, const int* tiles, const int* seg, int ldg,
                const float* G, const int* perm, const void* blobs, float alpha, int mode,
                int tile_rows, float* y, float* dG, void* stream);
/* out (10 entries): out[0] blob bytes, out[1] byte offset of the float section
 * (b1, b2, b3, wo, bo), out[2..7] element offsets of the forward (layers 1-3)
 * and backward (layers 3, 2, 1) fragment arrays, out[8] planes of the forward
 * arrays, out[9] planes of the backward arrays (2 = hi plane then lo plane). */
int tm_mlp_mfma_pack_info(int D, int H, int nh, int mode, long* out);

/* ---- electrostatics ----------------------------------------------------
 * Neutral charges q = qraw - mean_m(qraw) and dipole terms
 * (TFMolInstanceDirect.py:5274-5279); xq = float32 (x, y, z, q) (may be NULL
 * when cell_pos is given).  With mol_start the molecule means are formed in
 * the launch (every block sums its molecules in one fixed order, FP64) and
 * written to qmean when it is not NULL (molecules with atoms); with
 * mol_start = NULL qmean is the input (tm_mol_sum_f32 with mean = 1).  With
 * cell_pos (tm_cell_build's) the same values and elem are also written in
 * cell order to xqc (float4 [N]) and elc (int [N]), the copies tm_coulomb
 * sweeps (it then takes xq = NULL). */
int tm_charges(int N, const int* mol, const int* mol_start, double* qmean, const float* qraw,
               const double* pos, float* xq, double* q, double* qx, const int* cell_pos,
               const int* elem, float* xqc, int* elc, void* stream);

/* DSF-elu Coulomb (RawSymFunc.py:1307-1359) and Grimme-C6 vdW
 * (RawSymFunc.py:1361-1414) over all pairs within rc_ee, evaluated from the
 * cell grid (or the molecule).  consts (host f64[7]) = alpha', R_sr, R_lr, ZZ,
 * YY, elu_shift, elu_alpha in bohr units; c6/r0 host f64[nele].  The grid
 * arguments are those of tm_cell_build (NULL cell_atoms: molecule sweeps).  Outputs
 * phi[i] = sum_j K(r_ij) q_j, grad (dE/dx at fixed q, Ha/A), eatom[2i] =
 * q_i phi_i / 2, eatom[2i+1] = vdW energy share.  With a cell grid xqc
 * (float4 [N]) and elc (int [N]) hold the cell-ordered copies of xq and elem:
 * gathered here from xq, or, when xq is NULL, already written by tm_charges. */
int tm_coulomb(int N, const float* xq, const int* elem, const int* mol, const int* mol_start,
               const double* box, const int* cell_start, const int* cell_atoms, const int* cell_of,
               const int* cell_pos, const int* molc, int nx, int ny, int nz, double csx, double csy,
               double csz, double ox, double oy, double oz, int periodic, const double* consts,
               double rc_ee, double width,
               int add_ecc, int has_vdw, int nele, const double* c6, const double* r0, float* xqc,
               int* elc, double* phi, double* grad, double* eatom, int* flags, void* stream);

/* Tables of the pair functions tm_coulomb_tab reads (host only, no device
 * work): nint intervals of width h = rc_ee / (nint - 1) (A), interval k
 * spanning [k h - off, (k + 1) h - off] with off in [0, h) chosen so that a
 * node lands on the elu matching radius; float4 cubic pieces (c0, c1, c2, c3)
 * of c0 + c1 t + c2 t^2 + c3 t^3, t in [0, 1), fitted in float64 to the values
 * and r-derivatives at the interval ends.  scale (host double[2]) receives
 * 1/h and off/h: r maps to the table coordinate r / h + off / h.  Table 0 is
 * the DSF-elu kernel K(r) (RawSymFunc.py:1307-1359; zero when add_ecc is 0);
 * tables 1.. are the vdW energies V(r) of the element pairs (a <= b) in the
 * order (0,0), (0,1), ..., (0,nele-1), (1,1), ... (RawSymFunc.py:1361-1414,
 * with its double A->bohr scaling; zero when has_vdw is 0).  out: host
 * float[4 * (1 + nele (nele + 1) / 2) * nint]; consts, width, c6, r0 as for
 * tm_coulomb. */
int tm_coulomb_table(const double* consts, double rc_ee, double width, int add_ecc, int has_vdw,
                     int nele, const double* c6, const double* r0, int nint, float* out,
                     double* scale);

/* The tm_coulomb sweep for one system on a cell grid (molc NULL), from the
 * cell-ordered xqc / elc that tm_charges writes, with the pair functions
 * from tm_coulomb_table's tables (device copy; at most 3072 float4 entries,
 * nint >= 16; inv_h, u0 = its scale) and float32 lane sums added in float64.
 * Same outputs as tm_coulomb (phi, grad, eatom by atom index; flags as its
 * range overflow); every centre's sums are formed in an order fixed by the
 * inputs. */
int tm_coulomb_tab(int N, const double* box, const int* cell_start, const int* cell_atoms,
                       const int* cell_of, int nx, int ny, int nz, double csx, double csy,
                       double csz, double ox, double oy, double oz, int periodic, double rc_ee,
                       int nele, const float* table, int nint, double inv_h, double u0,
                       const float* xqc, const int* elc, double* phi, double* grad, double* eatom,
                       int* flags, void* stream);

/* u_i = phi_i - mean_m(phi): dE_cc/dqraw_i through the neutrality step. */
int tm_charge_upstream(int N, const int* mol, const double* phi, const double* phimean, float* u,
                       void* stream);

/* Per-molecule column sums (or means, mean != 0) of an [N, C] array:
 * the sparse_reduce_sum / reduce_sum per molecule of RawSymFunc.py:1354-1359
 * and TFMolInstanceDirect.py:5212, in a fixed FP64 order.  scratch (may be
 * NULL; 1 + ceil(N/1024)*C doubles, zero-filled before its first use) lets one
 * large molecule (B == 1, n_total = N > 2048) be summed by many blocks in one
 * launch: its first word is a completion counter that each call leaves at
 * zero, so calls sharing a scratch buffer must not overlap. */
int tm_mol_sum_f32(int B, const int* mol_start, int C, const float* in, int ldin, double* out,
                   int mean, double* scratch, int n_total, void* stream);
int tm_mol_sum_f64(int B, const int* mol_start, int C, const double* in, int ldin, double* out,
                   int mean, double* scratch, int n_total, void* stream);

/* Several per-molecule column sums in one launch, as tm_mol_sum_f64 per
 * source: the closing reductions of one evaluation (per-molecule BP energy,
 * Coulomb and vdW energies, dipole; TFMolInstanceDirect.py:5212, 5278-5279,
 * RawSymFunc.py:1354-1359).  Host arrays of nsrc <= 4 entries: source k is an
 * [N, ld[k]] device array (float64 when f64[k], else float32) whose first
 * cols[k] columns are summed into out[k] (device, [B, cols[k]]); at most 16
 * columns in all.  scratch as for tm_mol_sum_f64 with C = total columns. */
int tm_mol_sums(int B, const int* mol_start, int n_total, int nsrc, const void* const* in,
                const int* f64, const int* cols, const int* ld, double* const* out,
                double* scratch, void* stream);

/* tm_mol_sums with no scratch argument and no allocation: the many-block
 * path's partials and counter live in a library-owned static device array
 * (512 KiB per GPU, held until the process ends), set up on the first call on
 * each device.  No counterpart in the reference: it is for callers that
 * cannot allocate (HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  n_total above 4,194,304 atoms returns TM_ERR_CAPACITY.  Every
 * call on a GPU shares that GPU's array, so calls that overlap on one GPU
 * (other streams, other threads) race; results are bit-identical to
 * tm_mol_sums with a workspace. */
int tm_mol_sums_static(int B, const int* mol_start, int n_total, int nsrc, const void* const* in,
                       const int* f64, const int* cols, const int* ld, double* const* out,
                       void* stream);

/* ---- integrators (TensorMol/Simulations/SimpleMD.py) --------------------
 * VelocityVerletStep (SimpleMD.py:14-38) split into drift and kick;
 * NoseThermostat.step (SimpleMD.py:104-129) with one eta per molecule;
 * FIRE (Bitzek et al. 2006; no reference counterpart, the optimiser of the
 * serving workload): tm_fire_update works on atoms [0, N) of molecules
 * mol0 .. mol0+B-1; state[4b] = (dt, alpha, steps with P > 0, active),
 * mix[3b] scratch; dots_sum[3b] = per-molecule (F.v, F.F, v.v). */
int tm_vv_drift(int N, double dt, const double* v, const double* a, double* x, void* stream);
int tm_vv_kick(int N, double dt, const double* f, const double* m, double* v, double* a,
               double* ke, void* stream);
/* tm_vv_kick with the forces of step k, then tm_vv_drift of step k + 1, in one
 * launch (bit-identical to the two calls in turn): a loop of steps that starts
 * with one drift runs each step as evaluate + kick-drift. */
int tm_vv_kick_drift(int N, double dt, const double* f, const double* m, double* v, double* a,
                     double* ke, double* x, void* stream);
int tm_nh_drift(int N, double dt, const int* mol, const double* eta, const double* m, double* x,
                const double* v, const double* a, double* vh, double* ke2, void* stream);
int tm_nh_eta(int B, const int* mol_start, double dt, const double* kT, const double* Q,
              const double* ke2sum, double* eta, void* stream);
int tm_nh_kick(int N, double dt, const int* mol, const double* eta, const double* f,
               const double* m, const double* vh, double* v, double* a, double* ke, void* stream);
int tm_fire_dots(int N, const double* f, const double* v, double* dots, void* stream);
int tm_fire_update(int N, int B, const int* mol, int mol0, const double* dots_sum, double* state,
                   double* mix, const double* f, const double* m, double* v, double* x,
                   double dtmax, int nmin, double finc, double fdec, double alpha0,
                   double falpha, double maxstep, void* stream);

#ifdef __cplusplus
}
#endif
#endif
