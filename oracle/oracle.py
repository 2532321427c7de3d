"""FP64 CPU restatement of the TensorMol-0.1 BP+EE path (the oracle).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.

Every function cites the reference file:line it restates (paths relative to the
TensorMol repository root).  Arithmetic is float64 throughout; forces come from
torch-CPU float64 autograd of the restated energy, which the tests check
against central finite differences.

Pinning (what checks each function, so that a dropped term, wrong sign or
swapped index fails):
  nlist_naive / pair_sets     brute-force set definition, ulp-adversarial
                              distances around rc, symmetry of the full list,
                              counts vs the half list (tests/test_oracle.py)
  pairs_triples_with_ele      ordering invariants + hand-built 4-atom case
  sym_radial / sym_angular    closed forms on 2- and 3-atom systems,
                              rotation/translation/permutation invariance
  bp_energy (MLP)             torch.nn.functional.linear + softplus
  dsf kernel                  C1 continuity at R_sr, zero value and slope at
                              R_lr, scipy erfc closed form
  forces                      central finite differences; sum F = 0; torque 0
  vdw kernel                  large-r limit -sqrt(C6i C6j)/r^6, zero at r -> 0,
                              C1 switch ending at r_A = Poly_Width/BOHRPERA
                              (the double A->bohr scaling), i/j symmetry
  md steps (vv_step)          energy conservation, time reversibility
  nose_step                   Q -> infinity reduces to vv_step; the extended
                              energy's drift vanishes with dt (tests/test_drivers.py)
  fire_step                   convergence on a harmonic bowl, the P <= 0 reset,
                              the dt growth after nmin downhill steps, the
                              per-atom step cap
  descriptor_jvp              central finite differences of the descriptors
  force_matching_loss         central finite differences in the weights;
                              zero at the teacher's own weights
TensorFlow's own summation order and last-ulp acos/erfc cannot be pinned
without TensorFlow (absent here); they sit far below the stated tolerances.
"""
import math

import numpy as np
import torch
from scipy.special import erfc as _erfc

# ---------------------------------------------------------------- constants
# PhysicalData.py:30-35,54 and TMParams.py:26-38,140-154
BOHRPERA = 1.889725989
KJPERHARTREE = 2625.499638
JOULEPERHARTREE = KJPERHARTREE * 1000.0
IDEALGASR = 8.3144621
PI_LITERAL = 3.14159265359          # RawSymFunc.py:935-936,1738
TWO_OVER_SQRTPI = 1.1283791671      # RawSymFunc.py:1336, Util.py:176
INFINITESIMAL = 1e-27               # RawSymFunc.py:901,911,1731
ONESCALAR = 1.0 - 0.0000000000000001  # RawSymFunc.py:914
NL_EPS = 0.0000000000001            # MolEmb.cpp:1217

AMU = {1: 1.00794, 6: 12.0107, 7: 14.0067, 8: 15.9994}   # PhysicalData.py:24
C6_COFF = {1: 0.14, 6: 1.75, 7: 1.23, 8: 0.70}           # PhysicalData.py:26
VDW_RADIUS = {1: 1.001, 6: 1.452, 7: 1.397, 8: 1.342}    # PhysicalData.py:25


def atomic_mass_kg_per_mol(z):
    """ATOMICMASSES = 0.000999977*ATOMICMASSESAMU (PhysicalData.py:54)."""
    return 0.000999977 * np.array([AMU[int(a)] for a in np.atleast_1d(z)])


# ---------------------------------------------------------- neighbour lists
def _pair_dist_ok(xi, xj, rc, box=None):
    """Predicate of MolEmb.cpp:1214-1218: sqrt(dx*dx+dy*dy+dz*dz)+1e-13 < rc.

    numpy evaluates (dx*dx + dy*dy) + dz*dz with correctly rounded ops and no
    contraction, as gcc does for the reference on x86-64.  With a box the
    displacement is the minimum image (the tessellation at Periodic.py:131-168
    gives the same set when every box edge exceeds 2*rc).
    """
    dx = xi[..., 0] - xj[..., 0]
    dy = xi[..., 1] - xj[..., 1]
    dz = xi[..., 2] - xj[..., 2]
    if box is not None:
        dx = dx - box[0] * np.round(dx / box[0])
        dy = dy - box[1] * np.round(dy / box[1])
        dz = dz - box[2] * np.round(dz / box[2])
    d = np.sqrt(dx * dx + dy * dy + dz * dz) + NL_EPS
    return d < rc


def nlist_naive(x, rc, nreal=None, doperms=1, box=None, chunk=1024):
    """Set restatement of MolEmb.Make_NListNaive (MolEmb.cpp:1180-1247).

    Returns a list (length nreal) of sorted int64 arrays: for doperms=1 every j
    with the pair predicate true (full list, centres i < nreal); for
    doperms=0 each unordered pair stored once under min(i, j)
    (MolEmb.cpp:1220-1228).  The reference's order inside a row follows its
    x-sorted sweep; every consumer re-sorts, so rows are returned sorted.
    """
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = x.shape[0]
    nreal = n if nreal is None else int(nreal)
    out = [None] * nreal
    idx = np.arange(n)
    for s in range(0, nreal, chunk):
        e = min(nreal, s + chunk)
        ok = _pair_dist_ok(x[s:e, None, :], x[None, :, :], rc, box)
        ok[np.arange(e - s), np.arange(s, e)] = False
        for r in range(e - s):
            i = s + r
            js = idx[ok[r]]
            if not doperms:
                js = js[js > i]
            out[i] = js
    return out


def pair_array(x, rc, nreal=None, doperms=1, box=None, mol=0):
    """NeighborList.buildPairs with molind (Neighbors.py:75-115): (P,3) [m,i,j]."""
    rows = nlist_naive(x, rc, nreal, doperms, box)
    cnt = np.array([len(r) for r in rows], dtype=np.int64)
    i = np.repeat(np.arange(len(rows)), cnt)
    j = np.concatenate(rows) if len(rows) else np.zeros(0, np.int64)
    m = np.full_like(i, mol)
    return np.stack([m, i, j.astype(np.int64)], axis=1)


def element_pairs(eles):
    """eles_pairs = [(e_a, e_b) for a <= b] (TFMolInstanceDirect.py:1263-1267)."""
    out = []
    for a in range(len(eles)):
        for b in range(a, len(eles)):
            out.append((eles[a], eles[b]))
    return np.array(out, dtype=np.int64)


def pairs_triples_with_ele(xyzs, natom, Zs, eles, rr, ra, nreal=None, boxes=None):
    """NeighborListSet.buildPairsAndTriplesWithEleIndex (Neighbors.py:344-423).

    xyzs (B, maxN, 3), natom (B,), Zs (B, maxN).  Per molecule builds the full
    4.6 A pair list and the 3.1 A triple list of NeighborList.buildPairsAndTriples
    (Neighbors.py:117-201): every unordered {j, k} of the short list with the
    lower-Z atom first (ties keep the lower index first, :170-186).  Tags pairs
    with l = index of Z_j in eles (:370-372) and triples with the index of the
    unordered element pair (:369,373-377), sorts pairs by (m,i,l,j) and triples
    by (m,i,l,k,j) (:380-383) and ranks triples inside their (m,i,l) group
    (:395-420).

    Returns trpE (P,4) [m,i,j,l], trtE (T,5) [m,i,j,k,l], mil_jk (T,4)
    [m,i,l,rank] (float64 like the reference), jk_max.
    """
    eles = [int(e) for e in eles]
    epairs = element_pairs(eles)
    eidx = {e: a for a, e in enumerate(eles)}
    pidx = {}
    for p, (a, b) in enumerate(epairs):
        pidx[(int(a), int(b))] = p
        pidx[(int(b), int(a))] = p
    P, T = [], []
    B = xyzs.shape[0]
    for m in range(B):
        n = int(natom[m])
        nr = n if nreal is None else int(nreal[m])
        x = xyzs[m, :n]
        z = Zs[m, :n]
        box = None if boxes is None else boxes[m]
        prow = nlist_naive(x, rr, nr, 1, box)
        trow = nlist_naive(x, ra, nr, 1, box)
        for i in range(nr):
            for j in prow[i]:
                P.append((m, i, int(j), eidx[int(z[j])]))
            tj = trow[i]
            for aa in range(len(tj)):
                for bb in range(aa + 1, len(tj)):
                    j, k = int(tj[aa]), int(tj[bb])    # j < k
                    if z[j] > z[k]:
                        j, k = k, j
                    T.append((m, i, j, k, pidx[(int(z[j]), int(z[k]))]))
    trpE = np.array(P, dtype=np.int64).reshape(-1, 4)
    trtE = np.array(T, dtype=np.int64).reshape(-1, 5)
    if len(trpE):
        trpE = trpE[np.lexsort((trpE[:, 2], trpE[:, 3], trpE[:, 1], trpE[:, 0]))]
    if len(trtE):
        trtE = trtE[np.lexsort((trtE[:, 2], trtE[:, 3], trtE[:, 4], trtE[:, 1], trtE[:, 0]))]
    mil_jk = np.zeros((len(trtE), 4))
    if len(trtE):
        key = trtE[:, [0, 1, 4]]
        new = np.ones(len(trtE), dtype=bool)
        new[1:] = np.any(key[1:] != key[:-1], axis=1)
        start = np.maximum.accumulate(np.where(new, np.arange(len(trtE)), 0))
        mil_jk[:, :3] = key
        mil_jk[:, 3] = np.arange(len(trtE)) - start
    jk_max = float(mil_jk[:, 3].max()) if len(trtE) else 0
    return trpE, trtE, mil_jk, jk_max


def half_pairs(xyzs, natom, rc, boxes=None):
    """EE list NeighborListSet(..., DoPerms=False).buildPairs (Neighbors.py:262-282)."""
    out = []
    for m in range(xyzs.shape[0]):
        n = int(natom[m])
        box = None if boxes is None else boxes[m]
        out.append(pair_array(xyzs[m, :n], rc, n, 0, box, m))
    return np.concatenate(out, axis=0) if out else np.zeros((0, 3), np.int64)


# ------------------------------------------------------------- descriptors
class SymParams:
    """AN1 grid of SetANI1Param (TFMolInstanceDirect.py:3128-3164) and the
    defaults of TMParams.py:26-38."""

    def __init__(self, eles, rr=4.6, ra=3.1, eta=4.0, zeta=8.0, nr_r=32, nr_a=8, nth_a=8):
        self.eles = [int(e) for e in eles]
        self.nele = len(self.eles)
        self.epairs = element_pairs(self.eles)
        self.nelep = len(self.epairs)
        self.rr, self.ra, self.eta, self.zeta = rr, ra, eta, zeta
        self.nr_r, self.nr_a, self.nth_a = nr_r, nr_a, nth_a
        self.rs_r = np.array([rr * i / nr_r for i in range(nr_r)])
        self.rs_a = np.array([ra * i / nr_a for i in range(nr_a)])
        self.thetas = np.array([2.0 * math.pi * i / nth_a for i in range(nth_a)])
        self.D = self.nele * nr_r + self.nelep * nr_a * nth_a


def _diff(x, a, b, box):
    """DifferenceVectorsLinear (RawSymFunc.py:123-133): x[a] - x[b]; minimum image with a box."""
    d = x[a] - x[b]
    if box is not None:
        bt = torch.as_tensor(box, dtype=x.dtype)
        d = d - bt * torch.round((d / bt).detach())
    return d


def sym_radial(x, i, j, l, natom_total, sp, box=None):
    """TFSymRSet_Linear_WithEle (RawSymFunc.py:1696-1755), flat atoms.

    x (N,3) torch f64, pair index arrays i, j, element index l.  Returns
    (N, nele*nr_r) with column l*nr_r + s.
    """
    Rij = _diff(x, i, j, box)
    r = torch.sqrt((Rij * Rij).sum(1) + INFINITESIMAL)
    rs = torch.as_tensor(sp.rs_r, dtype=x.dtype)
    fac1 = torch.exp(-sp.eta * (r[:, None] - rs[None, :]) ** 2)
    fac2 = 0.5 * (torch.cos(PI_LITERAL * r / sp.rr) + 1.0)
    Gm = fac1 * fac2[:, None]
    out = torch.zeros(natom_total, sp.nele, sp.nr_r, dtype=x.dtype)
    out.index_put_((torch.as_tensor(i), torch.as_tensor(l)), Gm, accumulate=True)
    return out.reshape(natom_total, -1)


def sym_angular(x, i, j, k, p, natom_total, sp, box=None):
    """TFSymASet_Linear_WithEle (RawSymFunc.py:868-962), flat atoms.

    Returns (N, nelep*nth_a*nr_a) with column p*64 + t*8 + s.
    """
    Rij = _diff(x, i, j, box)
    Rik = _diff(x, i, k, box)
    rij = torch.sqrt((Rij * Rij).sum(1) + INFINITESIMAL)
    rik = torch.sqrt((Rik * Rik).sum(1) + INFINITESIMAL)
    c = (Rij * Rik).sum(1) / (rij * rik)
    c = torch.where(c >= 1.0, torch.full_like(c, ONESCALAR), c)
    c = torch.where(c <= -1.0, torch.full_like(c, -ONESCALAR), c)
    theta = torch.acos(c)
    th = torch.as_tensor(sp.thetas, dtype=x.dtype)
    rs = torch.as_tensor(sp.rs_a, dtype=x.dtype)
    fac1 = (2.0 ** (1.0 - sp.zeta)) * (1.0 + torch.cos(theta[:, None] - th[None, :])) ** sp.zeta
    fac2 = torch.exp(-sp.eta * (((rij + rik) / 2.0)[:, None] - rs[None, :]) ** 2)
    fac3 = 0.5 * (torch.cos(PI_LITERAL * rij / sp.ra) + 1.0)
    fac4 = 0.5 * (torch.cos(PI_LITERAL * rik / sp.ra) + 1.0)
    Gm = fac1[:, :, None] * fac2[:, None, :] * (fac3 * fac4)[:, None, None]
    out = torch.zeros(natom_total, sp.nelep, sp.nth_a * sp.nr_a, dtype=x.dtype)
    out.index_put_((torch.as_tensor(i), torch.as_tensor(p)), Gm.reshape(len(c), -1), accumulate=True)
    return out.reshape(natom_total, -1)


# --------------------------------------------------------------------- MLP
def activation(z, alpha):
    """sigmoid_with_param = log(1+exp(alpha z))/alpha (Util.py:200-201); alpha=1
    is tf.nn.softplus (TFInstance.py:117-118).  logaddexp avoids the overflow of
    the literal form and is the same function."""
    return torch.logaddexp(torch.zeros_like(z), alpha * z) / alpha


def mlp(Gx, layers, alpha):
    """One element branch of energy_inference / dipole_inference
    (TFMolInstanceDirect.py:5164-5212, 5219-5272): hidden layers act(h W + b),
    then a linear regression layer; dropout is the identity at evaluation."""
    h = Gx
    for (W, b) in layers[:-1]:
        h = activation(h @ W + b, alpha)
    W, b = layers[-1]
    return (h @ W + b)[:, 0]


# --------------------------------------------------------------- Coulomb
def dsf_consts(dsf_alpha=0.18, elu_width=4.6, ee_off=15.0):
    """DSF / DSF_Gradient at the elu matching radius (Util.py:172-192, used at
    TFMolInstanceDirect.py:4371-4372).  Returns (alpha_bohr, R_sr, R_lr,
    elu_shift, elu_alpha) in bohr units."""
    a = dsf_alpha / BOHRPERA
    R = elu_width * BOHRPERA
    Rc = ee_off * BOHRPERA
    XX = a * Rc
    ZZ = _erfc(XX) / Rc
    YY = TWO_OVER_SQRTPI * a * math.exp(-XX * XX) / Rc
    shift = _erfc(a * R) / R - ZZ + (R - Rc) * (ZZ / Rc + YY)
    grad = -((_erfc(a * R) / R / R + TWO_OVER_SQRTPI * a * math.exp(-a * R * a * R) / R) - (ZZ / Rc + YY))
    return a, R, Rc, shift, grad


def dsf_kernel(r, consts):
    """Pair kernel K(r) of TFCoulombEluSRDSFLR (RawSymFunc.py:1307-1359), r in bohr."""
    a, R_sr, R_lr, shift, ea = consts
    XX = a * R_lr
    ZZ = float(_erfc(XX)) / R_lr
    YY = TWO_OVER_SQRTPI * a * math.exp(-XX * XX) / R_lr
    # torch.special.erfc matches scipy's erfc to ~1 ulp in float64
    LR = torch.special.erfc(a * r) / r - ZZ + (r - R_lr) * (ZZ / R_lr + YY)
    LR = torch.where(r > R_lr, torch.zeros_like(LR), LR)
    SR = ea * (torch.exp(r - R_sr) - 1.0) + shift
    return torch.where(r > R_sr, LR, SR)


def vdw_kernel(r_bohr, c6i, c6j, r0i, r0j, poly_width=4.6):
    """TFVdwPolyLR pair kernel (RawSymFunc.py:1361-1414).  The caller passes
    bohr and the function scales by BOHRPERA again (:1377); R_begin =
    EECutoffOn*BOHRPERA = 0 for the DSF_elu instance."""
    r = r_bohr * BOHRPERA
    width = poly_width * BOHRPERA
    t = r / width
    cut1 = torch.where(t > 0.0, -t * t * (2.0 * t - 3.0), torch.zeros_like(t))
    cut = torch.where(t > 1.0, torch.ones_like(t), cut1)
    return -cut * torch.sqrt(c6i * c6j) / r ** 6 * 1.0 / (1.0 + 6.0 * (r / (r0i + r0j)) ** (-12.0))


def vdw_params(z):
    """C6 and R0 per atom (TFMolInstanceDirect.py:3764-3767), atomic units."""
    c6 = np.array([C6_COFF[int(a)] * (BOHRPERA * 10.0) ** 6.0 / JOULEPERHARTREE for a in z])
    r0 = np.array([VDW_RADIUS[int(a)] * BOHRPERA for a in z])
    return c6, r0


# ------------------------------------------------------- full evaluation
class Lists:
    """Flat (packed-atom) neighbour data for one evaluation."""

    def __init__(self, pi, pj, pl, ti, tj, tk, tp, ci, cj):
        self.pi, self.pj, self.pl = pi, pj, pl
        self.ti, self.tj, self.tk, self.tp = ti, tj, tk, tp
        self.ci, self.cj = ci, cj


def build_lists(x, z, mol_start, sp, rc_coul=15.0, boxes=None):
    """Flat lists for packed molecules (mol_start (B+1,)): the sorted pair and
    triple sets of pairs_triples_with_ele and the EE half list, mapped to
    global atom indices."""
    B = len(mol_start) - 1
    pis, pjs, pls, tis, tjs, tks, tps, cis, cjs = ([] for _ in range(9))
    for m in range(B):
        s, e = int(mol_start[m]), int(mol_start[m + 1])
        box = None if boxes is None else boxes[m]
        xm = x[s:e][None]
        zm = z[s:e][None]
        trpE, trtE, _, _ = pairs_triples_with_ele(xm, np.array([e - s]), zm, sp.eles, sp.rr, sp.ra,
                                                  boxes=None if box is None else [box])
        pis.append(trpE[:, 1] + s); pjs.append(trpE[:, 2] + s); pls.append(trpE[:, 3])
        tis.append(trtE[:, 1] + s); tjs.append(trtE[:, 2] + s); tks.append(trtE[:, 3] + s)
        tps.append(trtE[:, 4])
        hp = pair_array(x[s:e], rc_coul, e - s, 0, box)
        cis.append(hp[:, 1] + s); cjs.append(hp[:, 2] + s)
    cat = lambda a: np.concatenate(a).astype(np.int64) if a else np.zeros(0, np.int64)
    return Lists(*(cat(a) for a in (pis, pjs, pls, tis, tjs, tks, tps, cis, cjs)))


def evaluate(x_np, z, mol_start, sp, energy_nets, charge_nets, alpha=1.0, dsf_alpha=0.18,
             has_vdw=True, add_ecc=True, lists=None, boxes=None, need_forces=True,
             elu_width=4.6, ee_off=15.0, poly_width=4.6):
    """BP+EE energy, charges, dipole and gradient for packed molecules.

    Restates MolInstance_DirectBP_EE_ChargeEncode_Update_vdw_DSF_elu_Normalize_
    Dropout.evaluate (TFMolInstanceDirect.py:5684-5711): descriptors
    (RawSymFunc.py:2223-2265), charge nets + neutrality + dipole
    (TFMolInstanceDirect.py:5219-5279), DSF-elu Coulomb (RawSymFunc.py:1307-1359),
    energy nets (TFMolInstanceDirect.py:5164-5215), vdW (RawSymFunc.py:1361-1414)
    and tf.gradients (TFMolInstanceDirect.py:5761).  For a box the Coulomb and
    vdW sums run over minimum-image unordered pairs, which equals the halved
    full-list sums of the periodic graph (TFMolInstanceDirect.py:5824,5896).

    energy_nets / charge_nets: dict element -> [(W (in,out) f64, b (out,)) ...].
    Returns dict of float64 numpy arrays: Etotal, Ebp, Ebp_atom, Ecc, Evdw,
    dipole (e*bohr), charge, gradient (Ha/A, per atom), G.
    """
    z = np.asarray(z)
    mol_start = np.asarray(mol_start, dtype=np.int64)
    N = len(z)
    B = len(mol_start) - 1
    if lists is None:
        lists = build_lists(x_np, z, mol_start, sp, ee_off, boxes)
    mol_of = np.repeat(np.arange(B), np.diff(mol_start))
    box = None
    if boxes is not None:
        assert B == 1, "periodic oracle handles one box"
        box = np.asarray(boxes[0], dtype=np.float64)
    x = torch.tensor(x_np, dtype=torch.float64, requires_grad=need_forces)
    G = torch.cat([sym_radial(x, lists.pi, lists.pj, lists.pl, N, sp, box),
                   sym_angular(x, lists.ti, lists.tj, lists.tk, lists.tp, N, sp, box)], dim=1)
    eps = torch.zeros(N, dtype=torch.float64)
    qraw = torch.zeros(N, dtype=torch.float64)
    for e in sp.eles:
        sel = torch.as_tensor(np.nonzero(z == e)[0])
        if len(sel) == 0:
            continue
        eps = eps.index_put((sel,), mlp(G[sel], _t(energy_nets[e]), alpha))
        qraw = qraw.index_put((sel,), mlp(G[sel], _t(charge_nets[e]), alpha))
    molt = torch.as_tensor(mol_of)
    natom = torch.as_tensor(np.diff(mol_start), dtype=torch.float64)
    netq = torch.zeros(B, dtype=torch.float64).index_add(0, molt, qraw)
    q = qraw - (netq * (1.0 / natom))[molt]
    xb = x * BOHRPERA
    dip = torch.zeros(B, 3, dtype=torch.float64).index_add(0, molt, xb * q[:, None])
    ci, cj = torch.as_tensor(lists.ci), torch.as_tensor(lists.cj)
    Ecc = torch.zeros(B, dtype=torch.float64)
    Evdw = torch.zeros(B, dtype=torch.float64)
    if len(lists.ci):
        Rij = _diff(xb, ci, cj, None if box is None else box * BOHRPERA)
        r = torch.sqrt((Rij * Rij).sum(1) + INFINITESIMAL)
        if add_ecc:
            K = dsf_kernel(r, dsf_consts(dsf_alpha, elu_width, ee_off))
            Ecc = Ecc.index_add(0, molt[ci], q[ci] * q[cj] * K)
        if has_vdw:
            c6, r0 = vdw_params(z)
            c6 = torch.as_tensor(c6); r0 = torch.as_tensor(r0)
            V = vdw_kernel(r, c6[ci], c6[cj], r0[ci], r0[cj], poly_width)
            Evdw = Evdw.index_add(0, molt[ci], V)
    Ebp = torch.zeros(B, dtype=torch.float64).index_add(0, molt, eps)
    Etot = Ebp + Ecc + Evdw
    out = dict(Etotal=Etot.detach().numpy(), Ebp=Ebp.detach().numpy(), Ebp_atom=eps.detach().numpy(),
               Ecc=Ecc.detach().numpy(), Evdw=Evdw.detach().numpy(), dipole=dip.detach().numpy(),
               charge=q.detach().numpy(), qraw=qraw.detach().numpy(), G=G.detach().numpy())
    if need_forces:
        (g,) = torch.autograd.grad(Etot.sum(), x)
        out["gradient"] = g.numpy()
    return out


def _t(layers):
    return [(torch.as_tensor(W, dtype=torch.float64), torch.as_tensor(b, dtype=torch.float64))
            for (W, b) in layers]


# -------------------------------------------------------------- training
def descriptors(x, z, sp, lists, box=None):
    """G = concat(radial, angular) for packed atoms (RawSymFunc.py:2223-2265);
    x torch f64 (N, 3), differentiable."""
    N = len(z)
    return torch.cat([sym_radial(x, lists.pi, lists.pj, lists.pl, N, sp, box),
                      sym_angular(x, lists.ti, lists.tj, lists.tk, lists.tp, N, sp, box)], dim=1)


def bp_energies(x, z, mol_start, sp, layers, alpha, lists, box=None):
    """Per-molecule BP energy sum_i eps_i (energy_inference,
    TFMolInstanceDirect.py:5164-5212, without the electrostatic terms);
    layers: element -> [(W, b), ...] torch f64 tensors (may require grad)."""
    z = np.asarray(z)
    B = len(mol_start) - 1
    G = descriptors(x, z, sp, lists, box)
    eps = torch.zeros(len(z), dtype=torch.float64)
    for e in sp.eles:
        sel = torch.as_tensor(np.nonzero(z == e)[0])
        if len(sel):
            eps = eps.index_put((sel,), mlp(G[sel], layers[e], alpha))
    molt = torch.as_tensor(np.repeat(np.arange(B), np.diff(mol_start)))
    return torch.zeros(B, dtype=torch.float64).index_add(0, molt, eps)


def descriptor_jvp(x_np, z, sp, du, lists, box=None):
    """dG/dx . du (N, D) in float64: the forward-mode derivative of the
    descriptors along per-atom displacements du (N, 3)."""
    x = torch.tensor(x_np, dtype=torch.float64)
    v = torch.as_tensor(np.asarray(du, dtype=np.float64))
    _, jv = torch.autograd.functional.jvp(lambda y: descriptors(y, z, sp, lists, box), (x,), (v,))
    return jv.numpy()


def force_matching_loss(x_np, z, mol_start, sp, layers, alpha, E_ref, F_ref, w_e, w_f, lists):
    """L = sum_b w_e[b] (E_b - E*_b)^2 + sum_i w_f[b(i)] |F_i - F*_i|^2 with
    F = -dE/dx (Ha/A), built with create_graph so dL/dW exists.

    loss_op_EandG (TFMolInstanceDirect.py:4889-4903) is this loss with
    w_e = EnergyScalar/2 (maxatom/natom_b)^2 and w_f = GradScalar/2
    (maxatom/natom_b)^2 (its residuals are energy gradients, -(F - F*));
    mean squared errors, L = MSE(E) + 0.5 MSE(F), take w_e = 1/B and
    w_f = 0.5/(3N).  Returns (L, E, F) as torch f64."""
    x = torch.tensor(x_np, dtype=torch.float64, requires_grad=True)
    E = bp_energies(x, z, mol_start, sp, layers, alpha, lists)
    (g,) = torch.autograd.grad(E.sum(), x, create_graph=True)
    F = -g
    B = len(mol_start) - 1
    molt = torch.as_tensor(np.repeat(np.arange(B), np.diff(mol_start)))
    w_e = torch.as_tensor(np.asarray(w_e, dtype=np.float64))
    w_f = torch.as_tensor(np.asarray(w_f, dtype=np.float64))
    E_ref = torch.as_tensor(np.asarray(E_ref, dtype=np.float64))
    F_ref = torch.as_tensor(np.asarray(F_ref, dtype=np.float64))
    L = (w_e * (E - E_ref) ** 2).sum() + (w_f[molt][:, None] * (F - F_ref) ** 2).sum()
    return L, E, F


def layers_torch(layers, requires_grad=False):
    """element -> [(W, b)] numpy -> torch f64 leaves."""
    return {e: [(torch.tensor(np.asarray(W, dtype=np.float64), requires_grad=requires_grad),
                 torch.tensor(np.asarray(b, dtype=np.float64), requires_grad=requires_grad))
                for (W, b) in ls] for e, ls in layers.items()}


# ------------------------------------------------------------------- MD
def vv_step(x, v, a, m, dt, eandf):
    """VelocityVerletStep (SimpleMD.py:14-38): A, fs, kg/mol, forces J/mol/A."""
    x = x + v * dt + 0.5 * a * dt * dt
    e, f = eandf(x)
    a_new = 1e-10 * f / m[:, None]
    v = v + 0.5 * (a + a_new) * dt
    return x, v, a_new, e


def kinetic_energy(v, m):
    """KineticEnergy per atom in J/mol (SimpleMD.py:40-51)."""
    return 0.5 * np.dot(np.einsum("ia,ia->i", v, v) * 1e10, m) / len(m)


def nose_step(x, v, a, m, dt, eandf, eta, T, mddt):
    """NoseThermostat.step (SimpleMD.py:104-129).  Returns x, v, a, e, eta."""
    N = len(m)
    kT = IDEALGASR * 1e-10 * T
    tau = 20.0 * mddt * N
    Q = kT * tau * tau
    x = x + v * dt + 0.5 * (a - eta * v) * dt * dt
    vdto2 = v + 0.5 * (a - eta * v) * dt
    e, f = eandf(x)
    a_new = 1e-10 * f / m[:, None]
    ke = 0.5 * np.dot(np.einsum("ia,ia->i", v, v), m)
    etadto2 = eta + (dt / (2.0 * Q)) * (ke - ((3.0 * N + 1) / 2.0) * kT)
    kedto2 = 0.5 * np.dot(np.einsum("ia,ia->i", vdto2, vdto2), m)
    eta = etadto2 + (dt / (2.0 * Q)) * (kedto2 - ((3.0 * N + 1) / 2.0) * kT)
    v = (vdto2 + (dt / 2.0) * a_new) / (1 + (dt / 2.0) * eta)
    return x, v, a_new, e, eta


# ----------------------------------------------------------------- FIRE
def fire_step(x, v, f, m, state, dtmax, nmin=5, finc=1.1, fdec=0.5, alpha0=0.1, falpha=0.99,
              maxstep=0.1):
    """One FIRE iteration for one molecule (Bitzek, Koskinen, Gahler, Moseler,
    Gumbsch, PRL 97, 170201 (2006)); the reference has no FIRE.  Units as the
    MD drivers: x A, v A/fs, f J/mol/A, m kg/mol.  state = [dt, alpha, n_pos].

    With the power P = F.v: if P > 0 the velocity is mixed towards the force,
    v <- (1 - alpha) v + alpha |v| F/|F| (norms over the whole molecule), and
    after more than nmin such steps in a row dt grows by finc (up to dtmax) and
    alpha shrinks by falpha; otherwise dt shrinks by fdec, alpha resets and v
    is zeroed.  Then v <- v + dt a (a = 1e-10 F/m) and x <- x + dt v with each
    atom's displacement capped at maxstep A.  Returns (x, v, state)."""
    dt, alpha, npos = state
    P = float(np.sum(f * v))
    if P > 0.0:
        fn, vn = float(np.linalg.norm(f)), float(np.linalg.norm(v))
        v = (1.0 - alpha) * v + (alpha * vn / fn if fn > 0.0 else 0.0) * f
        npos += 1
        if npos > nmin:
            dt = min(dt * finc, dtmax)
            alpha *= falpha
    else:
        dt *= fdec
        alpha = alpha0
        npos = 0
        v = np.zeros_like(v)
    v = v + dt * 1e-10 * f / m[:, None]
    dx = dt * v
    n = np.linalg.norm(dx, axis=1)
    scale = np.where(n > maxstep, maxstep / np.where(n > 0, n, 1.0), 1.0)
    return x + dx * scale[:, None], v, [dt, alpha, npos]
