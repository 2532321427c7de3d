"""Seeded synthetic geometries for the BP+EE benchmark systems.

Host-side data generation only (numpy): rigid-water clusters and boxes, a
random-sequential-addition C/H/N/O sphere, and random-walk C/H/N/O chain
molecules.  Each generator returns (Z int32 (N,), xyz float64 (N,3) in A[, box]).
"""
import math

import numpy as np

OH = 0.96
HOH = math.radians(104.5)


def random_rotation(rng):
    """Uniform random rotation matrix from a unit quaternion (Shoemake 1992)."""
    u1, u2, u3 = rng.random(3)
    q = np.array([math.sqrt(1 - u1) * math.sin(2 * math.pi * u2),
                  math.sqrt(1 - u1) * math.cos(2 * math.pi * u2),
                  math.sqrt(u1) * math.sin(2 * math.pi * u3),
                  math.sqrt(u1) * math.cos(2 * math.pi * u3)])
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def _water(rng, o):
    h = OH * np.array([[math.sin(HOH / 2), math.cos(HOH / 2), 0.0],
                       [-math.sin(HOH / 2), math.cos(HOH / 2), 0.0]])
    R = random_rotation(rng)
    return np.vstack([o, o + h @ R.T])


def water_cluster(nwater=64, box=12.4, min_oo=2.6, seed=0):
    """Open cluster: O uniform in a cube with minimum O-O distance (rejection
    sampling), random orientations.  Atom order O, H, H per water."""
    rng = np.random.default_rng(seed)
    os_ = []
    while len(os_) < nwater:
        c = rng.random(3) * box
        if all(np.linalg.norm(c - o) >= min_oo for o in os_):
            os_.append(c)
    xyz = np.vstack([_water(rng, o) for o in os_])
    Z = np.tile(np.array([8, 1, 1], dtype=np.int32), nwater)
    return Z, xyz


def water_box(nside=16, edge=49.7, jitter=0.3, seed=1):
    """Periodic box of nside^3 rigid waters on a jittered cubic lattice.
    Positions are wrapped into [0, edge)."""
    rng = np.random.default_rng(seed)
    a = edge / nside
    xyz = []
    for ix in range(nside):
        for iy in range(nside):
            for iz in range(nside):
                o = (np.array([ix, iy, iz]) + 0.5) * a + rng.uniform(-jitter, jitter, 3)
                xyz.append(_water(rng, o))
    xyz = np.mod(np.vstack(xyz), edge)
    Z = np.tile(np.array([8, 1, 1], dtype=np.int32), nside ** 3)
    return Z, xyz, np.array([edge, edge, edge])


def chno_sphere(natom=24000, density=0.10, min_dist=1.0, seed=2,
                fractions=((1, 0.5), (6, 0.3), (7, 0.1), (8, 0.1))):
    """Random sequential addition in a sphere at the given number density."""
    rng = np.random.default_rng(seed)
    R = (3.0 * natom / (4.0 * math.pi * density)) ** (1.0 / 3.0)
    cell = min_dist
    ncell = int(math.ceil(2 * R / cell)) + 2
    grid = {}
    pts = []
    while len(pts) < natom:
        c = rng.uniform(-R, R, 3)
        if c @ c > R * R:
            continue
        key = tuple(((c + R) // cell).astype(int))
        ok = True
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for q in grid.get((key[0] + dx, key[1] + dy, key[2] + dz), ()):
                        d = c - pts[q]
                        if d @ d < min_dist * min_dist:
                            ok = False
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if not ok:
                break
        if ok:
            grid.setdefault(key, []).append(len(pts))
            pts.append(c)
    del ncell
    counts = [int(round(f * natom)) for _, f in fractions]
    counts[0] += natom - sum(counts)
    Z = np.concatenate([np.full(c, z, dtype=np.int32) for (z, _), c in zip(fractions, counts)])
    rng.shuffle(Z)
    return Z, np.array(pts)


def _unit(rng):
    v = rng.normal(size=3)
    return v / np.linalg.norm(v)


def _cross(a, b):
    return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])


def _unitv(a):
    n = math.sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2])
    return (a[0] / n, a[1] / n, a[2] / n), n


def _bond_dirs(rng, prev, prevprev, k):
    """k unit vectors for the next bond: bond angle U(105, 125) degrees to the
    previous bond and a uniform dihedral (uniform directions for the first bond)."""
    if prev is None:
        v = rng.normal(size=(k, 3))
        return v / np.linalg.norm(v, axis=1, keepdims=True)
    b1, _ = _unitv(tuple(prev))
    ref = tuple(prevprev) if prevprev is not None else tuple(rng.normal(size=3))
    n, nn = _unitv(_cross(b1, ref)) if np.linalg.norm(_cross(b1, ref)) > 1e-8 else \
        _unitv(_cross(b1, tuple(rng.normal(size=3))))
    m = _cross(n, b1)
    th = np.radians(rng.uniform(105.0, 125.0, k))[:, None]
    ph = rng.uniform(0.0, 2.0 * math.pi, k)[:, None]
    B1, Nn, M = np.array(b1), np.array(n), np.array(m)
    return -np.cos(th) * B1 + np.sin(th) * (np.cos(ph) * Nn + np.sin(ph) * M)


def chain_molecule(natom, rng, heavy=((6, 0.7), (7, 0.15), (8, 0.15)), min_nb=1.8, min_13=1.5,
                   tries=4):
    """Heavy-atom chain (bonds U(1.1, 1.5) A, bond angles U(105, 125) deg,
    random dihedrals) with pendant H at 1.09 A.  Non-bonded pairs (neither 1-2
    nor 1-3 neighbours) stay at least min_nb apart, 1-3 pairs at least min_13.
    Half the atoms (rounded up) are heavy; H are spread over the heavy atoms
    round robin.  Each placement tests `tries` candidates at once and keeps the
    first acceptable one."""
    nheavy = max(1, (natom + 1) // 2)
    nh = natom - nheavy
    for _attempt in range(200):
        pos = np.zeros((natom, 3))
        n = 1
        nbrs = [set()]

        def place(cands, host):
            nonlocal n
            d = cands[:, None, :] - pos[None, :n, :]
            d2 = (d * d).sum(-1)
            lim = np.full(n, min_nb * min_nb)
            lim[host] = 0.0
            lim[list(nbrs[host])] = min_13 * min_13
            ok = (d2 >= lim).all(axis=1)
            good = np.flatnonzero(ok)
            if len(good) == 0:
                return False
            pos[n] = cands[good[0]]
            nbrs.append({host})
            nbrs[host].add(n)
            n += 1
            return True

        failed = False
        for i in range(1, nheavy):
            prev = pos[i - 1] - pos[i - 2] if i >= 2 else None
            pp = pos[i - 2] - pos[i - 3] if i >= 3 else None
            placed = False
            for _ in range(32):
                c = pos[i - 1] + rng.uniform(1.1, 1.5, (tries, 1)) * _bond_dirs(rng, prev, pp, tries)
                if place(c, i - 1):
                    placed = True
                    break
            if not placed:
                failed = True
                break
        if failed:
            continue
        for h in range(nh):
            host = h % nheavy
            placed = False
            for _ in range(64):
                v = rng.normal(size=(tries, 3))
                c = pos[host] + 1.09 * v / np.linalg.norm(v, axis=1, keepdims=True)
                if place(c, host):
                    placed = True
                    break
            if not placed:
                failed = True
                break
        if failed:
            continue
        zs = [heavy[int(np.searchsorted(np.cumsum([f for _, f in heavy]), rng.random()))][0]
              for _ in range(nheavy)]
        Z = np.array(zs + [1] * nh, dtype=np.int32)
        return Z, pos[:n].copy()
    raise RuntimeError("chain_molecule: could not place atoms")


def request_sizes(n, lo=8, hi=256, seed=3):
    """Atom counts log-uniform in [lo, hi]."""
    rng = np.random.default_rng(seed)
    return np.floor(np.exp(rng.uniform(math.log(lo), math.log(hi + 1), n))).astype(int).clip(lo, hi)


def chain_set(sizes, seed=3):
    """A list of (Z, xyz) chain molecules of the given atom counts."""
    rng = np.random.default_rng(seed + 1000)
    return [chain_molecule(int(n), rng) for n in sizes]


def _chain_job(job):
    size, seed = job
    return chain_molecule(int(size), np.random.default_rng(seed))


def chain_molecules(jobs, workers=1):
    """chain_molecule for each (size, seed) job, each from its own generator,
    so the result does not depend on `workers` (spawned processes that import
    numpy and this module only)."""
    if workers > 1 and len(jobs) >= 64:
        import multiprocessing as mp
        with mp.get_context("spawn").Pool(workers) as pool:
            return pool.map(_chain_job, jobs, chunksize=max(1, len(jobs) // (8 * workers)))
    return [_chain_job(j) for j in jobs]


STREAM_KINDS = ((0.5, "sp", 1), (0.3, "nvt", 50), (0.2, "opt", 100))


def request_stream(n=4096, rate=2000.0, lo=8, hi=256, seed=3, kinds=STREAM_KINDS, workers=None):
    """Seeded serving stream: n requests, atom counts log-uniform in [lo, hi],
    random-walk C/H/N/O chains (each from its own seed), kinds (probability,
    name, steps) drawn with the given probabilities, Poisson arrivals at `rate`
    per second.  Returns a list of (Z, xyz, kind, steps, t_arrival, seed).
    Geometries are generated by `workers` spawned processes (default: the CPU
    count, at most 16); call this before the process initialises a GPU."""
    import os
    rng = np.random.default_rng(seed)
    sizes = np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi + 1), n))).astype(int).clip(lo, hi)
    gaps = rng.exponential(1.0 / rate, n)
    t = np.cumsum(gaps) - gaps[0]
    probs = np.array([k[0] for k in kinds])
    choice = rng.choice(len(kinds), size=n, p=probs / probs.sum())
    workers = workers if workers is not None else min(16, os.cpu_count() or 1)
    geos = chain_molecules([(int(sizes[r]), [seed, 1000, r]) for r in range(n)], workers)
    return [(geos[r][0], geos[r][1], kinds[choice[r]][1], kinds[choice[r]][2], float(t[r]),
             seed * 100000 + r) for r in range(n)]


def pack(mols):
    """Concatenate molecules into packed arrays: Z (N,), xyz (N,3), mol_start (B+1,)."""
    Z = np.concatenate([m[0] for m in mols]).astype(np.int32)
    xyz = np.concatenate([m[1] for m in mols]).astype(np.float64)
    ms = np.zeros(len(mols) + 1, dtype=np.int64)
    ms[1:] = np.cumsum([len(m[0]) for m in mols])
    return Z, xyz, ms


def maxwell_boltzmann(Z, T, seed, masses_kg_per_mol):
    """Velocities (A/fs) drawn from Maxwell-Boltzmann at T with zero net momentum.
    m in kg/mol: <v^2> per component = R T / m (m^2/s^2) = 1e-10 R T / m (A^2/fs^2)."""
    rng = np.random.default_rng(seed)
    R = 8.3144621
    sig = np.sqrt(1e-10 * R * T / masses_kg_per_mol)
    v = rng.normal(size=(len(Z), 3)) * sig[:, None]
    p = (v * masses_kg_per_mol[:, None]).sum(0) / masses_kg_per_mol.sum()
    return v - p
