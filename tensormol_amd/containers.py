"""Molecule containers of the drop-in API: Mol and MSet.

Mirrors the parts of TensorMol/Containers/Mol.py (Mol(atoms, coords),
NAtoms, properties, XYZ text I/O: Mol.py:7-23, 300-360) and
TensorMol/Containers/Sets.py (MSet.mols, ReadXYZ, MaxNAtoms, AtomTypes:
Sets.py:17-235) that the BP+EE scripts use.
"""
import os

import numpy as np

from .params import PARAMS

_SYMBOLS = {"H": 1, "He": 2, "Li": 3, "Be": 4, "B": 5, "C": 6, "N": 7, "O": 8, "F": 9, "Ne": 10,
            "Na": 11, "Mg": 12, "Al": 13, "Si": 14, "P": 15, "S": 16, "Cl": 17, "Ar": 18}
_NAMES = {v: k for k, v in _SYMBOLS.items()}


class Mol:
    """atoms: atomic numbers (uint8), coords: float64 (N, 3) in A."""

    def __init__(self, atoms_=None, coords_=None):
        self.atoms = np.zeros(0, dtype=np.uint8) if atoms_ is None else np.asarray(atoms_, dtype=np.uint8).copy()
        self.coords = np.zeros((0, 3)) if coords_ is None else np.asarray(coords_, dtype=np.float64).copy()
        self.properties = {}
        self.name = None

    def NAtoms(self):
        return self.atoms.shape[0]

    def AtomTypes(self):
        return np.unique(self.atoms)

    def Center(self, CenterOf="Atom", MomentOrder=1.0):
        """Mean of coords**MomentOrder over atoms, or its mass-weighted mean
        for CenterOf="Mass" (Mol.py:394-407)."""
        c = np.power(self.coords, MomentOrder)
        if CenterOf == "Mass":
            from .params import atomic_masses
            w = atomic_masses(self.atoms)
            return (w[:, None] * c).sum(0) / w.sum()
        return c.mean(0)

    @staticmethod
    def ParseProperties(s_):
        """Comment line 'Comment: ;;;energy E;;;...' -> {'energy': E}
        (Mol.py:268-285: only the lowercase 'energy' property is read here;
        the reference's binary 'Lattice' field is not portable text)."""
        props = {}
        if "Comment:" not in s_:
            return props
        for prop in s_.split("Comment:", 1)[1].split(";;;"):
            f = prop.split()
            if len(f) >= 2 and f[0] == "energy":
                props["energy"] = float(f[1])
        return props

    def PropertyString(self):
        """';;;name value' for every property (Mol.py:286-299)."""
        return "".join(f";;;{k} {v}" for k, v in self.properties.items())

    def FromXYZString(self, string):
        """XYZ block -> atoms, coords, properties (Mol.py:300-330)."""
        lines = string.split("\n")
        natoms = int(lines[0].split()[0])
        if len(lines[1].split()) > 1:
            try:
                self.properties = self.ParseProperties(lines[1])
            except ValueError:
                pass
        atoms, coords = [], []
        for line in lines[2:2 + natoms]:
            f = line.split()
            sym = f[0]
            atoms.append(int(sym) if sym.isdigit() else _SYMBOLS[sym])
            coords.append([float(f[1]), float(f[2]), float(f[3])])
        self.atoms = np.array(atoms, dtype=np.uint8)
        self.coords = np.array(coords, dtype=np.float64)

    def XYZString(self, comment="Comment: "):
        out = [str(self.NAtoms()), comment]
        for z, c in zip(self.atoms, self.coords):
            out.append(f"{_NAMES.get(int(z), str(int(z)))} {c[0]:.10f} {c[1]:.10f} {c[2]:.10f}")
        return "\n".join(out) + "\n"

    def WriteXYZfile(self, fpath=".", fname="mol", mode="a", wprop=False):
        os.makedirs(fpath, exist_ok=True)
        comment = "Comment: " + (self.PropertyString() if wprop else "")
        with open(os.path.join(fpath, fname + ".xyz"), mode) as f:
            f.write(self.XYZString(comment))

    def __str__(self):
        return self.XYZString()


class MSet:
    """A named set of molecules."""

    def __init__(self, name_="NaClH2O", path_=None, center_=True):
        self.name = name_
        self.path = PARAMS["sets_dir"] if path_ is None else path_
        self.mols = []
        self.center = center_

    def ReadXYZ(self, filename=None, xyz_type="mol"):
        fn = os.path.join(self.path, (filename or self.name) + ".xyz")
        txt = open(fn).read().split("\n")
        i = 0
        while i < len(txt):
            line = txt[i].strip()
            if not line:
                i += 1
                continue
            n = int(line.split()[0])
            m = Mol()
            m.FromXYZString("\n".join(txt[i:i + n + 2]))
            if self.center:
                m.coords = m.coords - m.coords.mean(0)
            self.mols.append(m)
            i += n + 2

    def MaxNAtoms(self):
        return max(m.NAtoms() for m in self.mols) if self.mols else 0

    def AtomTypes(self):
        return np.unique(np.concatenate([m.atoms for m in self.mols])) if self.mols else np.zeros(0)

    def NAtoms(self):
        return sum(m.NAtoms() for m in self.mols)
