"""Crystal structures from CIF files, the host input of the YAML applications.

The reference reads hosts with ``ase.io.iread(..., store_tags=True)`` and
``particles_from_ase`` (application/utils/particles.py:60-111):

* the cell comes from the ``_cell_length_*`` / ``_cell_angle_*`` items and is
  put in lower-triangular form (a along x, b in the xy plane,
  core/unitcell.py:509-528);
* sites are the ``_atom_site_fract_*`` rows expanded by the symmetry
  operations of ``_symmetry_equiv_pos_as_xyz`` (or
  ``_space_group_symop_operation_xyz``), with images that coincide modulo the
  lattice merged;
* labels are ``_atom_site_label`` (the species key of the force-field
  tables), charges ``_atom_type_partial_charge``, else ``_atom_site_charge``,
  else zero (particles.py:82-90);
* masses are the standard atomic weights of ``_atom_site_type_symbol``.

``make_supercell`` replicates the cell n_a x n_b x n_c times after wrapping the
sites into the cell (application/mcmc/data.py:365-372, core/unitcell.py:477-506).
This is a plain-text parser of the CIF 1.1 subset those files use; it runs no
code from the file.
"""

from __future__ import annotations

import math
import re
import shlex
from dataclasses import dataclass
from fractions import Fraction
from pathlib import Path

import numpy as np

# Standard atomic weights (IUPAC 2013/2016 conventional values, the table ASE
# uses for get_masses), u.
ATOMIC_MASSES = {
    "H": 1.008, "He": 4.002602, "Li": 6.94, "Be": 9.0121831, "B": 10.81, "C": 12.011,
    "N": 14.007, "O": 15.999, "F": 18.998403163, "Ne": 20.1797, "Na": 22.98976928,
    "Mg": 24.305, "Al": 26.9815385, "Si": 28.085, "P": 30.973761998, "S": 32.06,
    "Cl": 35.45, "Ar": 39.948, "K": 39.0983, "Ca": 40.078, "Sc": 44.955908,
    "Ti": 47.867, "V": 50.9415, "Cr": 51.9961, "Mn": 54.938044, "Fe": 55.845,
    "Co": 58.933194, "Ni": 58.6934, "Cu": 63.546, "Zn": 65.38, "Ga": 69.723,
    "Ge": 72.630, "As": 74.921595, "Se": 78.971, "Br": 79.904, "Kr": 83.798,
    "Rb": 85.4678, "Sr": 87.62, "Y": 88.90584, "Zr": 91.224, "Nb": 92.90637,
    "Mo": 95.95, "Ru": 101.07, "Rh": 102.90550, "Pd": 106.42, "Ag": 107.8682,
    "Cd": 112.414, "In": 114.818, "Sn": 118.710, "Sb": 121.760, "Te": 127.60,
    "I": 126.90447, "Xe": 131.293, "Cs": 132.90545196, "Ba": 137.327,
    "La": 138.90547, "Ce": 140.116, "Hf": 178.49, "W": 183.84, "Pt": 195.084,
    "Au": 196.966569, "Hg": 200.592, "Pb": 207.2, "Bi": 208.98040, "U": 238.02891,
}


@dataclass
class Structure:
    """A periodic structure: lower-triangular lattice (rows a, b, c, A),
    fractional coordinates, labels, element symbols, charges (e), masses (u)."""

    lattice: np.ndarray  # (3, 3)
    fractional: np.ndarray  # (n, 3)
    labels: tuple[str, ...]
    symbols: tuple[str, ...]
    charges: np.ndarray  # (n,)
    masses: np.ndarray  # (n,), NaN where the element is unknown

    @property
    def positions(self) -> np.ndarray:
        return self.fractional @ self.lattice

    @property
    def tril(self) -> np.ndarray:
        return self.lattice[[0, 1, 1, 2, 2, 2], [0, 0, 1, 0, 1, 2]]

    @property
    def volume(self) -> float:
        return float(abs(np.linalg.det(self.lattice)))

    def __len__(self) -> int:
        return len(self.labels)


def lower_triangular_cell(lengths, angles_deg) -> np.ndarray:
    """Rows a = (a, 0, 0), b = (b cos g, b sin g, 0), c in the upper half
    space; the lower-triangular form of core/unitcell.py:509-528."""
    a, b, c = (float(x) for x in lengths)
    al, be, ga = (math.radians(float(x)) for x in angles_deg)
    ca, cb, cg, sg = math.cos(al), math.cos(be), math.cos(ga), math.sin(ga)
    # cos/sin of exact right angles round to 0 / 1
    ca, cb, cg = (0.0 if abs(x) < 1e-15 else x for x in (ca, cb, cg))
    cy = (ca - cb * cg) / sg
    cz2 = 1.0 - cb * cb - cy * cy
    if cz2 <= 0:
        raise ValueError(f"cell angles {angles_deg} do not span a volume")
    return np.array([[a, 0.0, 0.0], [b * cg, b * sg, 0.0], [c * cb, c * cy, c * math.sqrt(cz2)]])


# ---------------------------------------------------------------- tokenising
def _strip_su(value: str) -> str:
    """'1.234(5)' -> '1.234' (standard uncertainty)."""
    return value.split("(", 1)[0]


def _number(value: str) -> float:
    v = _strip_su(value)
    if v in (".", "?"):
        return float("nan")
    return float(v)


def _tokens(lines):
    """CIF tokens: tags, values (quotes removed), 'loop_', 'data_*'; semicolon
    text fields become one value."""
    it = iter(lines)
    for raw in it:
        if raw.startswith(";"):
            text = [raw[1:]]
            for nxt in it:
                if nxt.startswith(";"):
                    break
                text.append(nxt)
            yield "value", "\n".join(text).strip()
            continue
        line = raw.strip()
        if not line or line.startswith("#"):
            continue
        lex = shlex.shlex(line, posix=True)
        lex.whitespace_split = True
        lex.commenters = "#"
        try:
            parts = list(lex)
        except ValueError:
            parts = line.split()
        for p in parts:
            low = p.lower()
            if low == "loop_":
                yield "loop", p
            elif low.startswith("data_"):
                yield "data", p
            elif p.startswith("_"):
                yield "tag", low
            else:
                yield "value", p


def parse_cif(text: str) -> tuple[dict[str, str], list[tuple[list[str], list[list[str]]]]]:
    """Items (tag -> value) and loops ([tags], rows) of the first data block."""
    items: dict[str, str] = {}
    loops: list[tuple[list[str], list[list[str]]]] = []
    toks = list(_tokens(text.splitlines()))
    i, seen_data = 0, False
    while i < len(toks):
        kind, val = toks[i]
        if kind == "data":
            if seen_data:
                break
            seen_data = True
            i += 1
        elif kind == "loop":
            i += 1
            tags = []
            while i < len(toks) and toks[i][0] == "tag":
                tags.append(toks[i][1])
                i += 1
            vals = []
            while i < len(toks) and toks[i][0] == "value":
                vals.append(toks[i][1])
                i += 1
            if tags:
                if len(vals) % len(tags):
                    raise ValueError(f"loop over {tags[0]}... has {len(vals)} values, "
                                     f"not a multiple of {len(tags)}")
                rows = [vals[k:k + len(tags)] for k in range(0, len(vals), len(tags))]
                loops.append((tags, rows))
        elif kind == "tag":
            if i + 1 < len(toks) and toks[i + 1][0] == "value":
                items[val] = toks[i + 1][1]
                i += 2
            else:
                items[val] = ""
                i += 1
        else:
            i += 1
    return items, loops


def parse_symop(op: str) -> tuple[np.ndarray, np.ndarray]:
    """'-x+1/2, y, z-1/4' -> (R, t) with x' = R x + t."""
    rows = [s.strip().lower().replace(" ", "") for s in op.split(",")]
    if len(rows) != 3:
        raise ValueError(f"bad symmetry operation {op!r}")
    R = np.zeros((3, 3))
    t = np.zeros(3)
    for r, expr in enumerate(rows):
        for sign, body in re.findall(r"([+-]?)([^+-]+)", expr):
            s = -1.0 if sign == "-" else 1.0
            m = re.fullmatch(r"(\d*\.?\d*(?:/\d+)?)\*?([xyz])", body)
            if m:
                coef = float(Fraction(m.group(1))) if m.group(1) else 1.0
                R[r, "xyz".index(m.group(2))] += s * coef
            else:
                t[r] += s * float(Fraction(body))
    return R, t


def _column(tags, rows, *names):
    for name in names:
        if name in tags:
            k = tags.index(name)
            return [row[k] for row in rows]
    return None


def _element(symbol: str) -> str:
    m = re.match(r"[A-Z][a-z]?", symbol.strip().capitalize())
    return m.group(0) if m else symbol


def read_cif(path: str | Path, symprec: float = 1e-4) -> Structure:
    """Read the first data block of a CIF file (see the module docstring)."""
    items, loops = parse_cif(Path(path).read_text())
    try:
        lengths = [_number(items[f"_cell_length_{x}"]) for x in "abc"]
        angles = [_number(items[f"_cell_angle_{x}"]) for x in ("alpha", "beta", "gamma")]
    except KeyError as e:
        raise ValueError(f"{path}: missing cell item {e}") from None
    lattice = lower_triangular_cell(lengths, angles)

    ops = [(np.eye(3), np.zeros(3))]
    for tags, rows in loops:
        col = _column(tags, rows, "_symmetry_equiv_pos_as_xyz", "_space_group_symop_operation_xyz")
        if col:
            ops = [parse_symop(c) for c in col]
    for key in ("_symmetry_equiv_pos_as_xyz", "_space_group_symop_operation_xyz"):
        if key in items and items[key]:
            ops = [parse_symop(items[key])]

    site = next(((t, r) for t, r in loops if "_atom_site_fract_x" in t), None)
    if site is None:
        raise ValueError(f"{path}: no _atom_site_fract_x loop")
    tags, rows = site
    frac = np.array([[_number(v) for v in c] for c in
                     (_column(tags, rows, f"_atom_site_fract_{x}") for x in "xyz")]).T
    labels = _column(tags, rows, "_atom_site_label")
    types = _column(tags, rows, "_atom_site_type_symbol")
    if types is None:
        types = labels
    if labels is None:
        labels = types
    if labels is None:
        raise ValueError(f"{path}: sites have neither a label nor a type symbol")
    symbols = [_element(t) for t in types]
    q_site = _column(tags, rows, "_atom_site_charge")
    charges = np.array([_number(v) for v in q_site]) if q_site else np.zeros(len(rows))
    for t2, r2 in loops:  # per-type partial charges take precedence
        qt = _column(t2, r2, "_atom_type_partial_charge")
        sym = _column(t2, r2, "_atom_type_symbol")
        if qt and sym:
            table = dict(zip(sym, (_number(v) for v in qt)))
            charges = np.array([table.get(t, q) for t, q in zip(types, charges)])
    charges = np.nan_to_num(charges)

    # symmetry expansion with coincident images merged
    out_f, out_i = [], []
    for i in range(len(frac)):
        seen: list[np.ndarray] = []
        for R, t in ops:
            f = (R @ frac[i] + t) % 1.0
            if any(np.all(np.abs((f - g + 0.5) % 1.0 - 0.5) < symprec) for g in seen):
                continue
            seen.append(f)
            out_f.append(f)
            out_i.append(i)
    fr = np.array(out_f).reshape(-1, 3)
    if len(ops) == 1 and np.allclose(ops[0][0], np.eye(3)) and not ops[0][1].any():
        fr = frac.copy()  # P1: keep the file's coordinates unwrapped
    masses = np.array([ATOMIC_MASSES.get(symbols[i], float("nan")) for i in out_i])
    return Structure(lattice, fr, tuple(labels[i] for i in out_i), tuple(symbols[i] for i in out_i),
                     charges[out_i].astype(np.float64), masses)


def wrap(structure: Structure) -> Structure:
    """Sites moved into [0, 1) fractional."""
    return Structure(structure.lattice, structure.fractional % 1.0, structure.labels,
                     structure.symbols, structure.charges, structure.masses)


def make_supercell(structure: Structure, replication) -> Structure:
    """n_a x n_b x n_c copies (an int means the same on every axis), sites
    wrapped first; copies are ordered image-major (all sites of image
    (0,0,0), then (0,0,1), ...) like core/unitcell.py:477-506."""
    rep = (replication,) * 3 if isinstance(replication, int) else tuple(int(r) for r in replication)
    if any(r < 1 for r in rep):
        raise ValueError(f"replication {rep} must be >= 1 on every axis")
    s = wrap(structure)
    grid = np.stack(np.meshgrid(*(np.arange(r) for r in rep), indexing="ij"), -1).reshape(-1, 3)
    frac = ((s.fractional[None, :, :] + grid[:, None, :]) / np.asarray(rep, float)).reshape(-1, 3)
    k = len(grid)
    return Structure(s.lattice * np.asarray(rep, float)[:, None], frac, s.labels * k,
                     s.symbols * k, np.tile(s.charges, k), np.tile(s.masses, k))
