"""Seeded synthetic proteins and templates for the benchmark workloads.

Proteins: a CA random walk with 3.8 A steps, self-avoiding at 4 A (rejection
resampling of the step direction), residue types uniform over the 20 amino
acids, each residue carrying its real PDB atom names (backbone N, CA, C, O and
the side chain of its type, 0-10 atoms); every atom after CA sits 1.5 A from
the atom it is bonded to. Coordinates are rounded to 3 decimals, as in PDB
files, so equal coordinates occur as they do in real data.

Templates: ``planted`` templates copy atoms of spatially close residues of a
molecule and jitter them with N(0, sigma); ``random`` ones draw residue types
and geometry. Match modes are 0, 1 and 3: mode 3 (any atom of the same
element) on carbons, mode 1 (same branch position, not carbon) on the
swappable carboxylate/amide/guanidinium atoms, mode 0 (exact name) otherwise.
"""

from __future__ import annotations

import math
from typing import List, Optional, Sequence, Tuple

import numpy as np

from ._model import Molecule, Template, TemplateAtom

# side-chain atoms per residue type, in PDB order, each bonded to the previous
# one listed or, for branches, to the atom named in _PARENT
SIDECHAINS = {
    "ALA": ["CB"],
    "ARG": ["CB", "CG", "CD", "NE", "CZ", "NH1", "NH2"],
    "ASN": ["CB", "CG", "OD1", "ND2"],
    "ASP": ["CB", "CG", "OD1", "OD2"],
    "CYS": ["CB", "SG"],
    "GLN": ["CB", "CG", "CD", "OE1", "NE2"],
    "GLU": ["CB", "CG", "CD", "OE1", "OE2"],
    "GLY": [],
    "HIS": ["CB", "CG", "ND1", "CD2", "CE1", "NE2"],
    "ILE": ["CB", "CG1", "CG2", "CD1"],
    "LEU": ["CB", "CG", "CD1", "CD2"],
    "LYS": ["CB", "CG", "CD", "CE", "NZ"],
    "MET": ["CB", "CG", "SD", "CE"],
    "PHE": ["CB", "CG", "CD1", "CD2", "CE1", "CE2", "CZ"],
    "PRO": ["CB", "CG", "CD"],
    "SER": ["CB", "OG"],
    "THR": ["CB", "OG1", "CG2"],
    "TRP": ["CB", "CG", "CD1", "CD2", "NE1", "CE2", "CE3", "CZ2", "CZ3", "CH2"],
    "TYR": ["CB", "CG", "CD1", "CD2", "CE1", "CE2", "CZ", "OH"],
    "VAL": ["CB", "CG1", "CG2"],
}
RESIDUES = sorted(SIDECHAINS)
# branch atoms bonded to an atom other than the previous one
_PARENT = {
    ("ARG", "NH2"): "CZ", ("ASN", "ND2"): "CG", ("ASP", "OD2"): "CG", ("GLN", "NE2"): "CD",
    ("GLU", "OE2"): "CD", ("HIS", "CD2"): "CG", ("HIS", "NE2"): "CD2", ("ILE", "CG2"): "CB",
    ("LEU", "CD2"): "CG", ("PHE", "CD2"): "CG", ("PHE", "CE2"): "CD2", ("PHE", "CZ"): "CE1",
    ("THR", "CG2"): "CB", ("TRP", "CD2"): "CG", ("TRP", "CE2"): "CD2", ("TRP", "CE3"): "CD2",
    ("TRP", "CZ3"): "CE3", ("TYR", "CD2"): "CG", ("TYR", "CE2"): "CD2", ("TYR", "CZ"): "CE1",
    ("VAL", "CG2"): "CB",
}
# atoms that a symmetric partner can replace (mode 1 in templates)
SYMMETRIC = {("ASP", "OD1"), ("ASP", "OD2"), ("GLU", "OE1"), ("GLU", "OE2"), ("ARG", "NH1"),
             ("ARG", "NH2"), ("ASN", "OD1"), ("ASN", "ND2"), ("GLN", "OE1"), ("GLN", "NE2")}

_SYM_TYPES = sorted({rn for rn, _ in SYMMETRIC})

CHAIN_IDS = "ABCDEFGHIJKLMNOPQRSTUVWXYZ"


def _unit(rng: np.random.Generator, n: int) -> np.ndarray:
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def ca_walk(rng: np.random.Generator, n: int, step: float = 3.8, avoid: float = 4.0,
            tries: int = 32) -> np.ndarray:
    """Self-avoiding CA walk: each step redraws its direction (up to ``tries``
    times, then keeps the last draw) until no earlier CA is closer than
    ``avoid``; the neighbour search uses a hash grid of cell ``avoid``."""
    pos = np.zeros((n, 3))
    grid = {}
    inv = 1.0 / avoid
    a2 = avoid * avoid

    def cell(p):
        return (int(math.floor(p[0] * inv)), int(math.floor(p[1] * inv)), int(math.floor(p[2] * inv)))

    def clash(p, last):
        cx, cy, cz = cell(p)
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for j in grid.get((cx + dx, cy + dy, cz + dz), ()):
                        if j == last:
                            continue
                        q = pos[j]
                        d0, d1, d2 = p[0] - q[0], p[1] - q[1], p[2] - q[2]
                        if d0 * d0 + d1 * d1 + d2 * d2 < a2:
                            return True
        return False

    grid.setdefault(cell(pos[0]), []).append(0)
    dirs = _unit(rng, n * 4)
    k = 0
    for i in range(1, n):
        cand = None
        for _ in range(tries):
            if k >= len(dirs):
                dirs = _unit(rng, n * 4)
                k = 0
            cand = pos[i - 1] + step * dirs[k]
            k += 1
            if not clash(cand, i - 1):
                break
        pos[i] = cand
        grid.setdefault(cell(cand), []).append(i)
    return pos


def _residue_topology():
    """Per residue type: atom names in PDB order, parent (local index, -1 for
    CA) and depth in the bond tree rooted at CA."""
    topo = {}
    for rn in RESIDUES:
        names = ["N", "CA", "C", "O"] + SIDECHAINS[rn]
        parent = []
        prev = "CA"
        for nm in names:
            if nm == "CA":
                parent.append(-1)
                continue
            if nm in ("N", "C", "CB"):
                par = "CA"
            elif nm == "O":
                par = "C"
            else:
                par = _PARENT.get((rn, nm), prev)
            if nm not in ("N", "C", "O"):
                prev = nm
            parent.append(names.index(par))
        depth = []
        for i in range(len(names)):
            d, j = 0, i
            while parent[j] >= 0:
                j = parent[j]
                d += 1
            depth.append(d)
        topo[rn] = (names, parent, depth)
    return topo


_TOPO = _residue_topology()


def protein(rng: np.random.Generator, n_residues: int, n_chains: int = 1, mol_id: str = "SYN1",
            ca: Optional[np.ndarray] = None) -> Molecule:
    """A synthetic protein of ``n_residues`` residues split into ``n_chains`` chains
    (residue numbers restart at 1 in every chain)."""
    if ca is None:
        ca = ca_walk(rng, n_residues)
    types = rng.integers(0, len(RESIDUES), size=n_residues)
    per_chain = int(math.ceil(n_residues / n_chains))
    counts = np.array([len(_TOPO[RESIDUES[t]][0]) for t in range(len(RESIDUES))])[types]
    first = np.concatenate([[0], np.cumsum(counts)[:-1]])
    n = int(counts.sum())
    res_of = np.repeat(np.arange(n_residues), counts)
    tlist = [RESIDUES[t] for t in types]
    names = [nm for rn in tlist for nm in _TOPO[rn][0]]
    par_flat = np.concatenate([np.array(_TOPO[rn][1]) for rn in tlist])
    dep_flat = np.concatenate([np.array(_TOPO[rn][2]) for rn in tlist])
    parent = np.where(par_flat >= 0, first[res_of] + par_flat, -1)
    xyz = np.zeros((n, 3))
    is_ca = par_flat < 0
    xyz[is_ca] = ca[res_of[is_ca]]
    bonds = _unit(rng, n)
    for d in range(1, int(dep_flat.max()) + 1):
        sel = np.nonzero(dep_flat == d)[0]
        xyz[sel] = xyz[parent[sel]] + 1.5 * bonds[sel]
    xyz = np.round(xyz, 3)
    chain_idx = np.minimum(res_of // per_chain, len(CHAIN_IDS) - 1)
    resseq = (res_of % per_chain + 1).astype(np.int32)
    enc_names = np.frombuffer(b"".join(_enc_name(nm) for nm in names), dtype=np.uint8).reshape(n, 4)
    res_codes = np.frombuffer(b"".join(rn.encode() + b"\0" for rn in RESIDUES), dtype=np.uint8).reshape(-1, 4)
    enc_res = res_codes[types[res_of]]
    ch = np.zeros((n, 2), dtype=np.uint8)
    ch[:, 0] = ord(" ")
    ch[:, 1] = np.frombuffer(CHAIN_IDS.encode(), dtype=np.uint8)[chain_idx]
    element = np.zeros((n, 2), dtype=np.uint8)
    element[:, 0] = ord("_")
    element[:, 1] = enc_names[:, 1]
    return Molecule.from_arrays(xyz, enc_names, np.ascontiguousarray(enc_res), ch, resseq,
                                np.arange(1, n + 1, dtype=np.int32), id=mol_id, element=element)


def _enc_name(s: str) -> bytes:
    raw = s.encode("ascii")
    if len(raw) < 4:
        raw = b"_" + raw
    return raw.ljust(4, b"_")


def protein_with_atoms(rng: np.random.Generator, n_atoms: int, n_chains: int = 1,
                       mol_id: str = "SYN1") -> Molecule:
    """A protein with close to ``n_atoms`` atoms (about 8.0 atoms per residue)."""
    return protein(rng, max(3, int(round(n_atoms / 8.05))), n_chains=n_chains, mol_id=mol_id)


# ---------------------------------------------------------------------------
# templates
# ---------------------------------------------------------------------------

def _mode_for(resname: str, name: str, rng: np.random.Generator, p_flexible: float) -> int:
    if rng.random() >= p_flexible:
        return 0
    if (resname, name) in SYMMETRIC:
        return 1
    if name.startswith("C"):
        return 3
    return 0


def _residue_atoms(mol: Molecule, first: int) -> List[int]:
    """Atom indices of the residue starting at atom ``first``."""
    key = (mol.chains[first, 1], mol.resseq[first])
    out = [first]
    i = first + 1
    while i < len(mol) and (mol.chains[i, 1], mol.resseq[i]) == key:
        out.append(i)
        i += 1
    return out


def _residue_starts(mol: Molecule) -> np.ndarray:
    same = (mol.chains[1:, 1] == mol.chains[:-1, 1]) & (mol.resseq[1:] == mol.resseq[:-1])
    return np.concatenate([[0], np.nonzero(~same)[0] + 1])


def planted_template(rng: np.random.Generator, mol: Molecule, n_residues: int, atoms_per_residue,
                     jitter: float = 0.3, radius: float = 10.0, p_flexible: float = 0.3,
                     tpl_id: str = "PLNT", starts: Optional[np.ndarray] = None,
                     p_symmetric: float = 0.0) -> Template:
    """Atoms of ``n_residues`` residues whose CA lie within ``radius`` of a random
    seed residue's CA, copied from ``mol`` with N(0, jitter) noise.

    With ``p_symmetric`` > 0, that fraction of the residues (on average) carry
    both atoms of a symmetric pair (carboxylate O, amide O/N, guanidinium N) in
    match mode 1, so each matches either atom of the pair: a residue of a type
    that has such a pair is made symmetric with probability
    ``p_symmetric / 0.25`` (5 of the 20 types have one)."""
    if starts is None:
        starts = _residue_starts(mol)
    ca_idx = np.array([s + 1 for s in starts])  # N, CA, ...: CA is the 2nd atom
    cas = mol.xyz[ca_idx]
    for _ in range(100):
        seed = int(rng.integers(len(starts)))
        d = np.linalg.norm(cas - cas[seed], axis=1)
        near = np.nonzero(d < radius)[0]
        near = near[near != seed]
        if len(near) >= n_residues - 1:
            break
    chosen = [seed] + list(rng.choice(near, size=min(len(near), n_residues - 1), replace=False))
    atoms = []
    p_sym_type = min(1.0, p_symmetric / 0.25)
    for r in chosen:
        res_atoms = _residue_atoms(mol, int(starts[r]))
        k = int(atoms_per_residue() if callable(atoms_per_residue) else atoms_per_residue)
        rn = bytes(mol.resnames[res_atoms[0]][:3]).decode()
        pair = [a for a in res_atoms if (rn, bytes(mol.names[a]).decode().strip("_")) in SYMMETRIC]
        if p_symmetric > 0.0 and pair and rng.random() < p_sym_type:
            # the symmetric pair, then other atoms of the residue
            rest = [a for a in res_atoms if a not in pair]
            extra = sorted(rng.choice(len(rest), size=min(max(k - 2, 0), len(rest)), replace=False))
            picked = sorted(pair + [rest[j] for j in extra])
            sym = True
        else:
            # side-chain atoms first (they carry the chemistry), then backbone
            side = [a for a in res_atoms if bytes(mol.names[a]) not in (b"_N__", b"_C__", b"_O__")]
            pool = side if len(side) >= k else res_atoms
            picked = [pool[j] for j in sorted(rng.choice(len(pool), size=min(k, len(pool)), replace=False))]
            sym = False
        for a in picked:
            nm = bytes(mol.names[a]).decode().strip("_")
            x, y, z = (mol.xyz[a] + rng.normal(scale=jitter, size=3)).tolist()
            mode = 1 if sym and a in pair else _mode_for(rn, nm, rng, p_flexible)
            atoms.append(TemplateAtom(
                chain_id=chr(mol.chains[a, 1]), residue_number=int(mol.resseq[a]),
                x=round(x, 3), y=round(y, 3), z=round(z, 3), residue_names=[rn], atom_names=[nm],
                distance_weight=0.0, match_mode=mode))
    return Template(atoms, id=tpl_id)


def random_template(rng: np.random.Generator, n_residues: int, atoms_per_residue, radius: float = 8.0,
                    p_flexible: float = 0.3, tpl_id: str = "RAND", p_symmetric: float = 0.0) -> Template:
    """Residue types drawn uniformly, residue centres uniform in a ball of
    ``radius``, atoms 1.5 A steps around their centre. A fraction
    ``p_symmetric`` of the residues is drawn from the types with a symmetric
    atom pair instead, and carries that pair in match mode 1."""
    atoms = []
    for r in range(n_residues):
        sym = p_symmetric > 0.0 and rng.random() < p_symmetric
        if sym:
            rn = _SYM_TYPES[int(rng.integers(len(_SYM_TYPES)))]
        else:
            rn = RESIDUES[int(rng.integers(len(RESIDUES)))]
        names = ["CA"] + SIDECHAINS[rn]
        centre = _unit(rng, 1)[0] * radius * rng.random() ** (1 / 3)
        k = int(atoms_per_residue() if callable(atoms_per_residue) else atoms_per_residue)
        if sym:
            pair = [j for j, nm in enumerate(names) if (rn, nm) in SYMMETRIC]
            rest = [j for j in range(len(names)) if j not in pair]
            extra = rng.choice(len(rest), size=min(max(k - 2, 0), len(rest)), replace=False)
            pick = sorted(pair + [rest[j] for j in extra])
        else:
            pair = []
            pick = sorted(rng.choice(len(names), size=min(k, len(names)), replace=False))
        p = centre.copy()
        for j in pick:
            p = p + 1.5 * _unit(rng, 1)[0]
            nm = names[j]
            mode = 1 if j in pair else _mode_for(rn, nm, rng, p_flexible)
            atoms.append(TemplateAtom(
                chain_id="A", residue_number=100 + r, x=round(p[0], 3), y=round(p[1], 3),
                z=round(p[2], 3), residue_names=[rn], atom_names=[nm], distance_weight=0.0,
                match_mode=mode))
    return Template(atoms, id=tpl_id)


def template_set(rng: np.random.Generator, mols: Sequence[Molecule], n_templates: int,
                 planted_fraction: float, residues: Tuple[int, int], atoms_per_residue: Tuple[int, int],
                 jitter: float = 0.3, p_flexible: float = 0.3, p_symmetric: float = 0.0) -> List[Template]:
    """``n_templates`` templates, a fraction planted in random molecules of ``mols``."""
    out = []
    starts = [_residue_starts(m) for m in mols]
    n_planted = int(round(planted_fraction * n_templates))
    for i in range(n_templates):
        nres = int(rng.integers(residues[0], residues[1] + 1))
        apr = lambda: int(rng.integers(atoms_per_residue[0], atoms_per_residue[1] + 1))
        if i < n_planted:
            mi = int(rng.integers(len(mols)))
            out.append(planted_template(rng, mols[mi], nres, apr, jitter=jitter, p_flexible=p_flexible,
                                        tpl_id=f"P{i:05d}", starts=starts[mi], p_symmetric=p_symmetric))
        else:
            out.append(random_template(rng, nres, apr, p_flexible=p_flexible, tpl_id=f"R{i:05d}",
                                       p_symmetric=p_symmetric))
    order = rng.permutation(len(out))
    return [out[i] for i in order]


# ---------------------------------------------------------------------------
# workloads
# ---------------------------------------------------------------------------

def workload(name: str, seed: int = 0, scale: float = 1.0):
    """(molecules, templates, query kwargs) of a named workload.

    ``plain-2k``   1 protein of 250 residues (~2,000 atoms) x 10 templates of
                   3 residues / 6 atoms, 3 planted.
    ``plain-8k``   1 protein of 8,000 atoms x 500 templates of 3-6 residues
                   x 2-3 atoms, 30% planted.
    ``batch-4k``   4,096 proteins, lognormal sizes (median 4,000 atoms, clipped
                   to 1,000-40,000) x 200 templates of 3-6 residues, 10% planted.
    ``library-64`` 64 assemblies of 100,000 atoms in 20 chains (residue numbers
                   restarting per chain) x 7,000 templates of 2-8 residues x
                   2-3 atoms, 10% planted, 20% of template residues carrying a
                   symmetric atom pair in mode 1; distance_cutoff =
                   max_dynamic_distance = 2.5, max_candidates 100,000.
    ``scale`` shrinks the molecule count of the batch and library workloads
    (tests). Unless stated, rmsd_threshold 2.0, distance_cutoff =
    max_dynamic_distance = 1.5, max_candidates 10,000.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    kw = dict(rmsd_threshold=2.0, distance_cutoff=1.5, max_dynamic_distance=1.5, max_candidates=10000)
    if name == "plain-2k":
        mol = protein(rng, 250, mol_id="SYN2")
        tpls = template_set(rng, [mol], 10, 0.3, (3, 3), (2, 2))
        return [mol], tpls, kw
    if name == "plain-8k":
        mol = protein_with_atoms(rng, 8000, mol_id="SYN8")
        tpls = template_set(rng, [mol], 500, 0.3, (3, 6), (2, 3))
        return [mol], tpls, kw
    if name == "batch-4k":
        n = max(1, int(round(4096 * scale)))
        sizes = np.clip(np.round(np.exp(rng.normal(math.log(4000), 0.6, size=n))), 1000, 40000).astype(int)
        mols = _batch_molecules(rng, sizes)
        tpls = template_set(rng, mols, 200, 0.1, (3, 6), (2, 3))
        return mols, tpls, kw
    if name == "library-64":
        n = max(1, int(round(64 * scale)))
        mols = [protein_with_atoms(rng, 100000, n_chains=20, mol_id=f"L{i:03d}") for i in range(n)]
        tpls = template_set(rng, mols, 7000, 0.1, (2, 8), (2, 3), p_symmetric=0.2)
        kw.update(distance_cutoff=2.5, max_dynamic_distance=2.5, max_candidates=100000)
        return mols, tpls, kw
    raise ValueError(f"unknown workload {name!r}")


def _batch_molecules(rng: np.random.Generator, sizes: Sequence[int]) -> List[Molecule]:
    """Molecules cut from a pool of long self-avoiding walks (any stretch of a
    self-avoiding walk is one), each with its own residue types and side chains."""
    need = [max(3, int(round(s / 8.05))) for s in sizes]
    pool_len = max(max(need), 6000)
    n_pool = max(4, min(32, len(sizes) // 64 + 1))
    pool = [ca_walk(rng, pool_len) for _ in range(n_pool)]
    mols = []
    for i, n in enumerate(need):
        w = pool[int(rng.integers(n_pool))]
        start = int(rng.integers(0, pool_len - n + 1))
        mols.append(protein(rng, n, mol_id=f"M{i:04d}", ca=w[start:start + n]))
    return mols


def random_rotations(rng: np.random.Generator, n: int) -> np.ndarray:
    """n rotation matrices, uniform over SO(3) (unit quaternions)."""
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def superposition_pairs(rng: np.random.Generator, n_sets: int, k_min: int = 3, k_max: int = 12,
                        noise: float = 0.5, collinear: float = 0.02, scatter: float = 2e-2, n_proteins: int = 4):
    """Point-set pairs for the differentiable RMSD: (a, b, offsets).

    Each query set a is ``k`` consecutive atoms (k uniform in [k_min, k_max])
    of a synthetic protein; its template b is a random rotation of it, moved
    to another place, plus N(0, noise) per coordinate. A fraction
    ``collinear`` of the sets are near-degenerate: k = 3 points on a line
    (spacing 1.5 A) with ``scatter`` A of scatter. Returns a and b as
    (points, 3) float64 and offsets as int64 (n_sets + 1).
    """
    pool = np.concatenate([protein_with_atoms(rng, 8000, mol_id=f"DIF{i}").xyz for i in range(n_proteins)])
    k = rng.integers(k_min, k_max + 1, size=n_sets)
    deg = rng.random(n_sets) < collinear
    k[deg] = 3
    offsets = np.zeros(n_sets + 1, dtype=np.int64)
    np.cumsum(k, out=offsets[1:])
    total = int(offsets[-1])
    set_of = np.repeat(np.arange(n_sets), k)
    local = np.arange(total) - offsets[set_of]
    start = rng.integers(0, len(pool) - k_max, size=n_sets)
    a = pool[start[set_of] + local].copy()
    # near-degenerate sets: three points on a random line
    if deg.any():
        dsel = np.nonzero(deg[set_of])[0]
        dirs = _unit(rng, n_sets)
        origin = pool[start]
        a[dsel] = (origin[set_of[dsel]] + 1.5 * local[dsel, None] * dirs[set_of[dsel]]
                   + rng.normal(scale=scatter, size=(len(dsel), 3)))
    R = random_rotations(rng, n_sets)
    shift = rng.normal(scale=10.0, size=(n_sets, 3))
    b = np.einsum("nij,nj->ni", R[set_of], a) + shift[set_of] + rng.normal(scale=noise, size=(total, 3))
    return np.ascontiguousarray(a), np.ascontiguousarray(b), offsets
