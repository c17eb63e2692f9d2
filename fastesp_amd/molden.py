"""MOLDEN wavefunction input (``fastesp/molden.py:1-169``).

``read_molden(path) -> Wavefunction`` reads what the ESP needs: atoms, the
contracted shells of [GTO] and the MOs of [MO] (energy, occupation, the
coefficient column).  The semantics follow the reference reader:

* section headers are ``[Name] rest`` lines; the first occurrence of a name
  counts;
* [Atoms] lines are ``symbol index Z x y z``; other lines are skipped;
* [GTO]: a two-field line ``atom 0`` selects the atom (1-based); a shell is
  ``letter nprim [scale]`` followed by nprim ``exponent coefficient`` lines,
  the coefficient as printed (it carries N_s(alpha, l), ``fastesp/basis.py:
  53-66``);
* ``spherical[l]`` records whether a [5D] / [7F] / [9G] header is present;
* [MO]: every ``Sym=`` line opens an orbital; ``Ene=``, ``Occup=`` and
  ``Spin=`` lines set its fields; each following line with two or more fields
  is ``index ... value``; MO columns are scattered into an (n_ao, n_mo)
  matrix, n_ao the largest index.

Two inputs the reference reads differently are handled here as the format
defines them: ``[Atoms] Angs`` coordinates are converted to Bohr (the
reference takes every coordinate as Bohr), and Fortran ``D`` exponents
(``1.0D+00``) are accepted.  Combined flag headers such as ``[5D7F]`` or
``[5D10F]`` set the flags they name.

``write_molden(path, wfn)`` writes the same subset (AU, pure d/f/g).
"""
from __future__ import annotations

import re
from pathlib import Path

import numpy as np

from .basis import BOHR, Shell, Wavefunction, Z_OF

_L_LETTERS = "spdfghijkl"
_HEADER = re.compile(r"^\[([^\]]+)\]\s*(.*)$")


def _num(tok: str) -> float:
    return float(tok.replace("D", "E").replace("d", "e"))


def _sections(lines):
    """{name: (first line index, rest of the header line)}, first occurrence wins."""
    out = {}
    for i, line in enumerate(lines):
        m = _HEADER.match(line.strip())
        if m:
            out.setdefault(m.group(1).strip(), (i, m.group(2).strip()))
    return out


def _body(lines, sec, name):
    if name not in sec:
        raise ValueError(f"MOLDEN file has no [{name}] section")
    i = sec[name][0] + 1
    out = []
    while i < len(lines) and not lines[i].lstrip().startswith("["):
        out.append(lines[i])
        i += 1
    return out


def _flags(sec) -> dict:
    """spherical[l] for l <= 4 from [5D], [7F], [9G] and combined headers."""
    names = set()
    for n in sec:
        names.update(re.findall(r"(\d+)([A-Za-z])", n))
    have = {f"{a}{b.upper()}" for a, b in names}
    return {l: f"{2 * l + 1}{'SPDFG'[l]}" in have for l in range(5)}


def read_molden(path) -> Wavefunction:
    lines = Path(path).read_text(errors="ignore").splitlines()
    sec = _sections(lines)

    # ---- atoms
    scale = 1.0
    if "Atoms" in sec and sec["Atoms"][1].lower().startswith("ang"):
        scale = 1.0 / BOHR
    symbols, coords = [], []
    for line in _body(lines, sec, "Atoms"):
        f = line.split()
        if len(f) < 6:
            continue
        try:
            int(f[1]), int(f[2])
            xyz = [_num(f[3]) * scale, _num(f[4]) * scale, _num(f[5]) * scale]
        except ValueError:
            continue
        symbols.append(f[0])
        coords.append(xyz)

    # ---- basis
    shells = []
    gto = _body(lines, sec, "GTO")
    atom = 0
    i = 0
    while i < len(gto):
        f = gto[i].split()
        i += 1
        if not f:
            continue
        if len(f) == 2 and f[0].lstrip("-").isdigit():
            atom = int(f[0]) - 1
            continue
        l = _L_LETTERS.index(f[0].lower())
        nprim = int(f[1])
        rows = [gto[i + k].split() for k in range(nprim)]
        i += nprim
        shells.append(Shell(atom=atom, l=l, exps=np.array([_num(r[0]) for r in rows]),
                            coeffs=np.array([_num(r[1]) for r in rows])))

    # ---- orbitals
    energies, occs, cols = [], [], []
    cur = None
    for line in _body(lines, sec, "MO"):
        t = line.strip()
        if t.startswith("Sym="):
            cur = {"e": 0.0, "o": 0.0, "idx": [], "val": []}
            cols.append(cur)
            continue
        if cur is None:
            continue
        if t.startswith("Ene="):
            cur["e"] = _num(t.split("=", 1)[1])
        elif t.startswith("Occup="):
            cur["o"] = _num(t.split("=", 1)[1])
        elif t.startswith("Spin=") or not t:
            continue
        else:
            f = t.split()
            if len(f) >= 2:
                cur["idx"].append(int(f[0]))
                cur["val"].append(_num(f[-1]))
    if not cols:
        raise ValueError("MOLDEN file has no orbitals in [MO]")
    n_ao = max(max(c["idx"]) for c in cols if c["idx"])
    C = np.zeros((n_ao, len(cols)))
    for j, c in enumerate(cols):
        if c["idx"]:
            C[np.asarray(c["idx"]) - 1, j] = c["val"]
        energies.append(c["e"])
        occs.append(c["o"])
    return Wavefunction(symbols=symbols, coords=np.asarray(coords, float).reshape(-1, 3),
                        shells=shells, mo_energy=np.asarray(energies, float),
                        mo_occ=np.asarray(occs, float), mo_coeff=C, spherical=_flags(sec))


def write_molden(path, wfn: Wavefunction, title: str = "fastesp_amd") -> Path:
    """MOLDEN text with [Atoms] AU, [GTO], [5D] [7F] [9G] and [MO] (pure
    functions, the MO columns of ``wfn.mo_coeff``)."""
    atoms = [sh.atom for sh in wfn.shells]
    if atoms != sorted(atoms):
        raise ValueError("write_molden: shells must be grouped by atom in atom order "
                         "(the MO coefficient order follows [GTO])")
    out = ["[Molden Format]", "[Title]", f" {title}", "[Atoms] AU"]
    for i, (s, xyz) in enumerate(zip(wfn.symbols, np.asarray(wfn.coords, float))):
        out.append(f"{s:<3s} {i + 1:5d} {int(Z_OF.get(s, 0)):3d} "
                   f"{xyz[0]:22.14f} {xyz[1]:22.14f} {xyz[2]:22.14f}")
    out.append("[GTO]")
    for a in range(len(wfn.symbols)):
        out.append(f"{a + 1:5d} 0")
        for sh in wfn.shells:
            if sh.atom != a:
                continue
            out.append(f"{_L_LETTERS[sh.l]} {len(sh.exps):4d} 1.00")
            for e, c in zip(np.asarray(sh.exps, float), np.asarray(sh.coeffs, float)):
                out.append(f"  {e:24.16e} {c:24.16e}")
        out.append("")
    out += ["[5D]", "[7F]", "[9G]", "[MO]"]
    C = np.asarray(wfn.mo_coeff, float)
    for j in range(C.shape[1]):
        out.append(f" Sym= {j + 1}a")
        out.append(f" Ene= {float(wfn.mo_energy[j]) if len(wfn.mo_energy) else 0.0:.10e}")
        out.append(" Spin= Alpha")
        out.append(f" Occup= {float(wfn.mo_occ[j]):.10f}")
        for i in range(C.shape[0]):
            out.append(f"{i + 1:6d} {C[i, j]:24.16e}")
    p = Path(path)
    p.write_text("\n".join(out) + "\n")
    return p
