"""Regenerate the electron scattering-factor coefficient tables.

The coefficients are published physical data:

* Lobato & Van Dyck (2014), Acta Cryst. A70, 636-649: five (a_i, b_i) pairs per
  element for f(g) = sum_i a_i (2 + b_i g^2) / (1 + b_i g^2)^2 (g in 1/Angstrom).
* International Tables for Crystallography Vol. C, table 4.3.2.3: five
  (a_i, b_i) pairs per element for f(g) = sum_i a_i exp(-b_i g^2 / 4).

The numbers must be identical to the ones diffsims uses, so this script reads
them from the reference's data modules *as text* (``ast.literal_eval``; nothing
is imported or executed) and writes them into this package's own JSON format:
``{"elements": [...], "Z": [...], "coeffs": [[a1, b1, ..., a5, b5], ...]}``.

Reference data modules (read as text):
  diffsims/utils/lobato_scattering_params.py:25-1058
  diffsims/utils/atomic_scattering_params.py:22-1005

Usage: python diffsims_amd/data/make_tables.py [/root/reference]
"""

import ast
import hashlib
import json
import os
import sys

# Periodic table symbols, Z = 1..103 (used to attach Z to each element).
SYMBOLS = (
    "H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca Sc Ti V Cr Mn Fe Co Ni Cu "
    "Zn Ga Ge As Se Br Kr Rb Sr Y Zr Nb Mo Tc Ru Rh Pd Ag Cd In Sn Sb Te I Xe Cs Ba "
    "La Ce Pr Nd Pm Sm Eu Gd Tb Dy Ho Er Tm Yb Lu Hf Ta W Re Os Ir Pt Au Hg Tl Pb "
    "Bi Po At Rn Fr Ra Ac Th Pa U Np Pu Am Cm Bk Cf Es Fm Md No Lr"
).split()

SOURCES = {
    "lobato": "diffsims/utils/lobato_scattering_params.py",
    "xtables": "diffsims/utils/atomic_scattering_params.py",
}


def _parse(path):
    """Return [(element, ((a1,b1),...)), ...] from a ``NAME = dict((...))`` module."""
    tree = ast.parse(open(path).read())
    for node in tree.body:
        if isinstance(node, ast.Assign) and isinstance(node.value, ast.Call):
            return ast.literal_eval(node.value.args[0])
    raise RuntimeError(f"no table found in {path}")


def main(ref_root):
    here = os.path.dirname(os.path.abspath(__file__))
    for name, rel in SOURCES.items():
        rows = _parse(os.path.join(ref_root, rel))
        elements, zs, coeffs = [], [], []
        for el, pairs in rows:
            flat = []
            for a, b in pairs:
                flat += [float(a), float(b)]
            # Entries with fewer than 5 pairs are zero-padded (a=0 contributes nothing).
            flat += [0.0] * (10 - len(flat))
            elements.append(el)
            zs.append(SYMBOLS.index(el) + 1 if el in SYMBOLS else 0)
            coeffs.append(flat)
        out = {"source": rel, "elements": elements, "Z": zs, "coeffs": coeffs}
        blob = json.dumps(out, indent=0)
        with open(os.path.join(here, f"scattering_{name}.json"), "w") as f:
            f.write(blob)
        print(name, len(elements), hashlib.sha256(blob.encode()).hexdigest()[:16])


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
