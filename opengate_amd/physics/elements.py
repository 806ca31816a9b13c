"""Per-element photon data: NIST mu/rho and mu_en/rho tables, Z/A, names.

The tables are the NIST (Hubbell & Seltzer) values that OpenGATE ships in
core/opengate_core/opengate_lib/GateMuDatabase.h:17-26 (database "NIST").
Element names follow the element table of opengate/data/PhotonAttenuation.py:7939
plus the spellings used by the Schneider table (e.g. "Phosphor").
"""
from functools import lru_cache
from pathlib import Path

import numpy as np

DATA = Path(__file__).resolve().parent.parent / "data"

# Schneider-table spellings that differ from the element table
# (opengate/geometry/materials.py uses the same alias list: elements_name_symbol).
_NAME_ALIASES = {"Phosphor": "P", "Phosphorus": "P", "Sulphur": "S", "Aluminum": "Al"}


@lru_cache(maxsize=None)
def _element_rows():
    rows = []
    for line in (DATA / "elements.csv").read_text().splitlines():
        if line.startswith("#") or not line.strip():
            continue
        z, sym, name, z_over_a = line.split(",")
        rows.append((int(z), sym, name, float(z_over_a)))
    return rows


@lru_cache(maxsize=None)
def symbol_to_z():
    d = {sym: z for z, sym, _, _ in _element_rows()}
    for z, sym, name, _ in _element_rows():
        d[name] = z
    for alias, sym in _NAME_ALIASES.items():
        d[alias] = d[sym]
    return d


def z_of(name_or_symbol):
    try:
        return symbol_to_z()[name_or_symbol]
    except KeyError:
        raise KeyError(f"unknown element '{name_or_symbol}'") from None


def symbol_of(z):
    return _element_rows()[z - 1][1]


def z_over_a(z):
    """Z/A in mol/g (PhotonAttenuation.py PropsEl column 0)."""
    return _element_rows()[z - 1][3]


def atomic_mass(z):
    """A in g/mol, derived from the tabulated Z/A."""
    return z / z_over_a(z)


@lru_cache(maxsize=None)
def nist_table(z):
    """(E [MeV], mu/rho [cm2/g], mu_en/rho [cm2/g]) for element z (1..92).

    Absorption edges appear as two consecutive rows with the same energy
    (value below the edge, then above), as in GateMuDatabase.h.
    """
    arr = _nist_all()
    sel = arr[:, 0] == z
    if not sel.any():
        raise KeyError(f"no NIST table for Z={z}")
    t = arr[sel]
    return t[:, 1].copy(), t[:, 2].copy(), t[:, 3].copy()


@lru_cache(maxsize=None)
def _nist_all():
    return np.loadtxt(DATA / "nist_mu_muen.csv", delimiter=",", comments="#")
