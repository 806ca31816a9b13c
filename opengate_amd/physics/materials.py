"""Material database: named NIST-style materials and mixtures by mass fraction.

Materials are identified by name, as in Geant4/GATE ("G4_WATER", ...). The
compositions below are the NIST/ICRU mass fractions; densities in g/cm3.
User materials (e.g. those created by the Schneider HU conversion,
opengate/geometry/materials.py:149-230) are added with `add_material_weights` or
`add_material_nb_atoms`, the calls the reference uses, and GATE material database
files (.db text: [Isotopes], [Elements], [Materials]) are read with `read_from_file`.
"""
import re
from dataclasses import dataclass, field

import numpy as np

from .elements import _element_rows, atomic_mass, z_of


@dataclass
class Material:
    name: str
    density: float  # g/cm3
    elements: list  # atomic numbers
    fractions: np.ndarray  # mass fractions, sum 1
    is_vacuum: bool = False
    extra: dict = field(default_factory=dict)


def _mk(name, density, comp, vacuum=False):
    zs = [z_of(s) for s in comp]
    w = np.array([comp[s] for s in comp], dtype=np.float64)
    w = w / w.sum()
    return Material(name, float(density), zs, w, vacuum)


# Mass fractions: NIST "Compositions of Materials" (ICRU Report 37/44/46 values).
_BUILTIN = {
    "G4_Galactic": (1e-25, {"H": 1.0}, True),
    "G4_AIR": (0.00120479, {"C": 0.000124, "N": 0.755268, "O": 0.231781, "Ar": 0.012827}, False),
    "G4_WATER": (1.0, {"H": 0.111894, "O": 0.888106}, False),
    "G4_BONE_CORTICAL_ICRP": (
        1.92,
        {"H": 0.034, "C": 0.155, "N": 0.042, "O": 0.435, "Na": 0.001, "Mg": 0.002,
         "P": 0.103, "S": 0.003, "Ca": 0.225},
        False,
    ),
    "G4_BONE_COMPACT_ICRU": (
        1.85,
        {"H": 0.063984, "C": 0.278, "N": 0.027, "O": 0.410016, "Mg": 0.002, "P": 0.07,
         "S": 0.002, "Ca": 0.147},
        False,
    ),
    "G4_LUNG_ICRP": (
        1.04,
        {"H": 0.105, "C": 0.083, "N": 0.023, "O": 0.779, "Na": 0.002, "P": 0.001,
         "S": 0.002, "Cl": 0.003, "K": 0.002},
        False,
    ),
    "G4_ADIPOSE_TISSUE_ICRP": (
        0.95,
        {"H": 0.114, "C": 0.598, "N": 0.007, "O": 0.278, "Na": 0.001, "S": 0.001, "Cl": 0.001},
        False,
    ),
    "G4_MUSCLE_SKELETAL_ICRP": (
        1.05,
        {"H": 0.102, "C": 0.143, "N": 0.034, "O": 0.71, "Na": 0.001, "P": 0.002, "S": 0.003,
         "Cl": 0.001, "K": 0.004},
        False,
    ),
    "G4_TISSUE_SOFT_ICRP": (
        1.03,
        {"H": 0.104472, "C": 0.23219, "N": 0.02488, "O": 0.630238, "Na": 0.00113,
         "Mg": 0.00013, "P": 0.00133, "S": 0.00199, "Cl": 0.00134, "K": 0.00199,
         "Ca": 0.00023, "Fe": 0.00005, "Zn": 0.00003},
        False,
    ),
    "G4_Al": (2.699, {"Al": 1.0}, False),
    "G4_Cu": (8.96, {"Cu": 1.0}, False),
    "G4_W": (19.3, {"W": 1.0}, False),
    "G4_Pb": (11.35, {"Pb": 1.0}, False),
    "G4_Ti": (4.54, {"Ti": 1.0}, False),
    "G4_SODIUM_IODIDE": (3.667, {"Na": 0.153373, "I": 0.846627}, False),
    "G4_PLEXIGLASS": (1.19, {"H": 0.080538, "C": 0.599848, "O": 0.319614}, False),
}


def _tag(text, tag):
    """Value of `tag= value` in one ';'-separated field, None when the field holds another
    tag (materials.py:276-281)."""
    k, _, v = text.partition("=")
    return v.strip() if k.strip() == tag else None


def _tag_with_unit(text, tag):
    """`tag= value unit` in Geant4 units (materials.py:284-291)."""
    from ..units import g4_units

    v = _tag(text, tag)
    if v is None:
        return None
    value, unit = v.split()[:2]
    if unit not in g4_units:
        raise ValueError(f"unknown unit '{unit}' in '{text.strip()}'")
    return float(value) * g4_units[unit]


@dataclass
class _DbElement:
    name: str
    z: int
    a: float  # g/mol


@dataclass
class _DbMaterial:
    name: str
    density: float  # g/cm3
    components: list  # (kind "el"/"mat", name, n or None, f or None)
    state: str = None


class MaterialDatabase:
    """Name -> Material. The photon path's part of opengate.geometry.materials.MaterialDatabase
    (materials.py:726-943): lookup by name (built-in NIST materials, then user materials,
    then materials read from database files, built on first use), add_material_weights,
    add_material_nb_atoms and read_from_file."""

    def __init__(self):
        self.materials = {}
        for name, (rho, comp, vac) in _BUILTIN.items():
            self.materials[name] = _mk(name, rho, comp, vac)
        self.filenames = []
        self.isotope_defs = {}  # name -> (Z, N, A g/mol)
        self.element_defs = {}  # name -> _DbElement
        self.material_defs = {}  # name -> _DbMaterial
        self.material_defs_by_filename = {}
        self._nist_symbols = {sym: z for z, sym, _, _ in _element_rows()}

    # ---------------------------------------------------------------- .db files
    def read_from_file(self, filename):
        """Reads a GATE material database file (materials.py:778-835). Lines are stripped,
        tabs become spaces, '#' starts a comment line; a material or isotopic element is
        followed by exactly n component lines (a line that is not a component is skipped,
        as the reference does). Later definitions of a name replace earlier ones."""
        filename = str(filename)
        self.filenames.append(filename)
        mine = self.material_defs_by_filename.setdefault(filename, {})
        with open(filename) as f:
            lines = [ln.strip().replace("\t", " ") for ln in f]
        section = None
        i = 0
        while i < len(lines):
            line = lines[i]
            i += 1
            if not line or line[0] == "#":
                continue
            head = line.split()[0]
            if head in ("[Isotopes]", "[Elements]", "[Materials]"):
                section = head[1:-1]
                continue
            if section is None:
                raise ValueError(f"{filename}: the file must start with [Isotopes], [Elements] or [Materials]")
            name, colon, rest = line.partition(":")
            if not colon:
                raise ValueError(f"{filename}: expected 'name:' in line '{line}'")
            fields = rest.split(";")
            if section == "Isotopes":
                self.isotope_defs[name] = (int(_tag(fields[0], "Z")), int(_tag(fields[1], "N")),
                                           _tag_with_unit(fields[2], "A") / _g_per_mole())
            elif section == "Elements":
                if "Z=" in line:
                    zf = float(_tag(fields[1], "Z"))
                    if zf != int(zf):
                        raise NotImplementedError(f"element {name}: Z = {zf} is not an integer (the photon "
                                                  "cross sections are tabulated per element)")
                    self.element_defs[name] = _DbElement(name, int(zf), _tag_with_unit(fields[2], "A") / _g_per_mole())
                if "n=" in line:
                    if len(fields) != 2:
                        raise ValueError(f"{filename}: error while parsing element {name}, line '{line}'")
                    n = int(_tag(fields[0], "n"))
                    isos = []
                    for _ in range(n):
                        ln = lines[i] if i < len(lines) else ""
                        i += 1
                        if ln.startswith("+iso"):
                            parts = re.split("[;,]", ln.split("+iso:")[1])
                            iso = _tag(parts[0], "name")
                            isos.append((name if iso == "auto" else iso, float(_tag(parts[1], "f"))))
                    self.element_defs[name] = ("isotopes", isos)
            else:
                if len(fields) not in (2, 3):
                    raise ValueError(f"{filename}: error while parsing material {name}, line '{line}'")
                d = _tag_with_unit(fields[0], "d")
                if not d:
                    raise ValueError(f"{filename}: material {name}: expected a density 'd=value unit'")
                n = int(_tag(fields[1], "n"))
                state = _tag(fields[2], "state") if len(fields) > 2 else None
                comps = []
                for _ in range(n):
                    ln = lines[i] if i < len(lines) else ""
                    i += 1
                    if ln.startswith("+mat:"):
                        parts = re.split("[;,]", ln.split("+mat:")[1])
                        f = _tag(parts[1], "f") if len(parts) == 2 else None
                        if f is None:
                            raise ValueError(f"{filename}: sub-material of {name} needs 'f=': '{ln}'")
                        comps.append(("mat", _tag(parts[0], "name"), None, float(f)))
                    elif ln.startswith("+el"):
                        parts = re.split("[;,]", ln.split("+el:")[1])
                        if len(parts) != 2:
                            raise ValueError(f"{filename}: expected 'name=' ; 'n=' or 'f=' in '{ln}'")
                        el = _tag(parts[0], "name")
                        el = name if el == "auto" else el
                        na = _tag(parts[1], "n")
                        if na:
                            comps.append(("el", el, int(na), None))
                        else:
                            f = float(_tag(parts[1], "f"))
                            if f != 0:
                                comps.append(("el", el, None, f))
                from ..units import g4_units

                m = _DbMaterial(name, d / g4_units.g_cm3, comps, state.lower() if state else None)
                self.material_defs[name] = m
                mine[name] = m
                self.materials.pop(name, None)  # a later file's definition is built anew

    def _element(self, name):
        """(Z, A g/mol) of an element name: NIST symbols first, then database elements
        (materials.py:917-931)."""
        if name in self._nist_symbols:
            z = self._nist_symbols[name]
            return z, atomic_mass(z)
        e = self.element_defs.get(name)
        if e is None:
            raise KeyError(f'Cannot find nor build element named "{name}"')
        if isinstance(e, _DbElement):
            return e.z, e.a
        # element made of isotopes: Z of its isotopes, A their abundance-weighted mass
        zs, num, den = set(), 0.0, 0.0
        for iso, f in e[1]:
            if iso not in self.isotope_defs:
                raise KeyError(f'Cannot find nor build isotope named "{iso}"')
            z, _, a = self.isotope_defs[iso]
            zs.add(z)
            num += f * a
            den += f
        if len(zs) != 1:
            raise ValueError(f"element {name}: isotopes of different Z {sorted(zs)}")
        return zs.pop(), num / den

    def _build(self, m):
        """Mass fractions of a database material (materials.py:668-723: elements by number
        of atoms, n_i A_i / sum n_j A_j, or by mass fraction; a sub-material adds its element
        fractions times f; the same element twice adds up). Fractions are used as given,
        without normalisation, as the built material does."""
        kinds = {("n" if n is not None else "f") for k, _, n, _ in m.components if k == "el"}
        if len(kinds) > 1:
            raise ValueError(f"material {m.name}: components by number of atoms and by mass fraction are mixed")
        frac = {}
        if kinds == {"n"}:
            za = [(self._element(el), n) for _, el, n, _ in m.components]
            total = sum(n * a for (_, a), n in za)
            for (z, a), n in za:
                frac[z] = frac.get(z, 0.0) + n * a / total
        else:
            for kind, nm, _, f in m.components:
                if kind == "el":
                    z = self._element(nm)[0]
                    frac[z] = frac.get(z, 0.0) + f
                else:
                    sub = self.get(nm)
                    for z, w in zip(sub.elements, sub.fractions):
                        frac[z] = frac.get(z, 0.0) + f * float(w)
        zs = list(frac)
        return Material(m.name, m.density, zs, np.array([frac[z] for z in zs], dtype=np.float64),
                        extra={"state": m.state, "source": "database"})

    def get_database_material_names(self, db=None):
        if db is None:
            return list(self.material_defs)
        if str(db) not in self.material_defs_by_filename:
            raise ValueError(f"The database '{db}' is not in the list of read database: {self.filenames}")
        return list(self.material_defs_by_filename[str(db)])

    def add_material_nb_atoms(self, name, elements, nb_atoms, density):
        """Material from element symbols and numbers of atoms (G4NistManager::
        ConstructNewMaterialNbAtoms; materials.py:845-853); density in Geant4 units."""
        from ..units import g4_units

        za = [(z_of(e), int(n)) for e, n in zip(elements, nb_atoms)]
        total = sum(n * atomic_mass(z) for z, n in za)
        comp = {}
        for z, n in za:
            comp[z] = comp.get(z, 0.0) + n * atomic_mass(z) / total
        zs = list(comp)
        self.materials[name] = Material(name, float(density) / g4_units.g_cm3, zs,
                                        np.array([comp[z] for z in zs], dtype=np.float64))
        return self.materials[name]

    def add_material_weights(self, name, elements, weights, density):
        """elements: symbols or names; weights: mass fractions; density in Geant4 units
        (g_cm3 = 6.24150907e18), as MaterialDatabase.add_material_weights of the reference."""
        from ..units import g4_units

        comp = {}
        for e, w in zip(elements, weights):
            comp[e] = comp.get(e, 0.0) + float(w)
        self.materials[name] = _mk(name, float(density) / g4_units.g_cm3, comp)
        return self.materials[name]

    def add_material(self, material: Material):
        self.materials[material.name] = material

    def __contains__(self, name):
        return name in self.materials or name in self.material_defs

    def get(self, name):
        m = self.materials.get(name)
        if m is not None:
            return m
        if name in self.material_defs:
            m = self._build(self.material_defs[name])
            self.materials[name] = m
            return m
        raise KeyError(f"unknown material '{name}'")

    def FindOrBuildMaterial(self, name):
        return self.get(name)


def _g_per_mole():
    from ..units import g4_units

    return g4_units["g/mole"]
