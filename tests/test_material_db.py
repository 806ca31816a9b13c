"""GATE material database files (opengate/geometry/materials.py:271-291, :396-943): the
reader and the mass fractions of the built materials, worked out by hand from the
committed GateMaterials.db (a data copy of opengate/contrib/GateMaterials.db). CPU only."""
import numpy as np
import pytest

import opengate_amd as gate
from opengate_amd.physics import xs
from opengate_amd.physics.materials import MaterialDatabase

DB_FILE = __import__("opengate_amd").__path__[0] + "/data/GateMaterials.db"


def fractions(m):
    return {int(z): float(w) for z, w in zip(m.elements, m.fractions)}


@pytest.fixture(scope="module")
def db():
    d = MaterialDatabase()
    d.read_from_file(DB_FILE)
    return d


def test_every_material_of_the_file_builds(db):
    names = db.get_database_material_names()
    assert len(names) == 115 and names == db.get_database_material_names(DB_FILE)
    for n in names:
        m = db.get(n)
        assert m.density > 0 and abs(m.fractions.sum() - 1.0) < 2e-3, n


def test_atoms_to_mass_fractions_by_hand(db):
    # Water: H n=2 (A 1.01), O n=1 (A 16.00): w_H = 2.02 / 18.02
    w = fractions(db.get("Water"))
    assert w[1] == pytest.approx(2.02 / 18.02, rel=1e-12) and w[8] == pytest.approx(16.00 / 18.02, rel=1e-12)
    assert db.get("Water").density == pytest.approx(1.0)
    # ABS: H 17 x 1.01, C 15 x 12.01, N 1 x 14.01 -> 17.17 + 180.15 + 14.01 = 211.33
    w = fractions(db.get("ABS"))
    assert w[6] == pytest.approx(180.15 / 211.33, rel=1e-12) and w[1] == pytest.approx(17.17 / 211.33, rel=1e-12)
    # Carbide: W (A 183.84) and C (12.01), one atom each
    assert fractions(db.get("Carbide"))[74] == pytest.approx(183.84 / 195.85, rel=1e-12)


def test_sub_material_and_units(db):
    # FR4_Copper = Au 0.01 + Ni 0.02 + Cu 0.21 + 0.76 x FR4 (Si 0.15, O 0.36, C 0.27, ...)
    w = fractions(db.get("FR4_Copper"))
    assert w[29] == pytest.approx(0.21) and w[79] == pytest.approx(0.01)
    assert w[14] == pytest.approx(0.76 * 0.15) and w[8] == pytest.approx(0.76 * 0.36) and w[6] == pytest.approx(0.76 * 0.27)
    assert len(w) == 3 + 11
    # mg/cm3; 'auto' names the element after the material
    assert db.get("Air").density == pytest.approx(1.29e-3)
    assert db.get("Vacuum").density == pytest.approx(1e-9)
    assert fractions(db.get("Beryllium")) == {4: 1.0} and db.get("Beryllium").density == pytest.approx(1.85)


def test_reader_rules(tmp_path):
    f = tmp_path / "m.db"
    f.write_text("[Isotopes]\nU235: Z= 92 ; N= 143 ; A= 235.044 g/mole\nU238: Z= 92 ; N= 146 ; A= 238.051 g/mole\n"
                 "[Elements]\nHydrogen: S= H ; Z= 1. ; A= 1.01 g/mole\nOxygen: S= O ; Z= 8. ; A= 16.00 g/mole\n"
                 "Uenr: n=2 ; S=U\n +iso: name=U235 ; f=0.2\n +iso: name=U238 ; f=0.8\n"
                 "[Materials]\n# comment\nW1: d=1 g/cm3 ; n=2\n\t+el: name=Hydrogen ; n=2\n\t+el: name=Oxygen ; n=1\n"
                 "Fuel: d=10 g/cm3 ; n=2 ; state=Solid\n +el: name=Uenr ; n=1\n +el: name=O ; n=2\n"
                 "Skip: d=1 g/cm3 ; n=2\n\n +el: name=Oxygen ; f=1\n"
                 "Zero: d=1 g/cm3 ; n=2\n +el: name=Oxygen ; f=1\n +el: name=Hydrogen ; f=0\n"
                 "W1: d=2 g/cm3 ; n=1\n +el: name=Oxygen ; n=1\n")
    d = MaterialDatabase()
    d.read_from_file(f)
    # a later definition replaces the earlier one
    assert fractions(d.get("W1")) == {8: 1.0} and d.get("W1").density == pytest.approx(2.0)
    # isotopic element: A = 0.2 x 235.044 + 0.8 x 238.051; 'O' is the NIST element
    fuel = fractions(d.get("Fuel"))
    a_u = 0.2 * 235.044 + 0.8 * 238.051
    from opengate_amd.physics.elements import atomic_mass

    assert fuel[92] == pytest.approx(a_u / (a_u + 2 * atomic_mass(8)), rel=1e-12)
    assert d.get("Fuel").extra["state"] == "solid"
    # exactly n component lines follow a material: the blank line takes one of the two slots
    assert fractions(d.get("Skip")) == {8: 1.0}
    # f = 0 components are dropped
    assert fractions(d.get("Zero")) == {8: 1.0}


def test_reader_errors(tmp_path):
    def read(text):
        f = tmp_path / "e.db"
        f.write_text(text)
        d = MaterialDatabase()
        d.read_from_file(f)
        return d

    with pytest.raises(ValueError, match="must start"):
        read("Water: d=1 g/cm3 ; n=1\n +el: name=H ; n=1\n")
    with pytest.raises(ValueError, match="mixed"):
        read("[Materials]\nM: d=1 g/cm3 ; n=2\n +el: name=H ; n=1\n +el: name=O ; f=0.5\n").get("M")
    with pytest.raises(KeyError, match="Cannot find nor build element"):
        read("[Materials]\nM: d=1 g/cm3 ; n=1\n +el: name=Unobtainium ; n=1\n").get("M")
    with pytest.raises(ValueError, match="density"):
        read("[Materials]\nM: d=0 g/cm3 ; n=1\n +el: name=H ; n=1\n")
    # n = 1 with a blank line first: the component line is read as a material of its own
    with pytest.raises(ValueError, match="density"):
        read("[Materials]\nM: d=1 g/cm3 ; n=1\n\n +el: name=H ; n=1\n")


def test_nb_atoms_material_and_facade(tmp_path):
    d = MaterialDatabase()
    bgo = d.add_material_nb_atoms("BGO", ["Bi", "Ge", "O"], [4, 3, 12], 7.13 * gate.g4_units.g_cm3)
    w = fractions(bgo)
    # NIST atomic masses 208.98, 72.63, 15.999: w_Bi = 835.92 / 1245.8
    assert w[83] == pytest.approx(835.92 / 1245.80, rel=2e-3) and bgo.density == pytest.approx(7.13)
    sim = gate.Simulation()
    sim.volume_manager.add_material_database(DB_FILE)
    with pytest.raises(ValueError, match="already"):
        sim.volume_manager.add_material_database(DB_FILE)
    m = sim.volume_manager.material_database.get("PMMA")
    t = xs.build_tables([m], 0.15)
    assert np.all(t.mu_total > 0)
