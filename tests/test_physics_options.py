"""Physics-list names and user limits (opengate/physics.py:39-130, :241-438; managers.py:677-
692, :1048-1102): names the reference cannot build are errors, and options that would make
the reference compute what the photon path does not raise. CPU only: the checks run before
any transport."""
import pytest

import opengate_amd as gate
from opengate_amd.managers import (EM_EXTENSIONS, PHYSICS_LIST_BASES, SPECIAL_BUILDERS, available_physics_lists,
                                   em_constructor_of)
from opengate_amd.physics.materials import MaterialDatabase

DB = MaterialDatabase()
MeV = gate.g4_units.MeV
keV = gate.g4_units.keV
SOFT = [DB.get(n) for n in ("G4_WATER", "G4_BONE_CORTICAL_ICRP", "G4_LUNG_ICRP", "G4_AIR")]
HEAVY = SOFT + [DB.get("G4_Pb")]


def pm(name=None):
    sim = gate.Simulation()
    if name is not None:
        sim.physics_manager.physics_list_name = name
    return sim, sim.physics_manager


def test_available_names_are_the_reference_set():
    names = available_physics_lists()
    # 11 constructors + 23 bases x (1 + 9 extensions) + 13 special builders
    assert len(names) == 11 + 23 * 10 + 13 == len(set(names))
    assert len(PHYSICS_LIST_BASES) == 23 and len(EM_EXTENSIONS) == 9 and len(SPECIAL_BUILDERS) == 13
    for n in ("QGSP_BERT_EMV", "QGSP_BERT_EMZ", "FTFP_INCLXX_HP_EMZ", "QGSP_BIC_HP_EMY", "G4EmStandardPhysics_option4",
              "Shielding_HPT", "QGSP_BERT__GS", "LBE"):
        assert n in names
        em_constructor_of(n)
    assert em_constructor_of("QGSP_BERT") == "G4EmStandardPhysics"
    assert em_constructor_of("QGSP_BIC_HP_EMY") == "G4EmStandardPhysics_option3"
    assert em_constructor_of("QGSP_BERT_EM0") == "G4EmStandardPhysics"
    # the longest suffix wins: "__LE" rather than a base ending in "_"
    assert em_constructor_of("FTFP_BERT__LE") == "G4EmLowEPPhysics"


@pytest.mark.parametrize("name", ["QGSP_BERT_EMQ", "QGSP_BERT_GS", "FOO", "G4EmStandardPhysics_option5",
                                  "Shielding_HP_EMZ"])
def test_unknown_list_names_raise(name):
    sim, p = pm(name)
    with pytest.raises(ValueError, match="Cannot find the physic list"):
        p.check_supported()


def test_dna_list_raises_when_set():
    sim, p = pm()
    with pytest.raises(ValueError, match="DNA"):
        p.physics_list_name = "G4EmDNAPhysics_option2"


def test_optical_and_polarised_physics_raise():
    for name in ("G4OpticalPhysics", "G4EmLivermorePolarizedPhysics"):
        sim, p = pm(name)
        with pytest.raises(NotImplementedError):
            p.check_supported()
    sim, p = pm()
    p.special_physics_constructors.G4OpticalPhysics = True
    with pytest.raises(NotImplementedError, match="optical"):
        p.check_supported()
    sim, p = pm()
    p.em_parameters.pixe = True
    with pytest.raises(NotImplementedError, match="pixe"):
        p.check_supported()


@pytest.mark.parametrize("name,fluo", [("QGSP_BERT_EMV", False), ("G4EmStandardPhysics", False),
                                       ("QGSP_BERT_EMZ", True), ("G4EmStandardPhysics_option3", True),
                                       ("QGSP_BIC_EMY", True), ("FTFP_BERT_LIV", True), ("QGSP_BERT_PEN", True),
                                       ("QBBC__LE", True), ("LBE", True), ("QGSP_BERT__GS", False)])
def test_fluorescence_by_list_default(name, fluo):
    sim, p = pm(name)
    assert p.fluorescence_on() is fluo
    # light materials: K x-rays stay within a voxel, same dose as the local deposit
    assert p.resolve_photon_cut(SOFT, 0.1405) == pytest.approx(2.0**-10)
    if fluo:
        with pytest.raises(NotImplementedError, match="Z=82 in G4_Pb"):
            p.resolve_photon_cut(HEAVY, 0.1405)
        p.em_parameters.fluo = False  # switched off explicitly: the reference run matches
        assert not p.fluorescence_on()
    p.resolve_photon_cut(HEAVY, 0.1405)
    p.em_parameters.fluo = True
    with pytest.raises(NotImplementedError, match="em_parameters.fluo"):
        p.resolve_photon_cut(HEAVY, 0.1405)


def test_user_limit_particles():
    sim, p = pm()
    assert p.user_limits_particles == ["all_charged"]
    p.user_limits_particles = "gamma"
    assert p.user_limits_particles == ["gamma"]
    p.user_limits_particles = ["gamma", "pion"]
    with pytest.raises(ValueError, match="unknown particles"):
        p.check_supported()
    with pytest.raises(AttributeError):
        p.set_user_limits_particles(["gamma"])


def test_gamma_min_ekine_is_the_photon_cut():
    # 'all_charged' (the default) puts no kill energy on photons (physics.py:109-114)
    sim, p = pm()
    p.set_min_ekine("world", 20 * keV)
    assert p.resolve_photon_cut(SOFT, 0.1405) == pytest.approx(2.0**-10)
    p.user_limits_particles = ["gamma"]
    assert p.resolve_photon_cut(SOFT, 0.1405) == pytest.approx(0.020)
    # a volume with a region of its own and no kill energy keeps its photons below 20 keV
    p.set_production_cut("ct", "electron", 1.0)
    with pytest.raises(NotImplementedError, match="differs between volumes"):
        p.resolve_photon_cut(SOFT, 0.1405)
    p.set_min_ekine("ct", 20 * keV)
    assert p.resolve_photon_cut(SOFT, 0.1405) == pytest.approx(0.020)
    # a kill energy below the tracking cut changes nothing
    sim, p = pm()
    p.user_limits_particles = ["gamma"]
    p.set_min_ekine("world", 0.5 * keV)
    assert p.resolve_photon_cut(SOFT, 0.1405) == pytest.approx(2.0**-10)


def test_electron_min_ekine_must_cover_the_highest_photon_energy():
    sim, p = pm()
    p.user_limits_particles = ["electron"]
    p.set_min_ekine("world", 100 * keV)
    with pytest.raises(NotImplementedError, match="transport electrons"):
        p.resolve_photon_cut(SOFT, 0.1405)
    p.set_min_ekine("world", 1 * MeV)
    assert p.resolve_photon_cut(SOFT, 0.1405) == pytest.approx(2.0**-10)
    assert not any("electrons are deposited" in w for w in sim.warnings)
    # a volume with a region of its own inherits nothing
    p.set_production_cut("ct", "gamma", 1.0)
    with pytest.raises(NotImplementedError, match="'ct': 0.0"):
        p.resolve_photon_cut(SOFT, 0.1405)
    # 'all' kills photons too: every photon below 1 MeV deposits where it starts
    sim, p = pm()
    p.user_limits_particles = ["all"]
    p.set_min_ekine("world", 1 * MeV)
    assert p.resolve_photon_cut(SOFT, 0.1405) == pytest.approx(1.0)
    # no electron limit: accepted, with the difference stated
    sim, p = pm()
    p.resolve_photon_cut(SOFT, 0.1405)
    assert any("electrons are deposited" in w for w in sim.warnings)


def test_volume_setters_reach_the_manager():
    sim = gate.Simulation()
    ct = sim.add_volume("Box", "ct")
    ct.set_min_ekine(5 * keV)
    ct.set_production_cut("gamma", 1.0)
    assert sim.physics_manager.min_ekine["ct"] == 5 * keV
    assert "ct" in sim.physics_manager.region_volumes
