"""HU -> label semantics (volumes.py:1144-1243) and the Schneider HU -> material
conversion (materials.py:61-230), CPU."""
import numpy as np
import pytest

from oracle import oracle as O
from opengate_amd.geometry import materials as M
from opengate_amd.geometry.volumes import ImageVolume
from opengate_amd.managers import Simulation
from opengate_amd.units import g4_units

VM = [[-1050, -950, "G4_AIR"], [-950, -120, "G4_LUNG_ICRP"], [-120, 19, "G4_ADIPOSE_TISSUE_ICRP"],
      [19, 80, "G4_WATER"], [120, 400.5, "G4_MUSCLE_SKELETAL_ICRP"], [400.5, 3000, "G4_BONE_CORTICAL_ICRP"]]


def brute_force_labels(hu, material, vm):
    """Independent statement of the rule: a value in [lo, hi) of a row gets that row's
    material label; values in no interval get label 0 (the volume's material)."""
    lut = {material: 0}
    nxt = 1
    for r in sorted(vm, key=lambda r: r[0]):
        if r[2] not in lut:
            lut[r[2]] = nxt
            nxt += 1
    out = np.zeros(hu.shape, np.uint16)
    for lo, hi, name in vm:
        out[(hu >= lo) & (hu < hi)] = lut[name]
    return out


def test_label_rule_exhaustive_int16_range():
    hu = np.arange(-1100, 3101, dtype=np.int16)
    ref = brute_force_labels(hu, "G4_WATER", VM)
    got, _ = O.label_image(hu, "G4_WATER", VM)
    np.testing.assert_array_equal(got, ref)
    # C restatement equals numpy
    lut = O.label_lut("G4_WATER", VM)
    bins, ls = O.label_bins(lut, VM)
    np.testing.assert_array_equal(O.hu_to_labels_c(hu, bins, ls), ref)


def test_label_lut_and_overlap():
    lut = O.label_lut("G4_AIR", VM)
    assert lut["G4_AIR"] == 0 and lut["G4_LUNG_ICRP"] == 1 and lut["G4_BONE_CORTICAL_ICRP"] == 5
    with pytest.raises(ValueError):
        O.label_lut("G4_AIR", [[0, 10, "A"], [5, 20, "B"]])


def test_image_volume_bins_equal_oracle():
    v = ImageVolume("ct")
    v.material = "G4_AIR"
    v.voxel_materials = VM
    lut = v.create_material_to_label_lut()
    assert lut == O.label_lut("G4_AIR", VM)
    b1, l1 = v.label_bins(lut)
    b2, l2 = O.label_bins(lut, VM)
    np.testing.assert_array_equal(b1, b2)
    np.testing.assert_array_equal(l1, l2)
    with pytest.raises(ValueError):
        v.voxel_materials = [[0, 10, "A"], [5, 20, "B"]]
        v.create_material_to_label_lut()


def test_schneider_conversion():
    sim = Simulation()
    data = __import__("opengate_amd").__path__[0] + "/data/"
    tol = 0.05 * g4_units.g_cm3
    vm, created = M.HounsfieldUnit_to_material(sim, tol, data + "Schneider2000MaterialsTable.txt",
                                               data + "Schneider2000DensitiesTable.txt")
    assert len(vm) == len(created)
    # contiguous half-open intervals covering [-1050, 3001)
    assert vm[0][0] == -1050 and vm[-1][1] == 3001
    assert all(abs(a[1] - b[0]) < 1e-9 for a, b in zip(vm[:-1], vm[1:]))
    # Air row is one interval; every sub-interval's density is the interpolated midpoint density
    assert vm[0][2].startswith("Air") and not vm[1][2].startswith("Air")
    dens = M.HU_read_density_table(data + "Schneider2000DensitiesTable.txt")
    db = sim.volume_manager.material_database
    for lo, hi, name in vm:
        exp = M.HU_linear_interpolate_densities(lo + (hi - lo) / 2, dens)
        assert db.get(name).density == pytest.approx(exp, rel=1e-12)
        assert db.get(name).fractions.sum() == pytest.approx(1.0)
    # sub-interval density steps never exceed the tolerance by more than the
    # discretisation of the table (n = ceil(ddiff / tol) sub-intervals per row)
    mats, _ = M.HU_read_materials_table(data + "Schneider2000MaterialsTable.txt")
    assert len(mats) == 27
    n_expected = 0
    for i, m in enumerate(mats):
        lo = m["HU"]
        hi = lo + 1 if i == len(mats) - 1 else mats[i + 1]["HU"]
        d0 = M.HU_linear_interpolate_densities(lo, dens)
        d1 = M.HU_linear_interpolate_densities(hi, dens)
        dd = M.HU_find_max_density_difference(lo, hi, d0, d1, dens)
        n = max(1, dd / 0.05)
        n_expected += 1 if ("Air" in m["name"] or n == 1) else int(np.ceil(n))
    assert len(vm) == n_expected


def test_read_voxel_materials(tmp_path):
    f = tmp_path / "vm.txt"
    f.write_text("# lo hi name\n-1000 -500 G4_AIR\n-400 100 G4_WATER\n100 2000 G4_BONE_CORTICAL_ICRP\n")
    vm = M.read_voxel_materials(f)
    assert vm == [[-1000, -500, "G4_AIR"], [-500, -400, "G4_AIR"], [-400, 100, "G4_WATER"],
                  [100, 2000, "G4_BONE_CORTICAL_ICRP"]]
