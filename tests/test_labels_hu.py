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
    """Rows worked out by hand from the reference's reading rule (materials.py:14-58)."""
    from opengate_amd.geometry.volumes import ImageVolume

    def rows(text, **kw):
        f = tmp_path / "vm.txt"
        f.write_text(text)
        return M.read_voxel_materials(f, **kw)

    # contiguous intervals, unsorted in the file, comments, a triple split across lines
    assert rows("# lo hi name\n100 2000 BONE\n-1000 -500 AIR # air\n-500\n100 WATER\n") == [
        [-1000, -500, "AIR"], [-500, 100, "WATER"], [100, 2000, "BONE"]]
    # a gap: def_mat fills [-500, -400], then WATER's row starts at -500 as well
    vm = rows("-1000 -500 G4_AIR\n-400 100 G4_WATER\n100 2000 G4_BONE_CORTICAL_ICRP\n")
    assert vm == [[-1000, -500, "G4_AIR"], [-500, -400, "G4_AIR"], [-500, 100, "G4_WATER"],
                  [100, 2000, "G4_BONE_CORTICAL_ICRP"]]
    # ... which the label table rejects as overlapping (volumes.py:1144-1168)
    v = ImageVolume("ct")
    v.voxel_materials = vm
    with pytest.raises(ValueError, match="Overlapping"):
        v.create_material_to_label_lut()
    # a bound of 0 counts as no bound: the gap [0, 10] is not filled and B starts at 0
    assert rows("-100 0 A\n10 20 B\n", def_mat="D") == [[-100, 0, "A"], [0, 20, "B"]]
    with pytest.raises(ValueError, match="not disjoint"):
        rows("0 10 A\n5 20 B\n-5 3 C\n")
    with pytest.raises(ValueError, match="Wrong interval"):
        rows("-10 -20 A\n")




def test_schneider_rows_worked_out_by_hand():
    """Rows of HounsfieldUnit_to_material (materials.py:149-230) derived by hand from the
    committed tables. Lung, HU [-950, -120): both ends lie between the density nodes
    (-1000, 0.00121) and (-98, 0.93), so d(h) = 0.00121 + (h + 1000) 0.92879 / 902, the
    density span is d(-120) - d(-950) = 0.92879 * 830 / 902 = 0.854652 g/cm3, at a tolerance
    of 0.1 n = 8.54652 -> 9 sub-intervals of 830 / 8.54652 = 97.1156 HU (the last one
    clipped at -120), and sub-interval j has the midpoint density d(-950) + (j + 1/2) 0.1 =
    0.052695 + 0.1 (j + 1/2). Marrow_Bone01, HU [120, 200): between the nodes (101, 1.0762)
    and (1600, 1.9642) the span is 0.888 * 80 / 1499 = 0.047392 < 0.05: one material, at
    the midpoint density 1.0762 + 0.888 * 59 / 1499 = 1.111151. The Lung weights (H 10.3,
    C 10.5, N 3.1, O 74.9, Na 0.2, P 0.2, S 0.3, Cl 0.3, K 0.2) sum to 100."""
    from types import SimpleNamespace

    from opengate_amd.physics.materials import MaterialDatabase
    from opengate_amd.units import g4_units

    data = __import__("opengate_amd").__path__[0] + "/data/"
    for tol in (0.1, 0.05):
        db = MaterialDatabase()
        sim = SimpleNamespace(volume_manager=SimpleNamespace(material_database=db))
        vm, created = M.HounsfieldUnit_to_material(sim, tol * g4_units.g_cm3, data + "Schneider2000MaterialsTable.txt",
                                                   data + "Schneider2000DensitiesTable.txt")
        lung = [r for r in vm if r[2].startswith("Lung_")]
        marrow = [r for r in vm if r[2].startswith("Marrow_Bone01_")]
        assert len(marrow) == 1 and marrow[0][:2] == [120, 200]
        assert db.get(marrow[0][2]).density == pytest.approx(1.111151, rel=2e-6)  # g/cm3
        if tol == 0.1:
            assert len(lung) == 9
            step = 830 / (0.92879 * 830 / 902 / 0.1)
            assert step == pytest.approx(97.1156, abs=1e-4)
            for j, (lo, hi, name) in enumerate(lung):
                assert lo == pytest.approx(-950 + j * step, abs=1e-9)
                assert hi == pytest.approx(min(-950 + (j + 1) * step, -120), abs=1e-9)
                if j < 8:
                    assert db.get(name).density == pytest.approx(0.052695 + 0.1 * (j + 0.5),
                                                                rel=1e-5)
            m = db.get(lung[0][2])
            frac = dict(zip(m.elements, m.fractions))
            assert frac[8] == pytest.approx(0.749) and frac[1] == pytest.approx(0.103) and frac[19] == pytest.approx(0.002)
            assert len(frac) == 9
        else:
            # 0.854652 / 0.05 = 17.09 -> 18 sub-intervals
            assert len(lung) == 18


def test_density_image_from_intervals():
    """create_density_img (materials.py:294-326): density of the material whose [hu0, hu1)
    interval holds the voxel, 0 in gaps and outside; geometry copied from the image."""
    import opengate_amd as gate

    sim = gate.Simulation()
    ct = sim.add_volume("Image", "ct")
    hu = np.array([[[-1001, -1000, -501], [-500, -400, 99]], [[100, 1999, 2000], [3000, -300, 0]]], dtype=np.int16)
    ct.image = gate.Image(hu, (1.0, 2.0, 3.0), origin=(1.0, 2.0, 3.0))
    ct.voxel_materials = [[-1000, -500, "G4_AIR"], [-400, 100, "G4_WATER"], [100, 2000, "G4_BONE_CORTICAL_ICRP"]]
    rho = ct.create_density_image()
    db = sim.volume_manager.material_database
    a, w, b = (db.get(n).density for n in ("G4_AIR", "G4_WATER", "G4_BONE_CORTICAL_ICRP"))
    exp = np.array([[[0, a, a], [0, w, w]], [[b, b, 0], [0, w, w]]], dtype=np.float32)
    np.testing.assert_array_equal(rho.array, exp)
    np.testing.assert_array_equal(rho.spacing, (1.0, 2.0, 3.0))
