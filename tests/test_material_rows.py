"""Materials that share an elemental composition share table rows (mu = density x row, the
layout that lets the Schneider HU conversion's many density steps run with small tables),
and uint16 label images (the reference's label image type, geometry/volumes.py:1237).

CPU pins: with power-of-two densities the FP32 product density x row is exact, so the
shared-row tables and the per-material tables give the same mu bit for bit and the oracle's
transport must give identical results with either; uint16 and uint8 labels of the same
image give identical runs; the majorant covers every material's mu as the transport forms
it; rows are grouped by composition."""
import numpy as np
import pytest

from oracle import oracle as O
from opengate_amd import workloads
from opengate_amd.physics import xs
from opengate_amd.physics.materials import Material, MaterialDatabase
from opengate_amd.transport import PhantomSpec, SourceSpec, TallySpec

DB = MaterialDatabase()


def scaled(base, rho, name):
    m = DB.get(base)
    return Material(name, float(rho), m.elements, m.fractions)


def water_family():
    # power-of-two densities only: density x row is then exact in FP32
    return [DB.get("G4_WATER"), scaled("G4_WATER", 2.0, "W2"), scaled("G4_WATER", 0.5, "W05"),
            scaled("G4_BONE_CORTICAL_ICRP", 2.0, "B2"), scaled("G4_BONE_CORTICAL_ICRP", 0.25, "B025")]


def layered_phantom(n_mat, shape=(20, 18, 16), spacing=2.5, seed=3):
    g = np.random.default_rng(seed)
    nz, ny, nx = shape
    # slabs of random materials (long enough runs of one label for many steps per voxel)
    lab = np.repeat(g.integers(0, n_mat, size=(nz, 1, 1)), ny, axis=1).repeat(nx, axis=2)
    lab[:, :, nx // 2:] = g.integers(0, n_mat, size=(nz, ny, nx - nx // 2))
    return lab


def test_rows_are_grouped_by_composition():
    mats = water_family()
    T = xs.build_tables(mats, 0.15)
    assert T.n_compositions == 2
    assert list(T.composition) == [0, 0, 0, 1, 1]
    Tp = xs.build_tables(mats, 0.15, share_compositions=False)
    assert Tp.composition is None and Tp.mu_total.shape[0] == 5
    # the same material set without repeats keeps a row per material
    T1 = xs.build_tables([DB.get("G4_WATER"), DB.get("G4_AIR")], 0.15)
    assert T1.composition is None


def test_power_of_two_densities_give_identical_mu_and_runs():
    mats = water_family()
    Ts = xs.build_tables(mats, 0.15)
    Tp = xs.build_tables(mats, 0.15, share_compositions=False)
    for m in range(len(mats)):
        np.testing.assert_array_equal(Ts.material_mu(m), Tp.mu_total[m])
    np.testing.assert_array_equal(Ts.mu_majorant, Tp.mu_majorant)
    ph = PhantomSpec(layered_phantom(len(mats)), (2.5, 2.5, 2.5))
    src = SourceSpec(pos_trans=(20.0, 22.5, 25.0), energy=0.1405)
    tal = TallySpec(size=(16, 18, 20), spacing=(2.5, 2.5, 2.5), dose=True)
    rs, cs = O.run(Ts, ph, src, tal, 20000, 5, threads=4)
    rp, cp = O.run(Tp, ph, src, tal, 20000, 5, threads=4)
    for k in ("histories", "steps", "deposits", "compton", "rayleigh", "photo", "escaped", "kn_rejections"):
        assert cs[k] == cp[k], k
    for k in ("edep", "edep2", "dose"):
        assert np.max(np.abs(rs[k] - rp[k])) <= 1e-12 * rp[k].max(), k
    # a density that is not a power of two changes mu at the last bit only
    m3 = mats + [scaled("G4_WATER", 1.3, "W13")]
    Ts3 = xs.build_tables(m3, 0.15)
    Tp3 = xs.build_tables(m3, 0.15, share_compositions=False)
    rel = np.abs(Ts3.material_mu(5).astype(np.float64) / Tp3.mu_total[5] - 1.0)
    assert rel.max() <= 2.0 ** -22


def test_shared_rows_interpolate_node_products():
    """With shared composition rows, material m's mu at E interpolates its node values
    density x row (FP32 products): at a node it is that product, between nodes the FP32
    interpolation of the two products; a per-material table of those products gives the
    same runs (the kernel's LDS expansion of shared rows relies on it)."""
    m3 = water_family() + [scaled("G4_WATER", 1.3, "W13")]
    Ts = xs.build_tables(m3, 0.15)
    nodes = xs.bucket_node_energies(Ts.n_buckets)
    prod = Ts.material_mu(5)
    for j in (0, 40, 200, Ts.n_buckets - 1):
        assert O.mu_at(Ts, 5, nodes[j]) == prod[j]
    j = 250
    e_mid = np.float32((np.float64(nodes[j]) + np.float64(nodes[j + 1])) / 2)
    v = O.mu_at(Ts, 5, e_mid)
    assert min(prod[j], prod[j + 1]) <= v <= max(prod[j], prod[j + 1])
    assert v == pytest.approx(0.5 * (float(prod[j]) + float(prod[j + 1])), rel=2e-7)
    # per-material tables built from the products run identically
    import copy

    Te = copy.deepcopy(Ts)
    Te.mu_total = np.stack([Ts.material_mu(m) for m in range(len(m3))]).astype(np.float32)
    Te.composition = None
    Te.density = Ts.density.copy()
    rows = [Ts.row_of(m) for m in range(len(m3))]
    for name in ("p_compton", "p_rayleigh", "mu_en", "ray_cdf", "incoh_s"):
        setattr(Te, name, getattr(Ts, name)[rows].copy())
    Te.mu_en = np.stack([np.float32(Ts.density[m]) * Ts.mu_en[rows[m]] for m in range(len(m3))]).astype(np.float32)
    ph = PhantomSpec(layered_phantom(len(m3)), (2.5, 2.5, 2.5))
    src = SourceSpec(pos_trans=(20.0, 22.5, 25.0), energy=0.1405)
    tal = TallySpec(size=(16, 18, 20), spacing=(2.5, 2.5, 2.5), dose=True)
    rs, cs = O.run(Ts, ph, src, tal, 20000, 5, threads=4)
    re, ce = O.run(Te, ph, src, tal, 20000, 5, threads=4)
    assert cs == ce
    np.testing.assert_array_equal(rs["edep"], re["edep"])


def test_uint16_labels_run_like_uint8_labels():
    mats = water_family()
    T = xs.build_tables(mats, 0.15)
    lab = layered_phantom(len(mats))
    p8 = PhantomSpec(lab.astype(np.uint8), (2.5, 2.5, 2.5))
    p16 = PhantomSpec(lab.astype(np.uint16), (2.5, 2.5, 2.5))
    assert p8.desc().label_bytes == 1 and p16.desc().label_bytes == 2
    src = SourceSpec(pos_trans=(20.0, 22.5, 25.0), energy=0.1405)
    tal = TallySpec(size=(16, 18, 20), spacing=(2.5, 2.5, 2.5))
    r8, c8 = O.run(T, p8, src, tal, 5000, 9)
    r16, c16 = O.run(T, p16, src, tal, 5000, 9)
    assert c8 == c16
    np.testing.assert_array_equal(r8["edep"], r16["edep"])


def test_phantom_spec_label_types():
    assert PhantomSpec(np.full((2, 2, 2), 300, np.int32), (1, 1, 1)).labels.dtype == np.uint16
    assert PhantomSpec(np.full((2, 2, 2), 3, np.int64), (1, 1, 1)).labels.dtype == np.uint8
    with pytest.raises(ValueError):
        PhantomSpec(np.full((2, 2, 2), 70000, np.int64), (1, 1, 1))


class _HostLabels:
    def hu_to_labels(self, hu, bins, labels_sorted):
        return labels_sorted[np.digitize(hu, bins)]


def test_schneider_materials_share_rows_and_fit_the_majorant():
    T, ph, src, tal, info = workloads.ct256_schneider(_HostLabels(), 0.01)
    assert info["n_materials"] == 287 and T.n_compositions < 40
    mu = np.stack([T.material_mu(m) for m in range(T.n_materials)]).astype(np.float64)
    # the majorant covers the materials that occur in the phantom, and only those: the
    # table's dental amalgam (HU >= 2300) is not in it and lies far above
    present = np.flatnonzero(np.bincount(ph.labels.ravel(), minlength=T.n_materials))
    assert T.majorant_materials == list(present)
    assert (mu[present] <= T.mu_majorant[None, :]).all()
    amalgam = [m for m, n in enumerate(info["names"]) if n.startswith("AmalgamTooth")]
    assert amalgam and not set(amalgam) & set(present)
    assert (mu[amalgam].max(0) > 10 * T.mu_majorant).any()
    # the majorant is the per-node maximum raised by 2^-16 (no looser)
    assert np.all(T.mu_majorant <= mu[present].max(0) * (1 + 2.0 ** -15))
    # the sub-materials of one Schneider row share its composition row
    names = info["names"]
    rows = {}
    for m, n in enumerate(names[1:], start=1):
        rows.setdefault(n.rsplit("_", 1)[0], set()).add(int(T.composition[m]))
    assert all(len(v) == 1 for v in rows.values())
