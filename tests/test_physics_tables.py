"""Cross-section model and FP32 tables vs closed forms and the NIST data (CPU)."""
import numpy as np
import pytest

from opengate_amd.physics import elements as el
from opengate_amd.physics import xs
from opengate_amd.physics.materials import MaterialDatabase

DB = MaterialDatabase()
RE2 = xs.R_E_CM**2


def test_klein_nishina_closed_forms():
    # Thomson limit and the k = 1 value 2 pi r_e^2 [2(4/3 - ln 3) + ln3/2 - 4/9]
    assert xs.sigma_kn_electron(1e-9) == pytest.approx(8 * np.pi / 3 * RE2, rel=1e-6)
    k1 = 2 * np.pi * RE2 * (2 * (4 / 3 - np.log(3)) + np.log(3) / 2 - 4 / 9)
    assert xs.sigma_kn_electron(xs.MEC2) == pytest.approx(k1, rel=1e-12)
    # the angular integral of dsigma/dmu reproduces the total
    mu = np.linspace(-1, 1, 200001)
    for e in (0.01, 0.1405, 1.0):
        tot = np.trapezoid(xs.kn_dsigma_dmu(e, mu), mu)
        assert tot == pytest.approx(xs.sigma_kn_electron(e), rel=1e-6)


def test_incoherent_limits():
    # S(q) -> Z at large q: high-energy incoherent -> Z sigma_KN; small at low energy
    for z in (1, 8, 20):
        assert xs.sigma_incoherent_atom(z, [20.0])[0] == pytest.approx(z * xs.sigma_kn_electron(20.0), rel=2e-3)
        assert xs.sigma_incoherent_atom(z, [1e-4])[0] < 0.2 * z * xs.sigma_kn_electron(1e-4)
    # exact one-electron relation S = 1 - F^2 for hydrogen
    x = np.geomspace(1e-3, 1e3, 50)
    assert np.allclose(xs.incoherent_function(1, x), 1 - xs.form_factor(1, x) ** 2)


def test_rayleigh_limits_and_integral():
    for z in (1, 8, 82):
        # Thomson: sigma_coh(E -> 0) = Z^2 8 pi/3 r_e^2
        assert xs.sigma_rayleigh_atom(z, [1e-7])[0] == pytest.approx(z * z * 8 * np.pi / 3 * RE2, rel=1e-5)
        # closed-form int_0^X F^2 against a numerical quadrature
        X = 50.0
        x = np.concatenate([[0.0], np.geomspace(1e-8, X, 200001)])
        num = np.trapezoid(xs.form_factor(z, x) ** 2, x)
        assert xs.form_factor_sq_integral(z, np.array([X]))[0] == pytest.approx(num, rel=1e-6)
    s = xs.sigma_rayleigh_atom(8, [0.01, 0.03, 0.1, 0.3])
    assert np.all(np.diff(s) < 0)


def test_nist_mixture_at_common_nodes():
    # water at energies that are nodes of both H and O: exact mass-fraction sum
    w = DB.get("G4_WATER")
    eh, muh, _ = el.nist_table(1)
    eo, muo, _ = el.nist_table(8)
    for e in (0.01, 0.03, 0.1, 0.15, 0.2):
        ih, io = np.where(eh == e)[0][0], np.where(eo == e)[0][0]
        exp = w.fractions[0] * muh[ih] + w.fractions[1] * muo[io]
        assert xs.mu_rho(w, e)[0] == pytest.approx(exp, rel=1e-12)
    # log-log interpolation between nodes lies between the neighbours
    v = xs.mu_rho(w, [0.0125])[0]
    assert xs.mu_rho(w, [0.015])[0] < v < xs.mu_rho(w, [0.01])[0]


def test_partition_sums_to_total():
    for name in ("G4_WATER", "G4_BONE_CORTICAL_ICRP", "G4_Pb", "G4_AIR"):
        m = DB.get(name)
        e = np.geomspace(0.001, 0.5, 40)
        mu, c, r, p, _ = xs.partial_mass_attenuation(m, e)
        assert np.allclose(c + r + p, mu, rtol=1e-12)
        assert (c >= 0).all() and (r >= 0).all() and (p >= 0).all()
    # photoelectric dominates at low energy in lead
    mu, c, r, p, _ = xs.partial_mass_attenuation(DB.get("G4_Pb"), [0.05])
    assert p[0] / mu[0] > 0.9


def test_fp32_tables():
    mats = [DB.get(n) for n in ("G4_WATER", "G4_BONE_CORTICAL_ICRP", "G4_AIR")]
    T = xs.build_tables(mats, 0.2)
    e_nodes = xs.bucket_node_energies(T.n_buckets).astype(np.float64)
    assert e_nodes[0] == 2.0**-10 and e_nodes[-1] > 0.2
    # node energies are exactly the float-bit bucket grid
    bits = e_nodes.astype(np.float32).view(np.uint32)
    assert np.all(np.diff(bits >> xs.SHIFT) == 1) and np.all((bits & ((1 << xs.SHIFT) - 1)) == 0)
    for i, m in enumerate(mats):
        ref = xs.mu_rho(m, e_nodes) * m.density / 10.0
        assert np.allclose(T.mu_total[i], ref, rtol=2e-7)
    assert np.all(T.mu_majorant[None, :] >= T.mu_total)
    assert np.all(T.p_compton <= T.p_rayleigh) and np.all(T.p_rayleigh <= 1.0 + 1e-7)
    assert np.all(np.diff(T.ray_cdf, axis=1) >= 0) and np.allclose(T.ray_cdf[:, -1], 1.0)
    assert T.ray_x[0] == 0 and np.all(np.diff(T.ray_x) > 0)
    kk = e_nodes[-1] * xs.K_PER_MEV
    assert T.ray_x[-1] >= 4 * kk * kk


def test_majorant_over_the_materials_that_occur():
    """build_tables(majorant_materials=...): the majorant is the per-node maximum over the
    named materials only (x (1 + 2^-16)); all of them lie under it and a denser material
    left out does not."""
    mats = [DB.get("G4_WATER"), DB.get("G4_BONE_CORTICAL_ICRP"), DB.get("G4_Pb")]
    full = xs.build_tables(mats, 0.15)
    sub = xs.build_tables(mats, 0.15, majorant_materials=[0, 1])
    np.testing.assert_array_equal(full.mu_total, sub.mu_total)
    exp = np.maximum(sub.mu_total[0], sub.mu_total[1]).astype(np.float64) * (1 + 2.0**-16)
    np.testing.assert_array_equal(sub.mu_majorant, exp.astype(np.float32))
    assert np.all(sub.mu_total[:2] <= sub.mu_majorant)
    assert np.all(sub.mu_total[2] > sub.mu_majorant)
    assert np.all(full.mu_total[2] <= full.mu_majorant) and sub.majorant_materials == [0, 1]
    with pytest.raises(ValueError):
        xs.build_tables(mats, 0.15, majorant_materials=[])
