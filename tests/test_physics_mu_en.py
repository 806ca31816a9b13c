"""The transport model's energy-transfer coefficient against the reference's NIST
energy-absorption coefficients (GateMuDatabase.h, mixed by mass fraction), 10-150 keV, for
the materials of the workloads.

mu_tr of the model = photoelectric (all of E given locally; fluorescence is off) + Compton
with the mean electron energy under Klein-Nishina x S(q) (quadrature). NIST mu_en is below
mu_tr by the fluorescence energy that escapes; for cortical bone (Ca, P) the K-shell part
is P_K w_K E_Ka / E (Ca: 0.89 x 0.163 x 3.69 keV; P: 0.9 x 0.063 x 2.01 keV) of the
photoelectric share, added back here. Measured deviation of the model: within 2 % below
20 keV, down to -9.5 % at 40-60 keV for the soft tissues (the photoelectric part is the
remainder mu - sigma_incoh - sigma_coh, a small difference of large terms there, so the
few-percent error of the Thomas-Fermi scattering functions moves into it), within 4 % for bone once the K x-rays are counted (4.3 % above at 10 keV
without them). The tolerance (-11 % .. +5 %) encodes that; a
free-electron Compton model (S = Z) fails it."""
from types import SimpleNamespace

import numpy as np
import pytest

from opengate_amd.physics import elements as el
from opengate_amd.physics import xs
from helpers import DB

NAMES = ["G4_WATER", "G4_BONE_CORTICAL_ICRP", "G4_LUNG_ICRP", "G4_ADIPOSE_TISSUE_ICRP", "G4_MUSCLE_SKELETAL_ICRP",
         "G4_AIR"]
ENERGIES = [0.010, 0.015, 0.020, 0.030, 0.040, 0.050, 0.060, 0.080, 0.100, 0.150]
K_FLUO_KEV = {20: 0.89 * 0.163 * 3.69, 15: 0.90 * 0.063 * 2.01}  # P_K w_K E_Ka [keV]


def model_mu_tr_fraction(mat, e, free_electron=False):
    """(mu_tr / mu, photoelectric share / mu) of the model at energies e [MeV]."""
    mu, compt, rayl, photo, muen = xs.partial_mass_attenuation(mat, e)
    y, wq = xs._Y, xs._YW
    cos = 1.0 - y
    num = np.zeros_like(e)
    den = np.zeros_like(e)
    for z, w in zip(mat.elements, mat.fractions):
        n_atoms = w / el.atomic_mass(z)
        kk = e[:, None] * xs.K_PER_MEV
        x = 2 * kk * kk * y[None, :]
        s = z if free_electron else z - xs.form_factor(z, x) ** 2 / z
        kn = xs.kn_dsigma_dmu(e[:, None], cos[None, :])
        r = 1.0 / (1.0 + e[:, None] / xs.MEC2 * y[None, :])
        num += n_atoms * np.sum(kn * s * (1.0 - r) * wq[None, :], axis=1)
        den += n_atoms * np.sum(kn * s * wq[None, :], axis=1)
    f_tr = num / den
    if free_electron:
        # free electrons: sigma_incoh = Z sigma_KN, the photoelectric remainder shrinks
        compt_f = np.zeros_like(e)
        for z, w in zip(mat.elements, mat.fractions):
            compt_f += w * xs.N_A / el.atomic_mass(z) * z * xs.sigma_kn_electron(e)
        photo = np.maximum(mu - compt_f - rayl, 0.0)
        compt = compt_f
    return (photo + compt * f_tr) / mu, photo, mu, muen


def fluorescence_escape(mat, e, photo):
    """Energy per unit mu [fraction of mu] carried away by the K x-rays of Ca and P after
    photoabsorption: sum over those elements of (their photoelectric coefficient in the
    mixture) x P_K w_K E_Ka / E."""
    out = np.zeros_like(e)
    for z, w in zip(mat.elements, mat.fractions):
        if z in K_FLUO_KEV:
            pure = SimpleNamespace(elements=[z], fractions=np.array([1.0]))
            out += w * xs.partial_mass_attenuation(pure, e)[3] * K_FLUO_KEV[z] / (e * 1000.0)
    return out


@pytest.mark.parametrize("name", NAMES)
def test_model_energy_transfer_tracks_nist_mu_en(name):
    mat = DB.get(name)
    e = np.array(ENERGIES)
    f_tr, photo, mu, muen = model_mu_tr_fraction(mat, e)
    nist_tr = (muen + fluorescence_escape(mat, e, photo)) / mu
    rel = f_tr / nist_tr - 1.0
    assert np.all(rel > -0.11) and np.all(rel < 0.05), dict(zip(ENERGIES, np.round(rel, 3)))
    # below 20 keV the photoelectric effect dominates and the model is within 2 %
    low = e <= 0.0151
    assert np.all(np.abs(rel[low]) < 0.02)


def test_bone_without_fluorescence_correction_fails_the_low_energy_pin():
    """At 10 keV nearly every photon is absorbed; Ca K x-rays carry 4.3 % of the energy
    away in NIST mu_en, which the model (no fluorescence) deposits locally."""
    mat = DB.get("G4_BONE_CORTICAL_ICRP")
    e = np.array([0.010])
    f_tr, photo, mu, muen = model_mu_tr_fraction(mat, e)
    assert f_tr[0] / (muen[0] / mu[0]) - 1.0 > 0.03
    assert abs(f_tr[0] / ((muen[0] + fluorescence_escape(mat, e, photo)[0]) / mu[0]) - 1.0) < 0.005


def test_free_electron_compton_model_fails_the_pin():
    mat = DB.get("G4_WATER")
    e = np.array(ENERGIES)
    f_tr, photo, mu, muen = model_mu_tr_fraction(mat, e, free_electron=True)
    rel = f_tr / (muen / mu) - 1.0
    assert rel.min() < -0.11
