"""Geant4 system of units (mm, MeV, ns = 1), as exposed by opengate.g4_units."""
import math


class _Units(dict):
    def __getattr__(self, k):
        try:
            return self[k]
        except KeyError:
            raise AttributeError(k) from None


def _make():
    u = _Units()
    u.update(mm=1.0, millimeter=1.0, um=1e-3, nm=1e-6, cm=10.0, centimeter=10.0, m=1000.0, meter=1000.0, km=1e6)
    u.update(mm2=1.0, cm2=100.0, m2=1e6, mm3=1.0, cm3=1000.0, m3=1e9)
    u.update(rad=1.0, radian=1.0, mrad=1e-3, deg=math.pi / 180.0, degree=math.pi / 180.0)
    u.update(ns=1.0, nanosecond=1.0, s=1e9, second=1e9, ms=1e6, us=1e3, ps=1e-3, min=60e9, h=3600e9)
    u.update(MeV=1.0, eV=1e-6, keV=1e-3, GeV=1e3, TeV=1e6, PeV=1e9)
    e_charge = 1.602176634e-19
    joule = 1.0 / (e_charge * 1e6)  # MeV per joule in internal units -> 1 J = 6.2415e12 MeV
    u.update(joule=joule)
    kg = joule * 1e18 / 1e6  # joule * s^2 / m^2 with s = 1e9, m = 1e3
    u.update(kg=kg, kilogram=kg, g=kg * 1e-3, gram=kg * 1e-3, mg=kg * 1e-6)
    u.update(g_cm3=u["g"] / u["cm3"], mg_cm3=u["mg"] / u["cm3"], kg_m3=kg / u["m3"])
    u.update(mole=1.0, g_mole=u["g"])
    # unit strings of material database files
    u.update({"g/cm3": u["g_cm3"], "mg/cm3": u["mg_cm3"], "kg/m3": u["kg_m3"], "g/mole": u["g"], "mg/mole": u["mg"],
              "kg/mole": kg})
    u.update(Gy=joule / kg, gray=joule / kg, mGy=1e-3 * joule / kg, uGy=1e-6 * joule / kg)
    u.update(Bq=1.0 / u["s"], becquerel=1.0 / u["s"], kBq=1e3 / u["s"], MBq=1e6 / u["s"], GBq=1e9 / u["s"])
    u.update(mCi=3.7e7 / u["s"], Ci=3.7e10 / u["s"], uCi=3.7e4 / u["s"])
    u.update(percent=0.01, perCent=0.01)
    return u


g4_units = _make()
