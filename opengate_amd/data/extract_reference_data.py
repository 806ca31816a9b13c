"""Regenerate the numeric data files of this package from the OpenGATE source tree.

The files written here are physics DATA, not code: NIST per-element mass
attenuation / mass energy-absorption tables (Hubbell & Seltzer), element Z/A
values, the Schneider 2000 HU tables and the radionuclide gamma lines. They are
read from the reference tree once, on the build machine, and committed as plain
text so that nothing at run time needs the reference.

Sources (reference tree):
  core/opengate_core/opengate_lib/GateMuDatabase.h:17-26   NIST mu, mu_en (Z=1..92)
  opengate/data/PhotonAttenuation.py:7939-8042             element Z/A (PropsEl)
  opengate/data/Schneider2000MaterialsTable.txt             HU -> composition
  opengate/data/Schneider2000DensitiesTable.txt             HU -> density
  opengate/data/rad_gamma_spectrum.json                     gamma lines
  opengate/contrib/GateMaterials.db                         GATE material database

Usage: python opengate_amd/data/extract_reference_data.py [/root/reference]
"""
import ast
import json
import re
import shutil
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent


def _c_array(text, name):
    m = re.search(re.escape(name) + r"\[\d+\]\s*=\s*\{(.*?)\};", text, re.S)
    if m is None:
        raise RuntimeError(f"array {name} not found")
    return [float(x) for x in m.group(1).replace("\n", " ").split(",") if x.strip()]


def extract_nist(ref):
    text = (ref / "core/opengate_core/opengate_lib/GateMuDatabase.h").read_text()
    counts = [int(x) for x in _c_array(text, "NIST_mu_muen_data_energyNumber")]
    data = _c_array(text, "NIST_mu_muen_data")
    rows = []
    k = 0
    for z in range(1, 93):
        for _ in range(counts[z]):
            e, mu, muen = data[k : k + 3]
            rows.append((z, e, mu, muen))
            k += 3
    assert k == len(data), (k, len(data))
    with open(HERE / "nist_mu_muen.csv", "w") as f:
        f.write("# NIST mass attenuation and mass energy-absorption coefficients\n")
        f.write("# (J.H. Hubbell and S.M. Seltzer), per element, Z=1..92.\n")
        f.write("# columns: Z, energy [MeV], mu/rho [cm2/g], mu_en/rho [cm2/g]\n")
        for z, e, mu, muen in rows:
            f.write(f"{z},{e:.5e},{mu:.5e},{muen:.5e}\n")


def extract_elements(ref):
    text = (ref / "opengate/data/PhotonAttenuation.py").read_text()
    start = text.index("PropsEl = np.array(")
    body = text[start + len("PropsEl = np.array(") :]
    body = body[: body.index(")  # last element")]
    rows = ast.literal_eval(body.strip())
    with open(HERE / "elements.csv", "w") as f:
        f.write("# element table: Z, symbol, name, Z/A [mol/g]\n")
        for z, r in enumerate(rows, start=1):
            f.write(f"{z},{r[2]},{r[3]},{float(r[0]):.5f}\n")


def copy_tables(ref):
    for name in ("Schneider2000MaterialsTable.txt", "Schneider2000DensitiesTable.txt"):
        shutil.copyfile(ref / "opengate/data" / name, HERE / name)
    shutil.copyfile(ref / "opengate/contrib/GateMaterials.db", HERE / "GateMaterials.db")
    d = json.loads((ref / "opengate/data/rad_gamma_spectrum.json").read_text())
    with open(HERE / "rad_gamma_lines.json", "w") as f:
        json.dump(d, f, indent=1)


def main():
    ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    extract_nist(ref)
    extract_elements(ref)
    copy_tables(ref)


if __name__ == "__main__":
    main()
