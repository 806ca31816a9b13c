"""HU -> material conversion and voxel material tables.

Restates the semantics of opengate/geometry/materials.py:
  read_voxel_materials            (:14-58)
  HU_read_materials_table         (:61-98)
  HU_read_density_table           (:101-113)
  HU_linear_interpolate_densities (:116-129)
  HU_find_max_density_difference  (:132-146)
  HounsfieldUnit_to_material      (:149-230)
  create_density_img              (:294-326)
"""
import math

import numpy as np

from ..physics.elements import symbol_of, z_of


def read_voxel_materials(filename, def_mat="G4_AIR"):
    """Text table of (hu_min, hu_max, material) triples -> [[lo, hi, name], ...], as the
    reference reads it (materials.py:14-58): the words of the file form triples in order
    (a word starting with '#' ends its line), the triples are sorted, and a gap before an
    interval adds a def_mat row [previous, lo] followed by the interval's row starting at
    `previous` as well (so a gap makes two rows overlap, which the label table then
    rejects); a previous bound of 0 counts as no bound. Overlapping intervals and an
    interval with lo > hi are errors."""
    triples, state, start, stop = [], 0, None, None
    with open(filename) as f:
        for line in f:
            for word in line.split():
                if word[0] == "#":
                    break
                if state == 0:
                    start, state = float(word), 1
                elif state == 1:
                    stop, state = float(word), 2
                else:
                    triples.append([start, stop, word])
                    state = 0
    triples.sort()
    rows = []
    previous = triples[0][0]
    for lo, hi, name in triples:
        if previous and previous > lo:
            raise ValueError(f"Error while reading {filename}\nIntervals are not disjoint: {previous} {[lo, hi, name]}")
        if lo > hi:
            raise ValueError(f"Error while reading {filename}\nWrong interval {[lo, hi, name]}")
        if not previous or previous == lo:
            rows.append([previous, hi, name])
        else:
            rows.append([previous, lo, def_mat])
            rows.append([previous, hi, name])
        previous = hi
    return rows


def HU_read_materials_table(file_mat):
    elements = ["HU"]
    materials = []
    section = None
    with open(file_mat) as f:
        for line in f:
            cur = None
            for i, word in enumerate(line.split()):
                if word[0] == "#":
                    break
                if word == "[Elements]":
                    section = "element"
                    break
                if word == "[/Elements]":
                    elements.append("name")
                    section = "table"
                    break
                if section is None:
                    break
                if section == "element":
                    elements.append(word)
                elif section == "table":
                    cur = cur if cur is not None else {}
                    if i == 0:
                        cur[elements[i]] = int(word)
                    elif i == len(elements) - 1:
                        cur[elements[i]] = word
                    else:
                        cur[elements[i]] = float(word)
            if cur:
                materials.append(cur)
    return materials, elements


def HU_read_density_table(file_density):
    out = []
    with open(file_density) as f:
        for line in f:
            w = line.split()
            if not w or w[0][0] == "#":
                continue
            out.append({"HU": float(w[0]), "density": float(w[1])})
    return out


def HU_linear_interpolate_densities(hu, densities):
    n = len(densities)
    i = 0
    while i < n and hu > densities[i]["HU"]:
        i += 1
    i -= 1
    if i < 0:
        return densities[0]["density"]
    if i >= n - 1:
        return densities[n - 1]["density"]
    a, b = densities[i], densities[i + 1]
    return (hu - a["HU"]) / (b["HU"] - a["HU"]) * (b["density"] - a["density"]) + a["density"]


def HU_find_max_density_difference(hu_min, hu_max, d_min, d_max, densities):
    n = len(densities)
    i = 0
    while i < n and hu_min > densities[i]["HU"]:
        i += 1
    j = 0
    while j < n and hu_max > densities[j]["HU"]:
        j += 1
    j -= 1
    for x in range(i, j):
        d_min = min(d_min, densities[x]["density"])
        d_max = max(d_max, densities[x]["density"])
    return d_max - d_min


def HounsfieldUnit_to_material(simulation, density_tolerance, file_mat, file_density):
    """Schneider HU intervals split into sub-intervals of density width <= tolerance.

    density_tolerance is in Geant4 density units (g_cm3 = 6.24150907e18). Returns
    (voxel_materials, created_material_names) and registers the materials in
    simulation.volume_manager.material_database (materials.py:149-230)."""
    from ..units import g4_units

    gcm3 = g4_units.g_cm3
    materials, elements = HU_read_materials_table(file_mat)
    densities = HU_read_density_table(file_density)
    elems = elements[1:-1]
    symbols = [symbol_of(z_of(e)) for e in elems]
    db = simulation.volume_manager.material_database
    voxel_materials, created = [], []
    num = 0
    last = len(materials) - 1
    for i, mat in enumerate(materials):
        hu_min = mat["HU"]
        hu_max = hu_min + 1 if i == last else materials[i + 1]["HU"]
        if hu_max <= hu_min:
            raise ValueError(f"Error, HU interval not valid: {mat}")
        dmin = HU_linear_interpolate_densities(hu_min, densities)
        dmax = HU_linear_interpolate_densities(hu_max, densities)
        ddiff = HU_find_max_density_difference(hu_min, hu_max, dmin, dmax, densities)
        n = max(1, ddiff * gcm3 / density_tolerance)
        if "Air" in mat["name"] or "AIR" in mat["name"]:
            n = 1
        htol = (hu_max - hu_min) / n
        if n != 1:
            n = int(math.ceil(n))
        for j in range(n):
            h1 = hu_min + j * htol
            h2 = min(hu_min + (j + 1) * htol, hu_max)
            d = HU_linear_interpolate_densities(h1 + (h2 - h1) / 2.0, densities)
            w = [mat[e] for e in elems]
            nz = [(a, s) for a, s in zip(w, symbols) if a > 0]
            tot = sum(a for a, _ in nz)
            name = f'{mat["name"]}_{num}'
            db.add_material_weights(name, [s for _, s in nz], [a / tot for a, _ in nz], d * gcm3)
            voxel_materials.append([h1, h2, name])
            created.append(name)
            num += 1
    return voxel_materials, created


def create_density_img(img_volume, material_database):
    """Density image [g/cm3] of an ImageVolume: each voxel gets the density of the material
    whose [hu0, hu1) interval holds its value, 0 outside every interval; same size and
    geometry as the image (materials.py:294-326)."""
    from ..image import Image

    img = img_volume.itk_image
    hu = np.asarray(img.array)
    rho = np.zeros(hu.shape, dtype=np.float32)
    for hu0, hu1, name in img_volume.voxel_materials:
        rho[(hu >= hu0) & (hu < hu1)] = material_database.get(name).density
    return Image(rho, img.spacing, img.origin, img.direction)
