"""Small shared builders for the tests (phantoms, tables)."""
from functools import lru_cache

import numpy as np

from opengate_amd.physics import xs
from opengate_amd.physics.materials import MaterialDatabase
from opengate_amd.transport import PhantomSpec, SourceSpec, TallySpec

DB = MaterialDatabase()


@lru_cache(maxsize=8)
def tables(names=("G4_WATER", "G4_BONE_CORTICAL_ICRP"), e_max=0.1405):
    return xs.build_tables([DB.get(n) for n in names], e_max)


def small_phantom(n=32, spacing=2.0):
    lab = np.zeros((n, n, n), np.uint8)
    q = n // 3
    lab[q:2 * q, q:2 * q, q:2 * q] = 1
    return PhantomSpec(lab, (spacing, spacing, spacing))


def centre(ph):
    nx, ny, nz = ph.shape_xyz
    return (nx * ph.spacing[0] / 2, ny * ph.spacing[1] / 2, nz * ph.spacing[2] / 2)


def aligned_tally(ph, **kw):
    return TallySpec(size=ph.shape_xyz, spacing=ph.spacing, **kw)
