"""opengate_amd: MI355X-native photon transport and voxel dose behind the OpenGATE API.

    import opengate_amd as gate
    sim = gate.Simulation()
    img = sim.add_volume("Image", "patient") ...
    dose = sim.add_actor("DoseActor", "dose") ...
    sim.run()

The transport runs in hand-written HIP kernels (libgatehip.so, gfx950) reached through
the C-ABI of include/gatehip.h; there is no CPU fallback.
"""
from .units import g4_units
from .managers import Simulation
from .image import Image, read_image, write_image, create_3d_image
from .geometry import materials as _materials

HounsfieldUnit_to_material = _materials.HounsfieldUnit_to_material
read_voxel_materials = _materials.read_voxel_materials

__all__ = ["g4_units", "Simulation", "Image", "read_image", "write_image", "create_3d_image",
           "HounsfieldUnit_to_material", "read_voxel_materials"]
