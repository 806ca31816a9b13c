"""Volumes of the photon path: the world and Box volumes (as mothers or homogeneous
phantoms) and ImageVolume (opengate/geometry/volumes.py:882-1294).

An ImageVolume is centred on its volume origin: voxel (0, 0, 0) has its centre at
-size*spacing/2 + spacing/2 (volumes.py:1289-1291). Its label image follows
create_material_to_label_lut / create_label_image (volumes.py:1144-1243); the digitize
runs on the GPU (gh_hu_to_labels).
"""
from pathlib import Path

import numpy as np

from ..base import GateObject
from ..image import Image, read_image


class VolumeBase(GateObject):
    user_info_defaults = {
        "mother": ("world", {"doc": "name of the mother volume"}),
        "material": ("G4_AIR", {"doc": "material name"}),
        "translation": ([0.0, 0.0, 0.0], {"doc": "centre relative to the mother [mm]"}),
        "rotation": (np.eye(3), {"doc": "rotation matrix relative to the mother"}),
        "color": ([1, 1, 1, 1], {"doc": "visualisation colour (unused)"}),
        "repeat": (None, {"doc": "repeated placements (not supported on the photon path)"}),
    }
    volume_type = "VolumeBase"
    runtime_attributes = ("volume_manager", "material_to_label_lut", "label_image")

    def __init__(self, name=None, **kw):
        super().__init__(name, **kw)
        self.volume_manager = None

    @property
    def type_name(self):
        return self.volume_type

    def _physics_manager(self):
        if self.volume_manager is None:
            raise RuntimeError(f"volume '{self.name}' is not in a simulation")
        return self.volume_manager.simulation.physics_manager

    # volumes.py:526-548 of the reference: per-volume physics settings create a region
    def set_production_cut(self, particle_name, value):
        self._physics_manager().set_production_cut(self.name, particle_name, value)

    def set_min_ekine(self, min_ekine):
        self._physics_manager().set_min_ekine(self.name, min_ekine)

    def set_max_step_size(self, max_step_size):
        self._physics_manager().set_max_step_size(self.name, max_step_size)


class BoxVolume(VolumeBase):
    user_info_defaults = {"size": ([10.0, 10.0, 10.0], {"doc": "full size [mm]"})}
    volume_type = "BoxVolume"


class ImageVolume(VolumeBase):
    user_info_defaults = {
        "voxel_materials": ([[-np.inf, np.inf, "G4_AIR"]], {"doc": "[[lo, hi, material], ...] half-open"}),
        "image": ("", {"doc": "path to the image (.mhd/.mha/.npy) or an opengate_amd.image.Image"}),
        "dump_label_image": (None, {"doc": "path where the label image is written"}),
    }
    volume_type = "ImageVolume"

    def __init__(self, name=None, **kw):
        super().__init__(name, **kw)
        self._image = None
        self.material_to_label_lut = None
        self.label_image = None

    # ---------------------------------------------------------------- image
    def load_input_image(self):
        if isinstance(self.image, Image):
            self._image = self.image
        elif isinstance(self.image, np.ndarray):
            self._image = Image(self.image)
        elif self.image:
            self._image = read_image(Path(self.image))
        else:
            raise ValueError(f"ImageVolume '{self.name}': no image set")
        return self._image

    @property
    def itk_image(self):
        if self._image is None:
            self.load_input_image()
        return self._image

    @property
    def size_pix(self):
        return self.itk_image.size

    @property
    def spacing(self):
        return np.array(self.itk_image.spacing)

    @property
    def native_translation(self):
        img = self.itk_image
        center = (img.size - 1.0) * img.spacing / 2.0
        return img.origin + img.direction @ center

    @property
    def native_rotation(self):
        return self.itk_image.direction.copy()

    # ---------------------------------------------------------------- labels
    def create_material_to_label_lut(self, material=None, voxel_materials=None):
        """label 0 = the volume's material, then unique materials in ascending lower bound
        order (volumes.py:1144-1168)."""
        voxel_materials = self.voxel_materials if voxel_materials is None else voxel_materials
        material = self.material if material is None else material
        lut = {material: 0}
        rows = sorted(voxel_materials, key=lambda r: r[0])
        lo = np.array([r[0] for r in rows], dtype=np.float64)
        hi = np.array([r[1] for r in rows], dtype=np.float64)
        if not (lo[1:] >= hi[:-1]).all():
            raise ValueError(f"Overlapping intervals in voxel_materials of volume {self.name}.")
        i = 1
        for r in rows:
            if r[2] not in lut:
                lut[r[2]] = i
                i += 1
        return lut

    def label_bins(self, lut=None):
        """(bins, labels_sorted) such that label = labels_sorted[digitize(HU, bins)]:
        lower bounds plus gap upper bounds, gaps and values below the first bound -> 0
        (volumes.py:1200-1243)."""
        lut = lut or self.create_material_to_label_lut()
        rows = sorted(self.voxel_materials, key=lambda r: r[0])
        bins = [r[0] for r in rows]
        extra = set([r[1] for r in rows]).difference(bins)
        bins.extend(extra)
        labels = [lut[r[2]] for r in rows] + [0] * len(extra)
        bins_sorted, labels_sorted = [], [0]
        for b, l in sorted(zip(bins, labels), key=lambda p: p[0]):
            bins_sorted.append(b)
            labels_sorted.append(l)
        return np.array(bins_sorted, dtype=np.float64), np.array(labels_sorted, dtype=np.uint16)

    def create_density_image(self):
        """Density [g/cm3] per voxel of the image (ImageVolume.create_density_image,
        volumes.py:1294-1295)."""
        from .materials import create_density_img

        return create_density_img(self, self.volume_manager.material_database)

    def create_label_image(self, engine):
        """uint16 label image on the GPU (gh_hu_to_labels)."""
        self.material_to_label_lut = self.create_material_to_label_lut()
        bins, labels_sorted = self.label_bins(self.material_to_label_lut)
        img = self.itk_image
        arr = img.array
        if arr.dtype != np.int16 and np.issubdtype(arr.dtype, np.integer) and arr.size and \
                arr.min() >= -32768 and arr.max() <= 32767:
            arr = arr.astype(np.int16)  # exact: the int16 kernel moves 2 bytes per voxel
        finite = np.clip(bins, -1e300, 1e300)
        lab = engine.hu_to_labels(arr, finite, labels_sorted)
        self.label_image = Image(lab, img.spacing, -(img.size * img.spacing) / 2.0 + img.spacing / 2.0)
        if self.dump_label_image:
            from ..image import write_image
            import json

            write_image(Image(lab, img.spacing, img.origin, img.direction), self.dump_label_image)
            Path(self.dump_label_image).with_suffix(".json").write_text(json.dumps(self.material_to_label_lut))
        return self.label_image
