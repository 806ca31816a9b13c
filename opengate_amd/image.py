"""Images with ITK geometry (array[z, y, x], spacing, origin at the first voxel centre,
direction) and MetaImage (.mhd/.raw, .mha) reading and writing, the formats the
DoseActor outputs use (opengate/image.py:500 write_itk_image)."""
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

_MET = {
    np.dtype(np.uint8): "MET_UCHAR", np.dtype(np.int8): "MET_CHAR", np.dtype(np.uint16): "MET_USHORT",
    np.dtype(np.int16): "MET_SHORT", np.dtype(np.uint32): "MET_UINT", np.dtype(np.int32): "MET_INT",
    np.dtype(np.float32): "MET_FLOAT", np.dtype(np.float64): "MET_DOUBLE",
}
_MET_INV = {v: k for k, v in _MET.items()}


@dataclass
class Image:
    array: np.ndarray  # [z, y, x]
    spacing: np.ndarray = field(default_factory=lambda: np.ones(3))
    origin: np.ndarray = field(default_factory=lambda: np.zeros(3))
    direction: np.ndarray = field(default_factory=lambda: np.eye(3))

    def __post_init__(self):
        self.spacing = np.asarray(self.spacing, dtype=np.float64)
        self.origin = np.asarray(self.origin, dtype=np.float64)
        self.direction = np.asarray(self.direction, dtype=np.float64).reshape(3, 3)

    @property
    def size(self):
        """(nx, ny, nz) like itk.size"""
        return np.array(self.array.shape[::-1], dtype=int)

    def copy_information(self, other):
        self.spacing = other.spacing.copy()
        self.origin = other.origin.copy()
        self.direction = other.direction.copy()

    def GetSpacing(self):
        return tuple(self.spacing)

    def GetOrigin(self):
        return tuple(self.origin)

    def GetDirection(self):
        return self.direction.copy()


def create_3d_image(size_xyz, spacing, origin=None, dtype=np.float64, fill=0):
    """opengate.image.create_3d_image (image.py:40): origin defaults to the
    centred convention -size*spacing/2 + spacing/2."""
    size_xyz = np.asarray(size_xyz, dtype=int)
    spacing = np.asarray(spacing, dtype=np.float64)
    if origin is None:
        origin = -size_xyz * spacing / 2.0 + spacing / 2.0
    arr = np.full(tuple(size_xyz[::-1]), fill, dtype=dtype)
    return Image(arr, spacing, np.asarray(origin, dtype=np.float64))


def write_image(img: Image, path):
    path = Path(path)
    if path.suffix == ".npy":
        np.save(path, img.array)
        return path
    if path.suffix not in (".mhd", ".mha"):
        path = path.with_suffix(".mhd")
    arr = np.ascontiguousarray(img.array)
    if arr.dtype not in _MET:
        arr = arr.astype(np.float64)
    hdr = [
        "ObjectType = Image", "NDims = 3", "BinaryData = True", "BinaryDataByteOrderMSB = False",
        "CompressedData = False",
        "TransformMatrix = " + " ".join(f"{v:.17g}" for v in img.direction.T.ravel()),
        "Offset = " + " ".join(f"{v:.17g}" for v in img.origin),
        "CenterOfRotation = 0 0 0", "AnatomicalOrientation = RAI",
        "ElementSpacing = " + " ".join(f"{v:.17g}" for v in img.spacing),
        "DimSize = " + " ".join(str(int(v)) for v in arr.shape[::-1]),
        f"ElementType = {_MET[arr.dtype]}",
    ]
    if path.suffix == ".mha":
        hdr.append("ElementDataFile = LOCAL")
        with open(path, "wb") as f:
            f.write(("\n".join(hdr) + "\n").encode())
            f.write(arr.tobytes())
    else:
        raw = path.with_suffix(".raw")
        hdr.append(f"ElementDataFile = {raw.name}")
        path.write_text("\n".join(hdr) + "\n")
        raw.write_bytes(arr.tobytes())
    return path


def read_image(path) -> Image:
    path = Path(path)
    if path.suffix == ".npy":
        return Image(np.load(path, allow_pickle=False))
    data = path.read_bytes()
    header = {}
    offset = 0
    for line in data.split(b"\n"):
        offset += len(line) + 1
        if b"=" not in line:
            continue
        k, v = line.decode().split("=", 1)
        header[k.strip()] = v.strip()
        if k.strip() == "ElementDataFile":
            break
    dims = [int(x) for x in header["DimSize"].split()]
    dt = _MET_INV[header["ElementType"]]
    if header.get("BinaryDataByteOrderMSB", "False") == "True":
        dt = dt.newbyteorder(">")
    n = int(np.prod(dims))
    if header["ElementDataFile"] == "LOCAL":
        buf = data[offset: offset + n * dt.itemsize]
    else:
        buf = (path.parent / header["ElementDataFile"]).read_bytes()
    arr = np.frombuffer(buf, dtype=dt, count=n).reshape(dims[::-1]).astype(dt.newbyteorder("="))
    spacing = [float(x) for x in header.get("ElementSpacing", "1 1 1").split()]
    origin = [float(x) for x in header.get("Offset", header.get("Origin", "0 0 0")).split()]
    tm = header.get("TransformMatrix")
    direction = np.array([float(x) for x in tm.split()]).reshape(3, 3).T if tm else np.eye(3)
    return Image(arr, spacing, origin, direction)
