"""ctypes mirror of include/gatehip.h (structures, enums) and the library loader.

The host side reaches the GPU only through these C entry points. If the compiled
library is missing the import of the runtime fails loudly: there is no CPU fallback.
"""
import ctypes as C
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libgatehip.so"

GH_OK, GH_EINVAL, GH_EHIP, GH_ESTATE, GH_ERANGE = 0, 1, 2, 3, 4
GEMM_NOALLOC_WS_BYTES = 16 << 20  # GH_GEMM_NOALLOC_WS_BYTES
POS = {"point": 0, "box": 1, "sphere": 2, "ellipsoid": 3, "disc": 4, "cylinder": 5, "phsp": 6, "voxels": 7,
       "surface_sphere": 8}
DIR = {"iso": 0, "momentum": 1, "focused": 2, "cos": 3, "beam2d": 4, "histogram": 5}
ENE = {"mono": 0, "spectrum_discrete": 1, "spectrum_histogram": 2, "range": 3, "gauss": 4}
TALLY_EDEP2, TALLY_DOSE, TALLY_DOSE2, TALLY_COUNTS, TALLY_ALIGNED, TALLY_TLE = 1, 2, 4, 8, 16, 32

FP = C.POINTER(C.c_float)
DP = C.POINTER(C.c_double)


class PhysicsDesc(C.Structure):
    _fields_ = [
        ("n_materials", C.c_int32), ("n_buckets", C.c_int32), ("n_x_nodes", C.c_int32), ("n_compositions", C.c_int32),
        ("density", FP), ("composition", C.POINTER(C.c_int32)), ("mu_total", FP), ("mu_majorant", FP),
        ("p_compton", FP), ("p_rayleigh", FP),
        ("mu_en", FP), ("ray_x", FP), ("ray_cdf", FP), ("incoh_s", FP),
        ("e_cut", C.c_float), ("reserved1", C.c_int32),
    ]


class PhantomDesc(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("world_material", C.c_int32),
        ("spacing", C.c_float * 3), ("world_lo", C.c_float * 3), ("world_hi", C.c_float * 3),
        ("label_bytes", C.c_int32), ("labels", C.c_void_p),
    ]


class SourceDesc(C.Structure):
    _fields_ = [
        ("pos_type", C.c_int32), ("dir_type", C.c_int32), ("ene_type", C.c_int32), ("n_energy", C.c_int32),
        ("pos_param", C.c_float * 4), ("pos_mat", C.c_float * 12), ("dir_mat", C.c_float * 9),
        ("cos_theta_min", C.c_float), ("cos_theta_max", C.c_float), ("phi_min", C.c_float), ("phi_max", C.c_float),
        ("momentum", C.c_float * 3), ("focus", C.c_float * 3),
        ("energy_mono", C.c_float), ("energy_max", C.c_float), ("weight", C.c_float), ("phi_full", C.c_int32),
        ("energy_value", FP), ("energy_width", FP), ("energy_cdf", FP),
        ("phsp_n", C.c_int64), ("phsp_start", C.c_int64), ("phsp_iso", C.c_int32), ("phsp_on_device", C.c_int32),
        ("phsp", FP * 8),
        ("vox_n", C.c_int64), ("vox_prob", FP), ("vox_alias", C.POINTER(C.c_int32)),
        ("vox_index", C.POINTER(C.c_int32)), ("vox_size", C.c_int32 * 3), ("reserved2", C.c_int32),
        ("vox_spacing", C.c_float * 3), ("pos_sigma", C.c_float * 2), ("energy_sigma", C.c_float),
        ("weight_sigma", C.c_float), ("dir_sigma", C.c_float * 2), ("surface_axis", C.c_float * 3),
        ("n_theta", C.c_int32), ("n_phi", C.c_int32), ("theta_value", FP), ("theta_width", FP),
        ("theta_cdf", FP), ("phi_value", FP), ("phi_width", FP), ("phi_cdf", FP),
    ]


class TallyDesc(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("flags", C.c_int32),
        ("corner", C.c_float * 3), ("spacing", C.c_float * 3),
        ("record_face", C.c_int32), ("record_capacity", C.c_uint32),
        ("detector_label_min", C.c_int32), ("tle_energy_min", C.c_float), ("tle_energy_max", C.c_float),
    ]


class SpectDesc(C.Structure):
    _fields_ = [
        ("center", C.c_float * 3), ("radius", C.c_float), ("n_proj", C.c_int32), ("nx", C.c_int32),
        ("ny", C.c_int32), ("angle0", C.c_float), ("spacing", C.c_float * 2), ("cos_theta_max", C.c_float),
        ("crystal_depth", C.c_float), ("pitch", C.c_float), ("patch_origin", C.c_float * 3), ("norm", C.c_float * 6),
    ]


BLUR_LAWS = {"none": 0, "Gaussian": 1, "InverseSquare": 2, "Linear": 3}


class Blur(C.Structure):
    _fields_ = [("law", C.c_int32), ("sigma", C.c_float), ("resolution", C.c_float), ("reference", C.c_float),
                ("slope", C.c_float)]


class Launch(C.Structure):
    _fields_ = [("block", C.c_int32), ("grid", C.c_int32), ("refill_interval", C.c_int32), ("chunk", C.c_int32),
                ("slow_min", C.c_int32), ("refill_min", C.c_int32), ("scatter_tries", C.c_int32)]


class AdamTensor(C.Structure):
    _fields_ = [("param", C.c_void_p), ("grad", C.c_void_p), ("m", C.c_void_p), ("v", C.c_void_p),
                ("param_bf16", C.c_void_p), ("rows", C.c_int32), ("cols", C.c_int32), ("valid_rows", C.c_int32),
                ("valid_cols", C.c_int32), ("decay", C.c_float), ("reserved", C.c_int32)]


class Counters(C.Structure):
    _fields_ = [
        ("histories", C.c_uint64), ("steps", C.c_uint64), ("deposits", C.c_uint64), ("compton", C.c_uint64),
        ("rayleigh", C.c_uint64), ("photo", C.c_uint64), ("escaped", C.c_uint64),
        ("world_interactions", C.c_uint64), ("cut_kills", C.c_uint64), ("sq_evictions", C.c_uint64),
        ("recorded", C.c_uint64), ("kn_rejections", C.c_uint64),
        ("e_deposited", C.c_double), ("e_outside_tally", C.c_double), ("e_escaped", C.c_double),
        ("e_world", C.c_double), ("kernel_ms", C.c_double),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


COUNTER_NAMES = [f for f, _ in Counters._fields_[:12]]
ENERGY_NAMES = ["e_deposited", "e_outside_tally", "e_escaped", "e_world"]


def fptr(a):
    """float32 C-contiguous numpy array -> float*"""
    if a is None:
        return FP()
    assert a.dtype == np.float32 and a.flags.c_contiguous, (a.dtype, a.flags)
    return a.ctypes.data_as(FP)


def dptr(a):
    if a is None:
        return DP()
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(DP)


# exported symbols of libgatehip.so, with (restype, argtypes)
SIGNATURES = {
    "gh_abi_version": (C.c_int, []),
    "gh_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "gh_destroy": (None, [C.c_void_p]),
    "gh_last_error": (C.c_char_p, [C.c_void_p]),
    "gh_set_physics": (C.c_int, [C.c_void_p, C.POINTER(PhysicsDesc)]),
    "gh_add_simulation": (C.c_int, [C.c_void_p, C.POINTER(PhantomDesc), C.POINTER(SourceDesc),
                                    C.POINTER(TallyDesc), C.POINTER(C.c_int32)]),
    "gh_clear_simulations": (C.c_int, [C.c_void_p]),
    "gh_set_source": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(SourceDesc)]),
    "gh_reset_tally": (C.c_int, [C.c_void_p, C.c_int32]),
    "gh_reset_phsp": (C.c_int, [C.c_void_p, C.c_int32]),
    "gh_run": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.c_uint64,
                         C.POINTER(Launch), C.POINTER(Counters)]),
    "gh_autotune": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.c_uint64, C.POINTER(Launch),
                              C.POINTER(C.c_double)]),
    "gh_get_tally": (C.c_int, [C.c_void_p, C.c_int32, DP, DP, DP, DP, DP]),
    "gh_get_phsp": (C.c_int, [C.c_void_p, C.c_int32, FP * 8, C.c_uint64, C.POINTER(C.c_uint64)]),
    "gh_get_event_edep": (C.c_int, [C.c_void_p, C.c_int32, FP, C.c_uint64]),
    "gh_hu_to_labels": (C.c_int, [C.c_void_p, C.POINTER(C.c_int16), C.c_int64, DP, C.c_int32,
                                  C.POINTER(C.c_uint16), C.POINTER(C.c_uint16)]),
    "gh_hu_to_labels_f64": (C.c_int, [C.c_void_p, DP, C.c_int64, DP, C.c_int32, C.POINTER(C.c_uint16),
                                      C.POINTER(C.c_uint16)]),
    "gh_device_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.c_char_p, C.c_int32]),
    "gh_gemm_bf16": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.c_void_p, C.c_int64, C.c_int64, C.c_int32,
                               C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p,
                               C.c_int64, C.c_int32, C.c_void_p]),
    "gh_gemm_bf16_noalloc": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.c_void_p, C.c_int64, C.c_int64,
                                       C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_int64,
                                       C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, C.c_void_p]),
    "gh_gaussian_bf16": (C.c_int, [C.c_uint64, C.c_uint64, C.c_int64, C.c_int32, C.c_int32, C.c_void_p,
                                   C.c_void_p]),
    "gh_exitplane_decode": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.c_void_p, C.c_int32, C.c_float, FP,
                                      C.c_float, FP, C.c_float, FP * 8, C.c_void_p]),
    "gh_alias_table": (C.c_int, [DP, C.c_int64, FP, C.POINTER(C.c_int32)]),
    "gh_mlp_forward_bf16": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                      C.POINTER(C.c_void_p), C.POINTER(FP), FP, C.c_int64, C.c_void_p]),
    "gh_adam_step": (C.c_int, [C.POINTER(AdamTensor), C.c_int32, C.c_float, C.c_float, C.c_float, C.c_float,
                               C.c_int32, C.c_void_p]),
    "gh_spect_pairs": (C.c_int, [FP * 8, C.c_int64, C.POINTER(SpectDesc), C.c_void_p, C.c_int64, C.c_void_p,
                                 C.c_void_p, C.c_void_p, C.POINTER(FP), C.c_void_p, C.c_void_p, C.c_void_p]),
    "gh_energy_windows": (C.c_int, [C.c_void_p, C.c_int64, C.c_uint64, C.c_uint64, C.POINTER(Blur), FP,
                                    C.c_int32, C.c_void_p, C.c_void_p]),
    "gh_arf_accumulate": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32,
                                    C.c_void_p, C.c_void_p, C.c_void_p]),
    "gh_counts_accumulate": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, C.c_int32,
                                       C.c_void_p, C.c_void_p, C.c_void_p]),
}

_LIB = None


def load_library(path=None):
    """Load libgatehip.so (built by opengate_amd.build). Raises if it is absent."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = Path(path) if path else LIB_PATH
    # One HIP runtime per process: torch bundles its own libamdhip64.so.7 / HSA runtime.
    # Loading torch first makes the dynamic linker resolve this library's libamdhip64.so.7
    # to torch's copy (same SONAME), so tensors and this library share one runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not p.exists():
        raise RuntimeError(
            f"{p} not found: the MI355X transport library is not built. "
            "Run `python -m opengate_amd.build` (hipcc, gfx950). There is no CPU fallback."
        )
    lib = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.gh_abi_version() != 6:
        raise RuntimeError("libgatehip.so ABI version mismatch")
    if path is None:
        _LIB = lib
    return lib
