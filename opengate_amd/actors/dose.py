"""DoseActor, SimulationStatisticsActor and the exit-face PhaseSpaceActor.

DoseActor follows opengate/actors/doseactors.py (VoxelDepositActor :34-287,
DoseActor :317-715): outputs edep (always), edep_squared, edep_uncertainty, dose,
dose_squared, dose_uncertainty, counts, density; `size`/`spacing` may be
"like_image_volume"; dose [Gy] = sum(edep/rho) / (Gy * voxel volume); squared images
are per-event sums squared; uncertainty = sqrt(var)/mean with
var = max(0, (S2/N - (S/N)^2)/(N-1)), N = number of events (dataitems.py:1956-1998,
utility.py:551-560). Output files get the suffix of each output inserted before the
extension (test041_dose_actor.py:81, actoroutput.py:206).
"""
import json
import time
from pathlib import Path

import numpy as np

from ..base import GateObject
from ..image import Image, write_image
from ..units import g4_units

GY_FACTOR = 1.602176634e-7  # 1 MeV / (1 g/cm3 * 1 mm3) in Gy


def calculate_variance(value, squared, n):
    """utility.py:551-560"""
    return np.clip((squared / n - np.power(value / n, 2)) / (n - 1), 0, None)


def variance_or_uncertainty(value, squared, n, which):
    """SingleItkImageWithVariance.get_variance_or_uncertainty (dataitems.py:1956-1998)."""
    value = np.asarray(value, dtype=np.float64)
    if not n > 1:
        return np.ones_like(value)
    if squared is None:
        return np.zeros_like(value)
    out = calculate_variance(value, np.asarray(squared, dtype=np.float64), n)
    if which in ("std", "uncertainty"):
        out = np.sqrt(out)
    if which == "uncertainty":
        out = np.divide(out, value / n, out=np.zeros_like(out), where=value != 0)
    return out


def mean_of_highest_values(values, count):
    """Mean of the `count` largest values (GateDoseActor::GetMeanOfHighestNValues,
    GateDoseActor.cpp:401-428; 0 for an empty image)."""
    v = np.asarray(values, dtype=np.float64).ravel()
    k = min(int(count), v.size)
    if k <= 0:
        return 0.0
    return float(np.partition(v, v.size - k)[v.size - k:].sum() / k)


def dose_actor_mean_uncertainty(edep, edep2, n_events, top_voxels_count, threshold):
    """Mean relative uncertainty of the stopping rule (GateDoseActor::ComputeMeanUncertainty,
    GateDoseActor.cpp:281-330): n = max(events, 2); over the voxels whose edep exceeds
    threshold x the mean of the top voxels, unc = sqrt(max(0, (S2/n - (S/n)^2) / (n - 1))) /
    (S/n), each clamped at 1; their mean, or 1 when no voxel qualifies."""
    n = float(max(n_events, 2))
    e = np.asarray(edep, dtype=np.float64).ravel()
    e2 = np.asarray(edep2, dtype=np.float64).ravel()
    sel = e > mean_of_highest_values(e, top_voxels_count) * threshold
    val = e[sel] / n
    sq = e2[sel] / n
    unc = (1.0 / (n - 1.0)) * (sq - val ** 2)
    unc = np.sqrt(np.maximum(unc, 0.0)) / val
    unc = np.minimum(unc, 1.0)
    total = float(unc.sum())
    return total / unc.size if (unc.size > 0 and total > 0) else 1.0


def next_uncertainty_check(u, goal, n_events, overshoot):
    """Events of the run at which the next check happens (GateDoseActor.cpp:272-274:
    (u / goal)^2 n overshoot, truncated to an integer)."""
    return int((u / goal) * (u / goal) * n_events * overshoot)


def insert_suffix_before_extension(path, suffix):
    p = Path(path)
    name = p.name
    for ext in (".nii.gz", ".mhd", ".mha", ".npy", ".nii"):
        if name.endswith(ext):
            return p.with_name(name[: -len(ext)] + "_" + suffix + ext)
    return p.with_name(name + "_" + suffix)


def sample_count_path(path):
    """Sidecar of an image output holding its number of samples: x_edep.mhd ->
    x_edep_samples.mhd.json (dataitems.py:235-245), read back when job outputs are merged."""
    p = insert_suffix_before_extension(Path(path), "samples")
    return p.with_suffix(Path(path).suffix + ".json")


def write_sample_count(path, n):
    """ItkImageDataItem.write sidecar (dataitems.py:1336-1340)."""
    sample_count_path(path).write_text(json.dumps({"number_of_samples": int(n)}))


def read_sample_count(path):
    """number_of_samples of an image output, or None without a sidecar (dataitems.py:1342-1354)."""
    p = sample_count_path(path)
    if not p.exists():
        return None
    return int(json.loads(p.read_text())["number_of_samples"])


class OutputItem:
    """User interface of one actor output (UserInterfaceToActorOutputImage, actoroutput.py:276)."""

    def __init__(self, actor, name, active=False, write_to_disk=True):
        self._actor = actor
        self.name = name
        self.active = active
        self.write_to_disk = write_to_disk
        self.output_filename = "auto"
        self.suffix = name
        self._data = None

    def get_data(self, which="merged", **kw):
        if self._data is None:
            raise RuntimeError(f"output '{self.name}' of actor '{self._actor.name}' has no data (inactive or not run)")
        return self._data

    @property
    def image(self):
        return self.get_data()

    def get_output_path(self, which="merged", **kw):
        sim = self._actor.simulation
        base = Path(sim.output_dir) if sim is not None else Path(".")
        if self.output_filename not in (None, "auto"):
            return base / self.output_filename
        fn = self._actor.output_filename
        if fn in (None, "auto", ""):
            fn = f"{self._actor.name}.mhd"
        return base / insert_suffix_before_extension(fn, self.suffix)


class ActorBase(GateObject):
    runtime_attributes = ("simulation", "number_of_samples", "counts", "raw", "data", "number_of_recorded")
    user_info_defaults = {
        "attached_to": ("world", {"doc": "volume the actor is attached to"}),
        "output_filename": ("auto", {"doc": "base file name; output suffixes are inserted"}),
        "write_to_disk": (True, {"doc": "write outputs at the end of the simulation"}),
        "keep_data_per_run": (False, {}),
        "priority": (100, {}),
        "filters": ([], {"doc": "not supported on the photon path"}),
    }

    def __init__(self, name=None, **kw):
        super().__init__(name, **kw)
        self.simulation = None

    @property
    def attached_to_name(self):
        a = self.attached_to
        return a if isinstance(a, str) else a.name


class DoseActor(ActorBase):
    user_info_defaults = {
        "size": ([10, 10, 10], {"doc": "voxels, or 'like_image_volume'"}),
        "spacing": ([1.0, 1.0, 1.0], {"doc": "mm, or 'like_image_volume'"}),
        "translation": ([0.0, 0.0, 0.0], {"doc": "grid centre relative to the attached volume"}),
        "rotation": (np.eye(3), {"doc": "rotation written into the output images' geometry; the "
                                        "scoring grid stays aligned with the attached volume, as "
                                        "GateDoseActor attaches it (GateDoseActor.cpp:83-118)"}),
        "repeated_volume_index": (0, {}),
        "hit_type": ("random", {"allowed_values": ("random", "pre", "post", "middle"),
                                "doc": "photon deposits are at interaction points; all values give the same result"}),
        "output_coordinate_system": ("local", {"allowed_values": ("local", "global", "attached_to_image", None)}),
        "score_in": ("material", {"doc": "'material' only"}),
        "uncertainty_goal": (None, {"doc": "stop when the mean relative uncertainty reaches this value"}),
        "uncertainty_first_check_after_n_events": (1e4, {}),
        "uncertainty_top_voxels_count": (10, {}),
        "uncertainty_voxel_edep_threshold": (0.7, {}),
        "uncertainty_overshoot_factor_N_events": (1.05, {}),
    }
    actor_type = "DoseActor"
    _OUTPUTS = ["edep", "edep_squared", "edep_uncertainty", "dose", "dose_squared", "dose_uncertainty", "counts",
                "density"]

    def __init__(self, name=None, **kw):
        object.__setattr__(self, "_outputs", {})
        super().__init__(name, **kw)
        for o in self._OUTPUTS:
            self._outputs[o] = OutputItem(self, o, active=(o == "edep"))
        self.number_of_samples = 0

    def __getattr__(self, k):
        outs = self.__dict__.get("_outputs")
        if outs is not None and k in outs:
            return outs[k]
        return super().__getattr__(k)

    @property
    def user_output(self):
        return self._outputs

    # ------------------------------------------------------------------ helpers
    def resolved_grid(self, volume):
        size = self.size
        spacing = self.spacing
        if isinstance(size, str):
            if size != "like_image_volume" or not hasattr(volume, "size_pix"):
                raise ValueError(f"size = {size!r} needs an ImageVolume")
            size = volume.size_pix
        if isinstance(spacing, str):
            if spacing != "like_image_volume" or not hasattr(volume, "size_pix"):
                raise ValueError(f"spacing = {spacing!r} needs an ImageVolume")
            spacing = volume.spacing
        return np.asarray(size, dtype=int), np.asarray(spacing, dtype=np.float64)

    def wants_squared(self):
        o = self._outputs
        return (o["edep_squared"].active or o["edep_uncertainty"].active or o["dose_squared"].active
                or o["dose_uncertainty"].active or self.uncertainty_goal is not None)

    def wants_dose(self):
        o = self._outputs
        return o["dose"].active or o["dose_squared"].active or o["dose_uncertainty"].active

    def output_geometry(self, size, spacing, volume, global_transform):
        """(origin, direction) of the output images (VoxelDepositActor.
        _update_output_coordinate_system, doseactors.py:170-232): None gives the grid centred
        at 0 without rotation; "local" the grid at `translation`, rotated by `rotation`;
        "attached_to_image" that placed by the image's native origin and direction;
        "global" that placed by the attached volume's world transform
        (`global_transform` = (R, t), get_transform_world_to_local)."""
        size = np.asarray(size, dtype=np.float64)
        spacing = np.asarray(spacing, dtype=np.float64)
        origin = -size * spacing / 2.0 + spacing / 2.0
        R = np.asarray(self.rotation, dtype=np.float64).reshape(3, 3)
        local = R @ origin + np.asarray(self.translation, dtype=np.float64)
        if self.output_coordinate_system is None:
            return origin, np.eye(3)
        if self.output_coordinate_system == "local":
            return local, R
        if self.output_coordinate_system == "attached_to_image":
            Rn = np.asarray(volume.native_rotation, dtype=np.float64)
            return Rn @ local + np.asarray(volume.native_translation, dtype=np.float64), Rn @ R
        Rg, tg = global_transform
        return Rg @ local + tg, Rg @ R

    def store_results(self, size, spacing, origin, edep, edep2, dose, dose2, counts, density, n_events,
                      direction=None):
        self.number_of_samples = int(n_events)
        o = self._outputs
        direction = np.eye(3) if direction is None else direction

        def img(a):
            return Image(np.ascontiguousarray(a), spacing, origin, direction)

        o["edep"]._data = img(edep)
        if edep2 is not None:
            o["edep_squared"]._data = img(edep2)
            o["edep_uncertainty"]._data = img(variance_or_uncertainty(edep, edep2, n_events, "uncertainty"))
        if dose is not None:
            o["dose"]._data = img(dose)
            if dose2 is not None:
                o["dose_squared"]._data = img(dose2)
                o["dose_uncertainty"]._data = img(variance_or_uncertainty(dose, dose2, n_events, "uncertainty"))
        if counts is not None:
            o["counts"]._data = img(counts)
        if density is not None:
            o["density"]._data = img(density)

    def write_outputs(self):
        if not self.write_to_disk:
            return
        for name, item in self._outputs.items():
            if item.active and item.write_to_disk and item._data is not None:
                p = item.get_output_path()
                p.parent.mkdir(parents=True, exist_ok=True)
                write_image(item._data, p)
                write_sample_count(p, self.number_of_samples)

    def __str__(self):
        return f"DoseActor {self.name}: {self.number_of_samples} events"


class SimulationStatisticsActor(ActorBase):
    """Counts runs, events, tracks and steps; PPS = events / duration
    (GateSimulationStatisticsActor.cpp:42-257, dataitems.py:468-485). Duration is the
    transport time (GPU launch to completion, including host/device copies)."""

    user_info_defaults = {
        "track_types_flag": (False, {}),
    }
    actor_type = "SimulationStatisticsActor"

    def __init__(self, name=None, **kw):
        super().__init__(name, **kw)
        self.counts = {"runs": 0, "events": 0, "tracks": 0, "steps": 0, "duration": 0.0, "init": 0.0,
                       "start_time": 0.0, "stop_time": 0.0, "track_types": {}}
        self.raw = {}

    @property
    def pps(self):
        d = self.counts["duration"] / g4_units.s
        return int(self.counts["events"] / d) if d > 0 else 0

    @property
    def tps(self):
        d = self.counts["duration"] / g4_units.s
        return int(self.counts["tracks"] / d) if d > 0 else 0

    @property
    def sps(self):
        d = self.counts["duration"] / g4_units.s
        return int(self.counts["steps"] / d) if d > 0 else 0

    def __str__(self):
        c = self.counts
        return (f"Runs {c['runs']}\nEvents {c['events']}\nTracks {c['tracks']}\nSteps {c['steps']}\n"
                f"Init {c['init'] / g4_units.s:.6f} s\nDuration {c['duration'] / g4_units.s:.6f} s\n"
                f"PPS {self.pps}\nTPS {self.tps}\nSPS {self.sps}")

    def write_outputs(self):
        if not self.write_to_disk or self.output_filename in (None, "auto", ""):
            return
        p = Path(self.simulation.output_dir) / self.output_filename
        p.parent.mkdir(parents=True, exist_ok=True)
        out = dict(self.counts)
        out["pps"] = self.pps
        out.update({k: v for k, v in self.raw.items() if isinstance(v, (int, float))})
        p.write_text(json.dumps(out, indent=1, default=float))


class TLEDoseActor(DoseActor):
    """Track-length estimator dose (TLEDoseActor, actors/doseactors.py:722-775;
    GateTLEDoseActor.cpp:200-300). On this path the estimator is the Woodcock collision
    estimator of kerma: every tentative collision of a photon inside the phantom scores
    w E mu_en(E)/mu_majorant(E), whose expectation per unit path is the track-length score
    w E mu_en L of the reference, and the electrons set in motion by such photons deposit
    nothing. mu_en/rho comes from the NIST table for both `database` values (EPDL tables
    are not available here). Threshold types 'None' (all photons) and 'energy' (photons at
    or below tle_threshold; above it, analog deposits) are supported."""

    user_info_defaults = {
        "energy_min": (0.0, {"doc": "photons at or below this energy deposit it locally and stop"}),
        "tle_threshold": (np.inf, {"doc": "energy below which the estimator applies (type 'energy')"}),
        "tle_threshold_type": ("None", {"allowed_values": ("None", "energy", "max range", "average range")}),
        "database": ("EPDL", {"allowed_values": ("EPDL", "NIST")}),
    }
    actor_type = "TLEDoseActor"

    def tle_energy_bounds(self):
        """(energy_min, energy_max) in MeV for the tally descriptor."""
        if self.score_in != "material":
            raise ValueError(f"TLEDoseActor cannot score in {self.score_in}, only 'material' is allowed.")
        t = self.tle_threshold_type
        if t in ("max range", "average range"):
            raise NotImplementedError(f"TLEDoseActor.tle_threshold_type '{t}' (needs electron CSDA ranges)")
        e_max = float(self.tle_threshold) / g4_units.MeV if t == "energy" else float("inf")
        return float(self.energy_min) / g4_units.MeV, e_max


FACES = {"-x": 0, "+x": 1, "-y": 2, "+y": 3, "-z": 4, "+z": 5, "all": 6}


class PhaseSpaceActor(ActorBase):
    """Records photons leaving the attached ImageVolume through one face (a subset of
    GatePhaseSpaceActor, GatePhaseSpaceActor.cpp:119, with steps_to_store='exiting').
    Positions are in the attached volume frame (PrePositionLocal_*), directions
    PreDirectionLocal_*, KineticEnergy, Weight; the recorded photon stops there."""

    user_info_defaults = {
        "exit_face": ("+z", {"allowed_values": tuple(FACES)}),
        "max_photons": (10_000_000, {}),
        "steps_to_store": ("exiting", {"allowed_values": ("exiting",)}),
        "attributes": (["PrePositionLocal", "PreDirectionLocal", "KineticEnergy", "Weight"], {}),
    }
    actor_type = "PhaseSpaceActor"

    def __init__(self, name=None, **kw):
        super().__init__(name, **kw)
        self.data = None
        self.number_of_recorded = 0

    def store(self, arrays, n_total, half_size):
        x, y, z, dx, dy, dz, e, w = arrays
        self.data = {"PrePositionLocal_X": x - half_size[0], "PrePositionLocal_Y": y - half_size[1],
                     "PrePositionLocal_Z": z - half_size[2], "PreDirectionLocal_X": dx, "PreDirectionLocal_Y": dy,
                     "PreDirectionLocal_Z": dz, "KineticEnergy": e, "Weight": w}
        self.number_of_recorded = int(n_total)

    def write_outputs(self):
        if not self.write_to_disk or self.data is None:
            return
        fn = self.output_filename if self.output_filename not in (None, "auto", "") else f"{self.name}.npz"
        p = Path(self.simulation.output_dir) / fn
        p.parent.mkdir(parents=True, exist_ok=True)
        np.savez(p.with_suffix(".npz"), **self.data)
