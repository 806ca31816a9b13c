"""GenericSource and PhaseSpaceSource with the reference's user parameters.

GenericSource: opengate/sources/generic.py:24-95 (position / direction / energy
parameter groups), :333-420 (source attributes); number of primaries `n` or
`activity` (sources/base.py:23-80). PhaseSpaceSource: opengate/sources/phspsources.py
:253-482 (keys, entry_start, global_flag, isotropic_direction, translate / rotate).
The ROOT reader of the reference (uproot) is replaced by numpy .npz / .npy-dict files
that hold the same branch names. VoxelSource: opengate/sources/voxelsources.py:13-136
(activity image, origin -size/2 spacing + translation + spacing/2 in the attached
volume, then the volume's global transform; GateSPSVoxelsPosDistribution.cpp:24-52).
"""
import numpy as np

from ..base import Box, GateObject
from ..units import g4_units


def _position_parameters():
    return Box(type="point", radius=0.0, sigma_x=0.0, sigma_y=0.0, size=[0.0, 0.0, 0.0],
               translation=[0.0, 0.0, 0.0], rotation=np.eye(3), confine=None, dz=None)


def _direction_parameters():
    return Box(type="iso", theta=[0.0, 180.0 * g4_units.deg], phi=[0.0, 360.0 * g4_units.deg],
               momentum=[0.0, 0.0, 1.0], focus_point=[0.0, 0.0, 0.0], sigma=[0.0, 0.0],
               angular_acceptance=Box(enabled=False), accolinearity_flag=False,
               accolinearity_fwhm=0.5 * g4_units.deg, histogram_theta_weights=[],
               histogram_theta_angles=[], histogram_phi_weights=[], histogram_phi_angles=[])


def _energy_parameters():
    return Box(type="mono", mono=0.0, sigma_gauss=0.0, is_cdf=False, min_energy=None, max_energy=None,
               spectrum_type=None, spectrum_weights=[], spectrum_energies=[], histogram_weight=[],
               histogram_energy=[], spectrum_energy_bin_edges=[], spectrum_histogram_interpolation=None)


class SourceBase(GateObject):
    """Number of primaries per run, or an activity, as the reference's sources
    (sources/base.py:23-80, resolve_and_validate_config :256-291) and its run loop
    (GateSourceManager::PrepareRunToStart / PrepareNextSource, GateSourceManager.cpp:313-359;
    GateVSource::PrepareNextTime / CalcNextTime / UpdateActivity, GateVSource.cpp:46-89)."""

    runtime_attributes = ("simulation",)
    user_info_defaults = {
        "attached_to": ("world", {"doc": "volume the source is attached to"}),
        "mother": (None, {"doc": "deprecated alias of attached_to"}),
        "number_of_primaries": (0, {"doc": "primaries per run: an int (one run) or one value per run interval"}),
        "n": (0, {"doc": "deprecated alias of number_of_primaries"}),
        "activity": (0.0, {"doc": "activity [Bq in Geant4 units] (exclusive with number_of_primaries)"}),
        "half_life": (-1, {"doc": "half-life [Geant4 time units]; -1: no decay"}),
        "start_time": (None, {"doc": "start of the source [Geant4 time units] (default: first run start)"}),
        "end_time": (None, {"doc": "end of the source [Geant4 time units] (default: last run end)"}),
    }

    def __init__(self, name=None, **kw):
        super().__init__(name, **kw)
        self.simulation = None

    def total_yield(self):
        """Factor of the activity for weighted spectra (1 here)."""
        return 1.0

    def _primaries_array(self):
        v = self.number_of_primaries
        if (np.isscalar(v) and not v) or v is None:
            v = self.n
        return np.atleast_1d(np.asarray(v if v is not None else 0, dtype=np.int64))

    def events_per_run(self, run_timing_intervals, rng):
        """Number of events of this source in each run interval. number_of_primaries: its
        value for the run if the source's [start_time, end_time) overlaps the interval
        (PrepareNextTime returns the start time, or the current time, while it lies before
        end_time). activity: a run that starts before start_time has its first event at
        start_time itself (PrepareNextTime returns the start time); then the events are the
        arrivals of a process of rate A(t) = A0 exp(-ln2 (t - start_time) / half_life) in the
        overlap, A0 = activity x the spectrum's total yield. The count is one draw of the
        Poisson law of the integrated rate with `rng` (numpy Generator). (CalcNextTime holds
        the rate of the last event over each exponential waiting time, which changes the
        mean count at first order in (ln2 / half_life) / A: 1e-11 for a 1 MBq Tc-99m source.)"""
        intervals = [(float(a), float(b)) for a, b in run_timing_intervals]
        npri = self._primaries_array()
        act = float(self.activity or 0.0)
        if act < 0 or np.any(npri < 0):
            raise ValueError(f"source '{self.name}': negative activity or number_of_primaries")
        if act == 0 and len(npri) != len(intervals):
            raise ValueError(f"source '{self.name}': number_of_primaries and run_timing_intervals do not have "
                             "the same length")
        if act > 0 and np.any(npri > 0):
            raise ValueError(f"source '{self.name}': cannot use both 'number_of_primaries' and 'activity'")
        if act == 0 and np.all(npri == 0):
            raise ValueError(f"source '{self.name}': set one of 'number_of_primaries' or 'activity'")
        start = intervals[0][0] if self.start_time is None else float(self.start_time)
        end = intervals[-1][1] if self.end_time is None else float(self.end_time)
        counts = []
        for k, (a, b) in enumerate(intervals):
            lo, hi = max(a, start), min(b, end)
            if not lo < hi:
                counts.append(0)
                continue
            if act == 0:
                counts.append(int(npri[k]))
                continue
            a0 = act / g4_units.Bq * self.total_yield()  # 1/s
            t0, t1 = (lo - start) / g4_units.s, (hi - start) / g4_units.s
            hl = float(self.half_life)
            if hl > 0:
                lam = np.log(2.0) / (hl / g4_units.s)
                mean = a0 / lam * (np.exp(-lam * t0) - np.exp(-lam * t1))
            else:
                mean = a0 * (t1 - t0)
            at_start = 1 if a < start else 0  # the event at start_time itself
            counts.append(at_start + int(rng.poisson(mean)))
        return counts

    def number_of_primaries_total(self, run_timing_intervals, rng):
        return int(sum(self.events_per_run(run_timing_intervals, rng)))

    @property
    def attached(self):
        return self.mother if self.mother else self.attached_to


class GenericSource(SourceBase):
    user_info_defaults = {
        "particle": ("gamma", {"doc": "only 'gamma' on the photon path"}),
        "weight": (-1, {"doc": "history weight (-1: 1)"}),
        "weight_sigma": (-1, {"doc": "if >= 0 (and weight > 0), the history weight is Gaussian around weight "
                                     "with this sigma (GateGenericSource.cpp:247-259)"}),
        "direction_relative_to_attached_volume": (False, {"doc": "rotate directions with the attached volume"}),
        "position": (_position_parameters(), {"doc": "position parameters"}),
        "direction": (_direction_parameters(), {"doc": "direction parameters"}),
        "energy": (_energy_parameters(), {"doc": "energy parameters"}),
        "polarization": ([], {"doc": "not supported"}),
    }
    source_type = "GenericSource"

    def total_yield(self):
        """Sum of the spectrum weights by which the reference multiplies the activity
        (GateGenericSource.cpp:544-574 spectrum_discrete: the weights; :583-632
        spectrum_histogram: weight x bin width; :530-542 histogram: the weights)."""
        e = self.energy
        if e.type == "spectrum_discrete":
            return float(np.sum(np.asarray(e.spectrum_weights, dtype=np.float64)))
        if e.type == "spectrum_histogram":
            w = np.asarray(e.spectrum_weights, dtype=np.float64)
            edges = np.asarray(e.spectrum_energy_bin_edges, dtype=np.float64)
            return float(np.sum(w * np.diff(edges)))
        if e.type == "histogram":
            return float(np.sum(np.asarray(e.histogram_weight, dtype=np.float64)))
        return 1.0


class PhaseSpaceSource(SourceBase):
    user_info_defaults = {
        "phsp_file": (None, {"doc": ".npz file (or dict of arrays) with the branches below"}),
        "position_key": ("PrePositionLocal", {"doc": "prefix of the _X/_Y/_Z position branches"}),
        "position_key_x": (None, {}), "position_key_y": (None, {}), "position_key_z": (None, {}),
        "direction_key": ("PreDirectionLocal", {"doc": "prefix of the direction branches"}),
        "direction_key_x": (None, {}), "direction_key_y": (None, {}), "direction_key_z": (None, {}),
        "energy_key": ("KineticEnergy", {}),
        "weight_key": ("Weight", {}),
        "PDGCode_key": ("PDGCode", {}),
        "particle": ("gamma", {}),
        "global_flag": (False, {"doc": "positions/directions are global (no attached-volume transform)"}),
        "isotropic_direction": (False, {"doc": "ignore the stored direction, emit isotropically"}),
        "translate_position": (False, {}),
        "rotate_direction": (False, {}),
        "position": (Box(translation=[0.0, 0.0, 0.0], rotation=np.eye(3)), {}),
        "entry_start": (0, {"doc": "first entry used"}),
        "batch_size": (10000, {"doc": "kept for compatibility"}),
        "generate_until_next_primary": (False, {}),
        "primary_lower_energy_threshold": (0.0, {}),
        "primary_PDGCode": (0, {}),
        "verbose_batch": (False, {}),
    }
    source_type = "PhaseSpaceSource"

    def load_arrays(self):
        """8 float32 arrays (x, y, z, dx, dy, dz, E, w) of the phase space file."""
        f = self.phsp_file
        data = dict(f) if isinstance(f, dict) else dict(np.load(f, allow_pickle=False))

        def key(prefix, axis, override):
            return override if override else f"{prefix}_{axis}"

        cols = [data[key(self.position_key, a, getattr(self, f"position_key_{a.lower()}"))] for a in "XYZ"]
        cols += [data[key(self.direction_key, a, getattr(self, f"direction_key_{a.lower()}"))] for a in "XYZ"]
        cols.append(data[self.energy_key])
        w = data.get(self.weight_key)
        cols.append(w if w is not None else np.ones(len(cols[0])))
        return [np.ascontiguousarray(c, dtype=np.float32) for c in cols]


class VoxelSource(GenericSource):
    """GenericSource whose positions follow a 3-D activity image (normalised to sum 1).
    Positions are uniform inside the sampled voxel; `position.rotation` is not applied,
    as in the reference (only the attached volume's rotation is)."""

    user_info_defaults = {
        "image": (None, {"doc": "activity image: file (.mhd/.mha/.npy) or opengate_amd.Image"}),
    }
    source_type = "VoxelSource"

    def activity_image(self):
        from ..image import Image, read_image

        if self.image is None:
            raise ValueError(f"VoxelSource '{self.name}': image is not set")
        img = self.image if isinstance(self.image, Image) else read_image(self.image)
        return np.asarray(img.array, dtype=np.float64), np.asarray(img.spacing, dtype=np.float64)
