"""Simulation facade (opengate/managers.py: Simulation :1805, add_volume :2597,
add_source :2607, add_actor :2610, run :2818; PhysicsManager :695; VolumeManager :1435).

Only the photon-transport-and-dose path is available: the world, Box and Image
volumes, GenericSource / VoxelSource / PhaseSpaceSource / GANSource with gamma
primaries, the gamma physics options, DoseActor, TLEDoseActor,
SimulationStatisticsActor, an exit-face PhaseSpaceActor, and SPECT detection (a
SPECTHeadVolume gantry with the digitizer chain ending in DigitizerProjectionActor,
ARFActor, ARFTrainingDatasetActor).
Anything else raises NotImplementedError instead of silently doing something else.
"""
import os
from pathlib import Path

import numpy as np

from .actors.digitizers import (ARFActor, ARFTrainingDatasetActor, DigitizerAdderActor, DigitizerBlurringActor,
                                DigitizerEnergyWindowsActor, DigitizerHitsCollectionActor, DigitizerProjectionActor)
from .actors.dose import DoseActor, PhaseSpaceActor, SimulationStatisticsActor, TLEDoseActor
from .base import Box, GateObject
from .geometry.spect import SPECTHeadVolume
from .geometry.volumes import BoxVolume, ImageVolume
from .physics.materials import MaterialDatabase
from .sources.gan import GANSource
from .sources.generic import GenericSource, PhaseSpaceSource, VoxelSource
from .units import g4_units

VACUUM_MATERIALS = ("G4_Galactic", "Vacuum", "G4_VACUUM")


class PhysicsManager(GateObject):
    runtime_attributes = ("simulation", "min_ekine", "production_cuts")
    user_info_defaults = {
        "physics_list_name": ("QGSP_BERT_EMV", {"doc": "any EM list: gamma processes are Compton, "
                                                       "photoelectric and Rayleigh in every case"}),
        "global_production_cuts": (Box(all=0.7 * g4_units.mm, gamma=None, electron=None, positron=None,
                                       proton=None), {}),
        "em_parameters": (Box(fluo=None, auger=None, auger_cascade=None, pixe=None, deexcitation_ignore_cut=None),
                          {}),
        "user_limits_particles": (Box(all=False, all_charged=True, gamma=False, electron=False, positron=False,
                                      proton=False), {}),
        "energy_range_min": (None, {}),
        "energy_range_max": (None, {}),
        "apply_cuts": (True, {}),
        "enable_decay": (False, {}),
        "photon_tracking_cut": (2.0 ** -10 * g4_units.MeV,
                                {"doc": "photons below this energy deposit locally (>= 2^-10 MeV)"}),
    }

    def __init__(self, simulation):
        super().__init__("physics_manager")
        self.simulation = simulation
        self.min_ekine = {}
        self.production_cuts = {}

    def set_production_cut(self, volume_name, particle_name, value):
        self.production_cuts[(volume_name, particle_name)] = value

    def set_min_ekine(self, volume_name, min_ekine):
        self.min_ekine[volume_name] = min_ekine

    def set_max_step_size(self, volume_name, max_step_size):
        pass  # photon transport with Woodcock tracking has no step limit to apply

    def check_supported(self):
        em = self.em_parameters
        for k in ("fluo", "auger", "pixe"):
            if em.get(k):
                raise NotImplementedError(f"em_parameters.{k} = True: atomic de-excitation is not simulated; "
                                          "the photon path deposits binding energy locally")


class VolumeManager:
    def __init__(self, simulation):
        self.simulation = simulation
        self.material_database = MaterialDatabase()
        self.volumes = {}
        self.world = BoxVolume("world", size=[3000.0, 3000.0, 3000.0], material="G4_AIR")
        self.world.volume_manager = self
        self.volumes["world"] = self.world

    def add_volume(self, volume_type, name=None):
        if isinstance(volume_type, GateObject):
            v = volume_type
        else:
            t = volume_type.lower().replace("volume", "")
            if t == "box":
                v = BoxVolume(name)
            elif t == "image":
                v = ImageVolume(name)
            elif t == "specthead":
                v = SPECTHeadVolume(name)
            else:
                raise NotImplementedError(f"volume type '{volume_type}' is not part of the photon dose path "
                                          "(Box, Image and SPECTHead volumes are)")
        if v.name in self.volumes:
            raise ValueError(f"a volume named '{v.name}' already exists")
        v.volume_manager = self
        self.volumes[v.name] = v
        return v

    def add_material_database(self, filename):
        raise NotImplementedError("Gate .db material files are not read; use "
                                  "material_database.add_material_weights")

    def global_transform(self, name):
        """(R, t) of a volume in the world frame, composing the mother chain."""
        R = np.eye(3)
        t = np.zeros(3)
        v = self.volumes[name]
        chain = []
        while v.name != "world":
            chain.append(v)
            v = self.volumes[v.mother if isinstance(v.mother, str) else v.mother.name]
        for v in reversed(chain):
            Rv = np.asarray(v.rotation, dtype=np.float64)
            t = R @ np.asarray(v.translation, dtype=np.float64) + t
            R = R @ Rv
        return R, t


class Simulation:
    def __init__(self, name="simulation"):
        self.name = name
        self.volume_manager = VolumeManager(self)
        self.physics_manager = PhysicsManager(self)
        self.sources = {}
        self.actors = {}
        self.random_seed = "auto"
        self.number_of_threads = 1
        self.output_dir = "."
        self.run_timing_intervals = [[0, 1 * g4_units.s]]
        self.verbose_level = "INFO"
        self.running_verbose_level = 0
        self.progress_bar = False
        self.visu = False
        self.g4_verbose = False
        self.check_volumes_overlap = False
        self.device = 0
        # first history id of this run (a job of a split campaign; see opengate_amd.jobs)
        self.first_history = 0
        # persistent-kernel launch parameters: {} (library defaults), a dict of gh_launch
        # fields, or "auto" (gh_autotune on a trial of the run)
        self.launch = {}
        self.warnings = []
        self.last_run = None
        self.engine = None

    @property
    def world(self):
        return self.volume_manager.world

    def add_volume(self, volume_type, name=None):
        return self.volume_manager.add_volume(volume_type, name)

    def get_volume(self, name):
        return self.volume_manager.volumes[name]

    def add_source(self, source_type, name):
        if name in self.sources:
            raise ValueError(f"a source named '{name}' already exists")
        if source_type == "GenericSource":
            s = GenericSource(name)
        elif source_type == "PhaseSpaceSource":
            s = PhaseSpaceSource(name)
        elif source_type == "GANSource":
            s = GANSource(name)
        elif source_type == "VoxelSource":
            s = VoxelSource(name)
        else:
            raise NotImplementedError(f"source type '{source_type}' is not part of the photon dose path")
        s.simulation = self
        self.sources[name] = s
        return s

    def get_source(self, name):
        return self.sources[name]

    def add_actor(self, actor_type, name):
        if name in self.actors:
            raise ValueError(f"an actor named '{name}' already exists")
        table = {"DoseActor": DoseActor, "TLEDoseActor": TLEDoseActor,
                 "SimulationStatisticsActor": SimulationStatisticsActor, "PhaseSpaceActor": PhaseSpaceActor,
                 "DigitizerHitsCollectionActor": DigitizerHitsCollectionActor,
                 "DigitizerAdderActor": DigitizerAdderActor, "DigitizerBlurringActor": DigitizerBlurringActor,
                 "DigitizerEnergyWindowsActor": DigitizerEnergyWindowsActor,
                 "DigitizerProjectionActor": DigitizerProjectionActor, "ARFActor": ARFActor,
                 "ARFTrainingDatasetActor": ARFTrainingDatasetActor}
        if actor_type not in table:
            raise NotImplementedError(f"actor type '{actor_type}' is not part of the photon dose path")
        a = table[actor_type](name)
        a.simulation = self
        self.actors[name] = a
        return a

    def get_actor(self, name):
        return self.actors[name]

    def resolve_seed(self):
        if self.random_seed == "auto":
            self.random_seed = int.from_bytes(os.urandom(8), "little") & ((1 << 63) - 1)
        return int(self.random_seed)

    def get_output_path(self, p):
        return Path(self.output_dir) / p

    def run(self, start_new_process=False):
        from .engine import SimulationEngine

        Path(self.output_dir).mkdir(parents=True, exist_ok=True)
        self.engine = SimulationEngine(self)
        self.last_run = self.engine.run()
        return self.last_run
