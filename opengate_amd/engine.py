"""Run driver: turns a Simulation into C-ABI descriptors and runs it on the GPU.

Plays the role of opengate/engines.py (SimulationEngine.run_engine :1554,
initialize :1721, start_and_stop :1632) for the photon path:
  * the phantom volume (one ImageVolume, or a homogeneous Box) becomes a label grid
    in the phantom frame p_ph = R^T (p_world - t) + L/2 (box [0, L));
  * materials of the labels (+ the world, unless vacuum) get FP32 tables;
  * every source becomes a SourceSpec in that frame; sources run one after the
    other into the same tally with disjoint history ranges of the RNG;
  * the DoseActor grid becomes the tally, dose/uncertainty are derived in FP64 as
    DoseActor.EndOfRunActionMasterThread and get_variance_or_uncertainty do.
"""
import dataclasses
import time

import numpy as np

from . import spect

from .actors.dose import FACES, GY_FACTOR, DoseActor, PhaseSpaceActor, SimulationStatisticsActor, TLEDoseActor
from .actors.digitizers import ARFTrainingDatasetActor
from .geometry.spect import SPECTHeadVolume
from .geometry.volumes import BoxVolume, ImageVolume
from .physics import xs
from .sources.gan import GANSource
from .sources.generic import GenericSource, PhaseSpaceSource, VoxelSource
from .transport import (GpuEngine, PhantomSpec, SourceSpec, TallySpec, angle_histogram, energy_user_histogram,
                        spectrum_histogram, spectrum_lines)
from .units import g4_units

VACUUM = ("G4_Galactic", "Vacuum", "G4_VACUUM")


def _rot(m):
    return np.asarray(m, dtype=np.float64).reshape(3, 3)


class SimulationEngine:
    def __init__(self, sim):
        self.sim = sim
        self.engine = None

    # ------------------------------------------------------------------ geometry
    def phantom_volume(self):
        vols = [v for n, v in self.sim.volume_manager.volumes.items()
                if n != "world" and not isinstance(v, SPECTHeadVolume)]
        images = [v for v in vols if isinstance(v, ImageVolume)]
        if len(images) > 1:
            raise NotImplementedError("one ImageVolume per simulation")
        if images:
            other = [v for v in vols if v is not images[0] and images[0].mother != v.name]
            if other:
                raise NotImplementedError(f"volumes {[v.name for v in other]} besides the ImageVolume and its mothers")
            return images[0]
        if len(vols) == 1 and isinstance(vols[0], BoxVolume):
            return vols[0]
        if not vols:
            raise NotImplementedError("a phantom volume (ImageVolume or Box) is required")
        raise NotImplementedError("only one ImageVolume (with Box mothers) or one Box phantom")

    def build_geometry(self, vol):
        db = self.sim.volume_manager.material_database
        if isinstance(vol, ImageVolume):
            lab16 = vol.create_label_image(self.engine).array
            lut = vol.material_to_label_lut
            names = [None] * len(lut)
            for m, k in lut.items():
                names[k] = m
            spacing = vol.spacing
            size = vol.size_pix
        else:
            names = [vol.material]
            spacing = np.asarray(vol.size, dtype=np.float64)
            size = np.array([1, 1, 1])
            lab16 = np.zeros((1, 1, 1), np.uint16)
        if len(names) > 65535:
            raise ValueError("more than 65535 materials in one phantom (uint16 labels)")
        mats = [db.get(n) for n in names]
        world = self.sim.world
        world_index = -1
        if world.material not in VACUUM:
            if world.material in names:
                world_index = names.index(world.material)
            else:
                mats.append(db.get(world.material))
                names.append(world.material)
                world_index = len(mats) - 1
        R, t = self.sim.volume_manager.global_transform(vol.name)
        L = size * spacing
        self.R, self.t, self.L, self.size, self.spacing = R, t, L, size, spacing
        wlo = whi = (0.0, 0.0, 0.0)
        if world_index >= 0:
            if not np.allclose(R, np.eye(3)):
                raise NotImplementedError("a rotated phantom needs a vacuum world (world.material = 'G4_Galactic')")
            half = np.asarray(world.size, dtype=np.float64) / 2
            wlo = tuple(-half - t + L / 2)
            whi = tuple(half - t + L / 2)
        ph = PhantomSpec(np.asarray(lab16, dtype=np.uint16), tuple(spacing), world_index, wlo, whi)
        return ph, mats, names

    def to_phantom(self, p_world):
        return self.R.T @ (np.asarray(p_world, dtype=np.float64) - self.t) + self.L / 2

    # ------------------------------------------------------------------ sources
    def source_spec(self, src, n):
        vm = self.sim.volume_manager
        Ra, ta = vm.global_transform(src.attached if isinstance(src.attached, str) else src.attached.name)
        if isinstance(src, GANSource):
            # GateGANSource.cpp:227-254: position R_g (R_l p + t_l) + t_g, direction R_g d
            seed = int(src.seed) if src.seed is not None else int(self._seed)
            arrs = src.generate(n, seed, device=f"cuda:{self.sim.device}")
            Rl = _rot(src.position.rotation)
            tl = np.asarray(src.position.translation, dtype=np.float64)
            return SourceSpec(pos_type="phsp", pos_rot=self.R.T @ Ra @ Rl,
                              pos_trans=tuple(self.R.T @ (Ra @ tl + ta - self.t) + self.L / 2),
                              dir_rot=self.R.T @ Ra, phsp=arrs)
        if isinstance(src, PhaseSpaceSource):
            arrs = src.load_arrays()
            Rp = _rot(src.position.rotation) if src.rotate_direction else np.eye(3)
            tp = np.asarray(src.position.translation, dtype=np.float64) if src.translate_position else np.zeros(3)
            Rpos = _rot(src.position.rotation) if src.translate_position else np.eye(3)
            if src.global_flag:
                Rg, tg = np.eye(3), np.zeros(3)
            else:
                Rg, tg = Ra, ta
            pos_rot = self.R.T @ Rg @ Rpos
            pos_trans = self.R.T @ (Rg @ tp + tg - self.t) + self.L / 2
            dir_rot = self.R.T @ Rg @ Rp
            return SourceSpec(pos_type="phsp", pos_rot=pos_rot, pos_trans=tuple(pos_trans), dir_rot=dir_rot,
                              phsp=arrs, phsp_start=int(src.entry_start), phsp_iso=bool(src.isotropic_direction))
        assert isinstance(src, GenericSource)
        if src.particle != "gamma":
            raise NotImplementedError(f"source '{src.name}': particle '{src.particle}' (photon path: gamma only)")
        pos = src.position
        Rl = _rot(pos.rotation)
        Rg = Ra @ Rl
        Tg = Ra @ np.asarray(pos.translation, dtype=np.float64) + ta
        kind = pos.type
        if isinstance(src, VoxelSource):
            # voxel corner frame: p_attached = p + (-size/2 spacing + translation), then the
            # attached volume's global transform (GateSPSVoxelsPosDistribution.cpp:44-51)
            act, sp = src.activity_image()
            size = np.array(act.shape[::-1], dtype=np.float64)
            corner = -size / 2 * sp + np.asarray(pos.translation, dtype=np.float64)
            Rg = Ra
            Tg = Ra @ corner + ta
            kind = "voxels"
            pp = (0.0, 0.0, 0.0, 0.0)
        elif kind == "point":
            pp = (0.0, 0.0, 0.0, 0.0)
        elif kind == "box":
            pp = tuple(pos.size) + (0.0,)
        elif kind == "sphere":
            pp = (pos.radius, 0.0, 0.0, 0.0)
        elif kind == "disc":
            pp = (pos.radius, 0.0, 0.0, 0.0)
        elif kind == "surface_sphere":
            pp = (pos.radius, 0.0, 0.0, 0.0)
        elif kind == "cylinder":
            pp = (pos.radius, pos.dz if pos.dz is not None else 0.0, 0.0, 0.0)
        elif kind == "ellipsoid":
            pp = tuple(np.asarray(pos.size, dtype=np.float64) / 2) + (0.0,)
        else:
            raise NotImplementedError(f"source position type '{kind}'")
        if pos.confine:
            raise NotImplementedError("position.confine")
        d = src.direction
        rel = bool(src.direction_relative_to_attached_volume)
        dir_rot = self.R.T @ (Rg if rel else np.eye(3))
        spec = SourceSpec(pos_type=kind, pos_param=pp, pos_rot=self.R.T @ Rg,
                          pos_trans=tuple(self.R.T @ (Tg - self.t) + self.L / 2), dir_rot=dir_rot,
                          weight=float(src.weight) if src.weight is not None and src.weight > 0 else 1.0)
        if kind == "voxels":
            spec.activity = act
            spec.activity_spacing = tuple(sp)
        sx, sy = float(pos.get("sigma_x", 0.0) or 0.0), float(pos.get("sigma_y", 0.0) or 0.0)
        if sx < 0 or sy < 0:
            raise ValueError(f"source '{src.name}': position.sigma_x / sigma_y must be >= 0")
        if (sx > 0 or sy > 0) and kind != "disc":
            raise NotImplementedError(f"source '{src.name}': position.sigma_x / sigma_y apply to disc positions")
        spec.pos_sigma = (sx, sy)
        # a Gaussian weight only around a positive weight, as GateGenericSource::GeneratePrimaries
        ws = src.weight_sigma
        if ws is not None and ws >= 0 and src.weight is not None and src.weight > 0:
            spec.weight_sigma = float(ws)
        if d.type in ("iso", "cos"):
            spec.dir_type = d.type
            spec.theta = (float(d.theta[0]), float(d.theta[1]))
            spec.phi = (float(d.phi[0]), float(d.phi[1]))
            # on a sphere surface the angles are taken in the tangent frame built with the
            # source rotation's third column (G4SPSPosDistribution::GeneratePointsOnSurface)
            spec.surface_axis = tuple(Rg[:, 2])
        elif d.type == "beam2d":
            sig = np.asarray(d.sigma, dtype=np.float64).ravel()
            if len(sig) != 2 or np.any(~np.isfinite(sig)) or np.any(sig < 0):
                raise ValueError(f"source '{src.name}': direction.sigma must be two values >= 0")
            if np.any(sig > 6.0):
                raise NotImplementedError(f"source '{src.name}': direction.sigma above 6 rad")
            spec.dir_type, spec.dir_sigma = "beam2d", (float(sig[0]), float(sig[1]))
        elif d.type == "histogram":
            # GateGenericSource.cpp:437-466; the angles outside [0, pi] / [0, 2 pi] are drawn
            # again by Geant4, which clipping the bins reproduces
            spec.dir_type = "histogram"
            spec.theta_table = angle_histogram(d.histogram_theta_angles, d.histogram_theta_weights, 0.0, np.pi)
            spec.phi_table = angle_histogram(d.histogram_phi_angles, d.histogram_phi_weights, 0.0, 2 * np.pi)
        elif d.type == "momentum":
            spec.dir_type = "momentum"
            spec.momentum = tuple(np.asarray(d.momentum, dtype=np.float64))
        elif d.type == "focused":
            spec.dir_type = "focused"
            f = np.asarray(d.focus_point, dtype=np.float64)
            if rel:
                f = Rg @ (f - np.asarray(pos.translation, dtype=np.float64)) + Tg
            spec.focus = tuple(self.to_phantom(f))
        else:
            raise NotImplementedError(f"direction type '{d.type}'")
        if getattr(d, "angular_acceptance", None) and d.angular_acceptance.get("enabled"):
            raise NotImplementedError("direction.angular_acceptance")
        e = src.energy
        if e.type == "mono":
            spec.ene_type, spec.energy = "mono", float(e.mono)
        elif e.type == "range":
            spec.ene_type, spec.energy, spec.energy_max = "range", float(e.min_energy), float(e.max_energy)
        elif e.type == "gauss":
            if float(e.sigma_gauss) < 0:
                raise ValueError(f"source '{src.name}': sigma_gauss must be >= 0")
            spec.ene_type, spec.energy, spec.energy_sigma = "gauss", float(e.mono), float(e.sigma_gauss)
        elif e.type == "histogram":
            # G4SPSEneDistribution "User" points (GateGenericSource.cpp:530-542)
            spec.ene_type = "spectrum_histogram"
            spec.energy_value, spec.energy_width, spec.energy_cdf = energy_user_histogram(
                e.histogram_energy, e.histogram_weight)
        elif e.type == "spectrum_discrete":
            spec.ene_type = "spectrum_discrete"
            spec.energy_value, _, spec.energy_cdf = spectrum_lines(e.spectrum_energies, e.spectrum_weights)
        elif e.type == "spectrum_histogram":
            if e.spectrum_histogram_interpolation not in (None, "None", "none"):
                raise NotImplementedError("spectrum_histogram_interpolation")
            spec.ene_type = "spectrum_histogram"
            spec.energy_value, spec.energy_width, spec.energy_cdf = spectrum_histogram(
                e.spectrum_energy_bin_edges, e.spectrum_weights)
        else:
            raise NotImplementedError(f"energy type '{e.type}'")
        if e.is_cdf:
            raise NotImplementedError("energy.is_cdf")
        return spec

    # ------------------------------------------------------------------ tally
    def tally_spec(self, vol, ph, n_total=0):
        doses = [a for a in self.sim.actors.values() if isinstance(a, DoseActor)]
        if len(doses) > 1:
            raise NotImplementedError("one DoseActor per simulation")
        recs = [a for a in self.sim.actors.values() if isinstance(a, PhaseSpaceActor)]
        if len(recs) > 1:
            raise NotImplementedError("one PhaseSpaceActor per simulation")
        rec_face, rec_cap = -1, 0
        if recs:
            r = recs[0]
            if r.attached_to_name != vol.name:
                raise NotImplementedError("the PhaseSpaceActor records photons leaving the phantom volume")
            if self.sim.world.material not in VACUUM:
                raise NotImplementedError("exit-face recording needs a vacuum world")
            rec_face, rec_cap = FACES[r.exit_face], int(r.max_photons)
        if self.spect is not None:
            if recs:
                raise NotImplementedError("a PhaseSpaceActor on the phantom together with SPECT projections")
            # every photon leaving the phantom box goes to the heads; at most one per history
            rec_face, rec_cap = FACES["all"], int(max(1, min(n_total, spect.RECORD_CHUNK)))
        self.dose = doses[0] if doses else None
        if self.dose is None:
            return TallySpec(size=tuple(self.size), spacing=tuple(self.spacing), edep2=False,
                             record_face=rec_face, record_capacity=rec_cap)
        a = self.dose
        if a.attached_to_name != vol.name:
            raise NotImplementedError("the DoseActor must be attached to the phantom volume")
        size, spacing = a.resolved_grid(vol)
        tr = np.asarray(a.translation, dtype=np.float64)
        aligned = (np.array_equal(size, self.size) and np.allclose(spacing, self.spacing) and not tr.any())
        corner = self.L / 2 + tr - size * spacing / 2
        self.tally_grid = (size, spacing)
        return TallySpec(size=tuple(int(x) for x in size), spacing=tuple(spacing), corner=tuple(corner),
                         aligned=aligned, edep2=a.wants_squared(), dose=a.wants_dose() and not aligned,
                         dose2=a.wants_dose() and a.wants_squared() and not aligned,
                         counts=a.counts.active, record_face=rec_face, record_capacity=rec_cap,
                         **self._tle_fields(a))

    @staticmethod
    def _tle_fields(a):
        if not isinstance(a, TLEDoseActor):
            return {}
        e_min, e_max = a.tle_energy_bounds()
        return {"tle": True, "tle_energy_min": e_min, "tle_energy_max": e_max}

    # ------------------------------------------------------------------ run
    def run(self):
        sim = self.sim
        t_init = time.time()
        sim.physics_manager.check_supported()
        seed = sim.resolve_seed()
        self._seed = seed
        if any(isinstance(a, ARFTrainingDatasetActor) for a in sim.actors.values()):
            return self._run_arf_training(seed, t_init)
        self.engine = GpuEngine(sim.device)
        vol = self.phantom_volume()
        ph, mats, names = self.build_geometry(vol)
        # events of each source in each run interval (GateSourceManager run loop): counts
        # per run for number_of_primaries, Poisson counts of the activity, drawn from a
        # generator keyed by the seed, the source and the run
        per_run = self.events_per_run(seed)
        specs = {}
        for s, counts in per_run:
            n = sum(counts)
            if n > 0:
                specs[s.name] = (s, self.source_spec(s, n))
        # launch plan: run by run, each source's events of that run
        runs = []
        run_of = []  # run interval of each launch-plan entry
        done = {}
        for k in range(len(sim.run_timing_intervals)):
            for s, counts in per_run:
                if counts[k] > 0:
                    src, spec = specs[s.name]
                    if spec.pos_type == "phsp":
                        # a phase-space source goes on from the row where the last run ended
                        spec = dataclasses.replace(spec, phsp_start=spec.phsp_start + done.get(s.name, 0))
                    runs.append((src, spec, counts[k]))
                    run_of.append(k)
                    done[s.name] = done.get(s.name, 0) + counts[k]
        if not runs:
            raise ValueError("no source with primaries (set source.number_of_primaries or source.activity)")
        self.events_per_source = {s.name: counts for s, counts in per_run}
        e_max = max(spec.max_energy() for _, spec, _ in runs)
        checked = list(mats)
        if spect.spect_actors(sim):
            from .learn.arf import HEAD_MATERIALS

            checked += [sim.volume_manager.material_database.get(n) for n in HEAD_MATERIALS]
        e_cut = max(sim.physics_manager.resolve_photon_cut(checked, e_max), 2.0 ** -10)
        # the Woodcock majorant covers the materials that occur in the phantom
        present = np.flatnonzero(np.bincount(ph.labels.ravel(), minlength=len(mats)))
        tables = xs.build_tables(mats, e_max * (1 + 1e-6), e_cut=e_cut, majorant_materials=present)
        self.spect = None
        n_total = sum(n for _, _, n in runs)
        if spect.spect_actors(sim):
            self.spect = spect.SpectProjections(self, seed, e_max, e_cut)
        tal = self.tally_spec(vol, ph, n_total)
        self.engine.set_physics(tables)
        sid = self.engine.add_simulation(ph, runs[0][1], tal)
        t_init = time.time() - t_init

        counters = None
        # history-id offset of this job: jobs with one random_seed and disjoint ranges of
        # history ids draw the streams of one long run
        first0 = int(getattr(sim, "first_history", 0))
        if first0 < 0:
            raise ValueError("first_history must be >= 0")
        first = first0
        self.launch = self._launch_parameters(runs)
        self.executed = []  # (SourceSpec, histories, first history) in launch order
        self.inputs = (tables, ph, tal)
        t0 = time.time()
        goal = self.dose is not None and self.dose.uncertainty_goal is not None
        for k, (src, spec, n) in enumerate(runs):
            if k > 0:
                self.engine.set_source(sid, spec)
            if goal and (k == 0 or run_of[k] != run_of[k - 1]):
                self._goal_begin_run(first)
            if goal and self._goal_stopped:
                continue  # the run was stopped at the goal: its other sources add nothing
            todo = n
            while todo > 0:
                chunk = min(todo, 2**32 - 1)
                if goal:
                    chunk = min(chunk, self._goal_chunk(first))
                if self.spect is not None:
                    chunk = min(chunk, tal.record_capacity)
                c = self.engine.run([chunk], seed, first_history=[first], **self.launch)
                if self.spect is not None:
                    rec, m = self.engine.get_phsp(sid, device=True)
                    self.spect.consume(rec, min(m, tal.record_capacity))
                    self.engine.reset_phsp(sid)
                self.executed.append((spec, chunk, first))
                counters = c if counters is None else {key: counters[key] + c[key] for key in c}
                first += chunk
                todo -= chunk
                if goal and self._goal_reached(first):
                    todo = 0
        duration = time.time() - t0
        n_events = first - first0
        self.seed = seed
        self.finish(vol, ph, tal, names, tables, counters, n_events, duration, t_init)
        return counters

    def events_per_run(self, seed):
        """[(source, [events in run k])] for the simulation's run_timing_intervals."""
        out = []
        for i, s in enumerate(self.sim.sources.values()):
            rng = np.random.Generator(np.random.PCG64([int(seed) & 0xFFFFFFFFFFFFFFFF, i, 0x5EED]))
            out.append((s, s.events_per_run(self.sim.run_timing_intervals, rng)))
        return out

    def _launch_parameters(self, runs):
        """Launch parameters of the persistent kernel: sim.launch as given (a dict of
        gh_launch fields, {} for the library defaults), or "auto" to let gh_autotune pick
        them on the first source, with trial runs of 1/1000 of the histories per candidate
        (40 candidates: about 4 % of the run; at least 2e5 and at most 2e7 histories each).
        The trials leave the tallies as they were."""
        launch = self.sim.launch
        if isinstance(launch, str):
            if launch != "auto":
                raise ValueError(f"Simulation.launch must be a dict or 'auto', not {launch!r}")
            n = sum(k for _, _, k in runs)
            trial = int(min(max(n // 1000, 200_000), 20_000_000))
            kw, _ = self.engine.autotune([trial])
            self.autotuned = kw
            return kw
        return dict(launch)

    def _run_arf_training(self, seed, t_init):
        """ARFTrainingDatasetActor simulation: head model only, no phantom."""
        sim = self.sim
        t0 = time.time()
        n = spect.run_arf_training(sim, seed)
        duration = time.time() - t0
        self.seed = seed
        for s in sim.actors.values():
            if isinstance(s, SimulationStatisticsActor):
                s.counts.update(runs=len(sim.run_timing_intervals), events=int(n), tracks=int(n), steps=0,
                                duration=duration * g4_units.s, init=(t0 - t_init) * g4_units.s)
                s.raw = {"histories": int(n)}
                s.write_outputs()
        return {"histories": int(n)}

    # uncertainty goal (GateDoseActor::EndOfEventAction / ComputeMeanUncertainty,
    # GateDoseActor.cpp:245-330): per run (the event count and the next check restart at
    # BeginOfRunActionMasterThread, :83-89), on the run's own part of the tallies
    def _goal_begin_run(self, first):
        t = self.engine.get_tally(0)
        self._goal_base = (t["edep"].copy(), t["edep2"].copy())
        self._goal_first = first
        self._goal_next = max(1, int(self.dose.uncertainty_first_check_after_n_events))
        self._goal_stopped = False

    def _goal_chunk(self, first):
        return int(max(1, self._goal_next - (first - self._goal_first)))

    def _goal_reached(self, first):
        from .actors.dose import dose_actor_mean_uncertainty, next_uncertainty_check

        a = self.dose
        n = first - self._goal_first
        if n < self._goal_next:
            return False
        t = self.engine.get_tally(0)
        edep = t["edep"] - self._goal_base[0]
        edep2 = t["edep2"] - self._goal_base[1]
        u = dose_actor_mean_uncertainty(edep, edep2, n, a.uncertainty_top_voxels_count,
                                        a.uncertainty_voxel_edep_threshold)
        self.goal_uncertainty = u
        self.goal_checks = getattr(self, "goal_checks", []) + [(n, u)]
        if u <= a.uncertainty_goal:
            self._goal_stopped = True
            return True
        self._goal_next = next_uncertainty_check(u, a.uncertainty_goal, n, a.uncertainty_overshoot_factor_N_events)
        return False

    def finish(self, vol, ph, tal, names, tables, counters, n_events, duration, t_init):
        sim = self.sim
        t = self.engine.get_tally(0)
        if self.dose is not None:
            a = self.dose
            size, spacing = self.tally_grid
            vvox = float(np.prod(spacing))
            origin, direction = a.output_geometry(size, spacing, vol, sim.volume_manager.global_transform(vol.name))
            edep = t["edep"]
            edep2 = t.get("edep2")
            dose = dose2 = density = None
            if tal.aligned:
                rho = tables.density.astype(np.float64)[ph.labels]
                density = rho
                if a.wants_dose():
                    k = GY_FACTOR / vvox
                    inv = np.divide(1.0, rho, out=np.zeros_like(rho), where=rho > 0)
                    dose = edep * inv * k
                    if edep2 is not None:
                        dose2 = edep2 * inv * inv * k * k
            elif a.wants_dose():
                k = GY_FACTOR / vvox
                dose = t["dose"] * k
                if "dose2" in t:
                    dose2 = t["dose2"] * k * k
            if a.density.active and density is None:
                raise NotImplementedError("density output needs the dose grid of the image")
            a.store_results(size, spacing, origin, edep, edep2, dose, dose2, t.get("counts"), density, n_events,
                            direction)
            a.write_outputs()
        if self.spect is not None:
            self.spect.finish()
        for r in sim.actors.values():
            if isinstance(r, PhaseSpaceActor):
                arrs, n = self.engine.get_phsp(0)
                r.store(arrs, n, self.L / 2)
                r.write_outputs()
        for s in sim.actors.values():
            if isinstance(s, SimulationStatisticsActor):
                s.counts.update(runs=len(sim.run_timing_intervals), events=int(n_events),
                                tracks=int(n_events + counters["compton"] + counters["photo"]),
                                steps=int(counters["steps"]), duration=duration * g4_units.s,
                                init=t_init * g4_units.s)
                s.raw = dict(counters)
                s.write_outputs()
        self.engine.close()
