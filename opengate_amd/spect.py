"""SPECT runs of the Simulation facade: projections of the photons that leave the phantom
(DigitizerProjectionActor through the transported head model, ARFActor through the learned
response) and ARF training datasets (ARFTrainingDatasetActor), for one SPECTHeadVolume.

The phantom run records every photon that leaves the phantom box (all faces) into a device
buffer; after each chunk of histories the recorded photons are paired with the heads and
projected (learn.arf.project), then the recorder is emptied (gh_reset_phsp). The history
chunks keep the recorder within a fixed size whatever the number of primaries.
"""
import math

import numpy as np

from .actors.digitizers import (FULL_RANGE, ARFActor, ARFTrainingDatasetActor, DigitizerAdderActor,
                                DigitizerBlurringActor, DigitizerEnergyWindowsActor, DigitizerHitsCollectionActor,
                                DigitizerProjectionActor, check_channel_overlap)
from .geometry.spect import SPECTHeadVolume
from .sources.generic import GenericSource
from .units import g4_units

# photons recorded per history chunk (8 float32 arrays: 32 B per photon)
RECORD_CHUNK = 1 << 25
CHAIN_TYPES = (DigitizerHitsCollectionActor, DigitizerAdderActor, DigitizerBlurringActor)


def derived_seed(seed, tag):
    """A Philox key for a detector stage, distinct from the phantom run's (splitmix64)."""
    z = (int(seed) + 0x9E3779B97F4A7C15 * (1 + int(tag))) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return z ^ (z >> 31)


def spect_actors(sim):
    return [a for a in sim.actors.values() if isinstance(a, (DigitizerProjectionActor, ARFActor))]


def head_volume(sim, name=None):
    heads = [v for v in sim.volume_manager.volumes.values() if isinstance(v, SPECTHeadVolume)]
    if name is not None:
        v = sim.volume_manager.volumes.get(name)
        if not isinstance(v, SPECTHeadVolume):
            raise NotImplementedError(f"'{name}' is not a SPECTHeadVolume (projection and ARF actors are attached "
                                      "to one)")
        return v
    if len(heads) != 1:
        raise NotImplementedError("one SPECTHeadVolume per simulation")
    return heads[0]


def chain_blur(sim, name):
    """Energy blurring of the digitizer chain ending in the collection `name`: the Blur of
    its DigitizerBlurringActor, or no blurring. The chain is hits -> (adder) ->
    (blurring) in any order with at most one blurring stage."""
    from .learn.arf import Blur

    blur = None
    seen = set()
    while name is not None and name not in seen:
        seen.add(name)
        a = sim.actors.get(name)
        if a is None:
            if name == "Hits":
                break
            raise ValueError(f"digi collection '{name}' is not produced by any actor")
        if isinstance(a, DigitizerBlurringActor):
            if blur is not None:
                raise NotImplementedError("more than one energy blurring stage in a digitizer chain")
            blur = a.blur()
        elif not isinstance(a, CHAIN_TYPES):
            raise NotImplementedError(f"actor '{name}' ({type(a).__name__}) in a digitizer chain")
        name = getattr(a, "input_digi_collection", None)
    return blur if blur is not None else Blur(law="none")


def projection_inputs(sim, names):
    """Windows (lo, hi) in MeV and the Blur of the inputs of a DigitizerProjectionActor:
    energy channels of DigitizerEnergyWindowsActors, or chain collections (every event
    with energy in the crystal)."""
    channels = {}
    for a in sim.actors.values():
        if isinstance(a, DigitizerEnergyWindowsActor):
            for nm, lo, hi in a.windows():
                channels[nm] = (lo, hi, a)
    windows, blurs = [], set()
    for nm in names:
        if nm in channels:
            lo, hi, wa = channels[nm]
            windows.append((lo, hi))
            blurs.add(chain_blur(sim, wa.input_digi_collection))
        elif nm == "Hits" or isinstance(sim.actors.get(nm), CHAIN_TYPES):
            windows.append(FULL_RANGE)
            blurs.add(chain_blur(sim, nm))
        else:
            raise ValueError(f"projection input '{nm}' is neither an energy channel nor a digitizer collection")
    if len(blurs) != 1:
        raise NotImplementedError("projection inputs with different energy blurring")
    return windows, blurs.pop()


class SpectProjections:
    """Projection consumers of one phantom run."""

    def __init__(self, engine, seed, e_max, e_cut=None):
        sim = engine.sim
        self.sim = sim
        self.actors = spect_actors(sim)
        names = {a.attached_to_name for a in self.actors}
        if len(names) != 1:
            raise NotImplementedError("projection and ARF actors must be attached to the same SPECTHeadVolume")
        self.head = head_volume(sim, names.pop())
        if not np.allclose(engine.R, np.eye(3)):
            raise NotImplementedError("SPECT projections need an unrotated phantom (gantry axis = phantom z)")
        if sim.world.material not in ("G4_Galactic", "Vacuum", "G4_VACUUM"):
            sim.warnings.append("SPECT: photons fly from the phantom box to the heads in vacuum "
                                f"(world material {sim.world.material} outside the phantom is not transported)")
        center = engine.to_phantom(self.head.rotation_center)
        half = engine.L / 2
        reach = np.hypot(np.maximum(np.abs(center[0] - 0.0), np.abs(center[0] - 2 * half[0])),
                         np.maximum(np.abs(center[1] - 0.0), np.abs(center[1] - 2 * half[1])))
        if self.head.radius <= reach:
            raise ValueError(f"SPECTHeadVolume {self.head.name}: radius {self.head.radius} mm does not clear the "
                             f"phantom box (corner at {reach:.1f} mm from the gantry axis)")
        self.device = f"cuda:{sim.device}"
        self.seed = derived_seed(seed, 1)
        self.head_sim = None
        self.items = []
        for a in self.actors:
            if isinstance(a, ARFActor):
                arf = a.load_model(self.device)
                if abs(arf.theta_max - float(self.head.acceptance_angle)) > 1e-6:
                    raise ValueError(f"ARFActor {a.name}: the network was trained for an acceptance angle of "
                                     f"{math.degrees(arf.theta_max):.4g} deg, the head has "
                                     f"{math.degrees(self.head.acceptance_angle):.4g} deg")
                heads = self.head.heads(center, a.image_size, a.image_spacing, arf.e_range, theta_max=arf.theta_max)
                self.items.append({"actor": a, "heads": heads, "arf": arf, "out": None, "size": a.image_size,
                                   "spacing": a.image_spacing, "dtype": np.float32})
            else:
                if a.physical_volume_index not in (-1, 0) or not np.allclose(a.detector_orientation_matrix, np.eye(3)):
                    raise NotImplementedError("DigitizerProjectionActor: physical_volume_index and "
                                              "detector_orientation_matrix other than the defaults")
                windows, blur = projection_inputs(sim, a.input_digi_collections)
                if self.head_sim is None:
                    from .learn.arf import HeadSimulator

                    self.head_sim = HeadSimulator(self.head.head_model(), device=sim.device, e_max=e_max,
                                                  e_cut=e_cut)
                from .learn.arf import E_RANGE

                heads = self.head.heads(center, a.size, a.spacing, E_RANGE)
                self.items.append({"actor": a, "heads": heads, "windows": windows, "blur": blur, "out": None,
                                   "size": a.size, "spacing": a.spacing, "dtype": np.float64})

    def consume(self, rec, n):
        """Project n recorded photons (8 device arrays) into every consumer."""
        from .learn.arf import project

        if n == 0:
            return
        for it in self.items:
            a = it["actor"]
            if "arf" in it:
                it["out"] = project(rec, n, it["heads"], arf=it["arf"], squared=a.wants_squared(), out=it["out"])
            else:
                it["out"] = project(rec, n, it["heads"], head_sim=self.head_sim, seed=self.seed,
                                    squared=a.wants_squared(), windows=it["windows"], blur=it["blur"],
                                    out=it["out"])

    def finish(self):
        import torch

        for it in self.items:
            a, out = it["actor"], it["out"]
            key = "arf" if "arf" in it else "plain"
            h = it["heads"]
            n_win = it["arf"].n_win if "arf" in it else len(it["windows"])
            shape = (h.n_proj * n_win, h.ny, h.nx)
            if out is None:
                counts = np.zeros(shape)
                sq = np.zeros(shape) if a.wants_squared() else None
            else:
                torch.cuda.synchronize()
                counts = out[key].cpu().numpy()
                sq = out[key + "2"].cpu().numpy() if out.get(key + "2") is not None else None
            a.number_of_pairs = 0 if out is None else int(out["pairs"])
            a.store_images(counts, sq, it["size"], it["spacing"], it["dtype"])
            a.write_outputs()
        if self.head_sim is not None:
            self.head_sim.close()


# ---------------------------------------------------------------------- ARF training set
def _energy_sampler(src, torch, g, device):
    """E (MeV) sampler of a GenericSource energy description and its support (lo, hi)."""
    from .transport import spectrum_histogram, spectrum_lines

    e = src.energy
    if e.type == "mono":
        E0 = float(e.mono)
        return (lambda n: torch.full((n,), E0, device=device)), (E0, E0)
    if e.type == "range":
        lo, hi = float(e.min_energy), float(e.max_energy)
        return (lambda n: lo + (hi - lo) * torch.rand(n, device=device, generator=g)), (lo, hi)
    if e.type == "spectrum_discrete":
        val, _, cdf = spectrum_lines(e.spectrum_energies, e.spectrum_weights)
        v = torch.as_tensor(val, device=device)
        c = torch.as_tensor(cdf, device=device)
        # first k with u < cdf[k] (GateSPSEneDistribution.cpp:149-156)
        return (lambda n: v[torch.searchsorted(c, torch.rand(n, device=device, generator=g), right=True)
                            .clamp_(max=len(val) - 1)]), (float(val.min()), float(val.max()))
    if e.type == "spectrum_histogram":
        lo_e, width, cdf = spectrum_histogram(e.spectrum_energy_bin_edges, e.spectrum_weights)
        lo_t, w_t, c = (torch.as_tensor(x, device=device) for x in (lo_e, width, cdf))

        def sample(n):
            k = torch.searchsorted(c, torch.rand(n, device=device, generator=g), right=True).clamp_(max=len(lo_e) - 1)
            return lo_t[k] + w_t[k] * torch.rand(n, device=device, generator=g)

        return sample, (float(lo_e.min()), float((lo_e + width).max()))
    raise NotImplementedError(f"energy type '{e.type}' for an ARF training source")


def run_arf_training(sim, seed):
    """ARFTrainingDatasetActor run: returns the number of primaries."""
    import torch

    from .learn.arf import HeadSimulator

    actors = [a for a in sim.actors.values() if isinstance(a, ARFTrainingDatasetActor)]
    if len(actors) != 1:
        raise NotImplementedError("one ARFTrainingDatasetActor per simulation")
    actor = actors[0]
    if list(actor.plane_axis) != [0, 1, 2]:
        raise NotImplementedError("ARFTrainingDatasetActor.plane_axis other than [0, 1, 2]")
    head = head_volume(sim, actor.attached_to_name)
    wa = sim.actors.get(actor.energy_windows_actor)
    if not isinstance(wa, DigitizerEnergyWindowsActor):
        raise ValueError(f"ARFTrainingDatasetActor {actor.name}: energy_windows_actor = "
                         f"{actor.energy_windows_actor!r} is not a DigitizerEnergyWindowsActor of the simulation")
    channels = wa.windows()
    check_channel_overlap(channels)
    windows = [(lo, hi) for _, lo, hi in channels]
    blur = chain_blur(sim, wa.input_digi_collection)
    rr = int(actor.russian_roulette)
    if rr < 1:
        raise ValueError("russian_roulette must be >= 1")
    theta_acc = float(head.acceptance_angle)
    device = f"cuda:{sim.device}"
    g = torch.Generator(device=device).manual_seed(int(derived_seed(seed, 2) & 0x7FFFFFFFFFFFFFFF))
    plans, e_lo, e_hi = [], math.inf, 0.0
    for i, s in enumerate(sim.sources.values()):
        rng = np.random.Generator(np.random.PCG64([int(seed) & 0xFFFFFFFFFFFFFFFF, i, 0x5EED]))
        n = s.number_of_primaries_total(sim.run_timing_intervals, rng)
        if n <= 0:
            continue
        if not isinstance(s, GenericSource) or s.particle != "gamma":
            raise NotImplementedError("ARF training sources are gamma GenericSources")
        if s.direction.type != "iso":
            raise NotImplementedError("ARF training sources have 'iso' directions (angles from the face normal)")
        t0, t1 = float(s.direction.theta[0]), min(float(s.direction.theta[1]), theta_acc)
        if not t1 > t0:
            raise ValueError(f"source {s.name}: its theta range does not meet the head acceptance angle")
        sample_e, (lo, hi) = _energy_sampler(s, torch, g, device)
        e_lo, e_hi = min(e_lo, lo), max(e_hi, hi)
        plans.append((n, sample_e, (t0, t1), (float(s.direction.phi[0]), float(s.direction.phi[1]))))
    if not plans:
        raise ValueError("no source with primaries (set source.number_of_primaries or source.activity)")
    e_range = (e_lo if e_lo < e_hi else 0.0, e_hi)
    from .learn.arf import HEAD_MATERIALS

    db = sim.volume_manager.material_database
    e_cut = max(sim.physics_manager.resolve_photon_cut([db.get(n) for n in HEAD_MATERIALS], e_hi), 2.0 ** -10)
    hs = HeadSimulator(head.head_model(), device=sim.device, e_max=e_hi, e_cut=e_cut)
    det_seed = derived_seed(seed, 3)
    rows = {"E": [], "Theta": [], "Phi": [], "window": []}
    first = 0
    counts = np.zeros(len(windows) + 1, dtype=np.int64)
    try:
        for n, sample_e, (t0, t1), (p0, p1) in plans:
            todo = int(n)
            while todo > 0:
                k = min(todo, 1 << 23)
                a = torch.rand(k, device=device, generator=g)
                b = torch.rand(k, device=device, generator=g)
                E = sample_e(k).float()
                ct = math.cos(t1) + (math.cos(t0) - math.cos(t1)) * torch.rand(k, device=device, generator=g)
                theta = torch.acos(ct.clamp(-1.0, 1.0))
                phi = p0 + (p1 - p0) * torch.rand(k, device=device, generator=g)
                phi = torch.remainder(phi + math.pi, 2.0 * math.pi) - math.pi
                rec = hs.photons(a, b, E, theta, phi)
                cls = hs.windows(rec, k, det_seed, first=first, windows=windows, blur=blur).long()
                counts += np.bincount(cls.cpu().numpy(), minlength=len(windows) + 1)
                keep = cls > 0
                if rr > 1:
                    keep |= torch.rand(k, device=device, generator=g) < 1.0 / rr
                rows["E"].append(E[keep].double().cpu().numpy())
                rows["Theta"].append(torch.rad2deg(theta[keep]).double().cpu().numpy())
                rows["Phi"].append(torch.rad2deg(phi[keep]).double().cpu().numpy())
                rows["window"].append(cls[keep].double().cpu().numpy())
                first += k
                todo -= k
    finally:
        hs.close()
    data = {k: np.concatenate(v) for k, v in rows.items()}
    data["channels"] = np.array([[lo, hi] for lo, hi in windows], dtype=np.float64)
    data["channel_names"] = np.array([nm for nm, _, _ in channels])
    data["acceptance_angle"] = np.float64(theta_acc)
    data["e_range"] = np.array(e_range, dtype=np.float64)
    data["russian_roulette"] = np.int64(rr)
    actor.data = data
    actor.number_of_rows = len(data["E"])
    actor.window_counts = counts
    actor.write_outputs()
    return first
