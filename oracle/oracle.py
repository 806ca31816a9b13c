"""Oracle of the photon transport + dose path (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this module,
and only as the checker or the CPU baseline. The product never imports it.

Contents
  * ctypes binding of liboracle.so (transport_ref.c): per-history CPU restatement
    of the transport, Philox streams, the deterministic log / sincos, HU -> labels;
  * numpy restatements:
      - label image (ImageVolume.create_label_image, geometry/volumes.py:1144-1243);
      - ITK TransformPhysicalPointToIndex contract for a voxel index (GateDoseActor.cpp:177,
        GateHelpersImage.txx:60-93);
      - variance / uncertainty (utility.py:551-560, actors/dataitems.py:1956-1998);
      - dose scaling (actors/doseactors.py:686-696);
      - FP64 Beer-Lambert line integral through a label volume (Siddon's method),
        the oracle of the uncollided primary fluence.

Pinning (tests/test_oracle_*.py, test_voxel_index.py, test_gpu_index.py): Philox against
the Random123 known-answer vectors; digitize against numpy.digitize; uncertainty against
the closed form of a known per-history distribution; Siddon chords against the clipped
segment length and Beer-Lambert against exp(-mu L) in homogeneous boxes and across a
plane interface; the table lookup mu_at against the table nodes and the FP64 mixture;
dose_gy against a hand-worked unit conversion; the cell index against exact rational
floor((p - c)/s) and the FP64 ITK formula on every face of small grids; the transport
samplers against Klein-Nishina / Thomson closed forms; energy conservation of whole runs.
"""
import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        from opengate_amd import _abi as A

        L = C.CDLL(str(LIB))
        L.ref_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.ref_stream_words.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.POINTER(C.c_uint32), C.c_int]
        L.ref_log.restype = C.c_float
        L.ref_log.argtypes = [C.c_float]
        L.ref_sincos.argtypes = [C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        DP = C.POINTER(C.c_double)
        L.ref_run.restype = C.c_int
        L.ref_run.argtypes = [C.POINTER(A.PhysicsDesc), C.POINTER(A.PhantomDesc), C.POINTER(A.SourceDesc),
                              C.POINTER(A.TallyDesc), C.c_uint32, C.c_uint64, C.c_uint64, C.c_uint64, C.c_int,
                              DP, DP, DP, DP, DP, DP, C.POINTER(A.FP * 8), C.c_uint64, C.POINTER(C.c_uint64),
                              A.FP, C.POINTER(C.c_uint64), C.POINTER(C.c_double)]
        L.ref_sample_compton.argtypes = [C.POINTER(A.PhysicsDesc), C.c_int, C.c_float, C.c_uint64, C.c_uint64,
                                         A.FP, A.FP]
        L.ref_sample_rayleigh.argtypes = [C.POINTER(A.PhysicsDesc), C.c_int, C.c_float, C.c_uint64, C.c_uint64,
                                          A.FP]
        L.ref_sample_source.argtypes = [C.POINTER(A.SourceDesc), C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint32,
                                        C.POINTER(A.FP * 8)]
        L.ref_mu_at.restype = C.c_float
        L.ref_mu_at.argtypes = [C.POINTER(A.PhysicsDesc), C.c_int, C.c_float]
        L.ref_cell_index.restype = C.c_int
        L.ref_cell_index.argtypes = [C.c_float, C.c_float, C.c_float, C.c_float]
        L.ref_hu_to_labels.argtypes = [C.POINTER(C.c_int16), C.c_int64, DP, C.c_int, C.POINTER(C.c_uint16),
                                       C.POINTER(C.c_uint16)]
        _lib = L
    return _lib


# ------------------------------------------------------------------ RNG / math
def philox4x32_10(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().ref_philox4x32_10(c, k, o)
    return [int(x) for x in o]


def stream_words(seed, history, sim, n):
    out = (C.c_uint32 * n)()
    lib().ref_stream_words(seed, history, sim, out, n)
    return np.frombuffer(out, dtype=np.uint32).copy()


def ref_log(x):
    return lib().ref_log(float(x))


def ref_sincos(x):
    s, c = C.c_float(), C.c_float()
    lib().ref_sincos(float(x), C.byref(s), C.byref(c))
    return s.value, c.value


# ------------------------------------------------------------------ transport
def run(tables, phantom, source, tally, n, seed, first=0, sim_id=0, threads=1, exact=False, record=False,
        event=False):
    """CPU restatement of gh_run for one simulation. Returns (tallies, counters)."""
    from opengate_amd import _abi as A
    from opengate_amd.transport import physics_desc

    pd = physics_desc(tables)
    phd, sd, td = phantom.desc(), source.desc(), tally.desc()
    shape = tuple(reversed([int(x) for x in tally.size]))
    out = {"edep": np.zeros(shape)}
    if tally.edep2:
        out["edep2"] = np.zeros(shape)
    if tally.dose:
        out["dose"] = np.zeros(shape)
    if tally.dose2:
        out["dose2"] = np.zeros(shape)
    if tally.counts:
        out["counts"] = np.zeros(shape)
    if exact:
        out["exact_edep2"] = np.zeros(shape)
    rec = None
    rec_cap = 0
    if record:
        rec_cap = int(tally.record_capacity)
        out["phsp"] = [np.zeros(rec_cap, dtype=np.float32) for _ in range(8)]
        rec = (A.FP * 8)(*[A.fptr(a) for a in out["phsp"]])
    ev = None
    if event:
        out["event_edep"] = np.zeros(int(n), dtype=np.float32)
        ev = A.fptr(out["event_edep"])
    rec_n = C.c_uint64(0)
    cnt = (C.c_uint64 * 12)()
    en = (C.c_double * 4)()
    rc = lib().ref_run(C.byref(pd), C.byref(phd), C.byref(sd), C.byref(td), int(sim_id), int(n), int(first),
                       int(seed), int(threads), A.dptr(out["edep"]), A.dptr(out.get("edep2")),
                       A.dptr(out.get("dose")), A.dptr(out.get("dose2")), A.dptr(out.get("counts")),
                       A.dptr(out.get("exact_edep2")), C.byref(rec) if rec is not None else None, rec_cap,
                       C.byref(rec_n), ev if ev is not None else A.FP(), cnt, en)
    if rc != 0:
        raise RuntimeError("ref_run failed")
    counters = {name: int(cnt[i]) for i, name in enumerate(A.COUNTER_NAMES)}
    for i, name in enumerate(A.ENERGY_NAMES):
        counters[name] = float(en[i])
    counters["recorded_total"] = int(rec_n.value)
    return out, counters


def sample_compton(tables, m, E, n, seed=1):
    from opengate_amd import _abi as A
    from opengate_amd.transport import physics_desc

    pd = physics_desc(tables)
    q = np.zeros(n, np.float32)
    c = np.zeros(n, np.float32)
    lib().ref_sample_compton(C.byref(pd), m, E, n, seed, A.fptr(q), A.fptr(c))
    return q, c


def sample_rayleigh(tables, m, E, n, seed=1):
    from opengate_amd import _abi as A
    from opengate_amd.transport import physics_desc

    pd = physics_desc(tables)
    c = np.zeros(n, np.float32)
    lib().ref_sample_rayleigh(C.byref(pd), m, E, n, seed, A.fptr(c))
    return c


def sample_source(source, n, seed=1, first=0, sim_id=0):
    """Primaries of histories first .. first + n - 1 as emitted by the source, before any
    transport: 8 float32 arrays (x, y, z, u, v, w, E, weight)."""
    from opengate_amd import _abi as A

    sd = source.desc()
    out = [np.zeros(n, np.float32) for _ in range(8)]
    ptrs = (A.FP * 8)(*[A.fptr(a) for a in out])
    lib().ref_sample_source(C.byref(sd), n, first, seed, sim_id, C.byref(ptrs))
    return out


def mu_at(tables, m, E):
    from opengate_amd.transport import physics_desc

    pd = physics_desc(tables)
    return lib().ref_mu_at(C.byref(pd), m, float(E))


def hu_to_labels_c(hu, bins, lut):
    hu = np.ascontiguousarray(hu, dtype=np.int16)
    bins = np.ascontiguousarray(bins, dtype=np.float64)
    lut = np.ascontiguousarray(lut, dtype=np.uint16)
    out = np.zeros(hu.shape, dtype=np.uint16)
    lib().ref_hu_to_labels(hu.ctypes.data_as(C.POINTER(C.c_int16)), hu.size, bins.ctypes.data_as(C.POINTER(C.c_double)),
                           len(bins), lut.ctypes.data_as(C.POINTER(C.c_uint16)), out.ctypes.data_as(C.POINTER(C.c_uint16)))
    return out


# ------------------------------------------------------------------ numpy restatements
def label_lut(material, voxel_materials):
    """ImageVolume.create_material_to_label_lut (volumes.py:1144-1168)."""
    lut = {material: 0}
    rows = sorted(voxel_materials, key=lambda r: r[0])
    lo = np.array([r[0] for r in rows])
    hi = np.array([r[1] for r in rows])
    if not (lo[1:] >= hi[:-1]).all():
        raise ValueError("overlapping intervals")
    i = 1
    for r in rows:
        if r[2] not in lut:
            lut[r[2]] = i
            i += 1
    return lut


def label_bins(lut, voxel_materials):
    """bins and labels_sorted of ImageVolume.create_label_image (volumes.py:1200-1243)."""
    rows = sorted(voxel_materials, key=lambda r: r[0])
    bins = [r[0] for r in rows]
    extra = set([r[1] for r in rows]).difference(bins)
    bins.extend(extra)
    labels = [lut[r[2]] for r in rows] + [0] * len(extra)
    bs, ls = [], [0]
    for b, l in sorted(zip(bins, labels), key=lambda p: p[0]):
        bs.append(b)
        ls.append(l)
    return np.array(bs, dtype=np.float64), np.array(ls, dtype=np.uint16)


def label_image(hu, material, voxel_materials):
    lut = label_lut(material, voxel_materials)
    bins, ls = label_bins(lut, voxel_materials)
    return ls[np.digitize(hu, bins=bins)], lut


def cell_index(p, corner, spacing):
    """The C restatement's voxel index of an FP32 coordinate: floor((p - corner) / spacing)
    for the real values of the FP32 inputs (ref_cell_index, the kernel's cell_index)."""
    p, c, s = np.float32(p), np.float32(corner), np.float32(spacing)
    return lib().ref_cell_index(float(p), float(c), float(s), float(np.float32(1.0) / s))


def itk_index(point, origin, spacing, size):
    """itk::ImageBase::TransformPhysicalPointToIndex (ITK 5.4, itkImageBase.h) for an identity
    direction, in FP64 as ITK computes it: c = PhysicalPointToIndex (p - origin) with the
    matrix inverse(diag(spacing)) = diag(1 / spacing), idx = RoundHalfIntegerUp(c) =
    floor(c + 0.5); inside iff 0 <= idx < size. The point and origin are those of the
    dose image (origin = centre of voxel 0), as GateDoseActor::GetVoxelPosition passes them
    (GateDoseActor.cpp:144-178)."""
    c = (np.asarray(point, dtype=np.float64) - np.asarray(origin, dtype=np.float64)) * (
        1.0 / np.asarray(spacing, dtype=np.float64))
    idx = np.floor(c + 0.5).astype(np.int64)
    inside = np.all((idx >= 0) & (idx < np.asarray(size)), axis=-1)
    return idx, inside


def uncertainty(S, S2, N):
    """get_variance_or_uncertainty('uncertainty') (dataitems.py:1956-1998)."""
    S = np.asarray(S, dtype=np.float64)
    if not N > 1:
        return np.ones_like(S)
    if S2 is None:
        return np.zeros_like(S)
    var = np.clip((np.asarray(S2) / N - (S / N) ** 2) / (N - 1), 0, None)
    std = np.sqrt(var)
    return np.divide(std, S / N, out=np.zeros_like(std), where=S != 0)


GY_PER_MEV_G_CM3_MM3 = 1.602176634e-7  # 1 MeV / (1 g/cm3 * 1 mm3) in Gy


def dose_gy(edep_mev, density_g_cm3, voxel_volume_mm3):
    """DoseActor dose: sum(edep / rho) / (Gy * V) (GateDoseActor.cpp:194-214, doseactors.py:686-696)."""
    return np.divide(edep_mev, density_g_cm3, out=np.zeros_like(edep_mev, dtype=np.float64),
                     where=density_g_cm3 > 0) * GY_PER_MEV_G_CM3_MM3 / voxel_volume_mm3


def siddon_path_lengths(p0, p1, shape_xyz, spacing):
    """Voxel indices and chord lengths of the segment p0 -> p1 through the grid
    [0, n*spacing) (Siddon 1985). FP64."""
    p0 = np.asarray(p0, dtype=np.float64)
    p1 = np.asarray(p1, dtype=np.float64)
    d = p1 - p0
    n = np.asarray(shape_xyz)
    s = np.asarray(spacing, dtype=np.float64)
    alphas = [0.0, 1.0]
    for i in range(3):
        if d[i] != 0:
            planes = np.arange(n[i] + 1) * s[i]
            a = (planes - p0[i]) / d[i]
            alphas.extend(a[(a > 0) & (a < 1)].tolist())
    alphas = np.unique(np.array(alphas))
    length = np.linalg.norm(d)
    mids = 0.5 * (alphas[1:] + alphas[:-1])
    pts = p0[None, :] + mids[:, None] * d[None, :]
    idx = np.floor(pts / s[None, :]).astype(np.int64)
    ok = np.all((idx >= 0) & (idx < n[None, :]), axis=1)
    seg = np.diff(alphas) * length
    return idx[ok], seg[ok]


def beer_lambert(p0, p1, labels_zyx, spacing, mu_per_label):
    """exp(-sum mu_label * chord) along p0 -> p1 (FP64)."""
    idx, seg = siddon_path_lengths(p0, p1, tuple(reversed(labels_zyx.shape)), spacing)
    lab = labels_zyx[idx[:, 2], idx[:, 1], idx[:, 0]]
    return float(np.exp(-np.sum(np.asarray(mu_per_label)[lab] * seg)))


# ------------------------------------------------------------------ source run loop
def source_events_reference(run_timing_intervals, start_time, end_time, activity, half_life, n_per_run, uniform):
    """Events of one source per run, following the reference's loop event by event
    (GateSourceManager::PrepareRunToStart / PrepareNextSource, GateSourceManager.cpp:313-359;
    GateVSource::PrepareNextTime, CalcNextTime, UpdateActivity, GateVSource.cpp:46-89).
    Times in seconds, activity in 1/s (already multiplied by the spectrum yield), half_life
    in s (<= 0: none), n_per_run: list (number_of_primaries) or None (activity);
    uniform(): a uniform draw in (0, 1). Pure-Python loop: small activities only."""
    counts = []
    lam = np.log(2.0) / half_life if half_life > 0 else 0.0
    for run, (a, b) in enumerate(run_timing_intervals):
        max_n = 0 if n_per_run is None else int(n_per_run[run])
        t = a  # fCurrentSimulationTime at the start of the run
        generated = 0
        while True:
            act = activity * np.exp(-lam * (t - start_time)) if lam > 0 else activity
            if not (max_n <= 0 or max_n > generated):
                break
            if t < start_time:
                nt = start_time
            elif t >= end_time:
                break
            else:
                nt = t if max_n > 0 else t - np.log(uniform()) / act
                if nt >= end_time:
                    break
            if not (a <= nt < b):  # PrepareNextSource keeps times inside the run interval
                break
            t = nt
            generated += 1
        counts.append(generated)
    return counts


# ------------------------------------------------------------------ uncertainty goal
def dose_mean_uncertainty_reference(edep, edep2, n_events, top_count, thresh):
    """GateDoseActor::ComputeMeanUncertainty and GetMeanOfHighestNValues
    (GateDoseActor.cpp:281-330, 401-428), voxel by voxel as the C++ iterates (a min-heap of
    the top values, then the thresholded mean). Pure Python: small images only."""
    import heapq

    flat = [float(v) for v in np.asarray(edep, dtype=np.float64).ravel()]
    flat2 = [float(v) for v in np.asarray(edep2, dtype=np.float64).ravel()]
    heap = []
    for v in flat:
        if len(heap) < top_count:
            heapq.heappush(heap, v)
        elif v > heap[0]:
            heapq.heapreplace(heap, v)
    mean_max = sum(heap) / len(heap) if heap else 0.0
    n = float(n_events)
    if n < 2.0:
        n = 2.0
    mean_unc, n_vox = 0.0, 0
    for v, v2 in zip(flat, flat2):
        if v > mean_max * thresh:
            v /= n
            n_vox += 1
            sq = v2 / n
            u = (1.0 / (n - 1.0)) * (sq - v ** 2)
            if u < 0:
                u = 0.0
            u = u ** 0.5 / v
            if u > 1:
                u = 1.0
            mean_unc += u
    if n_vox > 0 and mean_unc > 0:
        return mean_unc / n_vox
    return 1.0
