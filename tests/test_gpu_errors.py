"""Invalid inputs are rejected by the C-ABI before any kernel runs (host-side checks; the
context stays usable afterwards)."""
import numpy as np
import pytest

from opengate_amd.transport import GhError, PhantomSpec, SourceSpec, TallySpec

from helpers import aligned_tally, centre, small_phantom, tables

pytestmark = pytest.mark.gpu


def test_rejections_leave_the_context_usable(engine):
    T = tables()
    ph = small_phantom(8, 2.0)
    src = SourceSpec(pos_trans=centre(ph), energy=0.1)
    tal = aligned_tally(ph)
    engine.clear()
    with pytest.raises(GhError):
        engine.run([10], 1)  # no physics, no simulation
    engine.set_physics(T)
    bad = PhantomSpec(np.full((8, 8, 8), 5, np.uint8), (2.0, 2.0, 2.0))
    with pytest.raises(GhError, match="label"):
        engine.add_simulation(bad, src, tal)
    with pytest.raises(GhError):
        engine.add_simulation(ph, SourceSpec(pos_trans=centre(ph), energy=0.5), tal)  # beyond the tables
    with pytest.raises(GhError):
        engine.add_simulation(ph, src, TallySpec(size=(8, 8, 9), spacing=(2.0, 2.0, 2.0)))  # aligned, wrong size
    with pytest.raises(GhError):
        engine.add_simulation(ph, src, aligned_tally(ph, record_face=7, record_capacity=10))
    air = PhantomSpec(ph.labels, ph.spacing, world_material=0, world_lo=(-50.0,) * 3, world_hi=(66.0,) * 3)
    with pytest.raises(GhError, match="vacuum"):
        engine.add_simulation(air, src, aligned_tally(ph, record_face=4, record_capacity=10))
    with pytest.raises(GhError):
        engine.add_simulation(ph, src, aligned_tally(ph, tle=True, detector_label_min=1))
    bad_cdf = SourceSpec(pos_trans=centre(ph), ene_type="spectrum_discrete", energy_value=np.array([0.1, 0.05], np.float32),
                         energy_cdf=np.array([0.6, 0.5], np.float32))
    with pytest.raises(GhError):
        engine.add_simulation(ph, bad_cdf, tal)
    # a majorant over water only: a phantom holding bone is refused, water alone runs
    from helpers import DB
    from opengate_amd.physics import xs

    Tw = xs.build_tables([DB.get("G4_WATER"), DB.get("G4_BONE_CORTICAL_ICRP")], 0.1405, majorant_materials=[0])
    engine.set_physics(Tw)
    with pytest.raises(GhError, match="material 1 occurs in the phantom but its mu exceeds the majorant"):
        engine.add_simulation(ph, src, tal)
    water = PhantomSpec(np.zeros((8, 8, 8), np.uint8), (2.0, 2.0, 2.0))
    engine.add_simulation(water, src, aligned_tally(water))
    assert engine.run([1000], 1)["histories"] == 1000
    engine.clear()
    engine.set_physics(T)
    # TLE tallies need single-simulation launches
    engine.add_simulation(ph, src, aligned_tally(ph, tle=True))
    engine.add_simulation(ph, src, tal)
    with pytest.raises(GhError, match="single-simulation"):
        engine.run([100, 100], 1)
    # still usable
    engine.clear()
    engine.set_physics(T)
    engine.add_simulation(ph, src, tal)
    c = engine.run([1000], 1)
    assert c["histories"] == 1000


def test_gemm_and_kernel_argument_checks():
    import ctypes as C

    import torch

    from opengate_amd import _abi as A

    lib = A.load_library()
    x = torch.zeros(64, 64, dtype=torch.bfloat16, device="cuda")
    out = torch.zeros(64, 64, device="cuda")
    p = lambda t: C.c_void_p(t.data_ptr())  # noqa: E731
    assert lib.gh_gemm_bf16(p(x), 64, 1, p(x), 64, 1, 64, 64, 64, None, None, 0, None, None, 64, 0, None) != 0
    assert lib.gh_gemm_bf16(p(x), 64, 1, p(x), 64, 1, 64, 64, 64, None, None, 0, p(out), None, 64, 9, None) != 0
    assert lib.gh_gaussian_bf16(1, 0, 10, 9, 8, p(x), None) != 0  # more columns than the row pitch
    w = (C.c_void_p * 2)(x.data_ptr(), x.data_ptr())
    b = (A.FP * 2)(C.cast(p(out), A.FP), C.cast(p(out), A.FP))
    assert lib.gh_mlp_forward_bf16(p(x), 64, 64, 2, 8, 192, 4, w, b, C.cast(p(out), A.FP), 64, None) != 0  # width
    assert lib.gh_mlp_forward_bf16(p(x), 64, 64, 9, 8, 256, 4, w, b, C.cast(p(out), A.FP), 64, None) != 0  # layers
    win = (C.c_float * 2)(0.15, 0.12)  # empty window
    cls = torch.zeros(4, dtype=torch.int32, device="cuda")
    blur = A.Blur(A.BLUR_LAWS["InverseSquare"], 0.0, 0.1, 0.14, 0.0)
    assert lib.gh_energy_windows(p(out), 4, 1, 0, C.byref(blur), win, 1, p(cls), None) != 0
    good = (C.c_float * 2)(0.12, 0.15)
    assert lib.gh_energy_windows(p(out), 4, 1, 0, C.byref(A.Blur(7, 0.0, 0.1, 0.14, 0.0)), good, 1, p(cls),
                                 None) != 0  # unknown law
    assert lib.gh_energy_windows(p(out), 4, 1, 0, C.byref(A.Blur(1, -1.0, 0.0, 0.0, 0.0)), good, 1, p(cls),
                                 None) != 0  # negative sigma
    prob = np.zeros(3, np.float32)
    alias = np.zeros(3, np.int32)
    assert lib.gh_alias_table(A.dptr(np.array([1.0, -1.0, 2.0])), 3, A.fptr(prob), alias.ctypes.data_as(C.POINTER(C.c_int32))) != 0
    assert lib.gh_alias_table(A.dptr(np.zeros(3)), 3, A.fptr(prob), alias.ctypes.data_as(C.POINTER(C.c_int32))) != 0
