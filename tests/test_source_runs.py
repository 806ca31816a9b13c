"""Number of events per run of a source, as the reference's run loop decides it:
number_of_primaries per run interval (an int for one run, a list otherwise), activity
scaled by the spectrum's total yield, exponential waiting times with half-life decay, the
source's [start_time, end_time) window, and the reference's configuration errors
(sources/base.py:256-291; GateVSource.cpp:46-89; GateSourceManager.cpp:313-359;
GateGenericSource.cpp:544-632).

The oracle's event-by-event loop (oracle.source_events_reference) is pinned by the Poisson
law it must produce for a constant activity (mean and variance of the count) and by the
decay integral; the product's counts (one Poisson draw of the integrated activity per run)
are compared with it in distribution."""
import numpy as np
import pytest

from oracle import oracle as O
from opengate_amd import Simulation
from opengate_amd.sources.generic import GenericSource
from opengate_amd.units import g4_units

s_ = g4_units.s
Bq = g4_units.Bq


def source(**kw):
    src = GenericSource("src")
    for k, v in kw.items():
        setattr(src, k, v)
    return src


def ref_counts(trials, intervals_s, act, hl=-1.0, start=None, end=None, seed=1):
    g = np.random.default_rng(seed)
    uni = lambda: 1.0 - g.random()  # noqa: E731  (0, 1]
    st = intervals_s[0][0] if start is None else start
    en = intervals_s[-1][1] if end is None else end
    return np.array([O.source_events_reference(intervals_s, st, en, act, hl, None, uni) for _ in range(trials)])


def test_reference_loop_is_poisson_for_a_constant_activity():
    c = ref_counts(2000, [(0.0, 2.0)], 20.0)[:, 0]
    mean, var = 40.0, 40.0
    assert abs(c.mean() - mean) < 4 * np.sqrt(var / len(c))
    assert abs(c.var() / var - 1) < 0.15


def test_reference_loop_follows_the_decay_integral():
    hl = 1.5
    c = ref_counts(2000, [(0.0, 1.0), (1.0, 3.0)], 30.0, hl=hl)
    lam = np.log(2) / hl
    # the reference updates the rate at each event: first order in lam / A of the integral
    exp = [30.0 / lam * (1 - np.exp(-lam)), 30.0 / lam * (np.exp(-lam) - np.exp(-3 * lam))]
    for k in range(2):
        assert abs(c[:, k].mean() / exp[k] - 1) < 0.03


@pytest.mark.parametrize("act,hl", [(25.0, -1.0), (300.0, 2.0)])
def test_product_counts_match_the_reference_loop_in_distribution(act, hl):
    # the first run starts before start_time (one event at start_time itself); the decay
    # case keeps (ln2 / half_life) / A small, as real sources have it
    intervals = [(0.0, 1.0), (1.0, 2.5)]
    src = source(activity=act * Bq, half_life=hl * s_, start_time=0.3 * s_, end_time=2.2 * s_)
    g = np.random.Generator(np.random.PCG64(7))
    mine = np.array([src.events_per_run([(a * s_, b * s_) for a, b in intervals], g) for _ in range(2000)])
    ref = ref_counts(2000, intervals, act, hl=hl, start=0.3, end=2.2, seed=3)
    for k in range(2):
        m1, m2 = mine[:, k].mean(), ref[:, k].mean()
        se = np.sqrt(mine[:, k].var() / len(mine) + ref[:, k].var() / len(ref))
        assert abs(m1 - m2) < 4 * se, (k, m1, m2)
        assert abs(mine[:, k].var() / ref[:, k].var() - 1) < 0.2


def test_number_of_primaries_per_run_and_time_window():
    iv = [(0.0, 1.0 * s_), (1.0 * s_, 2.0 * s_), (2.0 * s_, 3.0 * s_)]
    src = source(number_of_primaries=[10, 20, 30])
    assert src.events_per_run(iv, None) == [10, 20, 30]
    src.start_time, src.end_time = 1.5 * s_, 2.0 * s_
    assert src.events_per_run(iv, None) == [0, 20, 0]
    uni = iter([0.5] * 100).__next__
    assert O.source_events_reference([(0, 1), (1, 2), (2, 3)], 1.5, 2.0, 0.0, -1, [10, 20, 30], uni) == [0, 20, 0]
    # one run: an int is enough; the deprecated alias n works too
    assert source(number_of_primaries=7).events_per_run([(0.0, 1.0 * s_)], None) == [7]
    assert source(n=5).events_per_run([(0.0, 1.0 * s_)], None) == [5]


def test_configuration_errors_as_the_reference():
    iv = [(0.0, 1.0 * s_), (1.0 * s_, 2.0 * s_)]
    with pytest.raises(ValueError, match="same length"):
        source(number_of_primaries=10).events_per_run(iv, None)
    with pytest.raises(ValueError, match="both"):
        source(number_of_primaries=[1, 2], activity=3 * Bq).events_per_run(iv, None)
    with pytest.raises(ValueError, match="one of"):
        source().events_per_run(iv[:1], None)
    # with two runs and nothing set, the length check comes first, as in the reference
    with pytest.raises(ValueError, match="same length"):
        source().events_per_run(iv, None)


def test_activity_is_scaled_by_the_spectrum_yield():
    src = source(activity=1000.0 * Bq)
    src.energy.type = "spectrum_discrete"
    src.energy.spectrum_energies = [0.11295, 0.20837]
    src.energy.spectrum_weights = [0.062, 0.1038]
    assert src.total_yield() == pytest.approx(0.1658)
    g = np.random.Generator(np.random.PCG64(1))
    c = np.array([src.events_per_run([(0.0, 1.0 * s_)], g)[0] for _ in range(400)])
    assert abs(c.mean() - 165.8) < 4 * np.sqrt(165.8 / 400)
    h = source(activity=10.0 * Bq)
    h.energy.type = "spectrum_histogram"
    h.energy.spectrum_energy_bin_edges = [0.01, 0.02, 0.05]
    h.energy.spectrum_weights = [2.0, 1.0]
    assert h.total_yield() == pytest.approx(2.0 * 0.01 + 1.0 * 0.03)


def test_simulation_plans_sources_run_by_run():
    sim = Simulation()
    sim.run_timing_intervals = [[0, 1 * s_], [1 * s_, 2 * s_]]
    a = sim.add_source("GenericSource", "a")
    a.number_of_primaries = [100, 0]
    b = sim.add_source("GenericSource", "b")
    b.activity = 50 * Bq
    from opengate_amd.engine import SimulationEngine

    eng = SimulationEngine.__new__(SimulationEngine)
    eng.sim = sim
    plan = dict((s.name, c) for s, c in SimulationEngine.events_per_run(eng, 12345))
    assert plan["a"] == [100, 0]
    assert len(plan["b"]) == 2 and 20 < sum(plan["b"]) < 180
    # the same seed draws the same counts
    assert dict((s.name, c) for s, c in SimulationEngine.events_per_run(eng, 12345)) == plan
