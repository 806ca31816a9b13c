"""Source shapes and spreads of the generic source, pinned on the oracle's emitted primaries
(CPU): the Gaussian energy (G4SPSEneDistribution "Gauss": mono + sigma N(0,1), a negative
draw gives 0; GateGenericSource.cpp:514-520), the disc's Gaussian spread (the "Beam" circle
with SetBeamSigmaInX/Y; GateGenericSource.cpp:343-366), the surface sphere
(GateGenericSource.cpp:353-356), and the Gaussian history weight
(GateGenericSource.cpp:247-259).

Each law is checked against its closed form with a KS or chi-square test on fixed seeds,
chosen so that a swapped sigma, a spread in the wrong axis, a missing clamp, a sphere
sampled in volume instead of on the surface or a weight spread around 1 instead of the
source weight fails. The facade mapping from GenericSource attributes and its errors are
checked without a GPU."""
import numpy as np
import pytest
from scipy import integrate, stats

from oracle import oracle as O
from opengate_amd.transport import SourceSpec


def _chi2_p(counts, probs):
    exp = probs / probs.sum() * counts.sum()
    keep = exp > 5
    chi = np.sum((counts[keep] - exp[keep]) ** 2 / exp[keep])
    return stats.chi2.sf(chi, keep.sum() - 1)


def test_box_muller_normals_are_standard_normal():
    """The Gaussian weight is weight + sigma n1 with n1 the first Box-Muller value: with
    weight 0 and sigma 1 the weights are the normals themselves. KS over 400 000 draws and
    the 3-sigma and 4-sigma tail masses."""
    src = SourceSpec(weight=0.0, weight_sigma=1.0)
    n = O.sample_source(src, 400_000, seed=3)[7].astype(np.float64)
    assert stats.kstest(n, "norm").pvalue > 1e-4
    for k in (3.0, 4.0):
        p = 2 * stats.norm.sf(k)
        got = np.mean(np.abs(n) > k)
        assert abs(got - p) < 5 * np.sqrt(p / len(n)), (k, got, p)


def test_gauss_energy_far_from_zero():
    src = SourceSpec(ene_type="gauss", energy=0.1405, energy_sigma=0.01)
    E = O.sample_source(src, 200_000, seed=5)[6].astype(np.float64)
    assert stats.kstest(E, "norm", args=(0.1405, 0.01)).pvalue > 1e-4
    # other fields untouched: a point source, isotropic, weight 1
    pos = O.sample_source(src, 1000, seed=5)
    assert np.all(pos[0] == 0) and np.all(pos[7] == 1.0)


def test_gauss_energy_negative_draws_give_zero():
    """mean 20 keV, sigma 20 keV: Phi(-1) = 15.9 % of the draws are negative and become 0;
    the positive part follows the normal law restricted to E > 0; nothing exceeds the
    8-sigma cut."""
    m, s = 0.02, 0.02
    src = SourceSpec(ene_type="gauss", energy=m, energy_sigma=s)
    E = O.sample_source(src, 200_000, seed=9)[6].astype(np.float64)
    assert E.min() == 0.0
    p0 = stats.norm.cdf(0, m, s)
    f0 = np.mean(E == 0)
    assert abs(f0 - p0) < 5 * np.sqrt(p0 * (1 - p0) / len(E))
    pos = E[E > 0]
    trunc = stats.truncnorm(-m / s, np.inf, loc=m, scale=s)
    assert stats.kstest(pos, trunc.cdf).pvalue > 1e-4
    assert E.max() <= m + 8 * s


def test_gauss_energy_sigma_zero_is_mono():
    src = SourceSpec(ene_type="gauss", energy=0.1405, energy_sigma=0.0)
    E = O.sample_source(src, 1000, seed=1)[6]
    assert np.all(E == np.float32(0.1405))
    assert src.max_energy() == pytest.approx(0.1405)
    assert SourceSpec(ene_type="gauss", energy=0.1, energy_sigma=0.01).max_energy() == pytest.approx(0.18)


def _disc_spread_cdf(R, s):
    """CDF of x = R a + s n for a point (a, b) uniform in the unit disc and n ~ N(0, 1):
    the disc's marginal density 2 sqrt(R^2 - t^2) / (pi R^2) convolved with N(0, s^2),
    integrated numerically on a grid."""
    lim = R + 8 * s
    xs = np.linspace(-lim, lim, 4001)
    t = np.linspace(-R, R, 2001)
    f = 2 * np.sqrt(np.clip(R * R - t * t, 0, None)) / (np.pi * R * R)
    pdf = np.array([np.trapezoid(f * stats.norm.pdf(x - t, scale=s), t) for x in xs]) if s > 0 else None
    cdf = integrate.cumulative_trapezoid(pdf, xs, initial=0.0)
    cdf /= cdf[-1]
    return lambda x: np.interp(x, xs, cdf)


def test_disc_gaussian_spread():
    """Radius 10 mm, sigma_x 3 mm, sigma_y 0.5 mm: x and y follow the disc marginal
    convolved with their own sigma (a swap of the two fails both), the spread stays in the
    source plane (z = 0), and x and y are uncorrelated."""
    R, sx, sy = 10.0, 3.0, 0.5
    src = SourceSpec(pos_type="disc", pos_param=(R, 0, 0, 0), pos_sigma=(sx, sy))
    x, y, z = (a.astype(np.float64) for a in O.sample_source(src, 100_000, seed=11)[:3])
    assert np.all(z == 0)
    assert stats.kstest(x, _disc_spread_cdf(R, sx)).pvalue > 1e-4
    assert stats.kstest(y, _disc_spread_cdf(R, sy)).pvalue > 1e-4
    # second moments: R^2 / 4 + sigma^2 per axis
    for v, s in ((x, sx), (y, sy)):
        exp = R * R / 4 + s * s
        assert abs(np.mean(v * v) / exp - 1) < 0.02
    assert abs(np.corrcoef(x, y)[0, 1]) < 0.015
    assert np.max(np.hypot(x, y)) > R  # the spread leaves the circle


def test_disc_without_spread_stays_in_the_circle_and_keeps_its_stream():
    """sigma 0 draws no normals: the primaries are those of a plain disc, and every point
    lies in the circle."""
    R = 7.0
    a = O.sample_source(SourceSpec(pos_type="disc", pos_param=(R, 0, 0, 0)), 20_000, seed=2)
    b = O.sample_source(SourceSpec(pos_type="disc", pos_param=(R, 0, 0, 0), pos_sigma=(0.0, 0.0)), 20_000, seed=2)
    for u, v in zip(a, b):
        np.testing.assert_array_equal(u, v)
    assert np.max(np.hypot(a[0], a[1])) <= R * (1 + 1e-6)


def test_disc_spread_is_rotated_and_translated_with_the_source():
    R, sx = 5.0, 2.0
    rot = np.array([[0, 0, 1], [0, 1, 0], [-1, 0, 0]], dtype=np.float64)  # source x -> phantom -z
    src = SourceSpec(pos_type="disc", pos_param=(R, 0, 0, 0), pos_sigma=(sx, 0.0), pos_rot=rot,
                     pos_trans=(1.0, 2.0, 3.0))
    x, y, z = (a.astype(np.float64) for a in O.sample_source(src, 50_000, seed=4)[:3])
    assert np.allclose(x, 1.0, atol=1e-5)
    assert np.mean((z - 3.0) ** 2) == pytest.approx(R * R / 4 + sx * sx, rel=0.03)
    assert np.mean((y - 2.0) ** 2) == pytest.approx(R * R / 4, rel=0.03)


def test_surface_sphere():
    """Radius 25 mm about (5, -3, 40): every point on the sphere, cos(theta) uniform on
    [-1, 1] and phi uniform (a volume sphere fails the radius, a hemisphere the cosine)."""
    R, c = 25.0, np.array([5.0, -3.0, 40.0])
    src = SourceSpec(pos_type="surface_sphere", pos_param=(R, 0, 0, 0), pos_trans=tuple(c))
    p = np.stack([a.astype(np.float64) for a in O.sample_source(src, 100_000, seed=13)[:3]], axis=1) - c
    r = np.linalg.norm(p, axis=1)
    assert np.allclose(r, R, rtol=2e-6)
    ct = p[:, 2] / r
    assert stats.kstest(ct, "uniform", args=(-1, 2)).pvalue > 1e-4
    phi = np.arctan2(p[:, 1], p[:, 0])
    counts, _ = np.histogram(phi, bins=24, range=(-np.pi, np.pi))
    assert _chi2_p(counts.astype(float), np.ones(24)) > 1e-4


def test_gaussian_weight_around_the_source_weight():
    """weight 2.5, sigma 0.4: N(2.5, 0.4) (spread around 1 fails); the energy draw before it
    is untouched (mono), and a gauss energy and the weight use independent normals."""
    src = SourceSpec(weight=2.5, weight_sigma=0.4)
    out = O.sample_source(src, 100_000, seed=21)
    assert np.all(out[6] == np.float32(0.1405))
    assert stats.kstest(out[7].astype(np.float64), "norm", args=(2.5, 0.4)).pvalue > 1e-4
    src = SourceSpec(ene_type="gauss", energy=0.1, energy_sigma=0.01, weight=1.0, weight_sigma=0.1)
    out = O.sample_source(src, 100_000, seed=22)
    assert abs(np.corrcoef(out[6], out[7])[0, 1]) < 0.015
    assert stats.kstest(out[6].astype(np.float64), "norm", args=(0.1, 0.01)).pvalue > 1e-4
    assert stats.kstest(out[7].astype(np.float64), "norm", args=(1.0, 0.1)).pvalue > 1e-4
    # a negative sigma means a fixed weight
    out = O.sample_source(SourceSpec(weight=2.5, weight_sigma=-1.0), 1000, seed=21)
    assert np.all(out[7] == 2.5)


def test_cosine_law_direction():
    """direction "cos" (G4SPSAngDistribution::GenerateCosineLawFlux): with theta in [0, 90 deg]
    mu = cos(theta) has density 2 mu, so mu^2 is uniform on [0, 1] (an iso sampler fails
    it); with theta in [20, 70 deg], sin^2(theta) is uniform between sin^2 20 and sin^2 70."""
    deg = np.pi / 180
    src = SourceSpec(dir_type="cos", theta=(0.0, 90 * deg))
    w = O.sample_source(src, 100_000, seed=31)[5].astype(np.float64)
    mu = -w
    assert mu.min() >= 0
    assert stats.kstest(mu * mu, "uniform").pvalue > 1e-4
    src = SourceSpec(dir_type="cos", theta=(20 * deg, 70 * deg), phi=(0.0, 90 * deg))
    u, v, w = (a.astype(np.float64) for a in O.sample_source(src, 100_000, seed=32)[3:6])
    s2 = 1 - w * w
    a, b = np.sin(20 * deg) ** 2, np.sin(70 * deg) ** 2
    assert stats.kstest(s2, "uniform", args=(a, b - a)).pvalue > 1e-4
    phi = np.arctan2(-v, -u)
    assert phi.min() >= -1e-6 and phi.max() <= 90 * deg + 1e-6
    assert stats.kstest(phi, "uniform", args=(0, 90 * deg)).pvalue > 1e-4


def _probe_hits(p, d, centre, r):
    """Number of rays p + t d (t > 0) that meet the sphere (centre, r)."""
    q = p - centre
    b = np.sum(q * d, axis=1)
    c = np.sum(q * q, axis=1) - r * r
    disc = b * b - c
    return int(np.sum((disc > 0) & ((-b + np.sqrt(np.maximum(disc, 0))) > 0)))


def test_surface_sphere_cosine_law_against_the_reference_probe_test():
    """The set-up of test010_generic_source_spherical_surface_isotropic.py: surface sphere of
    radius 199 mm, direction "cos" with theta in [0, 90 deg]: mu = -d . n follows the cosine
    law about the inward normal (mean 2/3, variance 1/18, mu^2 uniform), and each of the
    three 50 mm probe spheres is met by about 6300 of 1e5 rays, the number the reference
    test expects (rtol 0.1). An iso sampler meets the probes about 2.4 times less often."""
    deg = np.pi / 180
    src = SourceSpec(pos_type="surface_sphere", pos_param=(199.0, 0, 0, 0), dir_type="cos",
                     theta=(0.0, 90 * deg))
    out = O.sample_source(src, 100_000, seed=41)
    p = np.stack([a.astype(np.float64) for a in out[:3]], axis=1)
    d = np.stack([a.astype(np.float64) for a in out[3:6]], axis=1)
    n = p / np.linalg.norm(p, axis=1)[:, None]
    mu = -np.sum(d * n, axis=1)
    assert abs(mu.mean() - 2 / 3) < 0.005 and abs(mu.var() - 1 / 18) < 0.002
    assert stats.kstest(mu * mu, "uniform").pvalue > 1e-4
    for c in ([50.0, 70.0, -90.0], [-80.0, -60.0, 50.0], [0.0, 0.0, 0.0]):
        assert np.isclose(_probe_hits(p, d, np.array(c), 50.0), 6300, rtol=0.1), c
    iso = SourceSpec(pos_type="surface_sphere", pos_param=(199.0, 0, 0, 0), dir_type="iso")
    out = O.sample_source(iso, 100_000, seed=41)
    p = np.stack([a.astype(np.float64) for a in out[:3]], axis=1)
    d = np.stack([a.astype(np.float64) for a in out[3:6]], axis=1)
    assert _probe_hits(p, d, np.zeros(3), 50.0) < 3500


def test_surface_sphere_tangent_frame_azimuth():
    """With phi in [0, 60 deg] the tangential part of the direction lies between x' =
    unit(A x n) and the direction 60 deg from it towards y' = unit(x' x n) (momentum
    -(st cp, st sp, ct) in that frame), for the axis A = z and for a tilted A."""
    deg = np.pi / 180
    for axis in ((0.0, 0.0, 1.0), (0.3, -0.5, 0.2)):
        src = SourceSpec(pos_type="surface_sphere", pos_param=(10.0, 0, 0, 0), dir_type="iso",
                         theta=(0.0, 180 * deg), phi=(0.0, 60 * deg), surface_axis=axis)
        out = O.sample_source(src, 20_000, seed=43)
        p = np.stack([a.astype(np.float64) for a in out[:3]], axis=1)
        d = np.stack([a.astype(np.float64) for a in out[3:6]], axis=1)
        n = p / np.linalg.norm(p, axis=1)[:, None]
        A = np.asarray(axis) / np.linalg.norm(axis)
        xp = np.cross(A, n)
        xp /= np.linalg.norm(xp, axis=1)[:, None]
        yp = np.cross(xp, n)
        phi = np.arctan2(-np.sum(d * yp, axis=1), -np.sum(d * xp, axis=1))
        assert phi.min() > -1e-4 and phi.max() < 60 * deg + 1e-4
        assert stats.kstest(phi, "uniform", args=(0, 60 * deg)).pvalue > 1e-4


def test_beam2d_direction():
    """direction "beam2d" (GenerateBeamFlux): angles a ~ N(0, sx), b ~ N(0, sy), polar angle
    |(a, b)| about -z, azimuth the angle of (a, b): a = -u t / sin t, b = -v t / sin t with
    t = acos(-w) recover both normals (a swap of sx and sy fails)."""
    sx, sy = 0.02, 0.15
    src = SourceSpec(dir_type="beam2d", dir_sigma=(sx, sy))
    u, v, w = (a.astype(np.float64) for a in O.sample_source(src, 100_000, seed=51)[3:6])
    t = np.arccos(np.clip(-w, -1, 1))
    st = np.sqrt(u * u + v * v)
    f = np.where(st > 0, t / np.maximum(st, 1e-300), 1.0)
    a, b = -u * f, -v * f
    assert stats.kstest(a, "norm", args=(0, sx)).pvalue > 1e-4
    assert stats.kstest(b, "norm", args=(0, sy)).pvalue > 1e-4
    z = O.sample_source(SourceSpec(dir_type="beam2d", dir_sigma=(0.0, 0.0)), 100, seed=1)
    assert np.all(z[5] == -1.0) and np.all(z[3] == 0.0)


def test_histogram_direction_reference_tables():
    """The tables of test010_generic_source_angular_distribution.py: theta uniform on
    [80, 100] deg (one bin), phi piecewise uniform on the bins 60-70-80-100-110-120 deg with
    weights 0.3, 0.5, 1, 0.5, 0.3 (chi-square over the bins, KS inside the widest one)."""
    from opengate_amd.transport import angle_histogram

    deg = np.pi / 180
    th = angle_histogram(np.array([80, 100]) * deg, [1], 0, np.pi)
    edges = np.array([60, 70, 80, 100, 110, 120]) * deg
    wts = np.array([0.3, 0.5, 1, 0.5, 0.3])
    ph = angle_histogram(edges, wts, 0, 2 * np.pi)
    src = SourceSpec(dir_type="histogram", theta_table=th, phi_table=ph)
    u, v, w = (a.astype(np.float64) for a in O.sample_source(src, 200_000, seed=61)[3:6])
    t = np.arccos(np.clip(-w, -1, 1))
    assert stats.kstest(t, "uniform", args=(80 * deg, 20 * deg)).pvalue > 1e-4
    phi = np.arctan2(-v, -u)
    counts, _ = np.histogram(phi, bins=edges)
    assert _chi2_p(counts.astype(float), wts) > 1e-4
    inner = phi[(phi > 80 * deg) & (phi < 100 * deg)]
    assert stats.kstest(inner, "uniform", args=(80 * deg, 20 * deg)).pvalue > 1e-4


def test_histogram_direction_outside_the_angle_range_is_drawn_again():
    """Geant4 draws phi again while it lies outside [0, 2 pi]: bins [-30, 0, 30] deg with
    equal weights leave phi uniform on [0, 30] deg; a histogram with no weight inside the
    range is an error (Geant4 would loop for ever)."""
    from opengate_amd.transport import angle_histogram

    deg = np.pi / 180
    ph = angle_histogram(np.array([-30, 0, 30]) * deg, [1, 1], 0, 2 * np.pi)
    np.testing.assert_allclose(ph[0], [0.0], atol=1e-7)
    np.testing.assert_allclose(ph[1], [30 * deg], rtol=1e-6)
    th = angle_histogram(np.array([170, 200]) * deg, [1], 0, np.pi)
    np.testing.assert_allclose(th[0] + th[1], [np.pi], rtol=1e-6)
    src = SourceSpec(dir_type="histogram", theta_table=th, phi_table=ph)
    u, v, w = (a.astype(np.float64) for a in O.sample_source(src, 50_000, seed=62)[3:6])
    phi = np.arctan2(-v, -u)
    assert stats.kstest(phi, "uniform", args=(0, 30 * deg)).pvalue > 1e-4
    t = np.arccos(np.clip(-w, -1, 1))
    assert stats.kstest(t, "uniform", args=(170 * deg, 10 * deg)).pvalue > 1e-4
    with pytest.raises(ValueError, match="no weight"):
        angle_histogram(np.array([-30, -10]) * deg, [1], 0, 2 * np.pi)
    with pytest.raises(ValueError, match="increase"):
        angle_histogram(np.array([10, 5]) * deg, [1], 0, 2 * np.pi)


def test_user_energy_histogram():
    """energy type "histogram" (G4SPSEneDistribution "User", linear inverse of the running
    weight sum): points (20, 0.2), (30, 0.3), (50, 0.5) keV give a point mass 0.2 at 20 keV,
    0.3 uniform on (20, 30] and 0.5 uniform on (30, 50] keV."""
    from opengate_amd.transport import energy_user_histogram

    val, wid, cdf = energy_user_histogram([0.02, 0.03, 0.05], [0.2, 0.3, 0.5])
    src = SourceSpec(ene_type="spectrum_histogram", energy_value=val, energy_width=wid, energy_cdf=cdf)
    E = O.sample_source(src, 200_000, seed=71)[6].astype(np.float64)
    n = len(E)
    at = np.mean(E == np.float32(0.02))
    assert abs(at - 0.2) < 5 * np.sqrt(0.2 * 0.8 / n)
    mid = E[(E > np.float32(0.02)) & (E <= np.float32(0.03))]
    top = E[E > np.float32(0.03)]
    assert abs(len(mid) / n - 0.3) < 5 * np.sqrt(0.3 * 0.7 / n)
    assert abs(len(top) / n - 0.5) < 5 * np.sqrt(0.25 / n)
    assert stats.kstest(mid, "uniform", args=(0.02, 0.01)).pvalue > 1e-4
    assert stats.kstest(top, "uniform", args=(0.03, 0.02)).pvalue > 1e-4


# ---------------------------------------------------------------- facade mapping


def _engine_spec(configure):
    import opengate_amd as gate
    from opengate_amd.engine import SimulationEngine

    sim = gate.Simulation()
    src = sim.add_source("GenericSource", "s")
    src.particle = "gamma"
    src.energy.mono = 0.1405
    configure(src)
    eng = SimulationEngine(sim)
    # phantom frame = world frame
    eng.R, eng.t, eng.L = np.eye(3), np.zeros(3), np.zeros(3)
    return eng.source_spec(src, 10)


def test_facade_maps_gauss_spread_and_weight():
    MeV, mm = 1.0, 1.0

    def conf(s):
        s.position.type = "disc"
        s.position.radius = 4 * mm
        s.position.sigma_x = 1.5 * mm
        s.position.sigma_y = 0.5 * mm
        s.energy.type = "gauss"
        s.energy.mono = 0.2 * MeV
        s.energy.sigma_gauss = 0.01 * MeV
        s.weight = 3.0
        s.weight_sigma = 0.2

    spec = _engine_spec(conf)
    assert spec.pos_type == "disc" and spec.pos_sigma == (1.5, 0.5)
    assert spec.ene_type == "gauss" and spec.energy == pytest.approx(0.2) and spec.energy_sigma == pytest.approx(0.01)
    assert spec.weight == pytest.approx(3.0) and spec.weight_sigma == pytest.approx(0.2)

    def sphere(s):
        s.position.type = "surface_sphere"
        s.position.radius = 12.0

    spec = _engine_spec(sphere)
    assert spec.pos_type == "surface_sphere" and spec.pos_param[0] == 12.0

    # no Gaussian weight without a positive source weight (GateGenericSource.cpp:248)
    def w0(s):
        s.weight_sigma = 0.2

    assert _engine_spec(w0).weight_sigma < 0


def test_facade_rejects_bad_spreads():
    def neg_sigma(s):
        s.position.type = "disc"
        s.position.sigma_x = -1.0

    with pytest.raises(ValueError, match="sigma_x"):
        _engine_spec(neg_sigma)

    def box_sigma(s):
        s.position.type = "box"
        s.position.size = [1, 1, 1]
        s.position.sigma_y = 1.0

    with pytest.raises(NotImplementedError, match="disc"):
        _engine_spec(box_sigma)

    def neg_gauss(s):
        s.energy.type = "gauss"
        s.energy.sigma_gauss = -0.1

    with pytest.raises(ValueError, match="sigma_gauss"):
        _engine_spec(neg_gauss)


def test_facade_maps_direction_types_and_user_histogram():
    deg = np.pi / 180

    def cos(s):
        s.position.type = "surface_sphere"
        s.position.radius = 199.0
        s.position.rotation = np.array([[1, 0, 0], [0, 0, -1], [0, 1, 0]], dtype=float)
        s.direction.type = "cos"
        s.direction.theta = [0, 90 * deg]

    spec = _engine_spec(cos)
    assert spec.dir_type == "cos" and spec.theta == (0.0, 90 * deg)
    np.testing.assert_allclose(spec.surface_axis, [0, -1, 0], atol=1e-12)  # rotation's third column

    def beam(s):
        s.direction.type = "beam2d"
        s.direction.sigma = [0.01, 0.02]

    spec = _engine_spec(beam)
    assert spec.dir_type == "beam2d" and spec.dir_sigma == (0.01, 0.02)

    def hist(s):
        s.direction.type = "histogram"
        s.direction.histogram_theta_angles = [80 * deg, 100 * deg]
        s.direction.histogram_theta_weights = [1]
        s.direction.histogram_phi_angles = [60 * deg, 70 * deg, 80 * deg]
        s.direction.histogram_phi_weights = [0.3, 0.5]
        s.energy.type = "histogram"
        s.energy.histogram_energy = [0.02, 0.03, 0.05]
        s.energy.histogram_weight = [0.2, 0.3, 0.5]
        s.activity = 10.0

    spec = _engine_spec(hist)
    assert spec.dir_type == "histogram" and spec.ene_type == "spectrum_histogram"
    np.testing.assert_allclose(spec.theta_table[0], [80 * deg], rtol=1e-6)
    np.testing.assert_allclose(spec.phi_table[2], [0.375, 1.0], rtol=1e-6)
    np.testing.assert_allclose(spec.energy_width, [0.0, 0.01, 0.02], rtol=1e-6)
    assert spec.max_energy() == pytest.approx(0.05)

    import opengate_amd as gate

    sim = gate.Simulation()
    src = sim.add_source("GenericSource", "h")
    src.energy.type = "histogram"
    src.energy.histogram_energy = [0.02, 0.03]
    src.energy.histogram_weight = [1.0, 2.5]
    assert src.total_yield() == pytest.approx(3.5)

    # weight 0 is not a weight (GateGenericSource.cpp:248): the histories keep weight 1
    def w0(s):
        s.weight = 0.0

    assert _engine_spec(w0).weight == 1.0

    def bad_beam(s):
        s.direction.type = "beam2d"
        s.direction.sigma = [-0.1, 0.0]

    with pytest.raises(ValueError, match="sigma"):
        _engine_spec(bad_beam)

