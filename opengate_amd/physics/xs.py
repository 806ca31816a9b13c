"""Photon cross sections (FP64) and the FP32 tables the transport kernels read.

Physics model of this build (the reference takes these numbers from Geant4, which
is not available here; the transport reads tables, so Geant4-exported tables can
replace these through the same C-ABI call):

* total attenuation mu/rho: NIST element tables mixed by mass fraction on the
  union energy grid and interpolated log-log, a restatement of
  GateMaterialMuHandler::ConstructMaterial (GateMaterialMuHandler.cpp:159-254)
  and GateMuTable::GetMuOverRho (GateMuTables.cpp:72-100);
* incoherent (Compton): Klein-Nishina (Klein & Nishina 1929) weighted by the
  incoherent scattering function S(q) = Z - F(q)^2/Z of the same Thomas-Fermi
  atom (equal to the exact one-electron result for Z = 1); the kernels sample
  Klein-Nishina and accept with S(q)/Z;
* coherent (Rayleigh): Thomson cross section weighted by the squared atomic form
  factor of a Thomas-Fermi atom in Moliere's three-Yukawa approximation
  (Moliere 1947): F(q)/Z = sum_i a_i k_i^2/(k_i^2+q^2), a=(0.1,0.55,0.35),
  k_i = b_i/(0.88534 a0 Z^-1/3), b=(6,1.2,0.3);
* photoelectric: the remainder mu - incoherent - coherent per material; where
  the remainder would be negative the two scattering terms are scaled down so
  that the total always equals the NIST mixture value.

Energy grid of the GPU tables ("float-bit buckets"): node j has the FP32 energy
whose bit pattern is (BUCKET_BASE + j) << (23 - KBITS); there are 2**KBITS buckets
per octave starting at E_LO = 2**-10 MeV. A kernel finds the bucket of an
energy E from its bits and interpolates linearly in E inside the bucket.
"""
from dataclasses import dataclass
from functools import lru_cache

import numpy as np

from . import elements as el

# physical constants
R_E_CM = 2.8179403262e-13  # classical electron radius [cm]
N_A = 6.02214076e23
MEC2 = 0.51099895  # electron rest energy [MeV]
HBARC_MEV_A = 1.973269804e-3  # hbar c [MeV * Angstrom]
A0_A = 0.529177210903  # Bohr radius [Angstrom]
K_PER_MEV = 1.0 / HBARC_MEV_A  # photon wave number per MeV [1/Angstrom]

KBITS = 6
SHIFT = 23 - KBITS
E_LO_EXP = -10  # E_LO = 2**-10 MeV
BUCKET_BASE = ((127 + E_LO_EXP) << 23) >> SHIFT
# q^2 grid of the form-factor tables: node 0 is x = 0, node j >= 1 has the FP32
# value whose bits are (X_BUCKET_BASE + j - 1) << X_SHIFT (8 nodes per octave from
# 2**-12 A^-2).
X_KBITS = 3
X_SHIFT = 23 - X_KBITS
X_LO_EXP = -12
X_BUCKET_BASE = ((127 + X_LO_EXP) << 23) >> X_SHIFT

_MOLIERE_A = np.array([0.10, 0.55, 0.35])
_MOLIERE_B = np.array([6.0, 1.2, 0.3])


# ----------------------------------------------------------------------------
# total attenuation: NIST mixture (restatement of GateMaterialMuHandler)
# ----------------------------------------------------------------------------
def _loglog(xa, xb, ya, yb, x):
    # GateMaterialMuHandler.cpp:104-106
    return np.exp(np.log(ya) + np.log(yb / ya) / np.log(xb / xa) * np.log(x / xa))


def mixture_union_table(material):
    return _mixture_union_table_cached(tuple(material.elements), tuple(float(f) for f in material.fractions))


@lru_cache(maxsize=256)
def _mixture_union_table_cached(elements, fractions):
    from types import SimpleNamespace

    return _mixture_union_table(SimpleNamespace(elements=list(elements), fractions=np.array(fractions)))


def _mixture_union_table(material):
    """Union-grid mu/rho and mu_en/rho table of a material.

    Follows GateMaterialMuHandler::ConstructMaterial (GateMaterialMuHandler.cpp:159-254):
    energies = merged element grids (a duplicated edge energy is kept twice only when it
    is an edge of the element that contributes it), values = mass-fraction sum of the
    element values, interpolated log-log where an element has no node at that energy.
    """
    tabs = [el.nist_table(z) for z in material.elements]
    frac = material.fractions
    sizes = [len(t[0]) for t in tabs]
    n = len(tabs)
    nb_e = sum(sizes)
    energies = []
    index = [0] * n
    i = 0
    while i < nb_e:
        mt = 0
        while index[mt] >= sizes[mt]:
            mt += 1
        for j in range(mt + 1, n):
            if index[j] < sizes[j] and tabs[j][0][index[j]] < tabs[mt][0][index[mt]]:
                mt = j
        e = tabs[mt][0][index[mt]]
        keep = True
        if energies and e == energies[-1]:
            if not (index[mt] > 0 and tabs[mt][0][index[mt]] == tabs[mt][0][index[mt] - 1]):
                keep = False
        if keep:
            energies.append(e)
            i += 1
        else:
            nb_e -= 1
        index[mt] += 1
    energies = np.array(energies)
    mu = np.zeros(len(energies))
    muen = np.zeros(len(energies))
    index = [0] * n
    for i, ce in enumerate(energies):
        for j in range(n):
            e_t, mu_t, muen_t = tabs[j]
            if index[j] < sizes[j] - 1 and e_t[index[j]] < ce:
                index[j] += 1
            k = index[j]
            if e_t[k] == ce:
                mu[i] += frac[j] * mu_t[k]
                muen[i] += frac[j] * muen_t[k]
                if i != len(energies) - 1 and k + 1 < sizes[j] and e_t[k] == e_t[k + 1]:
                    index[j] += 1
            else:
                mu[i] += frac[j] * _loglog(e_t[k - 1], e_t[k], mu_t[k - 1], mu_t[k], ce)
                muen[i] += frac[j] * _loglog(e_t[k - 1], e_t[k], muen_t[k - 1], muen_t[k], ce)
    return energies, mu, muen


def lookup_loglog(table_e, table_v, energy):
    """GateMuTable::GetMuOverRho (GateMuTables.cpp:72-100): find the bracketing nodes in
    log E, interpolate log(value) linearly; outside the table the end-node value.
    Energies that coincide exactly with a node take that node's value."""
    le = np.log(table_e)
    lv = np.log(table_v)
    x = np.log(np.atleast_1d(np.asarray(energy, dtype=np.float64)))
    inf = np.clip(np.searchsorted(le, x, side="right") - 1, 0, len(le) - 2)
    sup = inf + 1
    inside = (le[inf] < x) & (x < le[sup])
    with np.errstate(divide="ignore", invalid="ignore"):
        interp = lv[inf] + (lv[sup] - lv[inf]) * (x - le[inf]) / (le[sup] - le[inf])
    below = x <= le[0]
    above = x >= le[-1]
    out = np.where(inside, interp, lv[inf])
    out = np.where(below, lv[0], out)
    out = np.where(above, lv[-1], out)
    return np.exp(out)


def mu_rho(material, energy):
    e, mu, _ = mixture_union_table(material)
    return lookup_loglog(e, mu, energy)


def muen_rho(material, energy):
    e, _, muen = mixture_union_table(material)
    return lookup_loglog(e, muen, energy)


# ----------------------------------------------------------------------------
# Klein-Nishina
# ----------------------------------------------------------------------------
def sigma_kn_electron(energy):
    """Klein-Nishina total cross section per free electron [cm2]."""
    k = np.asarray(energy, dtype=np.float64) / MEC2
    k = np.maximum(k, 1e-12)
    l = np.log1p(2 * k)
    small = k < 1e-3
    big = (1 + k) / k**2 * (2 * (1 + k) / (1 + 2 * k) - l / k) + l / (2 * k) - (1 + 3 * k) / (1 + 2 * k) ** 2
    # series near k -> 0: (4/3)(1 - 2k + 26k^2/5 - 133k^3/10)
    ser = 4.0 / 3.0 * (1 - 2 * k + 5.2 * k**2 - 13.3 * k**3)
    val = np.where(small, ser, big)
    return 2 * np.pi * R_E_CM**2 * val


def kn_dsigma_dmu(energy, mu):
    """Klein-Nishina dsigma/dcos(theta) per electron [cm2]."""
    k = energy / MEC2
    r = 1.0 / (1.0 + k * (1.0 - mu))
    return np.pi * R_E_CM**2 * r**2 * (r + 1.0 / r - (1.0 - mu**2))


# ----------------------------------------------------------------------------
# Rayleigh with the Thomas-Fermi (Moliere) form factor
# ----------------------------------------------------------------------------
def _moliere_c(z):
    b = 0.88534 * A0_A * z ** (-1.0 / 3.0)
    kappa = _MOLIERE_B / b
    return kappa**2  # [1/A^2]


def form_factor(z, x):
    """F(q, Z) with x = q^2 [1/A^2]."""
    c = _moliere_c(z)
    x = np.asarray(x, dtype=np.float64)[..., None]
    return z * np.sum(_MOLIERE_A * c / (c + x), axis=-1)


def form_factor_sq_integral(z, x):
    """G(X) = int_0^X F(x')^2 dx' (closed form for the three-Yukawa F)."""
    c = _moliere_c(z)
    a = _MOLIERE_A
    x = np.asarray(x, dtype=np.float64)
    out = np.zeros_like(x)
    for i in range(3):
        for j in range(3):
            ci, cj = c[i], c[j]
            if i == j:
                term = x / (ci * (ci + x))
            else:
                term = np.log((cj * (ci + x)) / (ci * (cj + x))) / (cj - ci)
            out = out + a[i] * a[j] * ci * cj * term
    return z * z * out


# quadrature over y = 1 - cos(theta) in [0, 2], panels geometric towards y = 0
_GL_X, _GL_W = np.polynomial.legendre.leggauss(12)
_Y_EDGES = np.concatenate([[0.0], np.geomspace(1e-10, 2.0, 160)])
_YA, _YB = _Y_EDGES[:-1], _Y_EDGES[1:]
_Y = (0.5 * (_YA + _YB)[:, None] + 0.5 * (_YB - _YA)[:, None] * _GL_X[None, :]).ravel()
_YW = (0.5 * (_YB - _YA)[:, None] * _GL_W[None, :]).ravel()


def _angular_integral(z, energy, kind):
    energy = np.atleast_1d(np.asarray(energy, dtype=np.float64))
    return _angular_integral_cached(int(z), energy.tobytes(), kind).copy()


@lru_cache(maxsize=1024)
def _angular_integral_cached(z, energy_bytes, kind):
    energy = np.frombuffer(energy_bytes, dtype=np.float64)
    out = np.empty_like(energy)
    mu = 1.0 - _Y
    for q0 in range(0, len(energy), 64):
        e = energy[q0 : q0 + 64, None]
        kk = e * K_PER_MEV
        x = 2 * kk * kk * _Y[None, :]
        f = form_factor(z, x)
        if kind == "rayleigh":
            integrand = np.pi * R_E_CM**2 * (1 + mu**2)[None, :] * f * f
        else:
            integrand = kn_dsigma_dmu(e, mu[None, :]) * (z - f * f / z)
        out[q0 : q0 + 64] = np.sum(integrand * _YW[None, :], axis=1)
    return out


def sigma_rayleigh_atom(z, energy):
    """sigma_coh per atom [cm2] = pi r_e^2 int_{-1}^{1} (1+mu^2) F(q)^2 dmu."""
    return _angular_integral(z, energy, "rayleigh")


def incoherent_function(z, x):
    """S(q, Z) = Z - F(q, Z)^2 / Z with x = q^2."""
    return z - form_factor(z, x) ** 2 / z


def sigma_incoherent_atom(z, energy):
    """sigma_incoh per atom [cm2] = int dsigma_KN/dmu * S(q(mu), Z) dmu, with q from
    the incident energy."""
    return _angular_integral(z, energy, "incoherent")


# ----------------------------------------------------------------------------
# per-material partial attenuation (cm2/g)
# ----------------------------------------------------------------------------
def partial_mass_attenuation(material, energy):
    """(mu_total, compton, rayleigh, photo, mu_en) mass coefficients [cm2/g] at energies."""
    energy = np.atleast_1d(np.asarray(energy, dtype=np.float64))
    e_u, mu_u, muen_u = mixture_union_table(material)
    mu = lookup_loglog(e_u, mu_u, energy)
    muen = lookup_loglog(e_u, muen_u, energy)
    compt = np.zeros_like(energy)
    rayl = np.zeros_like(energy)
    for z, w in zip(material.elements, material.fractions):
        n_atoms = w * N_A / el.atomic_mass(z)  # atoms per gram
        compt += n_atoms * sigma_incoherent_atom(z, energy)
        rayl += n_atoms * sigma_rayleigh_atom(z, energy)
    scat = compt + rayl
    photo = mu - scat
    scale = np.where(photo < 0, mu / scat, 1.0)
    compt *= scale
    rayl *= scale
    photo = np.maximum(photo, 0.0)
    photo = mu - compt - rayl
    photo = np.maximum(photo, 0.0)
    return mu, compt, rayl, photo, muen


# ----------------------------------------------------------------------------
# FP32 GPU tables
# ----------------------------------------------------------------------------
def bucket_node_energies(n_buckets):
    bits = (np.arange(n_buckets + 1, dtype=np.uint32) + BUCKET_BASE) << SHIFT
    return bits.view(np.float32)


def n_buckets_for(e_max):
    """Smallest whole number of octaves above E_LO covering e_max."""
    oct_ = int(np.ceil(np.log2(e_max) - E_LO_EXP + 1e-12))
    oct_ = max(oct_, 1)
    return oct_ << KBITS


def rayleigh_x_nodes(e_max):
    """q^2 nodes [1/A^2] covering q^2 <= 4 k(e_max)^2."""
    kk = e_max * K_PER_MEV
    xmax = 4 * kk * kk
    n_oct = int(np.ceil(np.log2(xmax) - X_LO_EXP + 1e-12))
    n = (n_oct << X_KBITS) + 1
    bits = (np.arange(n, dtype=np.uint32) + X_BUCKET_BASE) << X_SHIFT
    return np.concatenate([[0.0], bits.view(np.float32).astype(np.float64)])


@dataclass
class PhysicsTables:
    """FP32 arrays in the layout the C-ABI call gh_set_physics expects. Table rows ("R"):
    one per material (composition is None; mu_total, mu_en in 1/mm), or one per elemental
    composition shared by several materials (composition[m] = row of material m; mu_total,
    mu_en per unit density, 1/mm per g/cm3, and material m has mu = density[m] * row)."""

    material_names: list
    density: np.ndarray  # [M] float32 g/cm3
    n_buckets: int
    mu_total: np.ndarray  # [R, NB+1] float32
    mu_majorant: np.ndarray  # [NB+1] float32 1/mm
    p_compton: np.ndarray  # [R, NB+1] float32 cumulative
    p_rayleigh: np.ndarray  # [R, NB+1] float32 cumulative (compton + rayleigh)
    mu_en: np.ndarray  # [R, NB+1] float32
    ray_x: np.ndarray  # [NX] float32 q^2 nodes [1/A^2]
    ray_cdf: np.ndarray  # [R, NX] float32 normalised int F^2 dx
    incoh_s: np.ndarray  # [R, NX] float32 S(q)/Z of the row's composition
    e_cut: float
    e_max: float
    composition: np.ndarray = None  # [M] int32, or None (a row per material)
    majorant_materials: list = None  # materials the majorant covers (None: all)

    @property
    def n_materials(self):
        return len(self.material_names)

    @property
    def n_compositions(self):
        return 0 if self.composition is None else int(self.mu_total.shape[0])

    def material_mu(self, m):
        """[NB+1] mu nodes (1/mm) of material m, as FP32 products when rows are shared."""
        if self.composition is None:
            return self.mu_total[m]
        return np.float32(self.density[m]) * self.mu_total[self.composition[m]]

    def row_of(self, m):
        return m if self.composition is None else int(self.composition[m])


def _composition_key(mat):
    if mat.is_vacuum:
        return ("vacuum",)
    return tuple((int(z), float(w)) for z, w in zip(mat.elements, mat.fractions))


def build_tables(materials, e_max, e_cut=None, majorant_margin=2.0**-16, share_compositions=True,
                 majorant_materials=None):
    """Build the FP32 transport tables for `materials` (list of Material) up to e_max MeV.

    All arithmetic is FP64; the arrays are rounded to FP32 once at the end. When several
    materials share an elemental composition (Schneider HU materials at different
    densities) and share_compositions is set, the tables hold one row per composition for
    unit density; otherwise one row per material. The majorant node values are the
    per-node maximum of the mu (as the transport forms it in FP32) of every material, or of
    the materials with indices `majorant_materials` (those that occur in the phantoms: a
    CT table may hold dense materials, dental amalgam for one, that no voxel uses, and
    a majorant raised by them multiplies the Woodcock steps), raised by
    `majorant_margin` (relative) so that FP32 interpolation never exceeds it.
    """
    nb = n_buckets_for(e_max)
    e_nodes = bucket_node_energies(nb).astype(np.float64)
    keys = [_composition_key(m) for m in materials]
    rows, comp = {}, []
    for k in keys:
        comp.append(rows.setdefault(k, len(rows)))
    shared = share_compositions and len(rows) < len(materials)
    if shared:
        reps = [None] * len(rows)
        for m, c in zip(materials, comp):
            if reps[c] is None:
                reps[c] = m
        row_mats = reps
    else:
        row_mats = list(materials)
    r = len(row_mats)
    mu_t = np.zeros((r, nb + 1))
    pc = np.zeros((r, nb + 1))
    pr = np.zeros((r, nb + 1))
    mue = np.zeros((r, nb + 1))
    for i, mat in enumerate(row_mats):
        mu, compt, rayl, photo, muen = partial_mass_attenuation(mat, e_nodes)
        # shared rows are per unit density (the vacuum row keeps density 0)
        rho = 0.0 if mat.is_vacuum else (1.0 if shared else mat.density)
        mu_t[i] = mu * rho / 10.0  # 1/cm -> 1/mm
        mue[i] = muen * rho / 10.0
        pc[i] = compt / mu
        pr[i] = (compt + rayl) / mu
    density = np.array([0.0 if mm.is_vacuum else mm.density for mm in materials], dtype=np.float32)
    mu32 = mu_t.astype(np.float32)
    if shared:
        per_mat = np.stack([density[m] * mu32[comp[m]] for m in range(len(materials))])  # FP32 products
    else:
        per_mat = mu32
    sel = np.arange(len(materials)) if majorant_materials is None else np.asarray(sorted(set(majorant_materials)),
                                                                                     dtype=np.int64)
    if sel.size == 0:
        raise ValueError("majorant_materials is empty")
    maj = per_mat[sel].astype(np.float64).max(axis=0) * (1.0 + majorant_margin)
    maj = np.maximum(maj, np.finfo(np.float32).tiny)  # vacuum-only phantoms: a positive majorant
    xn = rayleigh_x_nodes(float(e_nodes[-1]))
    cdf = np.zeros((r, len(xn)))
    sfun = np.zeros((r, len(xn)))
    for i, mat in enumerate(row_mats):
        g = np.zeros(len(xn))
        s_num = np.zeros(len(xn))
        s_den = 0.0
        for z, w in zip(mat.elements, mat.fractions):
            n_atoms = w / el.atomic_mass(z)
            g += n_atoms * form_factor_sq_integral(z, xn)
            s_num += n_atoms * incoherent_function(z, xn)
            s_den += n_atoms * z
        cdf[i] = g / g[-1]
        sfun[i] = s_num / s_den
    if e_cut is None:
        e_cut = float(e_nodes[0])
    f32 = np.float32
    return PhysicsTables(
        material_names=[mm.name for mm in materials],
        density=density,
        n_buckets=nb,
        mu_total=mu32,
        mu_majorant=maj.astype(f32),
        p_compton=pc.astype(f32),
        p_rayleigh=pr.astype(f32),
        mu_en=mue.astype(f32),
        ray_x=xn.astype(f32),
        ray_cdf=cdf.astype(f32),
        incoh_s=sfun.astype(f32),
        e_cut=float(e_cut),
        e_max=float(e_nodes[-1]),
        majorant_materials=None if majorant_materials is None else [int(i) for i in sel],
        composition=np.array(comp, dtype=np.int32) if shared else None,
    )
