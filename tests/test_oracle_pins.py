"""CPU tests: pin the oracle's BSDF, microfacet, rough-transmittance and film
functions to the mathematics, and show that each check fails when the oracle
carries a plausible mistake.

Every check below is a closed form, an invariant or a limit that the
reference's formulas satisfy (file:line in each docstring); none of them
re-types the oracle's own formula.  For each check a twin test plants a slip
in the oracle (oracle/hk_oracle.hpp, enum Slip, through ho_set_slip) and
asserts that the same check then fails.
"""
import contextlib
import os

import numpy as np
import pytest

from hikari_amd import desc as D
from oracle import pyoracle as O
from tests.scenes_testlib import random_dirs


@contextlib.contextmanager
def planted(name):
    O.set_slip(name)
    try:
        yield
    finally:
        O.set_slip("none")


def _dir(theta_deg):
    t = np.radians(theta_deg)
    return np.array([np.sin(t), 0.0, np.cos(t)], np.float32)


def _sample(mats, index, wi, n, seed):
    rng = np.random.default_rng(seed)
    w = np.repeat(np.asarray(wi, np.float32)[None], n, 0)
    return O.bsdf_batch(np.stack(mats), index, w, w, np.ones(n, np.int32), rng.uniform(size=(n, 4)).astype(np.float32))


def fresnel_textbook(mu, eta):
    """Unpolarised reflectance of a dielectric interface (relative index eta)
    at incidence cosine mu > 0, from Snell's law and the s / p amplitudes
    (Born & Wolf); 1 beyond the critical angle."""
    mu = np.asarray(mu, np.float64)
    s2 = (1 - mu ** 2) / eta ** 2
    out = np.ones_like(mu)
    ok = s2 < 1
    ct, c = np.sqrt(1 - s2[ok]), mu[ok]
    rs = (c - eta * ct) / (c + eta * ct)
    rp = (eta * c - ct) / (eta * c + ct)
    out[ok] = 0.5 * (rs ** 2 + rp ** 2)
    return out


def fresnel_complex(mu, n, k):
    """Unpolarised reflectance of a conductor with complex index n + ik, from
    the complex Snell's law and the s / p Fresnel amplitudes."""
    N = n + 1j * k
    cos_t = np.sqrt(1 - (1 - mu ** 2) / N ** 2 + 0j)
    rs = (mu - N * cos_t) / (mu + N * cos_t)
    rp = (N * mu - cos_t) / (N * mu + cos_t)
    return 0.5 * (abs(rs) ** 2 + abs(rp) ** 2)


def fresnel_hemispherical(eta, n=200000):
    """2 int_0^1 F(mu) mu dmu: the cosine-weighted hemispherical average."""
    mu = (np.arange(n) + 0.5) / n
    return float(np.mean(fresnel_textbook(mu, eta) * 2 * mu))


# ----------------------------------------------------------------- Smith G1
def g1_identity_error(dist, alpha, theta_v, nw=3000, nphi=256):
    """Relative error of int G1(v, m) <v.m>+ D(m) dm = cos(theta_v), the
    defining identity of the Smith masking function (Walter et al. 2007,
    eq. 7-8; src/core/MicrofacetDistribution.cpp:79-112).  m is drawn from
    D(m) cos(theta_m) through the textbook CDFs of the two distributions,
    tan^2 = a^2 u / (1 - u) (GGX) and -a^2 log(1 - u) (Beckmann), on a
    midpoint grid in w = sqrt(1 - u) (which removes the grazing singularity)
    and phi."""
    w = (np.arange(nw) + 0.5) / nw
    phi = (np.arange(nphi) + 0.5) / nphi * 2 * np.pi
    W, P = np.meshgrid(w, phi, indexing="ij")
    t2 = alpha ** 2 * (1 - W * W) / (W * W) if dist == D.GGX else -alpha ** 2 * np.log(W * W)
    c = 1 / np.sqrt(1 + t2)
    s = np.sqrt(t2) * c
    m = np.stack([s * np.cos(P), s * np.sin(P), c], -1).reshape(-1, 3)
    v = np.array([np.sin(theta_v), 0.0, np.cos(theta_v)])
    _, g1, _, _ = O.microfacet(dist, alpha, alpha, m, np.repeat(v[None], len(m), 0), np.zeros((len(m), 2)))
    f = g1 * np.maximum(m @ v, 0.0) / m[:, 2] * 2 * W.reshape(-1)
    return abs(f.mean() / np.cos(theta_v) - 1)


def g1_worst(dist):
    return max(g1_identity_error(dist, a, np.radians(t)) for a in (0.05, 0.2, 0.5, 1.0) for t in (10, 45, 70, 85, 89))


# GGX's G1 is exact (the quadrature resolves it to ~2e-6); Beckmann's is the
# rational fit of Walter et al. (2007), accurate to about 0.35%.
G1_TOL = {D.GGX: 2e-5, D.BECKMANN: 4e-3}


@pytest.mark.parametrize("dist", [D.GGX, D.BECKMANN])
def test_smith_g1_identity(dist):
    assert g1_worst(dist) < G1_TOL[dist]


@pytest.mark.parametrize("dist,slip", [(D.GGX, "ggx_g1"), (D.BECKMANN, "beckmann_g1")])
def test_smith_g1_identity_catches_slip(dist, slip):
    with planted(slip):
        assert g1_worst(dist) > 10 * G1_TOL[dist]


# ----------------------------------------------------------------- conductor Fresnel
def conductor_fresnel_error():
    """FresnelConductor (src/core/Common.cpp:529-548, the real-arithmetic
    form) against the complex-index Fresnel equations."""
    worst = 0.0
    for eta, k in ((0.47, 3.98), (0.2, 3.9), (1.5, 0.0), (2.5, 1.2)):
        for mu in np.linspace(0.02, 1.0, 25):
            f = O.fresnel_conductor(mu, [eta] * 3, [k] * 3)[0]
            ref = fresnel_complex(mu, eta, k)
            worst = max(worst, abs(f - ref) / ref)
    return worst


def test_conductor_fresnel_is_complex_fresnel():
    assert conductor_fresnel_error() < 1e-12


def test_conductor_fresnel_catches_slip():
    with planted("conductor_fresnel"):
        assert conductor_fresnel_error() > 1e-2


# ----------------------------------------------------------------- rough conductor
ETA_AU, K_AU, KS = 0.47, (3.98, 2.39, 1.60), (0.9, 0.8, 0.7)


def rough_conductor_smooth_error():
    """RoughConductorBSDF::Sample (RoughConductorBSDF.cpp:54-90) at alpha =
    MIN_ALPHA: m -> n, so D G <m.wi> / (pdf cos_i) -> 1 and the weight is
    ks F_conductor(cos_i), F from the complex-index Fresnel equations."""
    worst = 0.0
    for dist in (D.GGX, D.BECKMANN):
        m = D.rough_conductor(alpha=0.0, dist=dist, int_ior=ETA_AU, ext_ior=1.0, k=K_AU, ks=KS)
        for th in (0, 30, 60, 80):
            wi = _dir(th)
            f = _sample([m], 0, wi, 20000, th)["f"].mean(0)
            exp = np.array(KS) * np.array([fresnel_complex(float(wi[2]), ETA_AU, k) for k in K_AU])
            worst = max(worst, float(np.max(np.abs(f / exp - 1))))
    return worst


def rough_conductor_albedo_z():
    """With F = 1 (k = 1e4) the rough conductor's albedo is the mean Sample()
    weight E[G1(wi) G1(wo) <m.wi> / (cos_m cos_i)], which the Smith identity
    bounds by 1 for every roughness and incidence; returns the largest
    z-score of (albedo - 1) over 2 x 4 x 5 settings (a missing shadowing term
    lifts grazing albedos to 1 / G1(wi) > 1)."""
    worst = -np.inf
    for dist in (D.GGX, D.BECKMANN):
        for a in (0.05, 0.2, 0.5, 1.0):
            m = D.rough_conductor(alpha=a, dist=dist, int_ior=1.0, ext_ior=1.0, k=1e4)
            for th in (0, 45, 75, 85, 88):
                f = _sample([m], 0, _dir(th), 40000, th)["f"][:, 0]
                worst = max(worst, (f.mean() - 1) / max(f.std() / np.sqrt(len(f)), 1e-9))
    return worst


def test_rough_conductor_smooth_limit_is_fresnel():
    assert rough_conductor_smooth_error() < 2e-4


def test_rough_conductor_smooth_limit_catches_slip():
    # (a missing shadowing term leaves this limit intact: the albedo bound
    # below catches that one)
    with planted("conductor_fresnel"):
        assert rough_conductor_smooth_error() > 1e-2


def test_rough_conductor_albedo_at_most_one():
    assert rough_conductor_albedo_z() < 4


def test_rough_conductor_albedo_catches_slip():
    with planted("rough_conductor_g"):
        assert rough_conductor_albedo_z() > 40


# ----------------------------------------------------------------- rough dielectric
def rough_dielectric_smooth_z():
    """RoughDielectricBSDF::Sample (RoughDielectricBSDF.cpp:57-147) at alpha =
    MIN_ALPHA: reflection is chosen with probability F(cos_i) at weight
    ksReflect, transmission with 1 - F at ksRefract eta_t^-2 (radiance
    scaling, 1/eta entering, eta leaving), each times the ratio D / D' of the
    sampled and the Walter-scaled distributions, whose mean is 1.  Returns
    the largest z-score of the two means against F and (1 - F) / eta^2."""
    worst = 0.0
    for dist in (D.GGX, D.BECKMANN):
        for eta in (1.5, 1 / 1.5):
            m = D.rough_dielectric(alpha=0.0, dist=dist, int_ior=eta, ext_ior=1.0)
            for th in (0, 30, 60):
                wi = _dir(th)
                r = _sample([m], 0, wi, 200000, th + 7)
                refl = r["wo"][:, 2] * wi[2] > 0
                F, _ = O.fresnel_dielectric(float(wi[2]), eta)
                for sel, exp in ((refl, F), (~refl, (1 - F) / eta ** 2)):
                    x = r["f"][:, 0] * sel
                    se = x.std() / np.sqrt(len(x))
                    worst = max(worst, 0.0 if (se == 0 and abs(x.mean() - exp) < 1e-12) else
                                abs(x.mean() - exp) / max(se, 1e-12))
    return worst


def test_rough_dielectric_smooth_limit():
    assert rough_dielectric_smooth_z() < 5


def test_rough_dielectric_smooth_limit_catches_slip():
    with planted("rough_dielectric_eta"):
        assert rough_dielectric_smooth_z() > 50


# ----------------------------------------------------------------- plastic, rough plastic
def plastic_kd0_error():
    """kd = 0: the specular sampling weight is 1 (PlasticBSDF.cpp:212-216,
    RoughPlasticBSDF.cpp:261-287), so every Sample() is the specular lobe and
    the weight is ks F_dielectric(cos_i) (PlasticBSDF.cpp:58-63; for the rough
    plastic at alpha = MIN_ALPHA, F D G / (4 cos_i) / pdf -> F,
    RoughPlasticBSDF.cpp:66-167)."""
    worst = 0.0
    for m in (D.plastic(int_ior=1.5, ext_ior=1.0, kd=0.0, ks=KS),
              D.rough_plastic(alpha=0.0, dist=D.GGX, int_ior=1.5, ext_ior=1.0, kd=0.0, ks=KS),
              D.rough_plastic(alpha=0.0, dist=D.BECKMANN, int_ior=1.5, ext_ior=1.0, kd=0.0, ks=KS)):
        for th in (0, 40, 70):
            wi = _dir(th)
            f = _sample([m], 0, wi, 20000, th)["f"].mean(0)
            F = fresnel_textbook(float(wi[2]), 1.5)
            worst = max(worst, float(np.max(np.abs(f / (np.array(KS) * F) - 1))))
    return worst


def plastic_albedo_error():
    """Plastic with a diffuse base, kd > 0: the mean Sample() weight is
        ks F(cos_i) + kd / (1 - Fdr_int) eta^-2 (1 - F(cos_i)) (1 - Fbar(eta))
    (PlasticBSDF.cpp:41-88: specular branch ks F / p, diffuse branch the
    cosine-sampled internal term divided by 1 - p), Fbar the cosine-weighted
    Fresnel average of the exit.  Fdr_int is the average internal
    reflectance: for plastic the published fits of
    ApproxFresnelDiffuseReflectance (Common.cpp:550-589, up to 0.6% off the
    quadrature), for rough plastic at alpha = MIN_ALPHA 1 - EvalDiffuse of
    the table (RoughPlasticBSDF.cpp:148-160, 1e-2 off from inside).  Returns
    the worst relative error beyond four standard errors."""
    worst = 0.0
    kd = np.array([0.2, 0.5, 0.8])
    for eta in (1.3, 1.5, 2.0):
        fbar_ext, fdr_int = fresnel_hemispherical(eta), fresnel_hemispherical(1 / eta)
        for m in (D.plastic(int_ior=eta, ext_ior=1.0, kd=tuple(kd), ks=KS),
                  D.rough_plastic(alpha=0.0, dist=D.BECKMANN, int_ior=eta, ext_ior=1.0, kd=tuple(kd), ks=KS)):
            for th in (0, 40, 70):
                wi = _dir(th)
                f = _sample([m], 0, wi, 200000, th + 3)["f"]
                fi = float(fresnel_textbook(float(wi[2]), eta))
                exp = np.array(KS) * fi + kd / (1 - fdr_int) / eta ** 2 * (1 - fi) * (1 - fbar_ext)
                se = f.std(0) / np.sqrt(len(f))
                worst = max(worst, float(np.max(np.maximum(np.abs(f.mean(0) - exp) - 4 * se, 0) / exp)))
    return worst


def test_plastic_without_diffuse_is_fresnel():
    assert plastic_kd0_error() < 1e-4


def test_plastic_albedo_closed_form():
    assert plastic_albedo_error() < 2e-2


def test_plastic_albedo_catches_slip():
    with planted("plastic_eta2"):
        assert plastic_albedo_error() > 0.1


@pytest.mark.parametrize("slip", ["plastic_fresnel", "rough_plastic_eta"])
def test_plastic_without_diffuse_catches_slip(slip):
    with planted(slip):
        assert plastic_kd0_error() > 1e-2


# ----------------------------------------------------------------- coating
NESTED = [D.diffuse((0.6, 0.5, 0.4)), D.rough_conductor(alpha=0.3, dist=D.GGX, k=(3.9, 2.4, 2.1)),
          D.rough_plastic(alpha=0.2, dist=D.BECKMANN, kd=0.4)]


def _pairs(seed, n=4000):
    rng = np.random.default_rng(seed)
    return (random_dirs(rng, n, upper=True), random_dirs(rng, n, upper=True),
            rng.uniform(size=(n, 4)).astype(np.float32))


def coating_index_one_difference():
    """A coating of index 1 without absorption (CoatingBSDF.cpp:39-174): no
    Fresnel reflection (FresnelDielectric returns 0 at eta = 1, Common.cpp:
    494-499), no bending, no specular sampling; Eval and Pdf (solid angle)
    and Sample() equal the nested BSDF's exactly."""
    wi, wo, u = _pairs(5)
    m = np.ones(len(wi), np.int32)
    worst = 0.0
    for nest in NESTED:
        mats = np.stack([D.coating(1, int_ior=1.0, ext_ior=1.0), nest])
        a = O.bsdf_batch(mats, 0, wi, wo, m, u)
        b = O.bsdf_batch(mats, 1, wi, wo, m, u)
        for k in ("eval", "pdf", "f", "wo", "measure"):
            worst = max(worst, float(np.max(np.abs(np.asarray(a[k], float) - np.asarray(b[k], float)))))
    return worst


def coating_absorption_error():
    """Index 1 with absorption sigma_a over thickness d: the nested BSDF
    attenuated by exp(-sigma_a d (1/|cos_i| + 1/|cos_o|)) (CoatingBSDF.cpp:
    82-89, 131-137), in Eval and in Sample()."""
    wi, wo, u = _pairs(6)
    m = np.ones(len(wi), np.int32)
    sig, th = np.array([0.3, 0.6, 1.2]), 0.7
    mats = np.stack([D.coating(1, int_ior=1.0, ext_ior=1.0, sigma_a=tuple(sig), thickness=th), NESTED[0]])
    a = O.bsdf_batch(mats, 0, wi, wo, m, u)
    b = O.bsdf_batch(mats, 1, wi, wo, m, u)
    att = np.exp(-sig[None] * th * (1 / wi[:, 2:3] + 1 / wo[:, 2:3]))
    e1 = np.max(np.abs(a["eval"] - b["eval"] * att) / np.maximum(b["eval"] * att, 1e-12))
    ok = np.abs(b["f"]).sum(1) > 0
    att_s = np.exp(-sig[None] * th * (1 / wi[:, 2:3] + 1 / np.abs(b["wo"][:, 2:3])))
    e2 = np.max(np.abs(a["f"][ok] - b["f"][ok] * att_s[ok]) / np.maximum(b["f"][ok] * att_s[ok], 1e-12))
    return max(float(e1), float(e2))


def coating_albedo_z():
    """A coating of index eta over a diffuse base of albedo a: Sample()
    returns ks R12 / p on the specular branch and a (1 - R12)(1 - R21) / (1 - p)
    on the other (CoatingBSDF.cpp:39-104), so its mean is
        ks R12(cos_i) + (1 - R12) a (1 - Fbar(1/eta)),
    Fbar the cosine-weighted average of the internal reflectance (total
    internal reflection included).  a = 0 isolates the specular branch.
    Returns the largest z-score against it."""
    worst = 0.0
    for eta in (1.3, 1.5, 2.0):
        fbar = fresnel_hemispherical(1 / eta)
        for alb in (0.0, 0.7):
            mats = [D.coating(1, int_ior=eta, ext_ior=1.0, ks=KS), D.diffuse(alb)]
            for th in (0, 40, 70):
                wi = _dir(th)
                f = _sample(mats, 0, wi, 200000, th)["f"]
                r12 = float(fresnel_textbook(float(wi[2]), eta))
                exp = np.array(KS) * r12 + (1 - r12) * alb * (1 - fbar)
                se = f.std(0) / np.sqrt(len(f))
                worst = max(worst, float(np.max(np.abs(f.mean(0) - exp) / np.maximum(se, 1e-12))))
    return worst


def test_coating_of_index_one_is_the_nested_bsdf():
    assert coating_index_one_difference() == 0.0


def test_coating_absorption_closed_form():
    assert coating_absorption_error() < 1e-5


def test_coating_absorption_catches_slip():
    with planted("coating_sigma"):
        assert coating_absorption_error() > 1e-2


def test_coating_albedo_closed_form():
    assert coating_albedo_z() < 5


def test_coating_albedo_catches_slip():
    with planted("coating_exit"):
        assert coating_albedo_z() > 50


# ----------------------------------------------------------------- rough coating
def rough_coating_limit_error():
    """RoughCoatingBSDF (RoughCoatingBSDF.cpp:129-221) at alpha = MIN_ALPHA
    against CoatingBSDF of the same index, absorption and base: the nested
    terms agree as the rough transmittance tends to 1 - F (T(cos_i) =
    1 - R12 and, by Fresnel reciprocity, T(cos_o) = 1 - R21 at the inner
    angle), with the same refracted directions and Jacobian.  Compared on
    Eval and Pdf (solid angle) for directions outside the specular lobe
    (tan theta_h > 0.5, where the GGX tail alpha^2 / tan^4 still adds ~1e-3)
    and away from grazing (cos > 0.1, where the tables' interpolation
    dominates).  Returns the worst relative error per distribution."""
    rng = np.random.default_rng(9)
    wi, wo = random_dirs(rng, 20000, upper=True), random_dirs(rng, 20000, upper=True)
    h = (wi + wo) / np.linalg.norm(wi + wo, axis=1, keepdims=True)
    keep = (np.sqrt(1 - h[:, 2] ** 2) / h[:, 2] > 0.5) & (wi[:, 2] > 0.1) & (wo[:, 2] > 0.1)
    wi, wo = wi[keep], wo[keep]
    m, u = np.ones(len(wi), np.int32), np.zeros((len(wi), 4), np.float32)
    worst = {D.GGX: 0.0, D.BECKMANN: 0.0}
    for eta in (1.3, 1.5, 2.0):
        for dist in (D.GGX, D.BECKMANN):
            for sig in (0.0, 0.4):
                mats = np.stack([D.rough_coating(2, alpha=0.0, dist=dist, int_ior=eta, ks=KS, sigma_a=sig,
                                                 thickness=1.0),
                                 D.coating(2, int_ior=eta, ks=KS, sigma_a=sig, thickness=1.0), D.diffuse((0.7, 0.5, 0.3))])
                a = O.bsdf_batch(mats, 0, wi, wo, m, u)
                b = O.bsdf_batch(mats, 1, wi, wo, m, u)
                e = max(float(np.max(np.abs(a["eval"] - b["eval"]) / np.abs(b["eval"]))),
                        float(np.max(np.abs(a["pdf"] - b["pdf"]) / np.abs(b["pdf"]))))
                worst[dist] = max(worst[dist], e)
    return worst


ROUGH_COATING_TOL = {D.GGX: 1e-2, D.BECKMANN: 2e-4}


def test_rough_coating_smooth_limit_is_coating():
    w = rough_coating_limit_error()
    assert all(w[d] < ROUGH_COATING_TOL[d] for d in w), w


def test_rough_coating_smooth_limit_catches_slip():
    with planted("rough_coating_t"):
        w = rough_coating_limit_error()
        assert all(w[d] > 10 * ROUGH_COATING_TOL[d] for d in w), w


# ----------------------------------------------------------------- rough transmittance
RFT_FILES = ("BeckmannRFTData.bin", "GGXRFTData.bin")


def eval_diffuse_error(inside):
    """RoughTransmittance::EvalDiffuse (RoughTransmitance.cpp:114-150) at
    alpha = 0, on the full table and after SetEta: the diffuse transmittance
    of a smooth interface, 1 - Fbar(eta), Fbar the cosine-weighted average of
    the Fresnel reflectance (for eta < 1 with total internal reflection)."""
    worst = 0.0
    for fn in RFT_FILES:
        path = os.path.join(D.DATA_DIR, fn)
        for eta in ((1 / 1.2, 1 / 1.5, 1 / 2.0) if inside else (1.05, 1.2, 1.5, 2.0, 3.0)):
            exp = 1 - fresnel_hemispherical(eta)
            for stage in (0, 1):
                worst = max(worst, abs(O.rft_eval_diffuse(path, stage, eta, 0.0) - exp))
    return worst


def transmittance_error():
    """RoughTransmittance::Eval (RoughTransmitance.cpp:50-112) at alpha = 0:
    1 - F_dielectric(cos, eta) of a smooth interface (eta > 1)."""
    worst = 0.0
    for fn in RFT_FILES:
        path = os.path.join(D.DATA_DIR, fn)
        for eta in (1.05, 1.5, 2.5):
            for c in np.linspace(0.05, 1.0, 12):
                worst = max(worst, abs(O.rft_eval(path, 0, eta, 0.0, float(c)) - (1 - fresnel_textbook(c, eta))))
    return worst


def test_transmittance_smooth_limit():
    assert transmittance_error() < 2e-5


def test_eval_diffuse_smooth_limit():
    """From outside the table reproduces 1 - Fbar to 2e-5; from inside (eta < 1)
    to 1e-2: the table's cubic interpolation in cos^(1/4) smooths the total
    internal reflection cutoff (T - (1 - F) reaches 3e-2 at the critical
    angle of eta = 1/2)."""
    assert eval_diffuse_error(inside=False) < 2e-5
    assert eval_diffuse_error(inside=True) < 1e-2


def test_eval_diffuse_catches_slip():
    with planted("rft_diffuse"):
        assert eval_diffuse_error(inside=False) > 0.1 and eval_diffuse_error(inside=True) > 0.1


# ----------------------------------------------------------------- film splat
def film_splat_error():
    """ImageBlock::Put (src/core/Block.cpp:95-134) of one unit sample at a
    known film position (px, py) with the Gaussian filter (GaussianFilter.cpp:
    15-22) tabulated at 32 points on [0, r) (Block.cpp:17-24): pixel (x, y)
    receives w(|x - (px - 0.5)|) w(|y - (py - 0.5)|) (centres at +0.5) over
    ceil(c - r) <= x <= floor(c + r), with w(d) = table[int(d 32 / r)] and
    table[i] = max(0, exp(-(r i / 32)^2 / (2 s^2)) - exp(-r^2 / (2 s^2))),
    evaluated in FP32; nothing elsewhere.  Returns the worst absolute
    difference over three filters and four positions (two on pixel
    centres, two off)."""
    worst = 0.0
    W, H = 16, 14
    f32 = np.float32
    for r, sd in ((2.0, 0.5), (1.5, 0.4), (3.2, 1.0)):
        al = f32(-1.0) / (f32(2.0) * f32(sd) * f32(sd))
        pos = (f32(r) * np.arange(32, dtype=f32)) / f32(32)
        tab = np.append(np.maximum(f32(0), np.exp(al * pos * pos) - np.exp(al * f32(r) * f32(r))), f32(0))
        lookup = f32(32) / f32(r)
        for px, py in ((7.5, 6.5), (7.0, 6.0), (7.93, 6.21), (7.25, 6.999)):
            acc = O.film_splat(1, r, sd, W, H, np.array([[px, py]]), np.array([[1.0, 2.0, 3.0]]))
            cx, cy = f32(px) - f32(0.5), f32(py) - f32(0.5)
            exp = np.zeros((H, W))
            for y in range(H):
                for x in range(W):
                    if np.ceil(cx - r) <= x <= np.floor(cx + r) and np.ceil(cy - r) <= y <= np.floor(cy + r):
                        exp[y, x] = (float(tab[int(abs(f32(x) - cx) * lookup)]) *
                                     float(tab[int(abs(f32(y) - cy) * lookup)]))
            for ch, scale in ((3, 1.0), (0, 1.0), (2, 3.0)):
                worst = max(worst, float(np.max(np.abs(acc[..., ch] - scale * exp))))
    return worst


def test_film_splat_of_one_sample():
    assert film_splat_error() < 1e-6


def test_film_splat_catches_slip():
    with planted("film_centre"):
        assert film_splat_error() > 0.1


def test_rr_start_furnace_catches_slip():
    """The furnace check of the roulette start depth (tests/test_oracle.py)
    fails when the weight 1 / 0.95 is applied below the start depth too."""
    from tests.test_oracle import _furnace, lum
    sc = _furnace(0.8)
    sc.max_depth = 1000
    sc.rr_start_depth = 10
    with planted("rr_start_weight"):
        vals = O.pixel_samples(sc, 0, 0, 0, 40000)
    L = lum(vals[:, :3])
    assert abs(L.mean() - 5.0) > 4 * L.std() / np.sqrt(len(L)) + 1e-3, L.mean()
