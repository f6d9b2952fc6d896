"""BSDF pins for the oracle's restatement of material.glsl / microfacet.glsl:
closed forms, normalisation, Snell / Fresnel against the textbook, reciprocity."""
import numpy as np
import pytest

from zillumgl_amd import scene
from oracle import oracle as orc

N = np.array([0.0, 0.0, 1.0], np.float32)


def fib_sphere(n):
    i = np.arange(n) + 0.5
    z = 1 - 2 * i / n
    phi = np.pi * (1 + 5 ** 0.5) * i
    r = np.sqrt(1 - z * z)
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], 1).astype(np.float32)


def mat(**kw):
    return scene.Material(**kw).record()


def test_lambertian_closed_form():
    m = mat(base_color=(0.2, 0.5, 0.8))
    wo = np.array([0.3, 0.1, 0.9], np.float32)
    wo /= np.linalg.norm(wo)
    for wi in fib_sphere(50):
        f, pdf = orc.bsdf_eval(m, wo, wi, N)
        assert np.allclose(f, np.float32([0.2, 0.5, 0.8]) / np.float32(np.pi), rtol=1e-6)
        assert abs(pdf - max(wi[2], 0) / np.pi) < 1e-6


def test_lambertian_sampling_is_cosine_distributed():
    m = mat(base_color=(1, 1, 1))
    rng = np.random.default_rng(0)
    cos = []
    for _ in range(20000):
        u = rng.uniform(size=3).astype(np.float32)
        wi, pdf, f, flag = orc.bsdf_sample(m, N, N, u)
        cos.append(wi[2])
        # white furnace per sample: f * cos / pdf = albedo
        assert abs(f[0] * wi[2] / pdf - 1.0) < 1e-4
    c = np.sort(np.array(cos))
    # CDF of cos(theta) under cosine-weighted sampling is c^2 (Kolmogorov-Smirnov)
    ks = np.max(np.abs(np.arange(1, len(c) + 1) / len(c) - c ** 2))
    assert ks < 0.015


def test_ggx_normalisation():
    """GGX D (microfacet.glsl:30-40) satisfies  int D(m) cos(theta_m) dw_m = 1."""
    for alpha in (0.05, 0.2, 0.5, 0.9):
        th = (np.arange(200_000) + 0.5) / 200_000 * (np.pi / 2)
        d = np.array([orc.ggx(np.float32(np.cos(t)), alpha) for t in th[::10]])
        integral = 2 * np.pi * np.sum(d * np.cos(th[::10]) * np.sin(th[::10])) * (np.pi / 2) / len(d)
        assert abs(integral - 1) < 2e-3, (alpha, integral)


@pytest.mark.parametrize("kw", [
    dict(type=scene.MAT_PRINCIPLED, roughness=0.6, metallic=0.0, specular=0.5),
    dict(type=scene.MAT_PRINCIPLED, roughness=0.8, metallic=1.0, specular=0.5),
    dict(type=scene.MAT_METAL_WORKFLOW, roughness=0.7, metallic=0.5),
])
def test_pdf_matches_sampling_density(kw):
    """The mass of the pdf over the upper hemisphere equals the fraction of samples that
    land there (the BRDF lobes lose what the GGX lobe reflects below the horizon; the
    cosine lobe uses |n.wi|, material.glsl:628, hence the hemisphere domain)."""
    m = mat(**kw)
    wo = np.array([0.4, 0.0, 0.9165], np.float32)
    dirs = fib_sphere(60_000)
    dirs = dirs[dirs[:, 2] > 0]
    integral = sum(orc.bsdf_eval(m, wo, wi, N)[1] for wi in dirs) * 2 * np.pi / len(dirs)
    rng = np.random.default_rng(4)
    n = 20000
    up = 0
    for k in range(n):
        wi, pdf, f, flag = orc.bsdf_sample(m, wo, N, rng.uniform(size=3).astype(np.float32), seed=k + 1)
        up += wi[2] > 1e-6
    assert 0.6 < integral <= 1.02
    assert abs(up / n - integral) < 0.04, (up / n, integral)


def test_rough_dielectric_transmission_pdf_is_zero():
    """Reference quirk: dielectricPdf builds the refraction half vector as
    normalize(wo + wi * eta) (material.glsl:362-369), which points away from n, so
    ggxPdfWm returns 0 for every transmitted direction; dielectricSample instead flips h
    to n's side.  The restatement keeps this."""
    m = mat(type=scene.MAT_DIELECTRIC, roughness=0.5, ior=1.5)
    wo = np.array([0.4, 0.0, 0.9165], np.float32)
    dirs = fib_sphere(40_000)
    below = dirs[dirs[:, 2] < 0]
    vals = np.array([orc.bsdf_eval(m, wo, wi, N)[1] for wi in below])
    mass_below = vals.sum() * 2 * np.pi / len(below)
    rng = np.random.default_rng(5)
    n = 4000
    trans = sum(orc.bsdf_sample(m, wo, N, rng.uniform(size=3).astype(np.float32))[0][2] < 0 for _ in range(n))
    assert (vals == 0).mean() > 0.9
    assert mass_below < 0.02 and trans / n > 0.8


def test_concentric_disk_quirk_and_area_preservation():
    """toConcentricDisk (math.glsl:25-41): (0,0) maps to the centre, the input (0.5,0.5)
    gives 0/0 (NaN) as in the reference, and the map is area preserving."""
    assert np.array_equal(orc.concentric(0.0, 0.0), [0.0, 0.0])
    assert np.isnan(orc.concentric(0.5, 0.5)).all()
    rng = np.random.default_rng(9)
    pts = np.array([orc.concentric(*rng.uniform(size=2).astype(np.float32)) for _ in range(20000)])
    r2 = (pts ** 2).sum(1)
    assert r2.max() <= 1 + 1e-6
    ks = np.max(np.abs(np.arange(1, len(r2) + 1) / len(r2) - np.sort(r2)))
    assert ks < 0.015


def test_disney_reciprocity_upper_hemisphere():
    m = mat(type=scene.MAT_PRINCIPLED, base_color=(0.7, 0.3, 0.2), roughness=0.5, metallic=0.3, specular=0.5,
            sheen=0.5, subsurface=0.3)
    rng = np.random.default_rng(1)
    for _ in range(300):
        a, b = fib_sphere(2000)[rng.integers(0, 1000, 2)]
        f1, _ = orc.bsdf_eval(m, a, b, N)
        f2, _ = orc.bsdf_eval(m, b, a, N)
        assert np.allclose(f1, f2, rtol=2e-5, atol=1e-7)


def fresnel_textbook(ci, eta):
    if ci < 0:
        eta, ci = 1 / eta, -ci
    st = np.sqrt(max(0, 1 - ci * ci)) / eta
    if st >= 1:
        return 1.0
    ct = np.sqrt(1 - st * st)
    rs = (ci - eta * ct) / (ci + eta * ct)
    rp = (eta * ci - ct) / (eta * ci + ct)
    return 0.5 * (rs * rs + rp * rp)


def test_fresnel_matches_textbook():
    assert abs(orc.fresnel(1.0, 1.5) - ((1.5 - 1) / (1.5 + 1)) ** 2) < 1e-7
    for c in np.linspace(-1, 1, 401):
        assert abs(orc.fresnel(np.float32(c), 1.5) - fresnel_textbook(float(np.float32(c)), 1.5)) < 2e-6
    # total internal reflection from inside beyond the critical angle
    assert orc.fresnel(-0.3, 1.5) == 1.0


def test_smooth_dielectric_obeys_snell():
    m = mat(type=scene.MAT_DIELECTRIC, roughness=0.0, ior=1.5)
    for ci in [0.95, 0.7, 0.4, 0.1]:
        wo = np.array([np.sqrt(1 - ci * ci), 0.0, ci], np.float32)
        wi, pdf, f, flag = orc.bsdf_sample(m, wo, N, np.float32([0.999, 0.5, 0.5]))
        assert flag == 1 << 4  # SpecTrans
        sin_t = np.sqrt(1 - wi[2] ** 2)
        assert abs(sin_t - np.sqrt(1 - ci * ci) / 1.5) < 1e-5
        assert wi[2] < 0 and abs(f[0] - 1 / 1.5 ** 2) < 1e-6  # radiance scaling 1/eta^2
        wr, _, fr, flag_r = orc.bsdf_sample(m, wo, N, np.float32([0.0, 0.5, 0.5]))
        assert flag_r == 1 << 3 and np.allclose(wr, [-wo[0], 0, wo[2]], atol=1e-6)
