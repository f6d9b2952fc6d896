"""Pins of oracle functions that only a restatement could otherwise vouch for:
the rand() hash against the published form of Wang's hash, the Disney BRDF value
against Burley's published model, environment importance sampling against the table's
distribution and the pdf it returns, and the triple tracer's MIS weights against the
path tracer's image.  Each check is shown to fail on a plausible one-token mistake."""
import numpy as np
import pytest

from zillumgl_amd import scene
from oracle import oracle as orc

from test_integrators_cpu import primary_light_mask

M32 = np.uint64(0xFFFFFFFF)


# ---------------------------------------------------------------------------- hash
def wang_hash_published(x, s1=16, s2=4, s3=15):
    """Thomas Wang's 32-bit integer hash as published (key ^ 61 ^ key >> 16; key +=
    key << 3; key ^= key >> 4; key *= 0x27d4eb2d; key ^= key >> 15), in uint64 numpy with
    explicit wrap; random.glsl:5-13 writes the second step as a multiply by 9."""
    k = x.astype(np.uint64)
    k = (k ^ np.uint64(61)) ^ (k >> np.uint64(s1))
    k = (k + (k << np.uint64(3))) & M32
    k = k ^ (k >> np.uint64(s2))
    k = (k * np.uint64(0x27D4EB2D)) & M32
    k = k ^ (k >> np.uint64(s3))
    return k.astype(np.uint32)


def test_hash_known_answers_over_random_inputs():
    rng = np.random.default_rng(11)
    x = np.concatenate([rng.integers(0, 2 ** 32, 1 << 20, dtype=np.uint64).astype(np.uint32),
                        np.array([0, 1, 61, 0xFFFF, 0x10000, 0x10001, 0x7FFFFFFF, 0x80000000, 0xFFFFFFFF],
                                 np.uint32)])
    got = orc.hash_many(x)
    assert np.array_equal(got, wang_hash_published(x))
    # a one-token mistake in any shift changes almost every output
    for bad in (dict(s1=15), dict(s2=3), dict(s3=16)):
        assert (wang_hash_published(x, **bad) != got).mean() > 0.99, bad


# ---------------------------------------------------------------------------- Disney
def _dot(a, b):
    return np.sum(a * b, -1)


def disney_burley(rec, wo, wi, n, fd90_coef=2.0, ss_blend=True):
    """FP64 value of the principled BRDF from Burley 2012 ("Physically Based Shading at
    Disney", sections 5.3-5.5 and its BRDF-explorer shader) with the reference's choices
    (material.glsl:357-487, material_loader.glsl:28-46): roughness r = mix(0.0134, 1, r0),
    alpha = r^2; Schlick-GGX masking with k = alpha / 2 (clearcoat alpha 0.25);
    tint = baseColor / Rec.601 luminance; the clearcoat's Fresnel has baseColor as F0 and
    is weighted 0.25 * clearcoat; sheen is not scaled by (1 - metallic)."""
    base = rec[0:3].astype(np.float64)
    r0, ss, met, spec, stint = (float(v) for v in rec[3:8])
    sheen, shtint, cc, ccg = (float(v) for v in rec[8:12])
    r = 0.0134 * (1 - r0) + r0
    alpha = r * r
    cl, cv = _dot(n, wi), _dot(n, wo)
    if cl < 1e-6 or cv < 1e-6:
        return None
    h = wi + wo
    h = h / np.linalg.norm(h)
    cd = _dot(h, wi)
    ch = _dot(n, h)
    FL, FV = (1 - cl) ** 5, (1 - cv) ** 5
    fd90 = 0.5 + fd90_coef * r * cd * cd
    fd = (1 + (fd90 - 1) * FL) * (1 + (fd90 - 1) * FV)
    diffuse = base / np.pi * fd
    fss90 = r * cd * cd
    fss = (1 + (fss90 - 1) * FL) * (1 + (fss90 - 1) * FV)
    sss = base / np.pi * 1.25 * (fss * (1 / (cl + cv) - 0.5) + 0.5)
    if ss_blend:
        diffuse = diffuse * (1 - ss) + sss * ss
    lum = 0.299 * base[0] + 0.587 * base[1] + 0.114 * base[2]
    tint = base / lum if lum > 0 else np.ones(3)
    f0 = (np.ones(3) * (1 - stint) + tint * stint) * (0.08 * spec)
    f0 = f0 * (1 - met) + base * met
    F = f0 + (1 - f0) * (1 - _dot(h, wo)) ** 5
    D = alpha ** 2 / (np.pi * ((alpha ** 2 - 1) * ch * ch + 1) ** 2)

    def g1(c, a):
        k = a / 2
        return c / (c * (1 - k) + k)

    G = g1(cv, alpha) * g1(cl, alpha)
    specular = F * D * G / (4 * cl * cv)
    ac = 0.1 * (1 - ccg) + 0.001 * ccg
    Dc = (ac * ac - 1) / (2 * np.pi * np.log(ac) * (1 + (ac * ac - 1) * max(ch, 0) ** 2))
    Fc = base + (1 - base) * (1 - _dot(h, wo)) ** 5
    Gc = g1(cv, 0.25) * g1(cl, 0.25)
    coat = Fc * Dc * Gc / (4 * cl * cv)
    sh = (np.ones(3) * (1 - shtint) + tint * shtint) * (1 - cd) ** 5 * sheen
    return diffuse * (1 - met) + specular + 0.25 * cc * coat + sh


DISNEY_SETS = [
    dict(base_color=(0.7, 0.3, 0.2), roughness=0.5, metallic=0.0, specular=0.0),
    dict(base_color=(0.7, 0.3, 0.2), roughness=0.6, metallic=1.0, specular=0.5),
    dict(base_color=(0.2, 0.6, 0.4), roughness=0.35, metallic=0.3, specular=0.8, specular_tint=0.4, subsurface=0.6),
    dict(base_color=(0.9, 0.8, 0.1), roughness=0.8, metallic=0.0, specular=0.5, sheen=0.7, sheen_tint=0.3),
    dict(base_color=(0.4, 0.4, 0.9), roughness=0.45, metallic=0.1, specular=0.5, clearcoat=0.5,
         clearcoat_gloss=0.6),
    dict(base_color=(0.5, 0.5, 0.5), roughness=0.25, metallic=0.0, specular=0.5, subsurface=1.0),
]


def _disney_cases(kw, n=400, seed=0):
    rng = np.random.default_rng(seed)
    rec = scene.Material(type=scene.MAT_PRINCIPLED, **kw).record()
    nrm = np.array([0.0, 0.0, 1.0])
    out = []
    while len(out) < n:
        d = rng.normal(size=(2, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        d[:, 2] = np.abs(d[:, 2])
        wo, wi = d.astype(np.float32).astype(np.float64)
        if wo[2] < 0.05 or wi[2] < 0.05:
            continue
        f, _ = orc.bsdf_eval(rec, wo.astype(np.float32), wi.astype(np.float32), nrm.astype(np.float32))
        out.append((rec, wo, wi, nrm, f.astype(np.float64)))
    return out


@pytest.mark.parametrize("kw", DISNEY_SETS)
def test_disney_value_matches_burley(kw):
    for rec, wo, wi, nrm, f in _disney_cases(kw):
        ref = disney_burley(rec, wo, wi, nrm)
        assert np.allclose(f, ref, rtol=2e-4, atol=1e-6), (kw, wo, wi, f, ref)


def test_disney_reductions():
    """Normal incidence with specular, metallic, sheen, subsurface and clearcoat at 0:
    Burley's diffuse is exactly baseColor / pi (F_D90 terms vanish, Schlick F with F0 = 0
    is 0).  metallic = 1: exactly GGX D * Schlick F(baseColor) * G / (4 cos cos)."""
    n = np.float32([0, 0, 1])
    rec = scene.Material(type=scene.MAT_PRINCIPLED, base_color=(0.3, 0.6, 0.9), roughness=0.7, metallic=0.0,
                         specular=0.0).record()
    f, _ = orc.bsdf_eval(rec, n, n, n)
    assert np.allclose(f, np.float32([0.3, 0.6, 0.9]) / np.pi, rtol=1e-6)
    rec = scene.Material(type=scene.MAT_PRINCIPLED, base_color=(0.9, 0.5, 0.2), roughness=0.55, metallic=1.0,
                         specular=0.5).record()
    for rec_, wo, wi, nrm, f in _disney_cases(dict(base_color=(0.9, 0.5, 0.2), roughness=0.55, metallic=1.0,
                                                   specular=0.5), n=100, seed=3):
        r = 0.0134 * (1 - 0.55) + 0.55
        a = r * r
        h = (wo + wi) / np.linalg.norm(wo + wi)
        D = a * a / (np.pi * ((a * a - 1) * h[2] ** 2 + 1) ** 2)
        F = np.array([0.9, 0.5, 0.2]) + (1 - np.array([0.9, 0.5, 0.2])) * (1 - h @ wo) ** 5
        G = (wo[2] / (wo[2] * (1 - a / 2) + a / 2)) * (wi[2] / (wi[2] * (1 - a / 2) + a / 2))
        assert np.allclose(f, F * D * G / (4 * wo[2] * wi[2]), rtol=2e-4)


@pytest.mark.parametrize("bad", [dict(fd90_coef=1.0), dict(ss_blend=False)])
def test_disney_check_catches_planted_mistakes(bad):
    """The comparison above fails when the model has a wrong F_D90 coefficient or drops
    the subsurface blend: it has the power to see such a restatement error."""
    kw = DISNEY_SETS[2]
    cases = _disney_cases(kw, n=200)
    ok = [np.allclose(f, disney_burley(rec, wo, wi, nrm, **bad), rtol=2e-4, atol=1e-6)
          for rec, wo, wi, nrm, f in cases]
    assert np.mean(ok) < 0.5


# ---------------------------------------------------------------------------- environment IS
def env_scene(img):
    d = scene.SceneDesc()
    fl = scene.quad((-50, -50, 0), (50, -50, 0), (50, 50, 0), (-50, 50, 0), (0, 0, 1))
    sph = scene.uv_sphere((0.0, 0.0, 0.6), 0.5, 24, 13)
    d.objects = [(fl, scene.Material(base_color=(0.6, 0.6, 0.6))),
                 (sph, scene.Material(base_color=(0.8, 0.5, 0.3), roughness=0.6, specular=0.5,
                                      type=scene.MAT_PRINCIPLED))]
    d.env = img
    d.env_rotation = 0.7
    d.camera = scene.CameraDesc(pos=(0, -4, 1.5), angle=(0, -15, 0), fov=35)
    d.width = d.height = 24
    return scene.pack(d)


def gray_sky(w=64, h=32, sun_radius_deg=12.0, peak=40.0):
    """A smooth gray (r = g = b) sky with a broad sun lobe: Rec.601 and Rec.709
    luminance agree and the sum is large, so the pdf's only approximations are FP16
    storage and the int truncation of the sum."""
    img = scene.sky_env(w, h, seed=3, sun_theta=50.0, sun_phi=30.0, peak=peak, sun_radius_deg=sun_radius_deg)
    g = img.mean(-1, keepdims=True) * 20.0
    return np.ascontiguousarray(np.repeat(g, 3, -1).astype(np.float32))


def table_distribution(img):
    """The distribution the env table is built to sample (EnvironmentMap.cpp:19-49):
    Rec.709 luminance times sin(theta) of the row, normalised, from the FP32 image."""
    h, w, _ = img.shape
    lum = img.astype(np.float64) @ np.array([0.2126, 0.7152, 0.0722])
    wgt = lum * np.sin((np.arange(h) + 0.5) / h * np.pi)[:, None]
    return wgt / wgt.sum()


def test_env_sampling_chi2_against_table_distribution():
    img = scene.sky_env(64, 32, seed=5, sun_theta=60.0, sun_phi=200.0, peak=50.0, sun_radius_deg=8.0)
    ps = env_scene(img)
    n = 400_000
    u = np.random.default_rng(7).uniform(size=(n, 4)).astype(np.float32)
    _, _, tx = orc.env_sample(ps, u)
    h, w = img.shape[:2]
    counts = np.bincount(tx[:, 0] * w + tx[:, 1], minlength=h * w).astype(np.float64)
    p = table_distribution(img).ravel()
    big = p * n >= 20
    chi2 = np.sum((counts[big] - n * p[big]) ** 2 / (n * p[big]))
    dof = big.sum() - 1
    assert chi2 < dof + 5 * np.sqrt(2 * dof), (chi2, dof)
    # the same histogram against a row-uniform (no sin) or transposed reading fails
    lum = img.astype(np.float64) @ np.array([0.2126, 0.7152, 0.0722])
    q = (lum / lum.sum()).ravel()
    chi2_bad = np.sum((counts[big] - n * q[big]) ** 2 / (n * np.maximum(q[big], 1e-12)))
    assert chi2_bad > 10 * dof


@pytest.mark.parametrize("plant", [0, orc.VAR_PLANT_ENV_SIN])
def test_env_pdf_times_texel_solid_angle_is_table_probability(plant):
    """At a sampled texel centre the returned pdf times the texel's solid angle
    (2 pi / w)(pi / h) sin(theta) equals the table's probability of that texel (gray
    map: up to FP16 storage and the int-truncated sum); a pdf with an extra 1/sin(theta)
    Jacobian fails."""
    img = gray_sky()
    ps = env_scene(img)
    h, w = img.shape[:2]
    u = np.random.default_rng(8).uniform(size=(20_000, 4)).astype(np.float32)
    with orc.variant(plant):
        wi, pdf, tx = orc.env_sample(ps, u)
        pdf2 = orc.env_pdf(ps, wi)
    assert np.array_equal(pdf, pdf2)  # envSampleWi's pdf is envPdfLi at the direction
    p = table_distribution(img)[tx[:, 0], tx[:, 1]]
    sin_t = np.sin((tx[:, 0] + 0.5) / h * np.pi)
    ratio = pdf * (2 * np.pi / w) * (np.pi / h) * sin_t / p
    if plant:
        assert np.abs(ratio - 1).max() > 0.05
    else:
        assert np.abs(ratio - 1).max() < 2e-3, np.abs(ratio - 1).max()


def _nee_on_off(ps, spp=128, nb=4, depth=3):
    means = {}
    for nee in (True, False):
        m = []
        for b in range(nb):
            img, _ = orc.render_pixels(ps, orc.params(0, b * spp, spp, 1 + b * spp, depth, sample_light=nee))
            m.append(img.astype(np.float64).mean() / spp)
        means[nee] = (np.mean(m), np.std(m) / np.sqrt(nb))
    (a, ea), (b, eb) = means[True], means[False]
    return a, b, (a - b) / np.hypot(ea, eb)


def smooth_sky(w, h, sun_theta=55.0, sun_phi=120.0, peak=30.0, rad=10.0):
    """Sky gradient over a dark ground plus a Gaussian sun lobe, without texel noise."""
    v = (np.arange(h) + 0.5) / h
    u = (np.arange(w) + 0.5) / w
    T, P = np.meshgrid(v * np.pi, u * 2 * np.pi, indexing="ij")
    up = np.cos(T)
    k = np.clip(up, 0, 1)[..., None]
    sky = np.array([0.9, 0.9, 1.0]) * (1 - k) + np.array([0.25, 0.45, 1.0]) * k
    img = np.where((up >= 0)[..., None], sky, np.array([0.12, 0.11, 0.09]))
    d = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1)
    st, sp = np.radians(sun_theta), np.radians(sun_phi)
    sd = np.array([np.sin(st) * np.cos(sp), np.sin(st) * np.sin(sp), np.cos(st)])
    ang = np.arccos(np.clip(d @ sd, -1, 1))
    img = img + (peak * np.exp(-0.5 * (ang / np.radians(rad)) ** 2))[..., None] * np.array([1.0, 0.95, 0.85])
    return np.ascontiguousarray(img.astype(np.float32))


def test_env_nee_on_and_off_agree_on_a_nonconstant_map():
    """Next-event estimation with the env table and BSDF sampling alone estimate the same
    image on a sun + sky map (MIS weights use envPdfLi); with the planted 1/sin(theta)
    pdf the NEE estimate moves away.  The map is smooth: the reference samples texel
    centres without jitter (light.glsl:198-200), a midpoint rule whose error grows with
    texel-to-texel variation (1% on a map with lognormal texel noise)."""
    ps = env_scene(smooth_sky(64, 32))
    a, b, z = _nee_on_off(ps)
    assert abs(a / b - 1) < 0.005 and abs(z) < 4, (a, b, z)
    with orc.variant(orc.VAR_PLANT_ENV_SIN):
        a2, b2, z2 = _nee_on_off(ps)
    assert abs(a2 / b2 - 1) > 0.03 and abs(z2) > 6, (a2, b2, z2)


# ---------------------------------------------------------------------------- triple MIS
def _pt_and_triple(ps, depth, mask, spp=256, nb=4):
    pts, trs = [], []
    for b in range(nb):
        prm = dict(first_pass=b * spp, n_passes=spp, free_counter=1 + b * spp, max_depth=depth)
        pt, _ = orc.render_pixels(ps, orc.params(0, **prm))
        pts.append((pt.astype(np.float64) / spp)[mask].mean())
        cam, _ = orc.render_pixels(ps, orc.params(2, blocks_one_pass=1, **prm))
        lpt, _ = orc.render_light(ps, orc.params(2, blocks_one_pass=1, **prm))
        trs.append(((cam.astype(np.float64) + lpt) / spp)[mask].mean())
    return np.mean(pts), np.mean(trs), np.std(pts) / np.sqrt(nb), np.std(trs) / np.sqrt(nb)


@pytest.fixture(scope="module")
def c1_32():
    ps = scene.pack(scene.config1(32, 32))
    return ps, ~primary_light_mask(ps)


@pytest.mark.parametrize("depth", [1, 3])
def test_triple_estimates_the_path_traced_image(c1_32, depth):
    """The triple tracer's three strategies (s=0, s=1, t=1) partition each path: off the
    pixels that see an emitter directly, camera pass + light pass estimate the path
    tracer's image.  That holds once the camera pass weighs the t=1 strategy with the
    light's actual emission pdf (cos/pi, as lightSampleOneLe samples); the reference
    weighs it with lightPdfLe's uniform 1/(2 pi) (triple_path_pass_pt.glsl:119, :162),
    and is brighter by 3-6% on this scene.  A dropped reverse-pdf factor (depth > 1)
    moves the image by several percent."""
    ps, mask = c1_32
    with orc.variant(orc.VAR_EMIT_COS):
        pt, tr, ept, etr = _pt_and_triple(ps, depth, mask)
    assert abs(tr / pt - 1) < 0.005, (depth, pt, tr, ept, etr)
    pt0, tr0, _, _ = _pt_and_triple(ps, depth, mask)
    assert pt0 == pt and 1.02 < tr0 / pt0 < 1.08, (pt0, tr0)
    if depth > 1:
        with orc.variant(orc.VAR_EMIT_COS | orc.VAR_PLANT_PREV):
            ptp, trp, _, _ = _pt_and_triple(ps, depth, mask)
        assert abs(trp / ptp - 1) > 0.03, (ptp, trp)
