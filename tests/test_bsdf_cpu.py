"""Pins of the oracle's GGX BSDF and light samplers against mathematics that
does not restate them (CPU only).

The oracle (oracle/or_render.cpp) restates
  SampleGGXBoundedVNDF / GGXBoundedVNDF_PDF      core/render/sample.cuh:179-248
  GGXLambert_BRDF_Spectrum, GGX_D, GGX_G_Correlated, FresnelSchlick
                                                 core/render/brdf.cuh:114-129
  SampleBSDF / EvaluateBSDFWithCos / BSDF_PDF     core/render/material.cuh:280-387
  NEESampleSphereCapArea / NEEPDFOmegaFromHit_Sphere, the box sampler and
  NEESampleLight                                 core/nee/nee.cuh:35-377
and the kernels equal the oracle bit for bit, so these pins are what ties both
to the reference's mathematics:
  * the bounded-VNDF sampler's histogram follows GGXBoundedVNDF_PDF (chi^2 over
    a (cos theta, phi) grid integrated by quadrature), and the pdf over the
    upper hemisphere plus the sampler's below-horizon fraction is one;
  * closed forms of the bounded VNDF: alpha = 1 is the uniform hemisphere
    (pdf 1/2pi), normal incidence gives (1 + alpha^2) D(h) / 4 and never goes
    below the horizon (Eto and Tokuyoshi 2023, the GGX NDF in its tan form);
  * EvaluateBSDFWithCos equals the textbook Cook-Torrance GGX with the
    height-correlated Smith G2 = 1 / (1 + Lambda(v) + Lambda(l)) and Schlick's
    Fresnel (Heitz 2014), written here in the tan/Lambda form, not the
    oracle's; reciprocity; energy <= 1 for the metal lobe (the dielectric mix
    is not energy-conserving in the reference: its Lambert weight is
    1 - F(l.h)); a smooth metal reflects F(theta);
  * the GGX mixture: SampleBSDF's pdf equals BSDF_PDF at the drawn direction,
    its histogram follows BSDF_PDF, and its terminated fraction is what
    BSDF_PDF leaves outside the hemisphere;
  * sphere-cap NEE: the sampler's pdf equals NEEPDFOmegaFromHit at the drawn
    point, points lie on the visible cap with cos(theta) uniform in
    [r/d, 1] (Archimedes), the cap area recovered from the pdf is
    2 pi r^2 (1 - r/d), and the pdf integrates to one over the sphere's solid
    angle (ray casting in FP64); the box sampler and a two-light table
    integrate to one likewise.
Each statistical pin is shown to fail on a planted mistake (oracle_lib.planted).
"""
import math

import numpy as np
import pytest

import aperture as ap
import oracle_lib as ol
import scenes

N_UP = np.array([0.0, 0.0, 1.0])


def _view(cos_v):
    return np.array([math.sqrt(max(0.0, 1.0 - cos_v * cos_v)), 0.0, cos_v])


def _grid(nc=24, nphi=24, sub=10):
    """Quadrature points of the upper hemisphere in (cos theta, phi): bins of
    equal d(cos theta) d(phi) (= solid angle), sub x sub midpoints per bin."""
    c_edges = np.linspace(0.0, 1.0, nc + 1)
    p_edges = np.linspace(-math.pi, math.pi, nphi + 1)
    cc = (np.arange(nc * sub) + 0.5) / (nc * sub)
    pp = -math.pi + (np.arange(nphi * sub) + 0.5) * (2 * math.pi / (nphi * sub))
    C, P = np.meshgrid(cc, pp, indexing="ij")
    s = np.sqrt(1.0 - C * C)
    dirs = np.stack([s * np.cos(P), s * np.sin(P), C], -1).reshape(-1, 3)
    dw = (1.0 / (nc * sub)) * (2 * math.pi / (nphi * sub))
    return c_edges, p_edges, dirs, dw, (nc, sub, nphi, sub)


def _expected(pdf_values, shape, dw):
    nc, sc, nphi, sp = shape
    return pdf_values.astype(np.float64).reshape(nc, sc, nphi, sp).sum(axis=(1, 3)) * dw


def _histogram(dirs, c_edges, p_edges):
    up = dirs[:, 2] > 0.0
    d = dirs[up].astype(np.float64)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    h, _, _ = np.histogram2d(d[:, 2], np.arctan2(d[:, 1], d[:, 0]), bins=[c_edges, p_edges])
    return h


def _chi2_z(observed, expected, n):
    """Pearson chi^2 over the bins with >= 20 expected samples, as a z score
    (Wilson-Hilferty); plus the probability mass the quadrature puts in the
    hemisphere."""
    e = expected * n
    keep = e >= 20.0
    chi2 = float((((observed - e) ** 2) / np.maximum(e, 1e-30))[keep].sum())
    k = max(int(keep.sum()) - 1, 1)
    z = ((chi2 / k) ** (1.0 / 3.0) - (1.0 - 2.0 / (9.0 * k))) / math.sqrt(2.0 / (9.0 * k))
    return z, float(expected.sum())


def vndf_stats(cos_v, alpha, n=400_000, seed=3):
    """(chi^2 z score, |hemisphere mass + below-horizon fraction - 1| in
    standard errors, pairing error p99) of the bounded-VNDF sampler."""
    v = _view(cos_v)
    rnd = np.random.default_rng(seed).random((n, 3))
    dirs, pdfs = ol.ggx_vndf_sample(v, N_UP, alpha, rnd)
    c_edges, p_edges, qd, dw, shape = _grid()
    exp = _expected(ol.ggx_vndf_pdf(v, N_UP, qd, alpha), shape, dw)
    z, mass = _chi2_z(_histogram(dirs, c_edges, p_edges), exp, n)
    below = float((dirs[:, 2] <= 0.0).mean())
    se = math.sqrt(max(below * (1 - below), 1.0 / n) / n)
    norm_sigma = abs(mass + below - 1.0) / (se + 2e-4)   # 2e-4: quadrature allowance
    up = dirs[:, 2] > 0.0
    pair = np.abs(ol.ggx_vndf_pdf(v, N_UP, dirs[up], alpha) / pdfs[up] - 1.0)
    return z, norm_sigma, float(np.percentile(pair, 99))


VNDF_CASES = [(0.15, 0.2), (0.5, 0.5), (0.9, 0.8), (0.35, 0.1), (0.7, 0.35)]


@pytest.mark.parametrize("cos_v,alpha", VNDF_CASES)
def test_vndf_sampler_follows_its_pdf(cos_v, alpha):
    z, norm_sigma, pair99 = vndf_stats(cos_v, alpha)
    print(f"[vndf] cosV {cos_v} alpha {alpha}: chi2 z {z:.2f}, normalisation {norm_sigma:.2f} sigma, "
          f"pairing p99 {pair99:.2e}")
    assert z < 4.0
    assert norm_sigma < 4.0
    assert pair99 < 1e-4


@pytest.mark.parametrize("plant", [1, 5])
def test_vndf_pins_catch_planted_mistakes(plant):
    """1: the pdf's factor 2 dropped in sampler and pdf alike (pairing cannot
    see it, the normalisation does); 5: the stretched view vector left
    unnormalised (the histogram does)."""
    with ol.planted(plant):
        z, norm_sigma, _ = vndf_stats(0.5, 0.5, n=200_000)
    print(f"[vndf] plant {plant}: chi2 z {z:.1f}, normalisation {norm_sigma:.1f} sigma")
    assert z > 10.0 or norm_sigma > 10.0


def test_vndf_closed_forms():
    rnd = np.random.default_rng(5).random((50_000, 3))
    # alpha = 1: k = 0, t = 1, D = 1/pi, so the pdf is 1/(2 pi): uniform hemisphere.
    for cos_v in (0.1, 0.6, 1.0):
        dirs, pdfs = ol.ggx_vndf_sample(_view(cos_v), N_UP, 1.0, rnd)
        assert (dirs[:, 2] > 0).all()
        assert np.allclose(pdfs, 1.0 / (2.0 * math.pi), rtol=2e-6)
        c = np.sort(dirs[:, 2].astype(np.float64))
        ks = np.abs(c - (np.arange(len(c)) + 0.5) / len(c)).max()
        assert ks < 1.63 / math.sqrt(len(c)) + 1e-3, ks   # 1% KS level
    # Normal incidence: the bounded cap is exact, pdf = (1 + a^2) D(h) / 4.
    for alpha in (0.05, 0.3, 0.7):
        dirs, pdfs = ol.ggx_vndf_sample(N_UP, N_UP, alpha, rnd)
        assert (dirs[:, 2] > 0).all()
        h = dirs.astype(np.float64) + N_UP
        h /= np.linalg.norm(h, axis=1, keepdims=True)
        cos_h = h[:, 2]
        tan2 = (1.0 - cos_h ** 2) / cos_h ** 2
        d = alpha ** 2 / (math.pi * cos_h ** 4 * (alpha ** 2 + tan2) ** 2)
        ref = (1.0 + alpha ** 2) * d / 4.0
        rel = np.abs(pdfs / ref - 1.0)
        assert np.percentile(rel, 99) < 2e-4 and rel.max() < 5e-3, (alpha, rel.max())


def _textbook(v, l, roughness, metallic, specular, al):
    """Cook-Torrance GGX + Lambert in FP64: tan form of D, Lambda form of the
    height-correlated Smith G2, Schlick's Fresnel; returns f * cos(l)."""
    a = roughness * roughness
    h = v + l
    h /= np.linalg.norm(h, axis=1, keepdims=True)
    cos_h, cv, cl = h[:, 2], v[:, 2], l[:, 2]
    tan2h = (1 - cos_h ** 2) / cos_h ** 2
    D = a * a / (math.pi * cos_h ** 4 * (a * a + tan2h) ** 2)

    def lam(c):
        return (-1.0 + np.sqrt(1.0 + a * a * (1 - c * c) / (c * c))) / 2.0

    G2 = 1.0 / (1.0 + lam(cv) + lam(cl))
    f0 = 0.08 * specular * (1 - metallic) + al * metallic
    F = f0 + (1 - f0) * (1 - np.sum(l * h, axis=1)) ** 5
    return F * D * G2 / (4 * cv) + (1 - metallic) * (1 - F) * al * cl / math.pi


def _hemi(n, seed, cmin=0.05):
    g = np.random.default_rng(seed)
    c = g.uniform(cmin, 1.0, n)
    p = g.uniform(-math.pi, math.pi, n)
    s = np.sqrt(1 - c * c)
    return np.stack([s * np.cos(p), s * np.sin(p), c], -1)


def _eval_pairs(v, l, lam, albedo, **kw):
    """EvaluateBSDFWithCos for rows of (v, l): one oracle call per row."""
    out = np.zeros(len(v))
    for i in range(len(v)):
        f, _ = ol.bsdf_eval(ap.MAT_GGX, -v[i], N_UP, l[i:i + 1], lam=lam, albedo=albedo, **kw)
        out[i] = f[0]
    return out


@pytest.mark.parametrize("roughness,metallic", [(0.25, 0.0), (0.5, 0.5), (0.8, 1.0), (1.0, 0.2)])
def test_ggx_eval_equals_textbook(roughness, metallic):
    lam, albedo, specular = 560.0, (0.8, 0.6, 0.3), 0.5
    al = float(ol.interact(np.array([lam], np.float32), np.array([albedo], np.float32))[0])
    v = _hemi(400, 11)
    l = _hemi(400, 12)
    v32, l32 = v.astype(np.float32), l.astype(np.float32)
    v, l = v32.astype(np.float64), l32.astype(np.float64)
    got = _eval_pairs(v32, l32, lam, albedo, specular=specular, metallic=metallic, roughness=roughness)
    ref = _textbook(v, l, roughness, metallic, specular, al)
    rel = np.abs(got / ref - 1.0)
    print(f"[ggx] roughness {roughness} metallic {metallic}: max rel {rel.max():.2e}")
    assert rel.max() < 2e-4
    # Reciprocity: f(v, l) = f(l, v), i.e. (f cos_l) / cos_l symmetric.
    back = _eval_pairs(l32, v32, lam, albedo, specular=specular, metallic=metallic, roughness=roughness)
    assert np.allclose(got / l[:, 2], back / v[:, 2], rtol=2e-5)


def test_ggx_eval_pin_catches_alpha_slip():
    lam, albedo = 560.0, (0.8, 0.6, 0.3)
    al = float(ol.interact(np.array([lam], np.float32), np.array([albedo], np.float32))[0])
    v, l = _hemi(50, 13).astype(np.float32), _hemi(50, 14).astype(np.float32)
    with ol.planted(4):
        got = _eval_pairs(v, l, lam, albedo, specular=0.5, metallic=0.5, roughness=0.5)
    ref = _textbook(v.astype(np.float64), l.astype(np.float64), 0.5, 0.5, 0.5, al)
    assert np.abs(got / ref - 1.0).max() > 0.1


def directional_albedo(cos_v, roughness, metallic, albedo=(1, 1, 1), n=200_000, seed=4, lam=560.0):
    """E(v) = integral of f cos over the hemisphere, estimated with the
    oracle's own SampleBSDF and EvaluateBSDFWithCos (f / pdf, terminated
    draws count zero); mean and standard error."""
    rd = -_view(cos_v)
    dirs, pdfs, term = ol.bsdf_sample(ap.MAT_GGX, rd, N_UP, n, lam=lam, albedo=albedo, specular=0.5,
                                      metallic=metallic, roughness=roughness, seed=seed)
    f, _ = ol.bsdf_eval(ap.MAT_GGX, rd, N_UP, dirs, lam=lam, albedo=albedo, specular=0.5, metallic=metallic,
                        roughness=roughness)
    w = np.where(term | (pdfs <= 0), 0.0, f.astype(np.float64) / np.maximum(pdfs, 1e-30))
    return float(w.mean()), float(w.std() / math.sqrt(n))


def test_ggx_energy_and_furnace():
    """The metal lobe (metallic 1) reflects at most its albedo response. The
    dielectric mix is the reference's own non-conserving model: its Lambert
    term is weighted by 1 - F(l.h) per direction (brdf.cuh:126-127), not by
    1 - E_spec, so with albedo 1 it returns up to ~1.2 at grazing view; the
    estimate must then equal the textbook form of that same model."""
    lam = 560.0
    al = float(ol.interact(np.array([lam], np.float32), np.ones((1, 3), np.float32))[0])
    for cos_v in (0.2, 0.6, 0.95):
        for roughness in (0.3, 0.7, 1.0):
            for metallic in (0.0, 1.0):
                e, se = directional_albedo(cos_v, roughness, metallic)
                # quadrature of the textbook form over the hemisphere
                c_edges, p_edges, qd, dw, shape = _grid(32, 32, 8)
                v = np.broadcast_to(_view(cos_v), qd.shape)
                q = float(_textbook(v, qd, roughness, metallic, 0.5, al).sum() * dw)
                print(f"[ggx] E(cosV {cos_v}, r {roughness}, m {metallic}) = {e:.4f} +- {se:.4f}, textbook {q:.4f}")
                if metallic == 1.0:
                    assert e <= al + 4 * se
                assert abs(e - q) < 4 * se + 2e-3 * q
    # A smooth metal of albedo response al reflects Schlick's F at the mirror angle.
    for cos_v in (0.3, 0.8):
        e, se = directional_albedo(cos_v, 0.05, 1.0, n=50_000)
        f = al + (1 - al) * (1 - cos_v) ** 5
        assert abs(e - f) < 4 * se + 2e-3, (cos_v, e, f)


def mixture_stats(cos_v, roughness, metallic, n=400_000, seed=6):
    """(chi^2 z, normalisation sigma, pairing p99) of SampleBSDF's GGX
    mixture against BSDF_PDF."""
    rd = -_view(cos_v)
    kw = dict(lam=560.0, albedo=(0.7, 0.7, 0.7), specular=0.5, metallic=metallic, roughness=roughness)
    dirs, pdfs, term = ol.bsdf_sample(ap.MAT_GGX, rd, N_UP, n, seed=seed, **kw)
    c_edges, p_edges, qd, dw, shape = _grid()
    _, qpdf = ol.bsdf_eval(ap.MAT_GGX, rd, N_UP, qd, **kw)
    exp = _expected(qpdf, shape, dw)
    live = ~term
    z, mass = _chi2_z(_histogram(dirs[live], c_edges, p_edges), exp, n)
    t = float(term.mean())
    se = math.sqrt(max(t * (1 - t), 1.0 / n) / n)
    norm_sigma = abs(mass + t - 1.0) / (se + 2e-4)
    _, back = ol.bsdf_eval(ap.MAT_GGX, rd, N_UP, dirs[live], **kw)
    pair = np.abs(back / pdfs[live] - 1.0)
    return z, norm_sigma, float(np.percentile(pair, 99))


@pytest.mark.parametrize("cos_v,roughness,metallic", [(0.3, 0.5, 0.0), (0.8, 0.7, 1.0), (0.5, 0.4, 0.5)])
def test_ggx_mixture_follows_bsdf_pdf(cos_v, roughness, metallic):
    z, norm_sigma, pair99 = mixture_stats(cos_v, roughness, metallic)
    print(f"[mix] cosV {cos_v} r {roughness} m {metallic}: chi2 z {z:.2f}, normalisation {norm_sigma:.2f}, "
          f"pairing p99 {pair99:.2e}")
    assert z < 4.0 and norm_sigma < 4.0 and pair99 < 1e-4


def test_ggx_mixture_pin_catches_dropped_lambert_term():
    with ol.planted(3):
        z, norm_sigma, pair99 = mixture_stats(0.5, 0.5, 0.0, n=100_000)
    assert pair99 > 1e-2


# ---- lights ----

def sphere_light(center=(0.2, 0.5, -0.1), radius=0.15, le=(8.0, 6.0, 4.0)):
    s = scenes.spheres([center], [radius])
    s["shading_model"] = ap.MAT_LIGHT
    s["albedo"] = 0.0
    s["emissive"] = le
    return s


def _points_around(c, r, dists, n_dir, seed):
    g = np.random.default_rng(seed)
    out = []
    for d in dists:
        u = g.normal(size=(n_dir, 3))
        u /= np.linalg.norm(u, axis=1, keepdims=True)
        out.append(np.asarray(c) + d * r * u)
    return np.concatenate(out).astype(np.float32)


def test_sphere_cap_sampler_pairs_and_covers_the_cap():
    c, r = np.array([0.2, 0.5, -0.1]), 0.15
    objs = scenes.concat([sphere_light(c, r)])
    for k, d in enumerate((1.05, 1.5, 4.0, 30.0)):
        x = np.broadcast_to(c + d * r * np.array([0.36, -0.48, 0.8]), (1, 3)).astype(np.float32)
        pts = np.repeat(x, 20_000, axis=0)
        s = ol.nee_sample(objs, pts, seed=17 + k)
        assert s["valid"].all() and (s["object"] == 0).all()
        assert np.allclose(s["pdf_sampler"], s["pdf_hit"], rtol=2e-5)
        y = s["point"].astype(np.float64) - c
        assert np.allclose(np.linalg.norm(y, axis=1), r, rtol=2e-6)
        axis = (x[0] - c) / np.linalg.norm(x[0] - c)
        cos_t = (y / r) @ axis
        lo = r / np.linalg.norm(x[0].astype(np.float64) - c)
        assert cos_t.min() >= lo - 2e-6
        # Archimedes: an area-uniform point on the cap has cos(theta) uniform on [r/d, 1].
        u = np.sort((cos_t - lo) / (1 - lo))
        ks = np.abs(u - (np.arange(len(u)) + 0.5) / len(u)).max()
        assert ks < 1.63 / math.sqrt(len(u)) + 2e-4, (d, ks)
        # Cap area from the pdf: pdf_w = dist^2 / (cosY area) (one light: select prob 1).
        to = y + c - x[0]
        dist2 = np.sum(to * to, axis=1)
        cos_y = np.sum((y / r) * (-to / np.sqrt(dist2)[:, None]), axis=1)
        area = dist2 / (cos_y * s["pdf_sampler"])
        expect = 2 * math.pi * r * r * (1 - lo)
        rel = np.abs(area / expect - 1.0)   # FP32 pdf; at the rim cosY -> 0 amplifies its rounding
        assert np.median(rel) < 1e-5 and np.percentile(rel, 99) < 1e-4, (d, np.median(rel))


def _frame(w):
    a = np.array([1.0, 0, 0]) if abs(w[0]) < 0.9 else np.array([0, 1.0, 0])
    t1 = np.cross(w, a)
    t1 /= np.linalg.norm(t1)
    return t1, np.cross(w, t1)


def _ray_sphere(x, dirs, c, r):
    oc = x - c
    b = dirs @ oc
    disc = b * b - (oc @ oc - r * r)
    t = -b - np.sqrt(np.maximum(disc, 0.0))
    y = x + t[:, None] * dirs
    return y, (y - c) / r, disc > 0


def _cap_quadrature(objs, light, x, c, r, nu=4000, nphi=8):
    """Integral of NEEPDFOmegaFromHit over the solid angle the sphere (c, r)
    subtends from x, by quadrature in FP64: cos(beta) = cos_max + u^2 (which
    removes the 1/cos singularity of the pdf at the rim), midpoints in u,
    nphi azimuths; first intersections in FP64."""
    w = c - x
    d = np.linalg.norm(w)
    w = w / d
    cos_max = math.sqrt(1 - (r / d) ** 2)
    umax = math.sqrt(1 - cos_max)
    u = (np.arange(nu) + 0.5) * (umax / nu)
    ph = (np.arange(nphi) + 0.5) * (2 * math.pi / nphi)
    U, P = np.meshgrid(u, ph, indexing="ij")
    ct = cos_max + U * U
    st = np.sqrt(np.maximum(0.0, 1 - ct * ct))
    t1, t2 = _frame(w)
    dirs = ((st * np.cos(P))[..., None] * t1 + (st * np.sin(P))[..., None] * t2 + ct[..., None] * w).reshape(-1, 3)
    y, nrm, ok = _ray_sphere(x, dirs, c, r)
    pdf = ol.nee_pdf(objs, light, np.broadcast_to(x, y.shape), y, nrm).astype(np.float64) * ok
    jac = (2 * U * (umax / nu) * (2 * math.pi / nphi)).reshape(-1)
    return float((pdf * jac).sum())


def sphere_cap_integral(d_over_r):
    c, r = np.array([0.2, 0.5, -0.1]), 0.15
    objs = scenes.concat([sphere_light(c, r)])
    x = c + d_over_r * r * np.array([0.0, -0.6, 0.8])
    return _cap_quadrature(objs, 0, x, c, r)


@pytest.mark.parametrize("d_over_r", [1.02, 1.3, 3.0, 25.0])
def test_sphere_cap_pdf_integrates_to_one(d_over_r):
    est = sphere_cap_integral(d_over_r)
    print(f"[cap] d/r {d_over_r}: integral {est:.6f}")
    assert abs(est - 1.0) < 2e-3


def test_sphere_cap_pin_catches_wrong_area():
    with ol.planted(2):
        est = sphere_cap_integral(3.0)
        c, r = np.array([0.2, 0.5, -0.1]), 0.15
        objs = scenes.concat([sphere_light(c, r)])
        x = np.repeat((c + 3.0 * r * np.array([0.0, -0.6, 0.8]))[None].astype(np.float32), 1000, axis=0)
        s = ol.nee_sample(objs, x, seed=3)
    assert abs(est - 1.0) > 0.5
    assert np.allclose(s["pdf_sampler"], s["pdf_hit"], rtol=2e-5)   # pairing alone cannot see it


def _box_integral(objs, light, x, n, seed):
    """Integral of NEEPDFOmegaFromHit over the directions from x that hit the
    light's (axis-aligned) box, in FP64: uniform directions in the cone of
    the box's bounding sphere; (estimate, standard error)."""
    ob = objs[light]
    c = ob["center"].astype(np.float64)
    e = ob["extent"].astype(np.float64)
    w = c - x
    d = np.linalg.norm(w)
    w /= d
    cos_max = math.sqrt(1 - min(1.0, float(e @ e) / (d * d)))
    g = np.random.default_rng(seed)
    ct = g.uniform(cos_max, 1.0, n)
    ph = g.uniform(0, 2 * math.pi, n)
    st = np.sqrt(1 - ct * ct)
    t1, t2 = _frame(w)
    dirs = (st * np.cos(ph))[:, None] * t1 + (st * np.sin(ph))[:, None] * t2 + ct[:, None] * w
    with np.errstate(divide="ignore", invalid="ignore"):
        t0 = (c - e - x) / dirs
        t1_ = (c + e - x) / dirs
    tn = np.nanmax(np.minimum(t0, t1_), axis=1)
    tf = np.nanmin(np.maximum(t0, t1_), axis=1)
    ok = (tn <= tf) & (tn > 0)
    y = x + tn[:, None] * dirs
    ax = np.argmax(np.minimum(t0, t1_), axis=1)
    nrm = np.zeros((n, 3))
    nrm[np.arange(n), ax] = -np.sign(dirs[np.arange(n), ax])
    pdf = np.zeros(n)
    pdf[ok] = ol.nee_pdf(objs, light, np.broadcast_to(x, (int(ok.sum()), 3)), y[ok], nrm[ok])
    vals = pdf * 2 * math.pi * (1 - cos_max)
    return float(vals.mean()), float(vals.std() / math.sqrt(n))


def test_box_and_two_light_tables_integrate_to_one():
    """A box light and a sphere light in one table: each light's pdf
    integrates to its selection probability (area x luminance, scene.cu:1056-
    1080), the box by ray casting into its bounding cone, the sphere by
    quadrature; NEESampleLight picks them with those probabilities."""
    box = scenes.cubes([[0.0, 0.9, 0.0]], [0.25, 0.05, 0.15], emissive=(15.0, 15.0, 15.0))
    sph = sphere_light((0.5, 0.2, -0.3), 0.1)
    objs = scenes.concat([box, sph])
    x = np.array([0.1, -0.4, 0.2])
    w = np.array([8 * (0.25 * 0.05 + 0.25 * 0.15 + 0.05 * 0.15) * 15.0,
                  4 * math.pi * 0.01 * (8 * 0.2126 + 6 * 0.7152 + 4 * 0.0722)])
    p_box, p_sphere = w / w.sum()
    est, se = _box_integral(objs, 0, x, 2_000_000, 1)
    cap = _cap_quadrature(objs, 1, x, objs[1]["center"].astype(np.float64), float(objs[1]["extent"].max()))
    print(f"[lights] box {est:.5f} +- {se:.5f} (select {p_box:.5f}), sphere {cap:.6f} (select {p_sphere:.6f})")
    assert abs(est - p_box) < 4 * se and se < 1e-3
    assert abs(cap - p_sphere) < 1e-4
    s = ol.nee_sample(objs, np.repeat(x[None].astype(np.float32), 20_000, axis=0), seed=9)
    assert s["valid"].all()
    assert np.allclose(s["pdf_sampler"], s["pdf_hit"], rtol=2e-5)
    frac_sphere = float((s["object"] == 1).mean())
    assert abs(frac_sphere - p_sphere) < 4 * math.sqrt(p_sphere * (1 - p_sphere) / 20_000)
