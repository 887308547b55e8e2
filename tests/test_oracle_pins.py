"""Closed-form and statistical pins of oracle functions that the GPU path is
compared against: the rough-plastic light-sample permutation, the pbr and
rough-plastic BSDF samplers, the per-sample clamp and the env/analytic light
selection with the env map's marginal halving.  Each expectation below comes
from the mathematics of the reference shaders, not from the oracle."""
import ctypes as C
import json
import math

import numpy as np
import pytest

import asuna_amd as A
from oracle import oracle as O

EPS = 1e-3
E_DELTA, E_AREA = 1, 2
LUM = (0.2126, 0.7152, 0.0722)  # luminance, sun_and_sky.glsl:29-31


def _mat(kind, **kw):
    m = A.GpuMaterial()
    m.ior, m.roughness = 1.5, 0.5
    for f in ("diffuseTextureId", "roughnessTextureId", "metalnessTextureId", "radianceTextureId",
              "normalTextureId", "tangentTextureId", "opacityTextureId"):
        setattr(m, f, -1)
    m.type = kind
    for k, v in kw.items():
        if isinstance(v, (tuple, list)):
            getattr(m, k)[:] = v
        else:
            setattr(m, k, v)
    return m


def _fresnel_normal(eta):
    """Unpolarised dielectric Fresnel at normal incidence."""
    return ((eta - 1.0) / (eta + 1.0)) ** 2


# ------------------------------------------------- rough-plastic permutation
def test_rough_plastic_light_terms_follow_the_permuted_call_at_normal_incidence():
    """At V = L = N the half vector is N, both Smith terms are 1/2 and
    Fresnel is its normal-incidence value, so the shader's call
    eval(..., eta, fdrInt, ax, ay, invEta2) -- bound to the signature
    (..., eta, fdrInt, invEta2, ax, ay) -- gives
      F(eta) / (4 (pi ay invEta2 + EPS)) + (1 - F(eta))^2 kd' ax / pi,
    kd' = kd / (1 - kd fdrInt), and pdf(..., eta, ax, ay, ...) bound to
    (..., ax, ay, eta, ...) gives
      p D / (4 + EPS) + (1 - p) / pi,  D = 1 / (pi eta ax + EPS),
      p = F(ay) / (F(ay) + lum(kd) (1 - F(ay)))
    (brdf_rough_plastic.rchit:82-83, 113-114, 233-241).  ax, ay and 1/eta^2
    are all different, so any other binding changes both values."""
    ax, ay, ior, fdr = 0.1, 0.6, 1.4, 0.2
    kd = np.array([0.5, 0.4, 0.3])
    inv_eta2 = 1.0 / ior ** 2
    assert len({ax, ay, round(inv_eta2, 6)}) == 3
    m = _mat(5, diffuse=tuple(kd), anisoAlpha=(ax, ay), ior=ior, radiance=(fdr, 0.0, 0.0))
    n = O.f3((0.0, 0.0, 1.0))
    bw, pdf = (C.c_float * 3)(), C.c_float()
    O.lib().orc_rough_plastic_light_terms(C.byref(m), n, n, E_AREA, bw, C.byref(pdf))
    F = _fresnel_normal(ior)
    kdp = kd / (1.0 - kd * fdr)
    want_bw = F / (4.0 * (math.pi * ay * inv_eta2 + EPS)) + (1.0 - F) ** 2 * kdp * ax / math.pi
    np.testing.assert_allclose(np.array(bw), want_bw, rtol=3e-6)
    Fa = _fresnel_normal(ay)
    p = Fa / (Fa + float(np.dot(LUM, kd)) * (1.0 - Fa))
    D = 1.0 / (math.pi * ior * ax + EPS)
    want_pdf = p * D / (4.0 + EPS) + (1.0 - p) / math.pi
    assert abs(pdf.value - want_pdf) / want_pdf < 3e-6, (pdf.value, want_pdf)
    # the signature-order call would give clearly different values
    plain_bw = F / (4.0 * (math.pi * ax * ay + EPS)) + (1.0 - F) ** 2 * kdp * inv_eta2 / math.pi
    assert np.all(np.abs(plain_bw - want_bw) / want_bw > 0.05)
    p0 = F / (F + float(np.dot(LUM, kd)) * (1.0 - F))
    plain_pdf = p0 / ((math.pi * ax * ay + EPS) * (4.0 + EPS)) + (1.0 - p0) / math.pi
    assert abs(plain_pdf - want_pdf) / want_pdf > 0.05
    # a delta light's flags make the pdf zero (MIS weight 1)
    O.lib().orc_rough_plastic_light_terms(C.byref(m), n, n, E_DELTA, bw, C.byref(pdf))
    assert pdf.value == 0.0


# ------------------------------------------------------------------ samplers
def _samples(m, Vdir, n, seed=777):
    L = O.lib()
    s = C.c_uint32(seed)
    w, d, pdf, fl = (C.c_float * 3)(), (C.c_float * 3)(), C.c_float(), C.c_uint32()
    V = O.f3(Vdir)
    W, D, P, F = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n), np.zeros(n, np.uint32)
    for i in range(n):
        L.orc_bsdf_sample(C.byref(m), V, C.byref(s), w, d, C.byref(pdf), C.byref(fl))
        W[i], D[i], P[i], F[i] = tuple(w), tuple(d), pdf.value, fl.value
    return W, D, P, F


def _histogram_vs_pdf(m, Vdir, D, P, nt=8, npf=8):
    """Sampled directions (pdf > 0; the shaders reject the rest) binned in
    (cos theta, phi) against the pdf integrated over each bin."""
    L = O.lib()
    V = O.f3(Vdir)
    ok = P > 0
    n = len(P)
    ct = np.clip(D[ok, 2], 0, 1)
    ph = np.mod(np.arctan2(D[ok, 1], D[ok, 0]), 2 * np.pi)
    obs, _, _ = np.histogram2d(ct, ph, bins=[nt, npf], range=[[0, 1], [0, 2 * np.pi]])
    sub = 20
    exp = np.zeros((nt, npf))
    for i in range(nt):
        cs = (i + (np.arange(sub) + 0.5) / sub) / nt
        for j in range(npf):
            ps = (j + (np.arange(sub) + 0.5) / sub) / npf * 2 * np.pi
            acc = 0.0
            for cc in cs:
                st = math.sqrt(max(0.0, 1 - cc * cc))
                for p in ps:
                    acc += L.orc_bsdf_pdf(C.byref(m), O.f3((st * math.cos(p), st * math.sin(p), cc)), V, E_AREA)
            exp[i, j] = acc / (nt * sub) * (2 * np.pi / (npf * sub)) * n
    big = exp > 50
    z = (obs[big] - exp[big]) / np.sqrt(exp[big])
    assert np.all(np.abs(z) < 5.5), (obs, exp)
    assert abs(obs.sum() - exp.sum()) < 5 * math.sqrt(n) + 0.01 * n


def _hemisphere(fn, nth=160, nph=160):
    """Midpoint quadrature of fn(direction) over the upper hemisphere."""
    acc = 0.0
    for i in range(nth):
        th = (i + 0.5) / nth * (np.pi / 2)
        st, ct = math.sin(th), math.cos(th)
        for j in range(nph):
            ph = (j + 0.5) / nph * 2 * np.pi
            acc += fn((st * math.cos(ph), st * math.sin(ph), ct)) * st
    return acc * (np.pi / 2 / nth) * (2 * np.pi / nph)


def _eval(m, Vdir):
    L = O.lib()
    V = O.f3(Vdir)
    out = (C.c_float * 3)()

    def f(l):
        L.orc_bsdf_eval(C.byref(m), O.f3(l), V, E_AREA, out)
        return np.array(out, float)
    return f


@pytest.mark.parametrize("rough,metal", [(0.5, 0.0), (0.8, 1.0)])
def test_pbr_sampler_draws_its_pdf_and_weights_carry_the_extra_cosine(rough, metal):
    """brdf_pbr_metalness_roughness.rchit:107-130: a 0.2 / 0.8 choice between
    a cosine lobe and GGX reflection; the sampled directions follow pdf()
    (the lobe mixture), while the returned pdf is that of the chosen lobe
    alone and the weight is eval (which already holds N.L) times |cos| once
    more.  So E[weight / pdf] = 0.2 int eval cos + 0.8 int eval cos over the
    directions each lobe reaches = int eval(L) cos(L) dL, not int eval."""
    m = _mat(3, diffuse=(0.8, 0.6, 0.4), roughness=rough, metalness=metal, ior=1.5)
    Vdir = (0.6, 0.0, 0.8)
    W, D, P, F = _samples(m, Vdir, 60000)
    _histogram_vs_pdf(m, Vdir, D, P)
    ev = _eval(m, Vdir)
    ok = P > 0
    for i in np.flatnonzero(ok)[:200]:  # weight = eval(d) |d.z| sample by sample
        np.testing.assert_allclose(W[i], ev(D[i]) * abs(D[i, 2]), rtol=2e-5, atol=1e-9)
    est = (W[ok] / P[ok][:, None]).sum(0) / len(P)
    se = (W[ok] / P[ok][:, None]).std(0) / math.sqrt(len(P))
    want = _hemisphere(lambda l: ev(l) * l[2])
    plain = _hemisphere(lambda l: ev(l))
    assert np.all(np.abs(est - want) < 4 * se + 0.01 * want), (est, want, se)
    assert np.all(np.abs(plain - want) > 0.1 * want)  # the extra cosine is visible


@pytest.mark.parametrize("ax,ay", [(0.2, 0.2), (0.15, 0.5)])
def test_rough_plastic_sampler_draws_its_pdf_and_weights_integrate_eval(ax, ay):
    """brdf_rough_plastic.rchit:138-189 (called with its arguments in
    signature order): the specular lobe is chosen with p = Fo / (Fo +
    lum(kd)(1 - Fo)), the weight of each lobe is that lobe's part of eval
    and its pdf includes the lobe probability, so E[weight / pdf] =
    int eval(L) dL over the upper hemisphere, and the directions follow
    pdf()."""
    m = _mat(5, diffuse=(0.6, 0.5, 0.3), anisoAlpha=(ax, ay), ior=1.6, radiance=(0.3, 0.0, 0.0))
    Vdir = (0.5, 0.3, math.sqrt(1 - 0.34))
    W, D, P, F = _samples(m, Vdir, 60000, seed=4242)
    _histogram_vs_pdf(m, Vdir, D, P)
    ev = _eval(m, Vdir)
    ok = P > 0
    r = W[ok] / P[ok][:, None]
    est = r.sum(0) / len(P)
    se = r.std(0) / math.sqrt(len(P))
    want = _hemisphere(ev)
    assert np.all(np.abs(est - want) < 4 * se + 0.01 * want), (est, want, se)
    # both lobes were drawn
    assert (F[ok] == 1).sum() > 1000 and (F[ok] == 4).sum() > 1000


# ---------------------------------------------------------- end to end scenes
def _plane_scene(path, lights, env=None, albedo=0.5, depth=2):
    path.mkdir(parents=True, exist_ok=True)
    (path / "plane.obj").write_text("v -50 0 50\nv 50 0 50\nv 50 0 -50\nv -50 0 -50\nvn 0 1 0\n"
                                    "f 1//1 2//1 3//1\nf 1//1 3//1 4//1\n")
    js = {"state": {"path_tracing": {"spp": 1, "max_path_depth": depth}},
          "camera": {"type": "perspective", "film": {"resolution": [64, 48]}, "fov": 50},
          "materials": [{"name": "m", "type": "brdf_lambertian", "diffuse_reflectance": [albedo] * 3}],
          "lights": lights, "meshes": [{"name": "plane", "path": "plane.obj"}],
          "instances": [{"mesh": "plane", "material": "m"}],
          "shots": [{"type": "lookat", "eye": [0, 2.2, 3.0], "lookat": [0, 0, 0.6], "up": [0, 1, 0]}]}
    if env is not None:
        A.write_image(str(path / "env.exr"), env)
        js["envmap"] = {"path": "env.exr"}
    (path / "scene.json").write_text(json.dumps(js))
    return str(path / "scene.json")


def clamp_scene(path):
    """A lambertian plane under a rect light of radiance (1e6, 0, 2e6) facing
    down: every path's red and blue sums exceed 10 by far, green is 0."""
    return _plane_scene(path, [{"type": "rect", "radiance": [1e6, 0.0, 2e6], "position": [-0.5, 1.0, -0.5],
                                "v1": [0.5, 1.0, -0.5], "v2": [-0.5, 1.0, 0.5]}])


CLAMP_CROP = (24, 18, 40, 30)


def test_per_sample_clamp_at_ten(tmp_path):
    """raytrace.projective.rgen:147 clamps each sample's radiance to [0, 10]
    per channel before the weighted mean: every pixel of the crop is then
    exactly 10 in red and blue (the unclamped mean would be ~1e5) and 0 in
    green."""
    s = A.HostScene(clamp_scene(tmp_path / "clamp"))
    cam, pc = s.shot(0)
    img, _ = O.OracleScene(s.desc).render(cam, pc, 64, 48, 0, 8, crop=CLAMP_CROP)
    x0, y0, x1, y1 = CLAMP_CROP
    c = img[0, y0:y1, x0:x1, :3]
    np.testing.assert_allclose(c[..., 0], 10.0, rtol=1e-6)
    np.testing.assert_allclose(c[..., 2], 10.0, rtol=1e-6)
    assert np.all(c[..., 1] == 0.0)


def env_selection_scene(path, W=64, H=32):
    """A lambertian plane (albedo 0.5) lit by a constant env map of radiance
    1 and a distant light of radiance 2 along (0.3, 1, 0.2)."""
    env = np.ones((H, W, 4), np.float32)
    return _plane_scene(path, [{"type": "distant", "radiance": [2.0, 2.0, 2.0], "direction": [0.3, 1.0, 0.2]}],
                        env=env)


def env_selection_expectation(W=64, H=32, albedo=0.5, c=1.0, Ld=2.0, ld=(0.3, 1.0, 0.2)):
    """The estimator's mean on the plane, from the mathematics of
    rchit_layouts.glsl:118-167, sample_light.glsl:84-129,
    brdf_lambertian.rchit:70-151 and raytrace.default.rmiss:41-51 for a
    constant env of radiance c:
     * env and distant light are selected with probability 1/2 each and the
       sample is divided by it;
     * the env map's marginal lookup at u = 0 blends column 0 with the zero
       column W-1, so the sampled v is halved: theta is uniform on
       [0, theta_max], theta_max = (pi/2)(H-1)/H (the REPEAT wrap at the
       ends of the inverse-CDF table spreads its share uniformly over the
       same range) and phi uniform on [0, 2 pi (W-1)/W], i.e. a true
       density 1 / (theta_max phi_max sin theta), while the reported pdf is
       (1/W)(1/2H) W H / (2 pi^2 sin theta) = 1 / (4 pi^2 sin theta);
     * MIS with the power heuristic: the light sample weighs
       p_env^2 / (p_env^2 + p_bsdf^2) with p_bsdf = cos/pi, and a BSDF ray
       that escapes weighs p_bsdf^2 / (p_bsdf^2 + p_env^2), its throughput
       being (albedo cos/pi) / (cos/pi + EPS);
     * the distant light is a delta light (pdf 1, MIS weight 1)."""
    th_max = (math.pi / 2) * (H - 1) / H
    n = 20000
    th = (np.arange(n) + 0.5) / n
    # light sample part: (1/th_max) int_0^th_max w_nee (a/pi) cos c / (p_rep + EPS) dtheta
    t = th * th_max
    p_rep = 1.0 / (4 * math.pi ** 2 * np.sin(t))
    p_b = np.cos(t) / math.pi
    w_nee = p_rep ** 2 / (p_rep ** 2 + p_b ** 2)
    nee = np.mean(w_nee * albedo / math.pi * np.cos(t) * c / (p_rep + EPS))
    # BSDF part over the whole upper hemisphere
    t = th * (math.pi / 2)
    p_rep = 1.0 / (4 * math.pi ** 2 * np.sin(t))
    p_b = np.cos(t) / math.pi
    w_b = p_b ** 2 / (p_b ** 2 + p_rep ** 2)
    thr = (albedo * np.cos(t) / math.pi) / (np.cos(t) / math.pi + EPS)
    bsdf = 2 * math.pi * np.mean(p_b * thr * c * w_b * np.sin(t)) * (math.pi / 2)
    lv = np.array(ld) / np.linalg.norm(ld)
    dist = albedo / math.pi * lv[1] * Ld / (1 + EPS)
    return nee + bsdf + dist, nee, bsdf, dist


def test_env_and_analytic_light_selection_with_marginal_halving(tmp_path):
    """The FP64 oracle's plane radiance equals the mean derived above
    within 1 %; dropping the halving (true and reported pdfs equal), the
    1/2 selection or the env pdf in the miss MIS would each move it by far
    more than that."""
    W, H = 64, 32
    s = A.HostScene(env_selection_scene(tmp_path / "envsel", W, H))
    cam, pc = s.shot(0)
    assert pc.hasEnvMap == 1 and pc.numLights == 1
    crop = (24, 18, 40, 30)
    img, _ = O.OracleScene(s.desc).render(cam, pc, 64, 48, 0, 1024, crop=crop, precision=1)
    x0, y0, x1, y1 = crop
    got = img[0, y0:y1, x0:x1, :3].mean()
    want, nee, bsdf, dist = env_selection_expectation(W, H)
    assert abs(got - want) / want < 0.01, (got, want, nee, bsdf, dist)
    # an estimator without the halving would be unbiased for the env part:
    # albedo * c plus the distant light
    assert abs((0.5 + dist) - want) / want > 0.1
