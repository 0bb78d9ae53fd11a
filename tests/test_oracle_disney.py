"""Pins the oracle's Disney BSDF (shared/material/disney_bsdf.wgsl restated)
with invariants and special cases that a dropped term or a swapped operand
would break: sample/eval agreement, pdf normalisation, energy bound, the
Burley-diffuse and GGX-conductor reductions (closed forms in numpy)."""
import ctypes as C

import numpy as np
import pytest

from tests.scenes_util import material

D3 = C.c_double * 3


def _fns(oracle):
    ev = oracle.lib.or_disney_eval
    ev.restype = C.c_double
    ev.argtypes = [C.c_void_p, D3, D3, D3, C.c_int, D3, C.c_void_p]
    sm = oracle.lib.or_disney_sample
    sm.restype = C.c_double
    sm.argtypes = [C.c_void_p, D3, D3, C.c_int, C.c_double, C.c_double, C.c_double, D3, D3, C.c_void_p]
    return ev, sm


def evaluate(oracle, m, n, wo, wi, back=False, ncc=None):
    ev, _ = _fns(oracle)
    f = D3()
    c = None if ncc is None else D3(*ncc)
    pdf = ev(C.cast(C.pointer(m), C.c_void_p), D3(*n), D3(*wo), D3(*wi), int(back), f,
             None if c is None else C.cast(c, C.c_void_p))
    return np.array(f[:]), pdf


def sample(oracle, m, n, wo, r, back=False, ncc=None):
    _, sm = _fns(oracle)
    wi, f = D3(), D3()
    c = None if ncc is None else D3(*ncc)
    pdf = sm(C.cast(C.pointer(m), C.c_void_p), D3(*n), D3(*wo), int(back), r[0], r[1], r[2], wi, f,
             None if c is None else C.cast(c, C.c_void_p))
    return np.array(wi[:]), np.array(f[:]), pdf


def rand_dir(rng, n=None):
    v = rng.normal(size=3)
    v /= np.linalg.norm(v)
    if n is not None and v @ n < 0:
        v = -v
    return v


MATS = {
    "diffuse": dict(color=(0.8, 0.5, 0.3), roughness=1.0, specular=0.0),
    "plastic": dict(color=(0.2, 0.6, 0.9), roughness=0.3, specular=0.5, sheen=0.5, clearcoat=0.7,
                    clearcoat_roughness=0.2),
    "metal": dict(color=(0.9, 0.7, 0.4), metallic=1.0, roughness=0.25),
    "mixed": dict(color=(0.5, 0.5, 0.5), metallic=0.4, roughness=0.6, specular=0.8, sheen=0.3,
                  clearcoat=0.2),
}


@pytest.mark.parametrize("name", list(MATS))
def test_sample_matches_eval(oracle, name):
    m = material(**MATS[name])
    rng = np.random.default_rng(1)
    n = np.array([0.0, 0.0, 1.0])
    checked = 0
    for _ in range(300):
        wo = rand_dir(rng, n)
        wi, f, pdf = sample(oracle, m, n, wo, rng.random(3))
        if pdf <= 0:
            continue
        fe, pe = evaluate(oracle, m, n, wo, wi)
        assert abs(pdf - pe) <= 1e-6 * max(1.0, pe), (pdf, pe)
        np.testing.assert_allclose(f, fe, rtol=1e-6, atol=1e-9)
        checked += 1
    assert checked > 200


def test_glass_sample_matches_eval(oracle):
    m = material(color=(1, 1, 1), roughness=0.3, transmission=1.0, eta=1 / 1.5)
    rng = np.random.default_rng(2)
    n = np.array([0.0, 0.0, 1.0])
    checked = 0
    for back in (False, True):
        for _ in range(300):
            wo = rand_dir(rng, n)
            wi, f, pdf = sample(oracle, m, n, wo, rng.random(3), back)
            if pdf <= 1e-3:
                continue
            fe, pe = evaluate(oracle, m, n, wo, wi, back)
            assert abs(pdf - pe) <= 1e-5 * max(1.0, pe), (back, pdf, pe)
            np.testing.assert_allclose(f, fe, rtol=1e-5, atol=1e-9)
            checked += 1
    assert checked > 300


@pytest.mark.parametrize("name", ["diffuse", "plastic", "mixed"])
def test_pdf_normalised_over_hemisphere(oracle, name):
    m = material(**MATS[name])
    rng = np.random.default_rng(3)
    n = np.array([0.0, 0.0, 1.0])
    wo = np.array([0.3, -0.2, 0.93])
    wo /= np.linalg.norm(wo)
    N = 40000
    tot = 0.0
    for _ in range(N):
        wi = rand_dir(rng, n)
        tot += evaluate(oracle, m, n, wo, wi)[1]
    est = tot / N * 2 * np.pi
    # diffuse and sheen lobes integrate to 1; microfacet lobes lose what
    # reflects below the horizon
    assert 0.9 < est < 1.03, est
    if name == "diffuse":
        assert abs(est - 1) < 0.02


@pytest.mark.parametrize("metallic", [0.0, 1.0])
def test_white_furnace_energy_bound(oracle, metallic):
    m = material(color=(1, 1, 1), metallic=metallic, roughness=0.5, specular=0.5)
    rng = np.random.default_rng(4)
    n = np.array([0.0, 0.0, 1.0])
    wo = np.array([0.0, 0.6, 0.8])
    vals = []
    for _ in range(20000):
        wi, f, pdf = sample(oracle, m, n, wo, rng.random(3))
        vals.append(f[1] * abs(wi @ n) / pdf if pdf > 0 else 0.0)
    vals = np.array(vals)
    mean, se = vals.mean(), vals.std() / np.sqrt(len(vals))
    assert mean < 1.0 + 3 * se + 0.02, (mean, se)
    assert mean > 0.5


def test_burley_diffuse_closed_form(oracle):
    # metallic = specular = sheen = clearcoat = transmission = 0, roughness 1:
    # only the diffuse lobe, f = c/pi * (1 + (fd90-1)(1-cos_i)^5)(1 + (fd90-1)(1-cos_o)^5)
    c = np.array([0.7, 0.4, 0.2], np.float32).astype(np.float64)  # stored as f32
    m = material(color=tuple(c), roughness=1.0, specular=0.0)
    rng = np.random.default_rng(5)
    n = np.array([0.0, 0.0, 1.0])
    for _ in range(200):
        wo, wi = rand_dir(rng, n), rand_dir(rng, n)
        h = (wo + wi) / np.linalg.norm(wo + wi)
        cd = wi @ h
        fd90 = 0.5 + 2 * cd * cd * 1.0
        fl, fv = (1 - wi @ n) ** 5, (1 - wo @ n) ** 5
        expect = c / np.pi * (1 + (fd90 - 1) * fl) * (1 + (fd90 - 1) * fv)
        f, pdf = evaluate(oracle, m, n, wo, wi)
        np.testing.assert_allclose(f, expect, rtol=1e-9)
        assert abs(pdf - (wi @ n) / np.pi) < 1e-12


def test_ggx_conductor_closed_form(oracle):
    # metallic = 1, no clearcoat: Schlick-Fresnel GGX conductor with the
    # height-correlated Smith G (Walter 2007 D, Heitz 2014 Lambda).
    c = np.array([0.9, 0.6, 0.3], np.float32).astype(np.float64)  # stored as f32
    r = float(np.float32(0.4))
    m = material(color=tuple(c), metallic=1.0, roughness=r)
    a = max(0.001, r * r)
    rng = np.random.default_rng(6)
    n = np.array([0.0, 0.0, 1.0])

    def lam(v):
        t2 = (1 - v[2] ** 2) / v[2] ** 2
        return (-1 + np.sqrt(1 + a * a * t2)) / 2

    for _ in range(200):
        wo, wi = rand_dir(rng, n), rand_dir(rng, n)
        h = (wo + wi) / np.linalg.norm(wo + wi)
        D = 1 / (np.pi * a * a * h[2] ** 4 * (1 + (1 - h[2] ** 2) / h[2] ** 2 / (a * a)) ** 2)
        G = 1 / (1 + lam(wo) + lam(wi))
        F = c + (1 - c) * (1 - abs(wo @ h)) ** 5
        expect = F * D * G / (4 * wo[2] * wi[2])
        f, _ = evaluate(oracle, m, n, wo, wi)
        np.testing.assert_allclose(f, expect, rtol=1e-8)
