"""Pins of the oracle's ReSTIR reservoir reuse (CPU only).

The merges the reuse passes run, probed on hand-made reservoirs
(or_reservoir_probe):

- DiReservoir::combine (shared/restir/di_reservoir.wgsl:99-110): M adds up,
  the weight sum is sum p^ W M, the second input is kept with probability
  w2 / (w1 + w2), and W p^(kept) = wsum / M, so with p^ equal to the
  integrand the merged estimate is the M-weighted mean of the inputs'
  estimates; merging with an empty reservoir gives back the input's W.
- DiReservoir::combine_unbiased (di_reservoir.wgsl:112-133): Z counts only the
  inputs whose domain gives the kept sample a nonzero target, and on a
  two-domain toy with different supports the Z-normalised merge is unbiased
  where the 1/M one is not.
- The temporal caps: the previous DI reservoir's M is clamped to 20 M
  (restir_di_temporal.wgsl:159), the previous GI reservoir's to 30
  (restir_gi_temporal.wgsl:100).

Each pin fails under a mistake planted in the oracle (or_set_plant).
"""
import ctypes as C

import numpy as np
import pytest

from tests.test_oracle_rng import np_pcg

_P = C.c_void_p
EMPTY = 0xFFFFFFFF


def _probe(oracle, kind, a, b, extra, rng):
    fn = oracle.lib.or_reservoir_probe
    fn.restype = C.c_int
    fn.argtypes = [C.c_uint32, _P, _P, _P, _P, _P]
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    ex = np.ascontiguousarray(extra, np.float32)
    st = np.array([rng], np.uint32)
    out = np.zeros(8, np.float32)
    assert fn(kind, a.ctypes.data, b.ctypes.data, ex.ctypes.data, st.ctypes.data, out.ctypes.data) == 0
    return out, int(st[0])


def di(light, u, v, wsum, M, W, phat):
    r = np.array([0, u, v, wsum, M, W, phat], np.float32)
    r[0:1] = np.array([light], np.uint32).view(np.float32)
    return r


def di_light(r):
    return int(np.asarray(r[0:1], np.float32).view(np.uint32)[0])


class plant:
    """Plants one reuse mistake in the oracle for the duration of a block."""

    def __init__(self, oracle, kind):
        self.o, self.k = oracle, kind

    def __enter__(self):
        self.o.lib.or_set_plant.argtypes = [C.c_int]
        self.o.lib.or_set_plant(self.k)

    def __exit__(self, *exc):
        self.o.lib.or_set_plant(0)


def _draw(state):
    state = int(np_pcg(np.array([state], np.uint32))[0])
    return state, np.float32(state) / np.float32(4294967296.0)


# ---- combine ------------------------------------------------------------------

def _check_combine(oracle, cases):
    for a, b, rng in cases:
        out, _ = _probe(oracle, 0, a, b, [], rng)
        M1, M2 = float(a[4]), float(b[4])
        w1 = float(a[6]) * float(a[5]) * M1
        w2 = float(b[6]) * float(b[5]) * M2
        assert out[4] == np.float32(M1 + M2)
        np.testing.assert_allclose(out[3], w1 + w2, rtol=1e-6)
        kept = di_light(out)
        if w1 + w2 > 0:
            assert kept in (di_light(a), di_light(b))
            ph = a[6] if kept == di_light(a) else b[6]
            assert out[6] == ph
            # the contribution weight: W p^(kept) = wsum / M
            if ph > 1e-6:
                np.testing.assert_allclose(out[5] * out[6], (w1 + w2) / (M1 + M2), rtol=2e-6)
        else:
            assert kept == EMPTY and out[5] == 0.0


def test_combine_closed_form(oracle):
    rng = np.random.default_rng(11)
    cases = []
    for i in range(400):
        a = di(1, *rng.random(2), 0.0, float(rng.integers(1, 40)), rng.uniform(0, 3), rng.uniform(0, 2))
        b = di(2, *rng.random(2), 0.0, float(rng.integers(0, 40)), rng.uniform(0, 3), rng.uniform(0, 2))
        if i % 7 == 0:
            b[6] = 0.0  # the neighbour's sample has no target here
        cases.append((a, b, int(rng.integers(0, 2**32))))
    _check_combine(oracle, cases)


def test_combine_with_empty_reservoir_keeps_w(oracle):
    """Single-candidate case: merging with an empty reservoir (M = 0) keeps
    the input's sample, M and contribution weight."""
    rng = np.random.default_rng(12)
    for _ in range(200):
        a = di(7, 0.25, 0.5, 0.0, float(rng.integers(1, 33)), rng.uniform(0.01, 5), rng.uniform(0.01, 4))
        out, _ = _probe(oracle, 0, a, di(EMPTY, 0, 0, 0, 0.0, 0.0, 0.0), [], int(rng.integers(0, 2**32)))
        assert di_light(out) == 7 and out[4] == a[4]
        np.testing.assert_allclose(out[5], a[5], rtol=2e-6)


def test_combine_estimate_is_the_m_weighted_mean(oracle):
    """With the target equal to the integrand, the merged estimate f(y) W is
    deterministic: (M1 p1 W1 + M2 p2 W2) / (M1 + M2) whichever sample is kept."""
    rng = np.random.default_rng(13)
    for _ in range(300):
        M1, M2 = float(rng.integers(1, 30)), float(rng.integers(1, 600))
        a = di(3, 0.1, 0.2, 0.0, M1, rng.uniform(0.1, 2), rng.uniform(0.1, 2))
        b = di(4, 0.3, 0.4, 0.0, M2, rng.uniform(0.1, 2), rng.uniform(0.1, 2))
        out, _ = _probe(oracle, 0, a, b, [], int(rng.integers(0, 2**32)))
        want = (M1 * a[6] * a[5] + M2 * b[6] * b[5]) / (M1 + M2)
        np.testing.assert_allclose(out[6] * out[5], want, rtol=2e-6)


def test_combine_selection_frequency(oracle):
    """The second input is kept with probability w2 / (w1 + w2), drawn by the
    WRS rule u <= w / wsum on the reservoir's own RNG stream."""
    a = di(1, 0, 0, 0, 4.0, 1.0, 0.5)     # w1 = 2
    b = di(2, 0, 0, 0, 6.0, 1.0, 1.0)     # w2 = 6
    n, kept_b = 20000, 0
    for s in range(n):
        out, st = _probe(oracle, 0, a, b, [], s * 2654435761 % 2**32)
        kept_b += di_light(out) == 2
        # the merge consumes exactly two draws
        st0 = s * 2654435761 % 2**32
        st1, _ = _draw(st0)
        st2, u2 = _draw(st1)
        assert st == st2
        assert (di_light(out) == 2) == bool(u2 <= np.float32(6.0) / np.float32(8.0))
    p = kept_b / n
    assert abs(p - 0.75) < 4 * np.sqrt(0.75 * 0.25 / n)


def test_combine_planted_missing_inv_m_fails(oracle):
    with plant(oracle, 3):
        with pytest.raises(AssertionError):
            test_combine_estimate_is_the_m_weighted_mean(oracle)


# ---- combine_unbiased: Z ------------------------------------------------------

@pytest.mark.parametrize("vis", [(1, 1), (1, 0), (0, 1), (0, 0)])
def test_combine_unbiased_z_counts_domains_with_nonzero_target(oracle, vis):
    """Z = M1 [p1(kept) > 0] + M2 [p2(kept) > 0], W = wsum / (Z p^(kept))."""
    rng = np.random.default_rng(14 + vis[0] * 2 + vis[1])
    for _ in range(200):
        M1, M2 = float(rng.integers(1, 30)), float(rng.integers(1, 30))
        a = di(5, 0.1, 0.2, 0.0, M1, rng.uniform(0.1, 2), rng.uniform(0.1, 2))
        b = di(6, 0.3, 0.4, 0.0, M2, rng.uniform(0.1, 2), rng.uniform(0.1, 2))
        # the target of either sample at domain 1 is positive (its p^ is); at
        # domain 2 it is positive or zero per `vis` for a's / b's sample
        extra = [a[6], 0.7 if vis[0] else 0.0, b[6], 0.9 if vis[1] else 0.0]
        out, _ = _probe(oracle, 1, a, b, extra, int(rng.integers(0, 2**32)))
        wsum = float(a[6] * a[5] * M1 + b[6] * b[5] * M2)
        kept_a = di_light(out) == 5
        z = M1 + (M2 if vis[0 if kept_a else 1] else 0.0)
        assert out[4] == np.float32(M1 + M2)
        np.testing.assert_allclose(out[5] * out[6], wsum / z, rtol=2e-6)


def test_combine_unbiased_z_zero_gives_zero_w(oracle):
    a = di(5, 0.1, 0.2, 0.0, 3.0, 1.0, 1.0)
    b = di(6, 0.3, 0.4, 0.0, 4.0, 1.0, 1.0)
    out, _ = _probe(oracle, 1, a, b, [0.0, 0.0, 0.0, 0.0], 99)
    assert out[5] == 0.0 and out[4] == 7.0


def _two_domain_toy(oracle, n_trials, seed):
    """Two shading points over 8 light samples (uniform source pdf 1/8) with
    different supports: the pixel's target p1 is zero on samples 6, 7, the
    neighbour's p2 on samples 0, 1, 2.  Each point builds a RIS reservoir of M
    candidates for its own target; the pixel merges the neighbour's with
    combine_unbiased.  Returns the mean and standard error of f1(kept) W,
    whose expectation is sum_x f1(x) when the merge is unbiased (f1 = p1)."""
    p1 = np.array([1.0, 2.0, 0.5, 3.0, 1.5, 0.8, 0.0, 0.0], np.float64)
    p2 = np.array([0.0, 0.0, 0.0, 2.5, 1.0, 4.0, 2.0, 1.0], np.float64)
    rng = np.random.default_rng(seed)

    def ris(p, M):
        xs = rng.integers(0, 8, M)
        w = p[xs] * 8.0
        ws = w.sum()
        if ws <= 0:
            return EMPTY, 0.0, 0.0
        y = int(xs[rng.choice(M, p=w / ws)])
        return y, ws / (M * p[y]), p[y]

    M1, M2 = 3, 5
    est = np.empty(n_trials)
    for t in range(n_trials):
        y1, W1, _ = ris(p1, M1)
        y2, W2, _ = ris(p2, M2)
        a = di(y1, 0.0, 0.0, 0.0, float(M1), W1, p1[y1] if y1 != EMPTY else 0.0)
        # the neighbour's sample, its target re-evaluated at the pixel
        b = di(y2 if y2 != EMPTY else EMPTY, 0.5, 0.0, 0.0, float(M2), W2, p1[y2] if y2 != EMPTY else 0.0)
        extra = [p1[y1] if y1 != EMPTY else 0, p2[y1] if y1 != EMPTY else 0,
                 p1[y2] if y2 != EMPTY else 0, p2[y2] if y2 != EMPTY else 0]
        out, _ = _probe(oracle, 1, a, b, extra, int(rng.integers(0, 2**32)))
        k = di_light(out)
        est[t] = 0.0 if k == EMPTY else p1[k] * float(out[5])
    return est.mean(), est.std(ddof=1) / np.sqrt(n_trials), p1.sum()


def test_combine_unbiased_two_domain_toy(oracle):
    m, se, want = _two_domain_toy(oracle, 40000, 15)
    print(f"[reservoir] two-domain toy: mean {m:.4f} +- {se:.4f}, exact {want:.4f}")
    assert abs(m - want) < 4 * se


@pytest.mark.parametrize("kind", [1, 5])
def test_combine_unbiased_planted_z_errors_fail(oracle, kind):
    with plant(oracle, kind):
        m, se, want = _two_domain_toy(oracle, 40000, 15)
    assert abs(m - want) > 4 * se


# ---- temporal M caps ----------------------------------------------------------

def _temporal_di_cases(oracle, mcap):
    rng = np.random.default_rng(16)
    for _ in range(300):
        Mr = float(rng.integers(1, 34))
        Mp = float(rng.choice([rng.integers(1, 20), rng.integers(20, 2000)]))
        r = di(8, 0.1, 0.1, 0.0, Mr, rng.uniform(0.1, 2), rng.uniform(0.1, 2))
        pr = di(9, 0.2, 0.2, 0.0, Mp, rng.uniform(0.1, 2), rng.uniform(0.1, 2))
        ph = rng.uniform(0.1, 2)   # the previous sample's target at this pixel
        out, _ = _probe(oracle, 2, r, pr, [mcap, ph], int(rng.integers(0, 2**32)))
        Mc = min(Mp, mcap * Mr)
        assert out[4] == np.float32(Mr + Mc)
        wsum = float(r[6] * r[5] * Mr) + ph * float(pr[5]) * Mc
        np.testing.assert_allclose(out[3], wsum, rtol=2e-6)
        np.testing.assert_allclose(out[5] * out[6], wsum / (Mr + Mc), rtol=2e-6)
        if di_light(out) == 9:
            np.testing.assert_allclose(out[6], ph, rtol=1e-7)


@pytest.mark.parametrize("mcap", [20.0, 2.0])
def test_di_temporal_caps_previous_m(oracle, mcap):
    _temporal_di_cases(oracle, mcap)


def test_di_temporal_planted_missing_cap_fails(oracle):
    with plant(oracle, 2):
        with pytest.raises(AssertionError):
            _temporal_di_cases(oracle, 20.0)


def gi(x, rng_bits, M, W, wsum, phat):
    r = np.array([x[0], x[1], x[2], 0, M, W, wsum, phat], np.float32)
    r[3:4] = np.array([rng_bits], np.uint32).view(np.float32)
    return r


def _temporal_gi_cases(oracle, mcap):
    rng = np.random.default_rng(17)
    for _ in range(300):
        Mr = 1.0
        Mp = float(rng.integers(1, 400))
        r = gi([1, 2, 3], 111, Mr, rng.uniform(0.1, 2), 0.0, rng.uniform(0.1, 2))
        pr = gi([4, 5, 6], 222, Mp, rng.uniform(0.1, 2), 0.0, rng.uniform(0.1, 2))
        ph = rng.uniform(0.1, 2)
        out, _ = _probe(oracle, 3, r, pr, [mcap, ph], int(rng.integers(0, 2**32)))
        Mc = min(Mp, mcap)
        assert out[4] == np.float32(Mr + Mc)
        wsum = float(r[7] * r[5] * Mr) + ph * float(pr[5]) * Mc
        np.testing.assert_allclose(out[6], wsum, rtol=2e-6)
        np.testing.assert_allclose(out[5] * out[7], wsum / (Mr + Mc), rtol=2e-6)
        kept_prev = out[3:4].view(np.uint32)[0] == 222
        np.testing.assert_array_equal(out[:3], [4, 5, 6] if kept_prev else [1, 2, 3])


@pytest.mark.parametrize("mcap", [30.0, 3.0])
def test_gi_temporal_caps_previous_m(oracle, mcap):
    _temporal_gi_cases(oracle, mcap)


def test_gi_temporal_planted_missing_cap_fails(oracle):
    with plant(oracle, 4):
        with pytest.raises(AssertionError):
            _temporal_gi_cases(oracle, 30.0)
