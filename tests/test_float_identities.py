"""Float32 identities the ray walks rely on, checked exhaustively near the
boundaries and on random samples (numpy float32 arithmetic rounds like the
device's v_add_f32 / v_cmp_f32).

* The BORDER validity test of a trilinear sample at q (cell-centre
  coordinates): the bricked walks compare q >= -0.5 and q < t(d), with t(d)
  the smallest float whose sum with 0.5 reaches d (validity_upper: d - 0.5,
  except the largest float below 0.5 for d = 1), where the reference-shaped
  form is q + 0.5 >= 0 and q + 0.5 < d.  The two decide the same for every
  float q and grid extent d tried; the plain q < d - 0.5 does not for d = 1.
* The brick address's clamp: floor(q) + 1 clamped to [0, d] equals
  clamp(floor(q), -1, d - 1) + 1.
"""
import numpy as np

F = np.float32


def _around(x, n=4000):
    """The n floats on each side of float32 x, and x itself."""
    up, down = [F(x)], [F(x)]
    for _ in range(n):
        up.append(np.nextafter(up[-1], F(np.inf)))
        down.append(np.nextafter(down[-1], F(-np.inf)))
    return np.array(down[::-1] + up[1:], dtype=F)


def validity_upper(d):
    """The device helper of the same name, in float32."""
    t = F(d) - F(0.5)
    below = np.array([t], dtype=F).view(np.int32)[0] - 1
    below = np.array([below], dtype=np.int32).view(F)[0]
    return below if below + F(0.5) >= F(d) else t


def _check_validity(q, d):
    half = F(0.5)
    ref_lo = (q + half) >= F(0.0)
    ref_hi = (q + half) < F(d)
    new_lo = q >= F(-0.5)
    new_hi = q < validity_upper(d)
    assert np.array_equal(ref_lo, new_lo), d
    assert np.array_equal(ref_hi, new_hi), d


def test_validity_without_adds():
    rng = np.random.default_rng(3)
    for d in (1, 2, 3, 5, 7, 16, 48, 127, 128, 160, 480, 1000, 4095, 4096, 65536, 1 << 20):
        q = np.concatenate([_around(-0.5), _around(F(d) - F(0.5)), _around(0.0), _around(-1.0),
                            _around(F(d)), rng.uniform(-3.0, d + 3.0, 200000).astype(F)])
        _check_validity(q, d)


def test_plain_upper_bound_differs_for_extent_one():
    q = F(0.49999997)
    assert not (q + F(0.5) < F(1))  # rounds (to even) up to 1
    assert q < F(1) - F(0.5)
    assert validity_upper(1) == q


def test_brick_clamp():
    rng = np.random.default_rng(4)
    for d in (1, 4, 37, 128, 4096):
        q = np.concatenate([rng.uniform(-1e3, d + 1e3, 100000).astype(F), _around(-1.0, 200), _around(F(d), 200),
                            np.array([-1e30, 1e30, -np.inf, np.inf], dtype=F)])
        fl = np.floor(q)
        new = np.clip(fl + F(1.0), F(0.0), F(d))  # med3(floor(q) + 1, 0, d)
        old = np.clip(np.clip(fl, -1e9, 1e9).astype(np.int64), -1, d - 1) + 1
        assert np.array_equal(new.astype(np.int64), old), d
