"""Nearest-ommatidium table of the learned sampler (host code of the C ABI, no
GPU): nearestOmmatidia against brute force in FP64 with the same coordinates,
summation order and index tie-break."""
import numpy as np
import pytest

import compound_ray_amd as cra
from compound_ray_amd import eyerenderer as er
from compound_ray_amd import scenes

POS_SCALE = 0.01


@pytest.fixture(scope="module")
def hostlib():
    lib = cra.load_library()
    lib.setVerbosity(False)
    return lib


def brute_force(omm, k):
    o = np.asarray(omm, np.float32).astype(np.float64)
    d = o[:, 3:6]
    length = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])
    X = np.concatenate([d / length[:, None], o[:, 0:3] / POS_SCALE], axis=1)
    n = len(o)
    out = np.zeros((n, k), np.uint32)
    idx = np.arange(n)
    for i in range(n):
        d2 = np.zeros(n)
        for c in range(6):
            e = X[i, c] - X[:, c]
            d2 = d2 + e * e
        d2[i] = np.inf
        order = np.lexsort((idx, d2))[: min(k, n - 1)]
        out[i, : len(order)] = order
        out[i, len(order):] = i
    return out


def fibonacci_eye(n, radius=0.001, centre=(0.0, 0.0, 0.0), seed=0):
    rng = np.random.default_rng(seed)
    g = np.pi * (3.0 - np.sqrt(5.0))
    t = np.arange(n) + 0.5
    y = 1.0 - 2.0 * t / n
    r = np.sqrt(1.0 - y * y)
    a = np.stack([np.cos(g * t) * r, y, np.sin(g * t) * r], axis=1)
    omm = np.zeros((n, 8), np.float32)
    omm[:, 0:3] = a * radius + np.asarray(centre)
    omm[:, 3:6] = a * rng.uniform(0.5, 2.0, (n, 1))  # unnormalised axes
    omm[:, 6] = 0.0349
    return omm


@pytest.mark.parametrize("n", [300, 5000])
def test_single_eye_matches_brute_force(hostlib, n):
    omm = fibonacci_eye(n)
    got = er.nearest_ommatidia_np(hostlib, omm)
    assert np.array_equal(got, brute_force(omm, er.NEIGHBOURS))


def test_merged_eyes_stay_apart(hostlib):
    """Two eyes one metre apart with identical axes: every neighbour is an
    ommatidium of the same eye."""
    a = fibonacci_eye(400, centre=(0.0, 0.0, 0.0), seed=1)
    b = fibonacci_eye(400, centre=(1.0, 0.0, 0.0), seed=1)
    b[:, 3:6] = a[:, 3:6]
    omm = np.concatenate([a, b])
    got = er.nearest_ommatidia_np(hostlib, omm)
    assert np.array_equal(got, brute_force(omm, er.NEIGHBOURS))
    assert ((got < 400) == (np.arange(800) < 400)[:, None]).all()


def test_ties_and_small_eyes(hostlib):
    # coincident ommatidia: distances tie at 0, the lower indices win
    omm = np.tile(fibonacci_eye(1), (9, 1))
    got = er.nearest_ommatidia_np(hostlib, omm)
    assert np.array_equal(got[0], [1, 2, 3, 4, 5, 6])
    assert np.array_equal(got, brute_force(omm, er.NEIGHBOURS))
    # fewer than k others: the remaining slots repeat the ommatidium itself
    for n in (1, 2, 4):
        omm = fibonacci_eye(n)
        got = er.nearest_ommatidia_np(hostlib, omm)
        assert np.array_equal(got, brute_force(omm, er.NEIGHBOURS))
        assert (got[:, n - 1:] == np.arange(n)[:, None]).all()


def test_bee_eye(hostlib):
    """The bee-scale eye (5,500 ommatidia on an ellipsoidal patch)."""
    omm = scenes.config2_eyes()[0]
    got = er.nearest_ommatidia_np(hostlib, omm)
    assert np.array_equal(got, brute_force(omm, er.NEIGHBOURS))


def test_rejects_bad_k(hostlib):
    omm = fibonacci_eye(10)
    for k in (0, 17):
        with pytest.raises(RuntimeError):
            er.nearest_ommatidia_np(hostlib, omm, k=k)
