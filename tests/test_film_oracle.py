"""The oracle's Film (the image denoise_film filters) pinned without the
reference's code: pbrt's RGB<->XYZ matrices against the published sRGB/D65
white point and against each other, Film::UpdateImage's normalisation,
clamp, empty-pixel and scale rules on closed-form inputs, and the render
loop's tile lifetime (src/statistics/statpath.cpp:144,176,388): the tiles'
sums run over the whole render, so a frame fed in several calls is the
frame fed in one."""
import numpy as np

import oracle


def _samples(rgb_per_sample, H=1, W=1):
    spp = len(rgb_per_sample)
    s = np.zeros((spp, 5, H, W), np.float32)
    for j, rgb in enumerate(rgb_per_sample):
        for c in range(3):
            s[j, c] = rgb[c]
    return s


def test_rgb_to_xyz_white_point():
    # sRGB (1, 1, 1) is the D65 white: X = 0.950456, Y = 1, Z = 1.088754
    xyzw = oracle.film_xyz(oracle.film_add(_samples([(1.0, 1.0, 1.0)])))
    np.testing.assert_allclose(xyzw[:3, 0, 0], [0.950456, 1.0, 1.088754], rtol=2e-6)
    assert xyzw[3, 0, 0] == 1.0


def test_round_trip_is_the_identity():
    rng = np.random.default_rng(1)
    rgb = rng.uniform(0.0, 50.0, (64, 3)).astype(np.float32)
    s = np.zeros((1, 5, 8, 8), np.float32)
    s[0, :3] = rgb.T.reshape(3, 8, 8)
    film = oracle.film_image(oracle.film_add(s))
    # pbrt's 6-digit matrices are inverse to within 2e-6 per entry, so every
    # channel comes back within a few 1e-6 of the pixel's largest channel
    err = np.abs(film - s[0, :3]) / s[0, :3].max(axis=0)
    assert err.max() < 8e-6, err.max()


def test_mean_clamp_empty_and_scale():
    # mean of four samples; a negative channel mean is clamped to 0
    s = _samples([(1.0, -2.0, 3.0), (3.0, -2.0, 5.0), (2.0, 0.5, 4.0), (2.0, 0.5, 4.0)])
    film = oracle.film_image(oracle.film_add(s), scale=2.0)
    exact = np.array([2.0, -0.75, 4.0])
    assert film[1, 0, 0] == 0.0
    np.testing.assert_allclose(film[[0, 2], 0, 0], 2.0 * exact[[0, 2]], rtol=1e-5)
    # a pixel without samples keeps weight 0 and resolves to 0
    s2 = np.ones((3, 5, 1, 2), np.float32)
    xyzw = oracle.film_add(s2, counts=np.array([[3, 0]], np.int32))
    assert xyzw[3, 0, 1] == 0.0 and np.all(oracle.film_image(xyzw)[:, 0, 1] == 0.0)


def test_sanitised_samples_count_as_black():
    # statpath.cpp:313-332: a NaN sample becomes black but is still a sample
    s = _samples([(np.nan, 1.0, 1.0), (3.0, 3.0, 3.0)])
    film = oracle.film_image(oracle.film_add(s, sanitize=True))
    np.testing.assert_allclose(film[:, 0, 0], [1.5, 1.5, 1.5], rtol=3e-6)


def test_calls_accumulate():
    rng = np.random.default_rng(2)
    s = rng.uniform(0, 1, (6, 5, 3, 4)).astype(np.float32)
    a = oracle.film_add(s[3:], oracle.film_add(s[:3]))
    assert np.all(a[3] == 6.0)
    np.testing.assert_allclose(oracle.film_image(a), s[:, :3].mean(axis=0), rtol=5e-6)


def test_frame_in_calls_is_the_frame_in_one_call():
    """16 + 16 + 16 samples in three calls give the 48-sample frame bit for
    bit: the sums continue in sample order and XYZ is taken once from the
    running RGB sum.  Adding each call's own XYZ instead (a Film whose tiles
    restart every call) rounds differently on most pixels of this frame."""
    rng = np.random.default_rng(7)
    s = (rng.lognormal(-0.3, 0.8, (48, 5, 32, 32)) * rng.uniform(0.01, 3, (1, 5, 32, 32))).astype(np.float32)
    one = oracle.film_add(s)
    three = oracle.film_add(s[32:], oracle.film_add(s[16:32], oracle.film_add(s[:16])))
    assert np.array_equal(one.view(np.uint32), three.view(np.uint32))
    img1, img3 = oracle.film_image(one), oracle.film_image(three)
    assert np.array_equal(img1.view(np.uint32), img3.view(np.uint32))
    # the per-call-XYZ model this replaces: XYZ of each call's sum, added up
    parts = [oracle.film_xyz(oracle.film_add(s[k:k + 16])) for k in (0, 16, 32)]
    xyz_sum = parts[0].copy()
    for p in parts[1:]:
        xyz_sum = (xyz_sum + p).astype(np.float32)
    other = oracle.film_resolve(xyz_sum)
    assert np.mean(other != img1) > 0.3
