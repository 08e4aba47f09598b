"""The oracle's sampler against the reference's own CPU tests
(crates/vork-render/src/path_sampler.rs:81-261) and its known-answer values.

Pinned by: the published Joe-Kuo direction numbers and the first eight
points (path_sampler.rs:81-97); the net, decorrelation, convergence,
joint-uniformity and light-pick independence properties (path_sampler.rs:
99-261).  The hash constants (vork_sampling.slang:3-15, path_sampler.slang:
65-93) are pinned only by these properties: no reference executable output
of the scrambled bits exists outside a GPU run of the reference.
"""
import numpy as np

import oracle

U32 = np.uint32


def unit(bits):
    return oracle.sample_unit(bits)


def test_sobol_table_matches_published_direction_numbers():
    d = np.zeros(96, np.uint32)
    oracle.lib().vo_sobol_directions(oracle.ptr(d))
    d = d.reshape(3, 32)
    published = [0x80000000, 0xC0000000, 0x20000000, 0x50000000, 0xF8000000, 0x74000000, 0xA2000000, 0x93000000]
    assert d[2][:8].tolist() == published
    expected = [[0, 0, 0, 0], [4, 4, 4, 4], [2, 6, 6, 6], [6, 2, 2, 2], [1, 5, 3, 1], [5, 1, 7, 5], [3, 3, 5, 7],
                [7, 7, 1, 3]]
    for i, row in enumerate(expected):
        o = np.zeros(4, np.uint32)
        oracle.lib().vo_sobol(i, oracle.ptr(o))
        assert (o >> U32(29)).tolist() == row, i


def test_every_power_of_two_prefix_is_a_net():
    pairs = [(0, 1, 0), (0, 2, 1), (0, 3, 2), (1, 2, 1), (1, 3, 2), (2, 3, 1)]
    cases = [((0, 0), 0, range(0, 3)), ((913, 77), 0x5EED, range(5, 7)), ((0xFFFFFFFF, 3), 1, range(40, 41))]
    for (px, py), stream, groups in cases:
        for g in groups:
            for count in [1, 2, 4, 16, 64, 256, 1024]:
                pts = oracle.sample_group_bits(px, py, count, stream, g, n=count).astype(np.uint64)
                k = int(np.log2(count))
                for c in range(4):
                    cells = (pts[:, c] >> np.uint64(32 - k)) if k else np.zeros(count, np.uint64)
                    assert np.unique(cells).size == count, (px, py, g, count, c)
                for a, b, t in pairs:
                    if k < t:
                        continue
                    for split in range(0, k - t + 1):
                        ka, kb = split, k - t - split
                        ca = (pts[:, a] >> np.uint64(32 - ka)) if ka else np.zeros(count, np.uint64)
                        cb = (pts[:, b] >> np.uint64(32 - kb)) if kb else np.zeros(count, np.uint64)
                        hits = np.bincount(((ca << np.uint64(kb)) | cb).astype(np.int64), minlength=1 << (ka + kb))
                        assert np.all(hits == (1 << t)), (px, py, g, count, a, b, split)


def _group(px, py, count, stream, g):
    return unit(oracle.sample_group_bits(px, py, count, stream, g, n=count))


def test_groups_and_pixels_are_decorrelated():
    count, pixels = 64, 256
    series = []
    for px, py in [(10, 10), (11, 10), (10, 11)]:
        for g in [0, 1, 2, 5, 400, 4001]:
            v = _group(px, py, count, 7, g)
            series += [tuple(v[:, c].tolist()) for c in range(4)]
    assert len(set(series)) == len(series), "two series repeat"
    bound = 5.0 / np.sqrt(pixels * count)
    for ga, gb in [(0, 1), (1, 2), (1, 5), (2, 6), (17, 400)]:
        xs, ys = [], []
        for p in range(pixels):
            xs.append(_group(p, 3 * p + 1, count, 7, ga) - 0.5)
            ys.append(_group(p, 3 * p + 1, count, 7, gb) - 0.5)
        x, y = np.concatenate(xs).astype(np.float64), np.concatenate(ys).astype(np.float64)
        r = (x[:, :, None] * y[:, None, :]).mean(0) * 12.0
        assert np.abs(r).max() < bound, (ga, gb, r)


def test_smooth_integrand_converges_faster_than_monte_carlo():
    f = lambda x, y: np.sin(3 * x) * np.cos(2 * y + x) + x * y
    n = 2048
    g = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(g.astype(np.float32), g.astype(np.float32))
    exact = f(X.astype(np.float64), Y.astype(np.float64)).mean()

    def rms(count, grp):
        sq = 0.0
        for p in range(512):
            u = _group(p % 32, p // 32, count, 11, grp).astype(np.float64)
            sq += (f(u[:, 0], u[:, 1]).mean() - exact) ** 2
        return np.sqrt(sq / 512)

    for grp in [0, 3, 77]:
        e16, e256 = rms(16, grp), rms(256, grp)
        assert e16 / e256 > 16.0, (grp, e16, e256)


def _hash(x):
    return np.asarray([oracle.lib().vo_pixel_hash(int(v)) for v in np.atleast_1d(x)], np.uint32)


def test_every_coordinate_pair_is_jointly_uniform_over_seeds():
    draws = 256 * 200
    k = np.arange(draws, dtype=np.uint32)
    for count in [1, 2, 4, 16, 64]:
        with np.errstate(over="ignore"):
            h = _hash(k * U32(0x9E3779B9) ^ U32(count))
        index = _hash(h) % U32(count)
        g = _hash(h ^ U32(0x51ED)) % U32(50)
        u = unit(oracle.sample_group_bits_many(h & U32(4095), h >> U32(12), index, count, 0x5EED, g))
        nx = unit(oracle.sample_group_bits_many(h & U32(4095), h >> U32(12), index, count, 0x5EED, g + U32(1)))
        c = np.minimum((u * 16).astype(np.int64), 15)
        n = np.minimum((nx * 16).astype(np.int64), 15)
        tables = [c[:, a] * 16 + c[:, b] for a, b in [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]]
        tables.append(((c[:, 0] // 4) * 64 + (c[:, 1] // 4) * 16 + (c[:, 2] // 4) * 4 + c[:, 3] // 4))
        tables += [c[:, 0] * 16 + n[:, 0], c[:, 2] * 16 + n[:, 2]]
        expected = draws / 256
        for t in tables:
            obs = np.bincount(t, minlength=256)
            chi2 = ((obs - expected) ** 2 / expected).sum()
            assert chi2 < 255 + 6 * 22.6, (count, chi2)


def test_light_point_is_independent_of_light_pick():
    k = np.arange(40000, dtype=np.uint32)
    for count in [1, 2, 4, 16]:
        with np.errstate(over="ignore"):
            h = _hash(k ^ (U32(count) * U32(0x2545F491)))
        u = unit(oracle.sample_group_bits_many(h & U32(1023), h >> U32(10), _hash(h) % U32(count), count, 3, 2))
        picked = u[:, 2] < 0.5
        n = picked.sum()
        for c in (0, 1):
            p = (u[picked, c] < 0.5).mean()
            assert abs(p - 0.5) < 5 * np.sqrt(0.25 / n), (count, c, p)


def test_path_random_stream_properties():
    """path_sampler.slang:72: the hashed white-noise stream behind the
    data-dependent draws (light-sample roulette), started from
    screen_ray_hash(pixel, index, seed).  The reference holds no vectors for
    it, so its properties are pinned: draws are bucket centres in (0, 1),
    uniform over 256 buckets; successive draws, neighbouring pixels and
    neighbouring sample indices are uncorrelated; and the state update is
    injective (no collision among 2^22 states; a random map would give about
    2,000), so streams never merge."""
    import ctypes
    L = oracle.lib()
    n = 20000
    rng = np.random.default_rng(5)
    px, py, idx = rng.integers(0, 1920, n), rng.integers(0, 1080, n), rng.integers(0, 4096, n)
    draws = np.zeros((n, 4))
    nbr_x, nbr_i = np.zeros(n), np.zeros(n)
    for k in range(n):
        st = ctypes.c_uint32(L.vo_screen_ray_hash(int(px[k]), int(py[k]), int(idx[k]), 7))
        for j in range(4):
            draws[k, j] = L.vo_path_random(ctypes.byref(st))
        st = ctypes.c_uint32(L.vo_screen_ray_hash(int(px[k]) + 1, int(py[k]), int(idx[k]), 7))
        nbr_x[k] = L.vo_path_random(ctypes.byref(st))
        st = ctypes.c_uint32(L.vo_screen_ray_hash(int(px[k]), int(py[k]), int(idx[k]) + 1, 7))
        nbr_i[k] = L.vo_path_random(ctypes.byref(st))
    assert (draws > 0).all() and (draws < 1).all()
    assert np.all(((draws * 8388608.0) - 0.5) % 1.0 == 0)
    for j in range(4):
        obs = np.bincount((draws[:, j] * 256).astype(int), minlength=256)
        e = n / 256
        assert ((obs - e) ** 2 / e).sum() < 255 + 6 * 22.6, j
    bound = 5 / np.sqrt(n)
    for a, b in ((draws[:, 0], draws[:, 1]), (draws[:, 1], draws[:, 2]), (draws[:, 0], nbr_x), (draws[:, 0], nbr_i)):
        assert abs(np.corrcoef(a, b)[0, 1]) < bound
    x = np.arange(1 << 22, dtype=np.uint32) * np.uint32(2654435761)
    h = oracle.pixel_hash_many(x)
    assert np.unique(h).size == x.size


def test_sample_values_are_bucket_centres_in_open_interval():
    b = oracle.sample_group_bits(5, 9, 64, 1, 3, n=64)
    u = unit(b)
    assert (u > 0).all() and (u < 1).all()
    assert np.all(((u * 8388608.0) - 0.5) % 1.0 == 0)
