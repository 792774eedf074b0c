"""Host logic of the serving loop: the seeded generator and the request
stream (no device work)."""
import numpy as np

from sph07_amd.serving import SplitMix64, draw_stream, scene_of_size


def test_splitmix64_known_answers():
    # first outputs of SplitMix64 from state 0 (Vigna's splitmix64.c)
    g = SplitMix64(0)
    assert g.next_u64() == 0xE220A8397B1DCDAF
    assert g.next_u64() == 0x6E789E6AA1B965F4
    assert g.next_u64() == 0x06C45D188009454F


def test_stream_is_deterministic_and_follows_the_mix():
    a = draw_stream(seed=7, n=4000, rate=10.0)
    b = draw_stream(seed=7, n=4000, rate=10.0)
    assert a == b
    kinds = [k for k, *_ in a]
    frac = {k: kinds.count(k) / len(kinds) for k in set(kinds)}
    for k, w in (("poiseuille", 0.4), ("cavity", 0.3), ("droplet", 0.2), ("cylinder", 0.1)):
        assert abs(frac[k] - w) < 0.03, (k, frac[k])
    sizes = np.array([s for _, s, _, _ in a])
    assert sizes.min() >= 4000 and sizes.max() <= 500000
    # log-uniform: the median is the geometric mean of the range
    assert abs(np.log(np.median(sizes)) - 0.5 * np.log(4000 * 500000)) < 0.1
    steps = np.array([s for _, _, s, _ in a])
    assert steps.min() >= 500 and steps.max() <= 5000
    gaps = np.diff([t for *_, t in a])
    assert abs(gaps.mean() - 0.1) < 0.01          # Poisson rate 10/s
    assert abs(gaps.std() / gaps.mean() - 1.0) < 0.1


def test_scene_sizes_track_the_request():
    for kind in ("poiseuille", "cavity", "droplet", "cylinder"):
        d = scene_of_size(kind, 10000)
        assert 0.7 * 10000 < d.n_fluid < 1.4 * 10000, (kind, d.n_fluid)
