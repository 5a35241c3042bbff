"""CPU tests of the autotuner's coordinate search (no GPU): finds the optimum of a
separable cost, evaluates each configuration once, respects the pass limit."""
from __future__ import annotations

import pytest


@pytest.fixture(scope="module")
def search():
    from StochasticGSCudaBackend.autotune import coordinate_search
    return coordinate_search


def test_separable_optimum(search):
    space = {"a": [1, 2, 3, 4], "b": [10, 20, 30]}
    cost = lambda c: (c["a"] - 3) ** 2 + (c["b"] - 20) ** 2 / 100.0
    best, ms, trials = search(cost, space, {"a": 1, "b": 10})
    assert best == {"a": 3, "b": 20} and ms == 0.0
    keys = [tuple(sorted(c.items())) for c, _ in trials]
    assert len(keys) == len(set(keys))


def test_coupled_cost_needs_second_pass(search):
    # optimum (2, 2) is reachable only after b moves: cost punishes a=2 unless b=2
    def cost(c):
        return {(1, 1): 5, (2, 1): 6, (1, 2): 4, (2, 2): 1}[(c["a"], c["b"])]
    best, ms, _ = search(cost, {"a": [1, 2], "b": [1, 2]}, {"a": 1, "b": 1}, passes=2)
    assert best == {"a": 2, "b": 2} and ms == 1
    best1, ms1, _ = search(cost, {"a": [1, 2], "b": [1, 2]}, {"a": 1, "b": 1}, passes=1)
    assert ms1 <= 4


def test_min_gain_keeps_defaults_within_noise(search):
    # 0.5 % faster stays below a 1 % threshold, 5 % faster passes it
    space = {"a": [0, 1], "b": [0, 1]}
    cost = lambda c: 100.0 - 0.5 * c["a"] - 5.0 * c["b"]
    best, ms, _ = search(cost, space, {"a": 0, "b": 0}, min_gain=0.01)
    assert best == {"a": 0, "b": 1} and ms == 95.0
    best0, _, _ = search(cost, space, {"a": 0, "b": 0})
    assert best0 == {"a": 1, "b": 1}


def test_defaults_are_candidates(search):
    best, ms, trials = search(lambda c: 1.0, {"x": [5, 6]}, {"x": 0})
    assert best == {"x": 0} and trials[0][0] == {"x": 0}


def test_views_with_bins_threshold():
    # 16-pixel bins for framebuffers (W ss x H ss) up to the threshold, 32 above; -1 keeps
    # each view's own bin size (by default the library's choice)
    import torch
    from StochasticGSCudaBackend.autotune import views_with_bins
    from StochasticGSCudaBackend.batch import BatchView
    from StochasticGSCudaBackend.render import Settings

    def view(w, h, ss=1, bin_size=0):
        s = Settings(torch.eye(4), torch.zeros(3), w, h, 100.0, 100.0, w / 2, h / 2, 0.2, 100.0, False, ss, False,
                     False, False)
        return BatchView(s, spp=2, frame_idx=7, bin_size=bin_size)

    views = [view(256, 256), view(640, 480), view(640, 480, ss=2), view(1920, 1080, bin_size=64)]
    assert [v.bin_size for v in views_with_bins(views, -1)] == [0, 0, 0, 64]
    assert [v.bin_size for v in views_with_bins(views, 1 << 19)] == [16, 16, 32, 32]
    assert [v.bin_size for v in views_with_bins(views, 0)] == [32, 32, 32, 32]
    out = views_with_bins(views, 1 << 40)
    assert [v.bin_size for v in out] == [16, 16, 16, 16]
    # the other fields are kept and the input views are not modified
    assert all(o.spp == 2 and o.frame_idx == 7 and o.settings is v.settings for o, v in zip(out, views))
    assert [v.bin_size for v in views] == [0, 0, 0, 64]


def test_split_views_contiguous_balanced_groups():
    # min(groups, views) non-empty contiguous groups covering every view once, each boundary
    # where the pixel-samples before it come closest to its share (checked by brute force)
    import itertools

    import torch
    from StochasticGSCudaBackend.batch import BatchView, split_views
    from StochasticGSCudaBackend.render import Settings

    def view(w, h, spp, ss=1):
        s = Settings(torch.eye(4), torch.zeros(3), w, h, 100.0, 100.0, w / 2, h / 2, 0.2, 100.0, False, ss, False,
                     False, False)
        return BatchView(s, spp=spp)

    views = [view(256, 256, 1), view(640, 480, 4), view(1920, 1080, 16), view(100, 80, 2, ss=2), view(1280, 720, 1)]
    cost = [256 * 256, 640 * 480 * 4, 1920 * 1080 * 16, 200 * 160 * 2, 1280 * 720]
    for k in range(1, 8):
        groups = split_views(views, k)
        assert len(groups) == min(k, len(views)) and all(groups)
        assert [i for g in groups for i in g] == list(range(len(views)))
        if len(groups) == 2:
            best = min(range(1, len(views)), key=lambda b: abs(sum(cost[:b]) - sum(cost) / 2))
            assert groups[0] == list(range(best))
    assert split_views(views[:2], 2) == [[0], [1]]  # the larger view second still gets its own group
    assert split_views(views[:1], 3) == [[0]]
    assert list(itertools.chain(*split_views(views, 1))) == list(range(5))
