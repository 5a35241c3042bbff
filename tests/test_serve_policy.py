"""CPU tests of the serving loop's admission policy (no GPU): arrival order, a stream's
frames stay in order (a refused frame blocks its stream's later ones), streams with frames
in flight wait, chains of consecutive frames, the memory bound stops admission."""
from __future__ import annotations

from collections import deque

import numpy as np
import pytest


def _req(stream, frame=0):
    from StochasticGSCudaBackend.serve import Request
    r = Request(stream=stream, w2c=np.eye(4, dtype=np.float32), cam_position=np.zeros(3, np.float32), spp=1)
    r.frame = frame
    return r


@pytest.fixture(scope="module")
def admit():
    from StochasticGSCudaBackend.serve import admit
    return admit


def test_arrival_order_and_max_views(admit):
    pending = deque(_req(s) for s in range(6))
    chosen, rest = admit(pending, lambda r, k: True, lambda r, i, k: True, max_views=4)
    assert [r.stream for r in chosen] == [0, 1, 2, 3]
    assert [r.stream for r in rest] == [4, 5]


def test_chains_and_busy_streams_wait(admit):
    pending = deque([_req(0, 0), _req(0, 1), _req(1, 0), _req(2, 0), _req(1, 1), _req(0, 2)])
    busy = {2}
    # at most two frames of a stream per batch
    chosen, rest = admit(pending, lambda r, k: r.stream not in busy and k < 2, lambda r, i, k: True, max_views=8)
    assert [(r.stream, r.frame) for r in chosen] == [(0, 0), (0, 1), (1, 0), (1, 1)]
    # the skipped requests keep their arrival order
    assert [(r.stream, r.frame) for r in rest] == [(2, 0), (0, 2)]


def test_refused_frame_blocks_later_frames_of_its_stream(admit):
    pending = deque([_req(0, 0), _req(0, 1), _req(1, 0)])
    calls = []

    def can_take(r, k):
        calls.append((r.stream, r.frame, k))
        return not (r.stream == 0 and r.frame == 0)

    chosen, rest = admit(pending, can_take, lambda r, i, k: True, max_views=8)
    assert [(r.stream, r.frame) for r in chosen] == [(1, 0)]
    assert [(r.stream, r.frame) for r in rest] == [(0, 0), (0, 1)]
    assert (0, 1, 0) not in calls


def test_memory_bound_stops_admission(admit):
    pending = deque(_req(s) for s in range(5))
    budget = 3  # "bytes": each request costs one
    chosen, rest = admit(pending, lambda r, k: True, lambda r, i, k: i + 1 <= budget, max_views=8)
    assert [r.stream for r in chosen] == [0, 1, 2]
    assert [r.stream for r in rest] == [3, 4]


def test_request_larger_than_the_arena_is_an_error(admit):
    with pytest.raises(RuntimeError):
        admit(deque([_req(0)]), lambda r, k: True, lambda r, i, k: False, max_views=8)
