"""HIP-graph launches: a task that captures its launch once and replays it
must give the eager result on every replay, with eager work and
synchronisations between replays."""
import numpy as np
import pytest

import gpuothello2_amd as go2
import oracle_lib as ol

pytestmark = pytest.mark.gpu


def test_graph_replays_match_eager():
    b = np.concatenate([ol.generate(313, 3, 300, 11), ol.generate(314, 3, 200, 13)])
    want, _, _ = ol.solve_batch(b, 13, threads=16)
    other = ol.generate(315, 3, 200, 10)
    other_want, _, _ = ol.solve_batch(other, 10, threads=16)
    for tt in (0, 18):
        t = go2.Task(len(b), 13, go2.default_config(graph_launch=1, tt_log2_buckets=tt, tt_min_empties=9,
                                                    split_min_empties=9, time_limit_ms=100000))
        t.upload(go2.as_boards(b))
        for rep in range(4):
            t.launch()
            t.wait()
            s, st = t.download()
            assert np.array_equal(s, want), (tt, rep)
            assert st["positions"] == len(b)
            # eager work in between
            e, _ = go2.solve(go2.as_boards(other), 10)
            assert np.array_equal(e, other_want)
        # a new upload with another count re-captures
        t.upload(go2.as_boards(b[:123]))
        t.launch()
        s, _ = t.download()
        assert np.array_equal(s, want[:123])
        t.close()
