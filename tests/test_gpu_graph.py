"""The plain Nesterov loop replays its batches of evaluation slots from a
captured HIP graph (rp_nesterov_place).  A run that replays the graph many
times, with the state read back and synchronised between replays, must give
exactly the records and positions of the same run launched kernel by kernel
(RP_NO_GRAPH=1)."""
import os

import numpy as np
import pytest

from replace_amd import netgen
from replace_amd.capi import Placer

pytestmark = pytest.mark.gpu


def _run(nl, p, graph):
    if not graph:
        os.environ["RP_NO_GRAPH"] = "1"
    try:
        g = Placer(nl, p)
        g.initial_place()
        g.setup()
        code, st = g.nesterov_place()
        recs = g.records()
        pos = g.positions()
        g.close()
        return code, int(st.iterations), recs, pos
    finally:
        os.environ.pop("RP_NO_GRAPH", None)


@pytest.mark.parametrize("cells", [3000, 20000])
def test_graph_replay_matches_eager_launches(cells):
    nl = netgen.generate(cells, int(cells * 1.05), seed=4242 + cells, util=0.65, td=0.7)
    nb = netgen.next_pow2(int(np.ceil(np.sqrt(cells))))
    p = netgen.params_default(target_density=0.7, bin_grid_cnt_x=nb, bin_grid_cnt_y=nb, initial_place_max_iter=3)
    eager = _run(nl, p, graph=False)
    graph = _run(nl, p, graph=True)
    assert eager[0] == graph[0] and eager[1] == graph[1]
    assert eager[1] > 64  # many batches, so the graph was replayed many times
    assert eager[2] == graph[2]
    assert np.array_equal(eager[3][0], graph[3][0]) and np.array_equal(eager[3][1], graph[3][1])


def _run_env(nl, p, env):
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        g = Placer(nl, p)
        g.initial_place()
        g.setup()
        code, st = g.nesterov_place()
        recs = g.records()
        pos = g.positions()
        g.close()
        return code, int(st.iterations), recs, pos
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


@pytest.mark.parametrize("cells,per", [(3000, "1"), (20000, "1"), (20000, "2")])
def test_fused_scatter_wirelength_launch_matches_two_launches(cells, per):
    """The single-instance path runs the density scatter and the wirelength
    kernel as one launch (their blocks interleaved per XCD); the run must be
    exactly the one with the two kernels launched one after the other
    (RP_NO_FUSE=1), for both scatter shapes."""
    nl = netgen.generate(cells, int(cells * 1.05), seed=777 + cells, util=0.65, td=0.7)
    nb = netgen.next_pow2(int(np.ceil(np.sqrt(cells))))
    p = netgen.params_default(target_density=0.7, bin_grid_cnt_x=nb, bin_grid_cnt_y=nb, initial_place_max_iter=3)
    two = _run_env(nl, p, {"RP_NO_FUSE": "1", "RP_SCATTER_PER": per})
    one = _run_env(nl, p, {"RP_SCATTER_PER": per})
    assert two[0] == one[0] and two[1] == one[1] and two[1] > 50
    assert two[2] == one[2]
    assert np.array_equal(two[3][0], one[3][0]) and np.array_equal(two[3][1], one[3][1])
