"""Batched / served placement: every job's result (iteration records and
final positions) is bit-identical to running the same job alone through
rp_nesterov_place, whatever runs beside it, and jobs are admitted against
the HBM budget."""
import numpy as np
import pytest

from helpers import assert_tracks_trace, design
from replace_amd import netgen
from replace_amd.capi import Placer, Server

pytestmark = pytest.mark.gpu


def _alone(nl, p):
    g = Placer(nl, p)
    g.initial_place()
    g.setup()
    code, st = g.nesterov_place()
    return code, st, g.records(), g.positions()


def _jobs():
    jobs = []
    for name in ("simple01", "ar02", "incremental01"):
        jobs.append(design(name))
    for i, nc in enumerate((700, 2500, 9000)):
        nl = netgen.generate(nc, int(nc * 1.05), seed=500 + i, util=0.65, td=0.7)
        nb = netgen.next_pow2(int(np.ceil(np.sqrt(nc))))
        p = netgen.params_default(target_density=0.7, bin_grid_cnt_x=max(32, nb),
                                  bin_grid_cnt_y=max(32, nb), initial_place_max_iter=0)
        jobs.append((nl, p))
    return jobs


def test_batch_results_identical_to_alone():
    jobs = _jobs()
    alone = [_alone(nl, p) for nl, p in jobs]
    srv = Server(steps_per_poll=4)
    ids = [srv.submit(nl, p, 0.0) for nl, p in jobs]
    srv.run()
    for jid, (code, st, rec, pos) in zip(ids, alone):
        r = srv.result(jid)
        assert r.done == 1 and r.code == code
        assert srv.records(jid) == rec
        lx, ly = srv.positions(jid)
        assert np.array_equal(lx, pos[0]) and np.array_equal(ly, pos[1])
        assert r.stats.final_hpwl == st.final_hpwl and r.stats.iterations == st.iterations


def test_served_arrivals_under_tight_budget():
    jobs = _jobs()
    alone = [_alone(nl, p) for nl, p in jobs]
    # budget below two jobs' footprint: jobs queue and are admitted as others retire
    srv = Server(hbm_budget_bytes=48 << 20, steps_per_poll=2, setup_threads=2)
    ids = [srv.submit(nl, p, 0.002 * k) for k, (nl, p) in enumerate(jobs)]
    srv.run()
    starts = []
    for jid, (code, st, rec, pos) in zip(ids, alone):
        r = srv.result(jid)
        assert r.done == 1
        assert r.start_s >= r.arrival_s
        assert srv.records(jid) == rec
        lx, ly = srv.positions(jid)
        assert np.array_equal(lx, pos[0]) and np.array_equal(ly, pos[1])
        starts.append((r.admit_s, r.done_s))
    # admission is FIFO in arrival order
    assert [a for a, _ in starts] == sorted(a for a, _ in starts)
    assert srv.stats().rebuilds >= 2


def test_batch_with_graph_and_without_agree():
    jobs = _jobs()[3:]
    out = []
    for use_graph in (1, 0):
        srv = Server(use_graph=use_graph, steps_per_poll=3)
        ids = [srv.submit(nl, p) for nl, p in jobs]
        srv.run()
        out.append([srv.records(j) for j in ids])
    assert out[0] == out[1]


@pytest.mark.parametrize("cells,pins", [(256, 256), (512, 1024), (1024, 512)])
def test_every_tile_shape_identical_to_alone(cells, pins):
    """The tile shapes the tuner chooses between give bit-identical jobs."""
    jobs = _jobs()
    alone = [_alone(nl, p) for nl, p in jobs]
    srv = Server(steps_per_poll=4)
    for b in range(5):
        srv.set_tiles(b, cells, pins)
    ids = [srv.submit(nl, p, 0.0) for nl, p in jobs]
    srv.run()
    for jid, (code, st, rec, pos) in zip(ids, alone):
        assert srv.records(jid) == rec
        lx, ly = srv.positions(jid)
        assert np.array_equal(lx, pos[0]) and np.array_equal(ly, pos[1])
    assert all(t["scatter_cells"] == cells and t["wa_pins"] == pins for t in srv.tuning())


def test_tuner_times_every_shape_of_the_buckets_present():
    jobs = _jobs()
    srv = Server(steps_per_poll=4, autotune=1)
    ids = [srv.submit(nl, p, 0.0) for nl, p in jobs]
    srv.run()
    assert all(srv.result(j).done == 1 for j in ids)
    tiles = {t["bucket"]: t for t in srv.tuning()}
    for b in (0, 1, 2):  # 294-cell designs and the 700 / 2.5k / 9k-cell jobs
        t = tiles[b]
        assert t["tuned"] == 1 and t["jobs_sampled"] >= 1
        assert min(t["scatter_us"]) > 0 and min(t["wa_us"]) > 0
        assert t["scatter_cells"] == 256 << int(np.argmin(t["scatter_us"]))
        assert t["wa_pins"] == 256 << int(np.argmin(t["wa_us"]))


@pytest.mark.parametrize("batched_ip", [0, 1, 2])
def test_batched_initial_place_identical_to_alone(batched_ip):
    """With batched_ip 1 the service runs the initial place of every job in
    set-up in shared launches (one-workgroup solves of small systems two
    workgroups per job, grid solves through one (job, block) table); with 2
    only the grid-solver jobs (the 12000-cell one here) go to those rounds.
    Each job still computes exactly what it computes alone, so its placement
    is bit-identical to the plain path's."""
    jobs = []
    for i, (nc, ip) in enumerate(((1500, 6), (6000, 5), (12000, 4), (900, 1), (3000, 0))):
        nl = netgen.generate(nc, int(nc * 1.05), seed=700 + i, util=0.65, td=0.7)
        nb = netgen.next_pow2(int(np.ceil(np.sqrt(nc))))
        p = netgen.params_default(target_density=0.7, bin_grid_cnt_x=max(32, nb),
                                  bin_grid_cnt_y=max(32, nb), initial_place_max_iter=ip)
        jobs.append((nl, p))
    alone = [_alone(nl, p) for nl, p in jobs]
    srv = Server(steps_per_poll=4, batched_ip=batched_ip)
    ids = [srv.submit(nl, p, 0.0) for nl, p in jobs]
    srv.run()
    for jid, (code, st, rec, pos) in zip(ids, alone):
        r = srv.result(jid)
        assert r.done == 1 and r.code == code
        assert srv.records(jid) == rec
        lx, ly = srv.positions(jid)
        assert np.array_equal(lx, pos[0]) and np.array_equal(ly, pos[1])


def test_batched_reference_designs_track_golden_traces():
    """Reference designs placed by the service, checked against the
    reference's own golden logs (not only against the plain GPU path)."""
    names = ("simple01", "ar02", "incremental01", "simple02", "core01")
    srv = Server(steps_per_poll=4)
    ids = [srv.submit(*design(n), 0.0) for n in names]
    srv.run()
    for name, jid in zip(names, ids):
        r = srv.result(jid)
        assert r.done == 1 and r.code == 0 and r.stats.finished == 1
        assert_tracks_trace(name, srv.records(jid), r.stats.iterations, r.stats.final_overflow, False)


def test_batch_workload_sample_identical_to_alone():
    """Sixteen instances of the 256-instance batched workload (500-50k cells,
    own target densities and bin grids, initial place on), every one
    bit-identical in the batch and alone."""
    idx = [0, 1, 3, 5, 13, 17, 18, 31, 47, 64, 90, 111, 128, 150, 200, 255]
    jobs = [netgen.batch_instance(i) for i in idx]
    alone = [_alone(nl, p) for nl, p in jobs]
    srv = Server()
    ids = [srv.submit(nl, p, 0.0) for nl, p in jobs]
    srv.run()
    for jid, (code, st, rec, pos) in zip(ids, alone):
        r = srv.result(jid)
        assert r.done == 1 and r.code == code
        assert srv.records(jid) == rec
        lx, ly = srv.positions(jid)
        assert np.array_equal(lx, pos[0]) and np.array_equal(ly, pos[1])


def test_served_stream_with_a_large_job_identical_to_alone():
    """A served stream (staggered arrivals) holding a large job (200,915
    cells, 512 x 512 bins) among instances of the batched distribution: every
    job bit-identical to its plain run, and no job starts before it arrives."""
    jobs = [netgen.large_instance(4)] + [netgen.batch_instance(i) for i in (2, 7, 19, 40, 77, 133)]
    arrivals = [0.0, 0.05, 0.1, 0.15, 0.3, 0.45, 0.6]
    alone = [_alone(nl, p) for nl, p in jobs]
    srv = Server(steps_per_poll=8)
    ids = [srv.submit(nl, p, a) for (nl, p), a in zip(jobs, arrivals)]
    srv.run()
    for jid, (code, st, rec, pos) in zip(ids, alone):
        r = srv.result(jid)
        assert r.done == 1 and r.code == code and r.start_s >= r.arrival_s
        assert srv.records(jid) == rec
        lx, ly = srv.positions(jid)
        assert np.array_equal(lx, pos[0]) and np.array_equal(ly, pos[1])
