"""Request-loop scheduling on the CPU: admission by slots and HBM budget,
first-come first-served order, no admission before arrival, Poisson
arrivals.  Workers are simulated with threads that sleep."""
import threading
import time

import numpy as np
import pytest

from ddg_amd.serve import Admission, Request, dispatch, make_requests, poisson_arrivals


def _run(requests, slots, budget, footprint, service):
    adm = Admission(slots, budget)
    cond = threading.Condition()
    log = []          # (event, rid, t, busy, used)
    t0 = time.perf_counter()

    def clock():
        return time.perf_counter() - t0

    def start(req, nb, t_admit):
        with cond:
            log.append(("start", req.rid, t_admit, adm.busy, adm.used))

        def work():
            time.sleep(service(req))
            with cond:
                adm.release(nb)
                log.append(("end", req.rid, clock(), adm.busy, adm.used))
                cond.notify_all()
        threading.Thread(target=work, daemon=True).start()

    def wait(timeout):
        with cond:
            cond.wait(timeout=max(0.0, min(timeout, 0.05)))

    dispatch(requests, adm, footprint, start, wait, clock)
    deadline = time.time() + 10
    while adm.busy and time.time() < deadline:
        time.sleep(0.005)
    return adm, log


def test_slots_and_budget_never_exceeded():
    reqs = [Request(i, 16, 2, 0.0, i, arrival=0.0) for i in range(24)]
    fp = {r.rid: (3 if r.rid % 3 == 0 else 1) << 20 for r in reqs}
    adm, log = _run(reqs, slots=4, budget=5 << 20, footprint=lambda r: fp[r.rid],
                    service=lambda r: 0.01)
    starts = [e for e in log if e[0] == "start"]
    assert [e[1] for e in starts] == list(range(24))       # first come, first served
    assert adm.peak_busy <= 4 and adm.peak_used <= 5 << 20
    assert adm.peak_busy >= 2                              # requests did overlap
    assert adm.busy == 0 and adm.used == 0


def test_no_admission_before_arrival():
    reqs = [Request(i, 16, 2, 0.0, i, arrival=0.02 * i) for i in range(10)]
    _, log = _run(reqs, slots=8, budget=1 << 30, footprint=lambda r: 1,
                  service=lambda r: 0.001)
    for e in log:
        if e[0] == "start":
            assert e[2] >= reqs[e[1]].arrival


def test_late_arrivals_join_while_others_run():
    # a long request is running when short ones arrive: they are admitted
    # into the free slots without waiting for it (continuous batching)
    reqs = [Request(0, 16, 2, 0.0, 0, arrival=0.0)] + \
           [Request(i, 16, 2, 0.0, i, arrival=0.02) for i in range(1, 4)]
    _, log = _run(reqs, slots=4, budget=1 << 30, footprint=lambda r: 1,
                  service=lambda r: 0.3 if r.rid == 0 else 0.01)
    t_end0 = [e[2] for e in log if e[0] == "end" and e[1] == 0][0]
    for rid in (1, 2, 3):
        t_end = [e[2] for e in log if e[0] == "end" and e[1] == rid][0]
        assert t_end < t_end0


def test_oversized_request_rejected():
    adm = Admission(2, 100)
    with pytest.raises(ValueError):
        dispatch([Request(0, 16, 2, 0.0, 0)], adm, lambda r: 101, lambda *a: None,
                 lambda t: None, lambda: 0.0)
    with pytest.raises(ValueError):
        Admission(0, 100)
    with pytest.raises(RuntimeError):
        Admission(1, 100).release(1)


def test_poisson_arrivals_and_request_mix():
    reqs = make_requests(4000, seed=3)
    assert {r.n for r in reqs} == {16, 32, 64, 128}
    assert {r.N for r in reqs} == {2, 3, 4, 5, 6}
    assert all(0.0 <= r.k2 <= 50.0 for r in reqs)
    poisson_arrivals(reqs, rate=200.0, seed=4)
    t = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(t) >= 0)
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert abs(gaps.mean() - 1 / 200.0) < 0.1 / 200.0      # mean gap 1/rate within 10 %
    assert abs(gaps.std() / gaps.mean() - 1.0) < 0.1       # exponential: cv = 1
    again = make_requests(50, seed=3)[7]                    # seeded
    assert (again.n, again.N, again.k2, again.seed) == (reqs[7].n, reqs[7].N, reqs[7].k2,
                                                        reqs[7].seed)


def _run_groups(requests, slots, budget, footprint, service, max_group):
    from ddg_amd.serve import dispatch_groups
    adm = Admission(slots, budget)
    cond = threading.Condition()
    log = []          # (group rids, t_admit)
    t0 = time.perf_counter()

    def clock():
        return time.perf_counter() - t0

    def start(group, nb, t_admit):
        with cond:
            log.append(([r.rid for r in group], t_admit))

        def work():
            time.sleep(service(group))
            with cond:
                adm.release(nb)
                cond.notify_all()
        threading.Thread(target=work, daemon=True).start()

    def wait(timeout):
        with cond:
            cond.wait(timeout=max(0.0, min(timeout, 0.05)))

    dispatch_groups(requests, adm, footprint, start, wait, clock, key=lambda r: (r.n, r.N),
                    max_group=max_group)
    deadline = time.time() + 10
    while adm.busy and time.time() < deadline:
        time.sleep(0.005)
    return adm, log


def test_groups_take_arrived_requests_of_the_head_class():
    # all arrive at once: the head's class is grouped up to max_group, in
    # arrival order; heads go first come, first served
    classes = [(16, 2), (32, 3), (16, 2), (16, 2), (32, 3), (16, 2), (16, 2), (64, 4)]
    reqs = [Request(i, n, N, 0.0, i, arrival=0.0) for i, (n, N) in enumerate(classes)]
    _, log = _run_groups(reqs, slots=1, budget=1 << 30, footprint=lambda g: len(g),
                         service=lambda g: 0.005, max_group=3)
    assert [g for g, _ in log] == [[0, 2, 3], [1, 4], [5, 6], [7]]


def test_groups_never_hold_requests_not_yet_arrived():
    reqs = [Request(0, 16, 2, 0.0, 0, arrival=0.0), Request(1, 16, 2, 0.0, 1, arrival=0.2),
            Request(2, 16, 2, 0.0, 2, arrival=0.0)]
    _, log = _run_groups(reqs, slots=2, budget=1 << 30, footprint=lambda g: len(g),
                         service=lambda g: 0.01, max_group=8)
    assert [g for g, _ in log] == [[0, 2], [1]]
    assert log[1][1] >= 0.2


def test_groups_cut_to_the_budget():
    reqs = [Request(i, 16, 2, 0.0, i, arrival=0.0) for i in range(7)]
    adm, log = _run_groups(reqs, slots=2, budget=3, footprint=lambda g: len(g),
                           service=lambda g: 0.01, max_group=8)
    assert all(len(g) <= 3 for g, _ in log)
    assert sorted(r for g, _ in log for r in g) == list(range(7))
    assert adm.peak_used <= 3
