"""Served form: jobs submitted while others run are admitted by device-memory
budget and advanced together, and each result equals the same job run alone
(bit for bit: jobs are independent and keyed by their own seeds)."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("genome", "obj", "idx", "pql", "sorted", "front_of")

# (residues, m, N, H, seed, generations)
SPECS = [(60, 2, 32, 2, 21, 9), (140, 4, 64, 3, 22, 5), (37, 8, 16, 1, 23, 12), (250, 2, 96, 4, 24, 3),
         (90, 3, 48, 2, 25, 7), (45, 2, 24, 2, 26, 0)]


def build(cdp, spec):
    res, m, N, H, seed, G = spec
    T = cdp.tables
    return (T.synth_protein(res, seed), T.synth_tables(seed), m, N, H, 0.1, seed), G


def alone(cdp, spec):
    args, G = build(cdp, spec)
    j = cdp.Job(*args).init()
    if G:
        j.step(G)
    r = j.read()
    j.close()
    return r


def bytes_of(cdp, spec):
    res, m, N, H, _, _ = spec
    return cdp.device_bytes(res + 1, m, N, H)


def test_server_results_equal_alone(cdp):
    print("server mix", flush=True)
    want = [alone(cdp, s) for s in SPECS]
    srv = cdp.Server(1 << 34)
    tickets = []
    for i, s in enumerate(SPECS):  # arrivals while earlier jobs are resident
        args, G = build(cdp, s)
        tickets.append(srv.submit(*args, G))
        if i % 2:
            time.sleep(0.01)
    for t, w, s in zip(tickets, want, SPECS):
        assert srv.wait(t, 60.0) == cdp.SERVER_DONE, s
        r = srv.result(t)
        for k in KEYS:
            assert np.array_equal(r[k], w[k]), (s, k)
        tm = r["times"]
        assert 0.0 <= tm["queued"] <= tm["time_to_front"] and tm["resident"] >= 0.0
    st = srv.stats()
    assert st["queued"] == 0 and st["resident"] == 0 and st["bytes"] == 0
    assert st["designs"] == sum(s[2] * s[5] for s in SPECS)
    srv.close()


def test_server_admission_by_budget(cdp):
    print("server budget", flush=True)
    specs = SPECS[:5]
    budget = max(bytes_of(cdp, s) for s in specs) + min(bytes_of(cdp, s) for s in specs)
    want = [alone(cdp, s) for s in specs]
    srv = cdp.Server(budget)
    tickets = [srv.submit(*build(cdp, s)[0], s[5]) for s in specs]
    seen_queue, peak = 0, 0
    deadline = time.time() + 60.0
    while time.time() < deadline:
        st = srv.stats()
        assert st["bytes"] <= budget
        seen_queue = max(seen_queue, st["queued"])
        peak = max(peak, st["bytes"])
        if st["queued"] == 0 and st["resident"] == 0:
            break
        time.sleep(0.001)
    assert seen_queue > 0  # the budget held requests back
    for t, w, s in zip(tickets, want, specs):
        assert srv.wait(t, 60.0) == cdp.SERVER_DONE
        r = srv.result(t)
        for k in KEYS:
            assert np.array_equal(r[k], w[k]), (s, k)
    srv.close()


def test_server_rejects_job_over_budget(cdp):
    args, G = build(cdp, SPECS[1])
    srv = cdp.Server(bytes_of(cdp, SPECS[1]) - 1)
    with pytest.raises(cdp.CdError, match="CD_E_BUDGET"):
        srv.submit(*args, G)
    with pytest.raises(cdp.CdError):
        srv.result(12345)  # unknown ticket
    srv.close()
