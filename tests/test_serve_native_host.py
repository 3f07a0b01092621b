"""The library's serving loop (pixelflow_amd/csrc/pf_serve.cpp) on the CPU, over a
stand-in batch (tests/native/serve_fake_batch.cpp): the same loop code the GPU build runs,
linked with a batch whose arena use is cells x 100 bytes and whose steps are instant.  A
gate in the stand-in holds every step call until the test grants it, so the interleaving
of the worker with submissions is fixed, and every batch call is logged.

Checked on every run, from the log: each step call advances the resident cases by
min(chunk, fewest steps any of them has left), no case is removed before its step count
(unless it failed), and a retired case's fields are downloaded before it is removed.
"""
import ctypes
import threading
import time

import numpy as np
import pytest

from pixelflow_amd import cases
from pixelflow_amd import solver as S

BYTES_PER_CELL = 100


@pytest.fixture
def L():
    import os
    if not os.path.exists(S.SERVE_STAND_IN_PATH):
        pytest.skip("stand-in serving library not built (python -m pixelflow_amd.build)")
    lib = S.serve_stand_in()
    lib.fake_reset(BYTES_PER_CELL)
    yield lib
    lib.fake_set_gate(0)


def _case(m, steps):
    """A 2D case of m x m/2 cells (m identifies the request in the stand-in's log)."""
    p = cases.CaseParams(dim=2, m=m, n=m // 2, width=1.0, height=0.5, time=1e-3, istep_max=steps, iter_max=5)
    return p, np.ones(p.cells)


def _bytes(m):
    return m * (m // 2) * BYTES_PER_CELL


def _log(L):
    """The whole call log (it grows while a worker runs, so read until it fits)."""
    cap = 1 << 16
    while True:
        buf = ctypes.create_string_buffer(cap)
        n = L.fake_log(buf, cap)
        if n < cap:
            return buf.value.decode().splitlines()
        cap = 2 * n + 1


def _check_log(log, steps_of_m, chunk, failed_m=()):
    """The invariants of the loop, replayed from the stand-in's log."""
    left, m_of = {}, {}
    downloaded = {}
    for line in log:
        w = line.split()
        if w[0] == "add":
            s, m = int(w[1]), int(w[2][2:])
            left[s], m_of[s] = steps_of_m[m], m
            downloaded[s] = False
        elif w[0] == "step":
            n, slots = int(w[1]), [int(x) for x in w[2:]]
            assert slots == sorted(left), line
            assert n == min(chunk, min(left[s] for s in slots)), (line, left)
            for s in slots:
                left[s] -= n
        elif w[0] == "download":
            downloaded[int(w[1])] = True
        elif w[0] == "remove":
            s = int(w[1])
            if m_of[s] not in failed_m:
                assert left[s] == 0, (line, left)
                assert downloaded[s], f"slot {s} removed before its fields were downloaded"
            del left[s], m_of[s]
    return log


def test_admission_in_arrival_order_by_free_memory(L):
    """Arena of 200 kB: a (80 kB) and b (88 kB) are admitted; c (180 kB) waits for memory and
    d (20 kB, which would fit beside a and b) waits behind it; c and d join once a and b
    have left.  Adds are held until all four are submitted, so the first admission sees the
    whole queue."""
    steps = {40: 4, 42: 2, 60: 2, 20: 1}
    arena = 200_000
    assert _bytes(40) + _bytes(42) + _bytes(20) <= arena < _bytes(40) + _bytes(60)
    L.fake_set_gate(1)
    L.fake_hold_adds(1)
    with S.NativeServer(_case(40, 1)[0], max_cases=4, arena_bytes=arena, chunk=2, library=L) as srv:
        try:
            ids = {m: srv.submit(*_case(m, k), k) for m, k in steps.items()}
            L.fake_hold_adds(0)
            assert L.fake_wait_blocked(10.0) == 1
            c = srv.counters()
            assert c["running"] == 2 and c["queued"] == 2, c
            assert [ln for ln in _log(L) if ln.startswith("add")] == ["add 0 m=40", "add 1 m=42"]
            while any(srv.poll(i)[0] in (S.SERVE_QUEUED, S.SERVE_RUNNING) for i in ids.values()):
                L.fake_grant_steps(1)
                for i in ids.values():
                    srv.wait(i, 0.05)
            for m, i in ids.items():
                st, done, its, _ = srv.poll(i)
                assert st == S.SERVE_DONE and done == steps[m], (m, st, done)
                assert its == 5 * steps[m]
                t_sub, t_adm, t_fin = srv.times(i)
                assert t_sub <= t_adm <= t_fin
        finally:
            L.fake_hold_adds(0)
            L.fake_set_gate(0)                  # a failed check must not leave the worker at the gate
    log = _check_log(_log(L), steps, chunk=2)
    adds = [ln for ln in log if ln.startswith("add")]
    assert [a.split()[2] for a in adds] == ["m=40", "m=42", "m=60", "m=20"]     # arrival order kept
    # c (m=60) is admitted only after both a and b were removed
    assert log.index("add 0 m=60") > max(log.index("remove 0"), log.index("remove 1"))


def test_results_are_downloaded_fields(L):
    """A done request's kept fields come from the batch's download of its slot."""
    with S.NativeServer(_case(40, 1)[0], max_cases=2, arena_bytes=1 << 20, chunk=3, library=L) as srv:
        a = srv.submit(*_case(40, 5), 5, fields=("u", "p"))
        assert srv.wait(a, 10.0) == S.SERVE_DONE
        u, p = srv.result(a, "u"), srv.result(a, "p")
        assert np.all(u == 10.0 * 0 + S.FIELDS["u"]) and np.all(p == 10.0 * 0 + S.FIELDS["p"])
        with pytest.raises(S.PixelFlowError):
            srv.result(a, "v")                                     # not kept
    _check_log(_log(L), {40: 5}, chunk=3)


def test_case_larger_than_the_arena_fails_and_the_next_one_runs(L):
    with S.NativeServer(_case(40, 1)[0], max_cases=2, arena_bytes=_bytes(40) + 100, chunk=2, library=L) as srv:
        big = srv.submit(*_case(60, 2), 2)
        small = srv.submit(*_case(40, 3), 3)
        assert srv.wait(big, 10.0) == S.SERVE_FAILED and "arena" in srv.error(big)
        assert srv.wait(small, 10.0) == S.SERVE_DONE
    _check_log(_log(L), {60: 2, 40: 3}, chunk=2)


def test_diverged_case_fails_and_the_others_finish(L):
    """The stand-in makes the m=30 case non-finite after its second step: it is retired as
    failed without a result; the other case runs to its step count."""
    L.fake_diverge(30, 2)
    with S.NativeServer(_case(40, 1)[0], max_cases=2, arena_bytes=1 << 20, chunk=1, library=L) as srv:
        bad = srv.submit(*_case(30, 6), 6)
        good = srv.submit(*_case(40, 6), 6)
        assert srv.wait(bad, 10.0) == S.SERVE_FAILED and "diverged" in srv.error(bad)
        assert srv.wait(good, 10.0) == S.SERVE_DONE
        assert srv.poll(bad)[1] == 2 and srv.poll(good)[1] == 6
        assert np.isnan(srv.worst(bad)["sor_error"])
        with pytest.raises(S.PixelFlowError):
            srv.result(bad, "u")
    _check_log(_log(L), {30: 6, 40: 6}, chunk=1, failed_m=(30,))


def test_destroy_fails_unfinished_requests_and_wakes_waiters(L):
    """A server destroyed with one request running (it would never finish) and one queued
    behind it (no memory): threads blocked in pf_serve_wait on both return with the
    requests failed, before destroy frees the server."""
    srv = S.NativeServer(_case(40, 1)[0], max_cases=2, arena_bytes=_bytes(40) + 100, chunk=1, library=L)
    running = srv.submit(*_case(40, 10 ** 12), 10 ** 12)
    queued = srv.submit(*_case(42, 3), 3)
    assert L.fake_wait_step_calls(3, 10.0) >= 3
    assert srv.poll(running)[0] == S.SERVE_RUNNING and srv.poll(queued)[0] == S.SERVE_QUEUED
    seen = {}
    h = srv._h
    calling = [threading.Event(), threading.Event()]

    def waiter(k, rid):
        st = ctypes.c_int()
        calling[k].set()
        L.pf_serve_wait(h, rid, -1.0, ctypes.byref(st))
        seen[rid] = st.value

    threads = [threading.Thread(target=waiter, args=(k, rid)) for k, rid in enumerate((running, queued))]
    for t in threads:
        t.start()
    # a wait entered after destroy would use a freed server: let both threads get into theirs
    assert all(e.wait(10.0) for e in calling)
    time.sleep(0.5)
    srv._h = None
    L.pf_serve_destroy(h)
    for t in threads:
        t.join(10.0)
        assert not t.is_alive(), "a waiter was still blocked after destroy returned"
    assert seen == {running: S.SERVE_FAILED, queued: S.SERVE_FAILED}
    assert _log(L)[-1] == "destroy"
