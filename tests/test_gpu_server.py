"""GPU tests of the serving loop (ccd_server_*): every result is the same alone and
in a concurrent mix of model types and sizes, admission honours the HBM budget,
and over-budget requests are refused."""
import numpy as np
import pytest

import _cases
import oracle
from cyclops_amd import _abi
from cyclops_amd import data as D
from cyclops_amd.control import create_control, create_prior

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def E():
    from cyclops_amd import engine
    assert engine.device_count() > 0, "no HIP device"
    return engine


def _mix():
    ctl = create_control(convergence_type="mittal", tolerance=1e-8)
    d0, _ = D.make_config(0, scale=1.2)   # 120,000 rows: a multi-workgroup persistent launch
    return [
        (d0, create_prior("laplace", 0.1), ctl),
        (_cases.logistic(K=2000, seed=11), create_prior("laplace", 0.5), ctl),
        (_cases.poisson(K=3000, seed=12), create_prior("normal", 1.0), ctl),
        (_cases.cox(K=1500, seed=13), create_prior("laplace", 0.5), ctl),
        (_cases.clr(n_strata=300, seed=14), create_prior("laplace", 1.0), ctl),
        (_cases.logistic(K=800, seed=15), create_prior("normal", 2.0), ctl),
    ]


def _overlap(stats):
    iv = sorted((s["start"], s["end"]) for s in stats)
    return any(iv[i + 1][0] < iv[i][1] for i in range(len(iv) - 1))


def test_results_same_alone_and_in_mix(E):
    mix = _mix()
    alone = []
    srv = E.Server(slots=1, hbm_budget_bytes=8e9)
    for i, (d, p, c) in enumerate(mix):
        print(f"[case] server alone request {i} model={d.model_type} K={d.K}", flush=True)
        alone.append(srv.poll(srv.submit(d, p, c), wait=True))
    srv.stop()

    print("[case] server mix of 6 requests, 4 slots", flush=True)
    srv = E.Server(slots=4, hbm_budget_bytes=8e9)
    ids = [srv.submit(d, p, c) for d, p, c in mix]
    together = [srv.poll(i, wait=True) for i in ids]
    srv.stop()

    for i, ((b1, r1, _), (b2, r2, _)) in enumerate(zip(alone, together)):
        assert np.array_equal(b1, b2), i
        assert r1["iterations"] == r2["iterations"] and r1["log_likelihood"] == r2["log_likelihood"]
    # the mix ran concurrently on the device
    assert _overlap([t[2] for t in together])
    # and the results are the oracle's
    for (d, p, c), (b, r, _) in zip(mix[1:4], together[1:4]):
        bo, ro = oracle.fit(d, p, c)
        assert r["iterations"] == ro["iterations"]
        assert np.max(np.abs(b - bo)) <= 1e-6 * max(np.max(np.abs(bo)), 1e-12)


def test_short_requests_do_not_wait_for_a_long_launch(E, monkeypatch):
    """A long fit held in one persistent launch (300 cycles, no convergence test) and
    short requests arriving while it runs: the short ones return long before it ends.
    Nothing in a request's path may wait on the device as a whole or on the null
    stream, which shares a hardware queue with some slot."""
    from cyclops_amd import _abi
    import time
    monkeypatch.setenv("CCD_CYCLES_PER_LAUNCH", "100000")
    big = D.make_request(2, max_rows=100_000, max_cols=10_000)["data"]   # LR, 73,412 x 1,414
    long_ctl = create_control(convergence_type="mittal", tolerance=0.0, max_iterations=300)
    small = [_cases.logistic(K=400, seed=40 + i) for i in range(3)]
    print(f"[case] server long launch K={big.K} J={big.J} + {len(small)} short requests", flush=True)
    srv = E.Server(slots=2, hbm_budget_bytes=8e9)
    try:
        lid = srv.submit(big, create_prior("laplace", 0.1), long_ctl, precision=_abi.FP32)
        time.sleep(0.5)
        sids = [srv.submit(d, create_prior("laplace", 1.0), create_control()) for d in small]
        short = [srv.poll(i, wait=True) for i in sids]
        _, lr_, ls = srv.poll(lid, wait=True)
    finally:
        srv.stop()
    assert lr_["iterations"] == 300
    assert ls["end"] - ls["start"] >= 1.0, ls   # the launch is long enough to tell
    for _, r, st in short:
        assert r["return_flag"] == 0
        assert st["start"] >= ls["start"], (st, ls)
        assert st["end"] < ls["end"] - 0.5, (st, ls)


def test_warm_start_on_many_sparse_columns_is_charged_and_exact(E):
    """Starting values on more than 32 sparse columns: the admission footprint adds the
    row-major copy the fit builds for xb, and the served result equals the same fit
    run alone."""
    d, _ = D.make_config(0, scale=0.1)   # 10,000 rows, 100 columns
    idx = [j for j in range(d.J) if int(d.formats[j]) in (1, 2)]   # SPARSE / INDICATOR
    assert len(idx) > 32
    b0 = np.zeros(d.J)
    b0[idx] = np.random.default_rng(7).normal(0, 0.2, len(idx))
    prior, ctl = create_prior("normal", 1.0), create_control(convergence_type="mittal", tolerance=1e-8)
    print(f"[case] server warm start on {len(idx)} sparse columns, K={d.K}", flush=True)
    eng = E.Engine(d, 64)
    try:
        b_alone, r_alone = eng.fit(prior, ctl, beta0=b0, mode="resident")
    finally:
        eng.close()
    srv = E.Server(slots=2, hbm_budget_bytes=8e9)
    try:
        cold = srv.poll(srv.submit(d, prior, ctl), wait=True)
        b, r, st = srv.poll(srv.submit(d, prior, ctl, beta0=b0), wait=True)
    finally:
        srv.stop()
    nnz = sum(int(d.col_ptr[j + 1] - d.col_ptr[j]) for j in idx)
    assert st["footprint_bytes"] - cold[2]["footprint_bytes"] == (d.K + 1) * 8 + nnz * (4 + 8)
    assert np.array_equal(b, b_alone)
    assert r["iterations"] == r_alone["iterations"] and r["log_likelihood"] == r_alone["log_likelihood"]


def _block(footprint):   # device block the server reserves for a request
    return ((int(footprint * 1.25) + (1 << 20)) + 255) & ~255


def test_budget_admits_one_at_a_time(E):
    mix = _mix()[1:5]
    srv = E.Server(slots=4, hbm_budget_bytes=8e9)
    blocks = sorted(_block(srv.poll(srv.submit(d, p, c), wait=True)[2]["footprint_bytes"]) for d, p, c in mix)
    srv.stop()
    assert blocks[-1] < blocks[0] + blocks[1]
    budget = (blocks[-1] + blocks[0] + blocks[1]) // 2   # holds any one request, no two
    print(f"[case] server budget {budget} bytes for blocks {blocks}", flush=True)
    srv = E.Server(slots=4, hbm_budget_bytes=budget)
    ids = [srv.submit(d, p, c) for d, p, c in mix]
    stats = [srv.poll(i, wait=True)[2] for i in ids]
    srv.stop()
    assert not _overlap(stats)   # any two together exceed the budget
    assert all(s["start"] >= s["submit"] for s in stats)


def test_over_budget_request_is_refused(E):
    with pytest.raises(E.CyclopsError):
        E.Server(slots=2, hbm_budget_bytes=1e4)   # not even one request fits
    srv = E.Server(slots=2, hbm_budget_bytes=4e6)
    with pytest.raises(E.CyclopsError):
        srv.submit(_cases.logistic(K=200_000, seed=3))
    with pytest.raises(E.CyclopsError):
        srv.poll(12345)
    srv.stop()


def test_poll_without_wait(E):
    srv = E.Server(slots=1, hbm_budget_bytes=8e9)
    rid = srv.submit(_cases.logistic(K=500, seed=4), create_prior("laplace", 1.0), create_control())
    out = srv.poll(rid, wait=False, release=False)
    assert out is None or out[1]["return_flag"] == 0
    b, r, st = srv.poll(rid, wait=True)
    assert r["return_flag"] == 0 and st["end"] >= st["start"] >= st["submit"]
    srv.stop()


def test_finished_requests_are_released(E):
    """ccd_server_release forgets a finished request: polling it afterwards is an unknown
    id; a request that still runs is refused with CCD_ERR_STATE."""
    import ctypes as C
    srv = E.Server(slots=1, hbm_budget_bytes=8e9)
    L = E.lib()
    long_ctl = create_control(convergence_type="mittal", tolerance=1e-12, max_iterations=400)
    slow = srv.submit(_cases.poisson(K=60_000, seed=8), create_prior("normal", 1.0), long_ctl)
    quick = srv.submit(_cases.logistic(K=500, seed=5), create_prior("laplace", 1.0), create_control())
    rc = L.ccd_server_release(srv.h, C.c_int64(quick))   # queued behind the slow one (one slot)
    assert rc in (_abi.ERR_STATE, 0)
    if rc == _abi.ERR_STATE:
        b, r, _ = srv.poll(quick, wait=True)   # reads and releases
        assert r["return_flag"] == 0
    with pytest.raises(E.CyclopsError, match="unknown request id"):
        srv.poll(quick, wait=True)
    srv.poll(slow, wait=True)
    srv.stop()


def test_packed_requests_equal_their_own_launches(E):
    """With pack > 1, queued requests of one model and precision whose resident fit is
    one workgroup share a persistent launch (ccd_fit_multi). Each result equals the
    result of its own launch bit for bit, members of a group start and end together,
    and a request of another model is never packed with them."""
    ctl = create_control(convergence_type="mittal", tolerance=1e-8)
    # a fit of 300 cycles (no convergence test) that holds the one slot for about a second
    long_ctl = create_control(convergence_type="mittal", tolerance=0.0, max_iterations=300)
    blocker = (D.make_request(2, max_rows=100_000, max_cols=10_000)["data"], create_prior("laplace", 0.1), long_ctl)
    reqs = [(_cases.logistic(K=500 + 300 * i, seed=60 + i), create_prior("laplace", 0.5 + 0.1 * i), ctl)
            for i in range(6)]
    reqs.insert(3, (_cases.poisson(K=900, seed=70), create_prior("normal", 1.0), ctl))
    print(f"[case] server packing: {len(reqs)} small requests behind a blocker, pack 4, one slot", flush=True)
    srv = E.Server(slots=1, hbm_budget_bytes=8e9)
    alone = [srv.poll(srv.submit(d, p, c), wait=True) for d, p, c in reqs]
    srv.stop()
    srv = E.Server(slots=1, hbm_budget_bytes=8e9, pack=4)
    first = srv.submit(*blocker)   # holds the slot while the others queue
    ids = [srv.submit(d, p, c) for d, p, c in reqs]
    packed = [srv.poll(i, wait=True) for i in ids]
    srv.poll(first, wait=True)
    srv.stop()
    for i, ((b1, r1, _), (b2, r2, _)) in enumerate(zip(alone, packed)):
        assert np.array_equal(b1, b2), i
        assert (r1["iterations"], r1["return_flag"], r1["log_likelihood"]) == \
               (r2["iterations"], r2["return_flag"], r2["log_likelihood"]), i
    starts = [round(p[2]["start"], 9) for p in packed]
    lr_starts = [s for s, (d, _, _) in zip(starts, reqs) if d.model_type == _abi.LOGISTIC]
    assert max(lr_starts.count(s) for s in lr_starts) >= 2, starts   # some logistic requests shared a launch
    pr_start = starts[3]
    assert pr_start not in lr_starts, starts                          # the Poisson one ran on its own
    for s in set(lr_starts):
        ends = {round(p[2]["end"], 9) for p, st in zip(packed, starts) if st == s}
        assert len(ends) == 1
