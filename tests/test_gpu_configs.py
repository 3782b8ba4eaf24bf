"""GPU parity on the benchmark generators and the reference's own fixtures.

- Config 1 (Poisson with offset, Zipf columns, Normal prior): the resident single fit
  and a 16-fit batch in FP64 against independent oracle fits (same cycles,
  coefficients within 1e-6, log-likelihood within 1e-9), and FP32 bulk against the
  FP64 oracle.
- Configs 0 (full size, L1 logistic), 2 (Cox) and 3 (CLR) likewise.
- The reference's fixture files through the HIP engine: the 17-variable logistic file
  against its stored SAS log-likelihood (test-largeBernoulli.R:20) and the oracle,
  infert and test1-clr (conditional logistic) and cox-tiny against the oracle.

FP32 bar (bulk arithmetic in FP32, prior step, reductions and convergence test in
FP64): every coefficient within 1e-4 relative of the FP64 oracle's, where relative
means |b32 - b64| / max(|b64|, floor) with floor = 1% of the largest |b64| (near-zero
coefficients are held to 1e-6 of the largest coefficient instead), and the
log-likelihood within 1e-7 relative.
"""
import os

import numpy as np
import pytest

import oracle
from cyclops_amd import _abi
from cyclops_amd import data as D
from cyclops_amd.control import create_control, create_prior

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FP32_COEF = 1e-4
FP32_LOGLIK = 1e-7
FLOOR = 1e-2


@pytest.fixture(scope="module")
def E():
    from cyclops_amd import engine
    assert engine.device_count() > 0, "no HIP device"
    return engine


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-12)


def _coef_rel(b32, b64):
    floor = FLOOR * max(float(np.max(np.abs(b64))), 1e-12)
    return float(np.max(np.abs(b32 - b64) / np.maximum(np.abs(b64), floor)))


def _ll_rel(a, b):
    return abs(a - b) / abs(b)


def _assert_fp64(bg, rg, bo, ro, what):
    assert rg["return_flag"] == ro["return_flag"], (what, rg, ro)
    assert rg["iterations"] == ro["iterations"], (what, rg, ro)
    assert _rel(bg, bo) <= 1e-6, (what, _rel(bg, bo))
    assert _ll_rel(rg["log_likelihood"], ro["log_likelihood"]) <= 1e-9, (what, rg, ro)


def _assert_fp32(bg, rg, bo, ro, what):
    ce = _coef_rel(bg, bo)
    le = _ll_rel(rg["log_likelihood"], ro["log_likelihood"])
    print(f"{what}: FP32 coef rel err {ce:.2e} (floor {FLOOR:g} x max|b|), loglik rel err {le:.2e}, "
          f"cycles {rg['iterations']} (FP64 oracle {ro['iterations']})", flush=True)
    assert ce <= FP32_COEF, (what, ce)
    assert le <= FP32_LOGLIK, (what, le)


# ---------------------------------------------------------------------------
# config 1: Poisson, offset, Zipf(1.1) columns, Normal prior
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def cfg1():
    d, _ = D.make_config(1, scale=0.03)   # 30,000 rows, 601 columns (offset + 600 Zipf indicators)
    return d


def test_config1_fp64_resident_matches_oracle(E, cfg1):
    print(f"[case] config 1 fp64 resident K={cfg1.K} J={cfg1.J}", flush=True)
    prior = create_prior("normal", 1.0)
    ctl = create_control(convergence_type="mittal", tolerance=1e-6)
    bo, ro = oracle.fit(cfg1, prior, ctl)
    for mode in ("resident", "plain"):
        bg, rg = E.fit_cyclops_model(cfg1, prior, ctl, precision=_abi.FP64, mode=mode)
        _assert_fp64(bg, rg, bo, ro, f"config 1 {mode}")


def _cfg1_batch():
    variances = [0.1, 0.5, 2.0, 10.0]
    var_list = [v for v in variances for _ in range(4)]
    held = [f for _ in variances for f in range(4)]
    return var_list, held


@pytest.mark.parametrize("precision", [_abi.FP64, _abi.FP32])
def test_config1_batch16_matches_oracle(E, cfg1, precision):
    """4 row folds x 4 Normal variances in one batch; every member against its own
    oracle fit (FP64: same cycles, 1e-6 / 1e-9; FP32: the FP32 bar)."""
    print(f"[case] config 1 batch of 16, precision {precision}", flush=True)
    rf = oracle.assign_folds(cfg1.K, None, 4, _abi.SELECT_BY_ROW, 1001)
    var_list, held = _cfg1_batch()
    tol = 1e-6 if precision == _abi.FP64 else 1e-8
    ctl = create_control(convergence_type="mittal", tolerance=tol)
    eng = E.Engine(cfg1, precision)
    try:
        betas, res = eng.fit_batch(var_list, "normal", ctl, held_out_fold=held, row_fold=rf)
    finally:
        eng.close()
    for i in range(len(var_list)):
        w = (rf != held[i]).astype(float)
        bo, ro = oracle.fit(cfg1, create_prior("normal", var_list[i]), ctl, weights=w)
        if precision == _abi.FP64:
            _assert_fp64(betas[i], res[i], bo, ro, f"config 1 member {i}")
        else:
            _assert_fp32(betas[i], res[i], bo, ro, f"config 1 member {i}")


def test_config1_fp32_resident_bar(E, cfg1):
    print(f"[case] config 1 fp32 resident", flush=True)
    prior = create_prior("normal", 1.0)
    ctl = create_control(convergence_type="mittal", tolerance=1e-8)
    bo, ro = oracle.fit(cfg1, prior, ctl)
    bg, rg = E.fit_cyclops_model(cfg1, prior, ctl, precision=_abi.FP32, mode="resident")
    _assert_fp32(bg, rg, bo, ro, "config 1 resident")


def test_stepped_batch_equals_fit_batch(E, cfg1):
    """The stepped interface gives fit_batch's coefficients and results bit for bit,
    whatever the cycles per step; done flags only ever turn on."""
    print("[case] stepped batch vs fit_batch", flush=True)
    rf = oracle.assign_folds(cfg1.K, None, 4, _abi.SELECT_BY_ROW, 1001)
    var_list, held = _cfg1_batch()
    ctl = create_control(convergence_type="mittal", tolerance=1e-6)
    eng = E.Engine(cfg1, _abi.FP32)
    try:
        ref_b, ref_r = eng.fit_batch(var_list, "normal", ctl, held_out_fold=held, row_fold=rf)
        for per_step in (1, 7):
            run = eng.begin_batch(var_list, "normal", ctl, held_out_fold=held, row_fold=rf)
            lay = run.layout()
            assert lay["teams"] * lay["fits_per_team"] >= len(var_list)
            # the engine runs no other whole fit while the batch is open
            with pytest.raises(E.CyclopsError, match="stepped batch"):
                eng.fit(create_prior("normal", 1.0), ctl, mode="resident")
            prev = np.zeros(len(var_list), dtype=bool)
            running, steps = len(var_list), 0
            while running:
                running, done = run.step(per_step)
                assert np.all(done[prev]), "a finished fit came back"
                assert running == int(np.sum(~done))
                prev = done
                steps += 1
                assert steps * per_step <= 1000
            b, r = run.end()
            assert np.array_equal(b, ref_b), per_step
            for x, y in zip(r, ref_r):
                assert (x["iterations"], x["return_flag"], x["log_likelihood"]) == \
                       (y["iterations"], y["return_flag"], y["log_likelihood"])
        # a batch can be abandoned part-way, and the engine is usable again
        run = eng.begin_batch(var_list, "normal", ctl, held_out_fold=held, row_fold=rf)
        run.step(2)
        run.abandon()
        b2, _ = eng.fit_batch(var_list, "normal", ctl, held_out_fold=held, row_fold=rf)
        assert np.array_equal(b2, ref_b)
    finally:
        eng.close()


# ---------------------------------------------------------------------------
# config 0 at full size: L1 logistic, FP64, 100,000 x 1,001
# ---------------------------------------------------------------------------
def test_config0_full_size_matches_oracle(E):
    d, _ = D.make_config(0)
    print(f"[case] config 0 full size K={d.K} J={d.J}", flush=True)
    prior = create_prior("laplace", 0.1)
    ctl = create_control(convergence_type="mittal", tolerance=1e-6)
    bo, ro = oracle.fit(d, prior, ctl)
    for mode in ("resident", "plain"):
        bg, rg = E.fit_cyclops_model(d, prior, ctl, precision=_abi.FP64, mode=mode)
        _assert_fp64(bg, rg, bo, ro, f"config 0 {mode}")


# ---------------------------------------------------------------------------
# configs 2 (Cox) and 3 (CLR): FP64 parity and the FP32 bar
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("cfg,scale,prior", [(2, 0.02, ("laplace", 0.05)), (3, 0.01, ("laplace", 0.1))])
def test_grouped_configs_fp64_and_fp32(E, cfg, scale, prior):
    d, _ = D.make_config(cfg, scale=scale)
    print(f"[case] config {cfg} at scale {scale}: K={d.K} J={d.J}", flush=True)
    p = create_prior(*prior)
    ctl = create_control(convergence_type="mittal", tolerance=1e-6)
    bo, ro = oracle.fit(d, p, ctl)
    bg, rg = E.fit_cyclops_model(d, p, ctl, precision=_abi.FP64, mode="resident")
    _assert_fp64(bg, rg, bo, ro, f"config {cfg} fp64")
    ctl8 = create_control(convergence_type="mittal", tolerance=1e-8)
    bo8, ro8 = oracle.fit(d, p, ctl8)
    b32, r32 = E.fit_cyclops_model(d, p, ctl8, precision=_abi.FP32, mode="resident")
    _assert_fp32(b32, r32, bo8, ro8, f"config {cfg} fp32")


# ---------------------------------------------------------------------------
# the reference's fixture files through the HIP engine
# ---------------------------------------------------------------------------
TIGHT = create_control(convergence_type="mittal", tolerance=1e-10)


def _fixture(name):
    if name == "17var":
        return D.read_generic_text(os.path.join(GOLDEN, "CCD_LOGISTIC_TEST_17var.txt"), "lr", first_field="row")
    if name == "infert":
        return D.read_generic_text(os.path.join(GOLDEN, "infert_ccd.txt"), "clr", first_field="stratum")
    if name == "test1-clr":
        return D.read_csv_clr(os.path.join(GOLDEN, "test1-clr.txt"))
    return D.read_generic_text(os.path.join(GOLDEN, "cox-tiny-test.txt"), "cox", first_field="time", dense=True)


@pytest.mark.parametrize("mode", ["resident", "plain"])
@pytest.mark.parametrize("name", ["17var", "infert", "test1-clr", "cox-tiny"])
def test_reference_fixtures_on_device(E, name, mode):
    d = _fixture(name)
    print(f"[case] fixture {name} {mode} K={d.K} J={d.J}", flush=True)
    prior = create_prior("none")
    bo, ro = oracle.fit(d, prior, TIGHT)
    bg, rg = E.fit_cyclops_model(d, prior, TIGHT, precision=_abi.FP64, mode=mode)
    assert rg["return_flag"] == ro["return_flag"] == _abi.SUCCESS
    assert _rel(bg, bo) <= 1e-6, (bg, bo)
    assert _ll_rel(rg["log_likelihood"], ro["log_likelihood"]) <= 1e-9
    if name == "17var":   # SAS: -2 log-likelihood = 1578.046 (test-largeBernoulli.R:20)
        assert rg["log_likelihood"] * 2 == pytest.approx(-1578.046, rel=1e-4)
    if name == "infert":  # clogit estimates used by test-smallCLR.R
        assert bg == pytest.approx([1.98587, 1.40901], rel=1e-4)


def test_create_reports_unsupported_models_like_the_factory(E):
    """A model type the engine does not implement (here NORMAL = 0, least squares) is
    the factory's nullptr, reported as "Invalid model type." with CCD_ERR_UNSUPPORTED;
    a precision outside PrecisionType's FP64 / FP32 is an invalid argument."""
    import ctypes as C
    import _cases
    d = _cases.logistic()
    L = E.lib()
    h = C.c_void_p()
    desc = d.desc(_abi.FP64)
    desc.model_type = 0
    assert L.ccd_create(C.byref(desc), 0, C.byref(h)) == _abi.ERR_UNSUPPORTED
    assert L.ccd_last_error().decode() == "Invalid model type."
    desc = d.desc(2)   # PrecisionType::FP16
    assert L.ccd_create(C.byref(desc), 0, C.byref(h)) == _abi.ERR_INVALID
    desc = d.desc(_abi.FP32_BITS)   # the bit count is accepted
    assert L.ccd_create(C.byref(desc), 0, C.byref(h)) == 0
    L.ccd_destroy(h)
