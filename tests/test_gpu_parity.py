"""GPU parity: the HIP engine against the CPU oracle on identical seeded
inputs, through the C ABI.  Tolerances (north star): FP64 solutions per node
within 1e-9 relative or 1e-12 absolute; integer/index results exact."""
import ctypes as C
import os

import numpy as np
import pytest

from spicier_amd import _abi as A
from spicier_amd import circuit as CK
from spicier_amd import engine as E

pytestmark = pytest.mark.gpu

RTOL, ATOL = 1e-9, 1e-12


def close(a, b, rtol=RTOL, atol=ATOL):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b) <= np.maximum(atol, rtol * np.abs(b))


@pytest.fixture(scope="module")
def solver():
    return E.HipBatchedSolver()


# ------------------------------------------------------------ BatchedLuSolver
def test_solve_batch_reference_known_answers(solver):
    # solver.rs:408-460, faer_solver.rs:119-192, batched_lu.rs:581-618
    r = solver.solve_batch([1, 0, 0, 1, 1, 0, 0, 1], [1, 2, 3, 4], 2, 2)
    assert r.singular_indices == [] and np.allclose(r.solutions, [1, 2, 3, 4], atol=1e-15)
    r = solver.solve_batch([2, 1, 1, 3], [5, 10], 2, 1)
    assert np.allclose(r.solution(0), [1, 3], atol=1e-12)
    r = solver.solve_batch([2, 1, 1, 3], [5, 5], 2, 1)
    assert np.allclose(r.solution(0), [2, 1], atol=1e-12)
    r = solver.solve_batch([1, 0, 0, 1, 1, 1, 2, 2], [1, 2, 1, 2], 2, 2)
    assert r.is_singular(1) and not r.is_singular(0) and r.num_solved() == 1
    assert np.all(r.solution(1) == 0)
    r = solver.solve_batch([], [], 3, 0)
    assert r.batch_size == 0 and r.solutions.size == 0
    with pytest.raises(E.InvalidDimension):
        solver.solve_batch([1.0, 2.0], [1.0], 2, 1)


@pytest.mark.parametrize("n", [1, 2, 5, 8, 13, 16, 31, 32, 33, 63, 64, 65, 100, 150])
def test_solve_batch_matches_oracle(solver, oracle, n):
    rng = np.random.default_rng(n)
    B = 300 if n <= 64 else 40
    m = rng.standard_normal((B, n, n)) + 0.2 * n * np.eye(n)
    # a few systems with exact pivot ties (MNA-like +-1 couplings) and a singular one
    m[1] = np.eye(n)
    if n >= 3:
        m[2, :, 0] = 0.0
        m[2, 0, 0] = 1.0
        m[2, 1, 0] = -1.0
        m[3] = 0.0
    b = rng.standard_normal((B, n))
    cm = np.ascontiguousarray(np.transpose(m, (0, 2, 1))).reshape(-1)  # column-major
    got = solver.solve_batch(cm, b.reshape(-1), n, B)
    want, sing, ns = oracle.solve_batch(cm, b.reshape(-1), n, B)
    assert got.singular_indices == [int(i) for i in np.nonzero(sing)[0]]
    assert np.all(close(got.solutions, want, rtol=1e-10, atol=1e-12))


def test_solve_batch_large_batch_splits_and_matches(solver, oracle):
    n, B = 6, 70000   # above the reference's 65,535 per-launch cap
    rng = np.random.default_rng(7)
    m = rng.standard_normal((B, n, n)) + 3 * np.eye(n)
    b = rng.standard_normal((B, n))
    cm = np.ascontiguousarray(np.transpose(m, (0, 2, 1))).reshape(-1)
    got = solver.solve_batch(cm, b.reshape(-1), n, B)
    want, sing, ns = oracle.solve_batch(cm, b.reshape(-1), n, B)
    assert ns == 0 and not got.singular_indices
    assert np.all(close(got.solutions, want, rtol=1e-10, atol=1e-12))


# ------------------------------------------------------------ parameter sampling
@pytest.mark.parametrize("gen", [A.SPX_GEN_HASH, A.SPX_GEN_CORNER, A.SPX_GEN_LINEAR])
def test_sample_params_match_oracle(oracle, gen):
    c = CK.inverter_chain() if gen == A.SPX_GEN_HASH else CK.cs_amp()
    B = 4096 if gen != A.SPX_GEN_CORNER else 2 ** len(c.varied_slots())
    got = E.sample_params(c, B, seed=42, generator=gen, sweep_offset=0)
    want = oracle.generate_params(c, B, seed=42, generator=gen)
    # uniform / corner / linear slots are exact IEEE arithmetic; Gaussian and
    # log-uniform slots go through exp/log/cos and agree to a few ulp
    assert np.all(close(got, want, rtol=4e-16, atol=0))
    exact = [i for i, s in enumerate(c.specs) if s[0] in (A.SPX_P_CONST, A.SPX_P_UNIFORM)]
    np.testing.assert_array_equal(got[:, exact], want[:, exact])


def test_sample_params_lcg_jump_ahead(oracle):
    c = CK.cs_amp(uniform_r=False)
    got = E.sample_params(c, 3000, seed=12345, generator=A.SPX_GEN_LCG, sweep_offset=17)
    want = oracle.generate_params(c, 3000, seed=12345, generator=A.SPX_GEN_LCG, sweep_offset=17)
    assert np.all(close(got, want, rtol=4e-16, atol=0))


# ------------------------------------------------------------ DC-OP sweeps
def _parity(got, want, tag):
    np.testing.assert_array_equal(got.status, want["status"], err_msg=tag)
    same_it = got.iterations == want["iterations"]
    assert same_it.mean() >= 0.999, (tag, np.count_nonzero(~same_it))
    ok = close(got.solutions, want["x"])
    # instances whose Newton stopped on a different iteration are compared at
    # the SPICE tolerance instead (a borderline Δx test)
    assert np.all(ok[same_it]), (tag, np.argwhere(~ok[same_it])[:5])
    if (~same_it).any():
        assert np.all(close(got.solutions[~same_it], want["x"][~same_it], rtol=1e-3, atol=1e-6))
    # the oracle has no FP32 path, so its flags never carry FP64_ESCALATED
    np.testing.assert_array_equal(got.flags & ~A.SPX_FLAG_FP64_ESCALATED, want["flags"],
                                  err_msg=tag)
    esc = np.count_nonzero(got.flags & A.SPX_FLAG_FP64_ESCALATED)
    assert esc <= 0.01 * len(got.flags), (tag, "FP32+IR escalated to FP64 on", esc)


@pytest.mark.parametrize("prec", [E.Precision.F64, E.Precision.F32_IR])
def test_cs_amp_sweep_matches_oracle(oracle, prec):
    c = CK.cs_amp()
    B = 1024
    cfg = E.nr_config(precision=prec)
    got = E.sweep_dcop(c, B, seed=42, config=cfg)
    params = oracle.generate_params(c, B, seed=42)
    assert np.all(close(got.params, params, rtol=4e-16, atol=0))
    want = oracle.sweep_dcop(c, got.params, A.default_nr_config(), nthreads=8)
    _parity(got, want, f"cs_amp {prec.name}")
    assert got.converged_count == B and np.all(got.residual <= 1)


@pytest.mark.parametrize("prec", [E.Precision.F64, E.Precision.F32_IR])
def test_inverter_chain_sweep_matches_oracle(oracle, prec):
    c = CK.inverter_chain()
    B = 2048
    got = E.sweep_dcop(c, B, seed=42, config=E.nr_config(precision=prec))
    want = oracle.sweep_dcop(c, got.params, A.default_nr_config(), nthreads=8)
    _parity(got, want, f"inverter_chain {prec.name}")
    assert np.all(got.accepted_mask())
    # node statistics over the converged instances
    st = oracle.stats_accumulate(got.solutions, got.status == A.SPX_CONVERGED)
    np.testing.assert_array_equal(got.node_stats[:, 0], st[:, 0])
    assert np.all(close(got.node_stats[:, 1:3], st[:, 1:3], rtol=1e-12, atol=1e-12))
    np.testing.assert_array_equal(got.node_stats[:, 3:], st[:, 3:])


def test_warm_start_parity(oracle):
    c = CK.inverter_chain()
    B = 512
    base = E.sweep_dcop(c, B)
    x0 = base.solutions + np.random.default_rng(3).normal(0, 2e-3, base.solutions.shape)
    for prec in (E.Precision.F64, E.Precision.F32_IR):
        got = E.sweep_dcop(c, B, config=E.nr_config(precision=prec), x0=x0)
        cfg = A.default_nr_config(warm_start=A.SPX_WARM_GIVEN)
        want = oracle.sweep_dcop(c, got.params, cfg, x0=x0, nthreads=8)
        _parity(got, want, f"warm {prec.name}")
        assert np.all(got.flags & A.SPX_FLAG_WARM_START)
        assert got.iterations.mean() < base.iterations.mean()


@pytest.mark.parametrize("limiter", [E.Limiter.Reference, E.Limiter.Pnjlim])
def test_diode_limiters_match_oracle(oracle, limiter):
    c = CK.Circuit("diode_mc")
    v = c.uniform(0.8, 5.0, name="V")
    r = c.loguniform(100.0, 10e3, name="R")
    isat = c.loguniform(1e-15, 1e-13, name="IS")
    c.vsource("in", "0", v)
    c.resistor("in", "a", r)
    c.diode("a", "0", isat, 1.0)
    B = 1024
    cfg = E.nr_config(limiter=limiter)
    got = E.sweep_dcop(c, B, config=cfg)
    want = oracle.sweep_dcop(c, got.params, A.default_nr_config(limiter=int(limiter)), nthreads=8)
    _parity(got, want, f"diode {limiter.name}")


def test_stepping_fallbacks_match_oracle(oracle):
    # a 2-iteration budget makes plain NR fail so Gmin, then source stepping run
    c = CK.cs_amp()
    B = 256
    for kw in ({"gmin_stepping": True, "source_stepping": True},
               {"gmin_stepping": False, "source_stepping": True}):
        cfg = E.nr_config(E.ConvergenceCriteria(max_iterations=3), **kw)
        got = E.sweep_dcop(c, B, config=cfg)
        ocfg = A.default_nr_config(max_iterations=3, gmin_stepping=int(kw["gmin_stepping"]))
        want = oracle.sweep_dcop(c, got.params, ocfg, nthreads=8)
        _parity(got, want, str(kw))
        assert np.all(got.flags & (A.SPX_FLAG_GMIN_STEPPING | A.SPX_FLAG_SOURCE_STEPPING))
    # every level fails: each failed level counts all its attempts
    # (newton.rs:300-315, 411-426), 1 + 5 + 5 x 5 = 31 iterations
    got = E.sweep_dcop(c, B, config=E.nr_config(E.ConvergenceCriteria(max_iterations=1)))
    want = oracle.sweep_dcop(c, got.params, A.default_nr_config(max_iterations=1), nthreads=8)
    np.testing.assert_array_equal(got.status, want["status"])
    np.testing.assert_array_equal(got.iterations, want["iterations"])
    np.testing.assert_array_equal(got.iterations, 31)


def test_linear_sweep_single_solve(oracle):
    # the reference's voltage-divider sweep (sweep.rs:286-394)
    c = CK.voltage_divider()
    got = E.sweep_dcop(c, 100, seed=42)
    want = oracle.sweep_dcop(c, got.params, A.default_nr_config())
    np.testing.assert_array_equal(got.iterations, 1)
    assert np.all(close(got.solutions, want["x"]))
    vout = got.node_voltages(1)
    assert 3 < vout.mean() < 7 and got.converged_count == 100   # sweep.rs:360-369
    st = got.statistics(1)
    assert st.count == 100 and abs(st.mean - vout.mean()) < 1e-12


def test_full_size_headline_properties():
    # 65,536 instances at n=64: every instance converged, passes the FP64 KCL
    # test, FP32+IR agrees with FP64 per node, and the statistics are the
    # moments of the returned solutions.
    c = CK.inverter_chain()
    B = 65536
    f64 = E.sweep_dcop(c, B, config=E.nr_config())
    ir = E.sweep_dcop(c, B, config=E.nr_config(precision=E.Precision.F32_IR))
    assert f64.converged_count == B and np.all(f64.accepted_mask())
    assert ir.converged_count == B and np.all(ir.accepted_mask())
    same = f64.iterations == ir.iterations
    assert same.mean() > 0.999
    assert np.all(close(ir.solutions[same], f64.solutions[same], rtol=1e-9, atol=1e-12))
    assert np.all(close(ir.solutions[~same], f64.solutions[~same], rtol=1e-3, atol=1e-6))
    np.testing.assert_array_equal(f64.node_stats[:, 0], B)
    assert np.allclose(f64.node_stats[:, 1], f64.solutions.sum(0), rtol=1e-12)


def test_fp32_handover_to_fp64_on_ill_conditioned_systems(oracle):
    # Nodes a and b joined by 1e4 S, each with only ~1e-9 S (plus a diode) to
    # ground: rounded to FP32, G + eps == G and the Jacobian is exactly
    # singular, so FP32 factors cannot serve and the instance must go to the
    # FP64 pass, come back flagged, and equal the FP64 oracle.
    c = CK.Circuit("fp32_singular")
    rleak = c.loguniform(0.9e9, 1.1e9, name="Rleak")
    c.isource("0", "a", 1e-6)
    c.resistor("a", "b", 1e-4)
    c.resistor("a", "0", rleak)
    c.resistor("b", "0", rleak)
    c.diode("b", "0", 1e-14, 1.0)
    B = 256
    got = E.sweep_dcop(c, B, config=E.nr_config(precision=E.Precision.F32_IR))
    want = oracle.sweep_dcop(c, got.params, A.default_nr_config(), nthreads=8)
    np.testing.assert_array_equal(got.status, want["status"])
    # the zero-start Jacobian has condition ~1e13, so two FP64 LUs that round
    # differently (fused vs separate multiply-add) agree to ~1e13 * 2^-53
    assert np.all(close(got.solutions, want["x"], rtol=1e-6, atol=1e-9))
    np.testing.assert_array_equal(got.flags & ~A.SPX_FLAG_FP64_ESCALATED, want["flags"])
    assert np.count_nonzero(got.flags & A.SPX_FLAG_FP64_ESCALATED) == B
