"""GPU parity of the trapezoidal, TR-BDF2 and adaptive transients
(transient/solver.rs:41-249, 461-676; companion.rs:41-150) against the
oracle's restatement on identical seeded parameters, through the C ABI, on
every kernel kind (register LU, shared-symbolic sparse with 64 and 256
threads, one lane per instance).  Bar: waveforms, times and final
solutions within 1e-9 relative (1e-12 absolute for waveforms); status, flags, iteration
counts, point counts and step counts equal.  Every instance is compared.
Adaptive grids branch on LTE comparisons, and FP32+IR Newton solves on dx
tests, so an instance whose grid or iteration count differs from the
oracle's by a rounding-level decision (at most 0.1 % of instances) is
compared at the SPICE tolerance (reltol 1e-3, vntol 1 uV) instead: its final
solution, and its waveforms interpolated onto the oracle's time grid."""
import math

import numpy as np
import pytest

from spicier_amd import _abi as A
from spicier_amd import circuit as CK
from spicier_amd import engine as E

pytestmark = pytest.mark.gpu

M = E.IntegrationMethod
RTOL, ATOL = 1e-9, 1e-12


def close(a, b, rtol=RTOL, atol=ATOL):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b) <= np.maximum(atol, rtol * np.abs(b))


def _cfg():
    return E.nr_config(limiter=E.Limiter.Pnjlim)


def _ocfg():
    return A.default_nr_config(limiter=A.SPX_LIMIT_PNJLIM)


def _check_fixed(got, want):
    np.testing.assert_array_equal(got.status, want["status"])
    np.testing.assert_array_equal(got.fail_step, want["fail_step"])
    np.testing.assert_array_equal(got.flags, want["flags"])
    np.testing.assert_array_equal(got.iterations, want["iterations"])
    ok = close(got.waveforms, want["waveforms"])
    assert np.all(ok), (np.argwhere(~ok)[:5], np.abs(got.waveforms - want["waveforms"]).max())
    assert np.all(close(got.solutions, want["x"]))


def spice_close(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b) <= 1e-3 * np.abs(b) + 1e-6


def _check_adaptive(got, want):
    np.testing.assert_array_equal(got.status, want["status"])
    same = ((got.num_points == want["num_points"]) & (got.total_steps == want["total_steps"])
            & (got.rejected_steps == want["rejected_steps"]))
    assert same.mean() >= 0.999, np.count_nonzero(~same)
    for i in np.flatnonzero(~same):  # another grid: compared at the SPICE tolerance
        n, m = got.num_points[i], want["num_points"][i]
        assert got.point_times[i, n - 1] == want["times"][i, m - 1]
        assert np.all(spice_close(got.solutions[i], want["x"][i])), i
        for p in range(got.waveforms.shape[2]):
            w = np.interp(want["times"][i, :m], got.point_times[i, :n], got.waveforms[i, :n, p])
            assert np.all(spice_close(w, want["waveforms"][i, :m, p])), (i, p)
    for i in np.flatnonzero(same):
        n = got.num_points[i]
        # step sizes scaled by sqrt(tol/lte) carry the LTE's rounding (the
        # device fuses multiply-adds), and t accumulates them
        assert np.all(close(got.point_times[i, :n], want["times"][i, :n], atol=0)), i
        assert np.all(close(got.waveforms[i, :n], want["waveforms"][i, :n])), i
        assert close(got.min_step[i], want["min_step"][i], atol=0)
        assert close(got.max_step[i], want["max_step"][i], atol=0)
    np.testing.assert_array_equal(got.iterations[same], want["iterations"][same])
    np.testing.assert_array_equal(got.flags[same], want["flags"][same])
    np.testing.assert_array_equal(got.fail_step[same], want["fail_step"][same])
    assert np.all(close(got.solutions[same], want["x"][same]))


@pytest.mark.parametrize("method", [M.Trapezoidal, M.TrBdf2])
@pytest.mark.parametrize("n", [8, 32, 64, 128, 256])
def test_ladder_methods_match_oracle(oracle, n, method):
    c = CK.ladder(n, seed=n)
    B = 32 if n <= 64 else 12
    h, steps, probes = 1e-12, 100, [0, n // 2, n - 2]
    got = E.transient_sweep(c, B, h, steps, probes=probes, seed=7, config=_cfg(), method=method)
    want = oracle.transient_ex(c, got.params, _ocfg(), h, steps, probes=probes,
                               method=int(method), nthreads=16)
    _check_fixed(got, want)
    assert np.all(got.accepted_mask())


def test_reference_rc_tests_on_gpu(oracle):
    # transient/mod.rs:100-265, 387-430 through the GPU path
    R, Cap, V = 1e3, 1e-6, 5.0
    c = CK.rc_step(R, Cap, V)
    out = c.node_names["out"] - 1
    v_tau = V * (1 - math.exp(-1))
    for method, fin, at_tau in ((M.BackwardEuler, 0.05, 0.2), (M.Trapezoidal, 0.05, 0.1),
                                (M.TrBdf2, 0.15, 0.6)):
        got = E.transient_sweep(c, 2, 10e-6, 500, probes=[out], seed=1, method=method)
        w = got.waveforms[0, :, 0]
        assert abs(w[-1] - 5.0) < fin and abs(w[100] - v_tau) < at_tau, method
        want = oracle.transient_ex(c, got.params, A.default_nr_config(), 10e-6, 500,
                                   probes=[out], method=int(method))
        _check_fixed(got, want)
    ad = E.AdaptiveTransientParams(5e-3, 1e-7, 1e-9, 1e-4, 1e-3, 1e-6, 4096)
    got = E.transient_sweep(c, 2, 0.0, 0, probes=[out], seed=1, adaptive=ad)
    n = got.num_points[0]
    assert abs(got.waveforms[0, n - 1, 0] - 5.0) < 0.05
    assert got.total_steps[0] < 200 and got.max_step[0] > 10 * ad.h_init
    assert got.point_times[0, n - 1] == ad.tstop
    want = oracle.transient_ex(c, got.params, A.default_nr_config(), probes=[out], adaptive=ad)
    _check_adaptive(got, want)


@pytest.mark.parametrize("n", [8, 48, 64, 128])
def test_adaptive_ladders_match_oracle(oracle, n):
    c = CK.ladder(n, seed=3 * n)
    B = 32 if n <= 64 else 12
    probes = [0, n // 2, n - 2]
    ad = E.AdaptiveTransientParams(tstop=1e-9, h_init=1e-13, h_min=1e-16, h_max=2e-11,
                                   reltol=1e-5, abstol=1e-8, max_points=3000)
    got = E.transient_sweep(c, B, 0.0, 0, probes=probes, seed=5, config=_cfg(), adaptive=ad)
    want = oracle.transient_ex(c, got.params, _ocfg(), probes=probes, adaptive=ad, nthreads=16)
    _check_adaptive(got, want)
    assert np.all(got.status == A.SPX_CONVERGED)
    assert np.all(got.point_times[np.arange(B), got.num_points - 1] == ad.tstop)
    assert np.all(got.rejected_steps > 0)  # the 10 ps edge forces step cuts


def test_adaptive_capacity_stops_instance(oracle):
    c = CK.ladder(16, seed=2)
    ad = E.AdaptiveTransientParams(tstop=1e-9, h_init=1e-13, h_min=1e-16, h_max=2e-11,
                                   reltol=1e-3, abstol=1e-6, max_points=20)
    got = E.transient_sweep(c, 8, 0.0, 0, probes=[0, 8], seed=1, config=_cfg(), adaptive=ad)
    want = oracle.transient_ex(c, got.params, _ocfg(), probes=[0, 8], adaptive=ad)
    assert np.all(got.status == A.SPX_FAILED) and np.all(got.num_points == 20)
    assert np.all(got.fail_step == 20)
    _check_adaptive(got, want)


@pytest.mark.parametrize("kind", [E.KernelKind.Dense, E.KernelKind.Sparse64,
                                  E.KernelKind.Sparse256, E.KernelKind.Lanes])
def test_every_kernel_kind_with_methods(oracle, kind):
    ad = E.AdaptiveTransientParams(tstop=5e-10, h_init=1e-13, h_min=1e-16, h_max=2e-11,
                                   reltol=1e-5, abstol=1e-8, max_points=2000)
    cls = [E.TranClass(CK.ladder(40, seed=4), 64, 1e-12, 60, probes=[0, 10, 38], seed=3,
                       method=M.TrBdf2),
           E.TranClass(CK.ladder(40, seed=6), 64, 0.0, 0, probes=[0, 10, 38], seed=9,
                       adaptive=ad)]
    got_kinds = E.transient_mixed_device(cls, _cfg(), graph=False, kernels=[kind, kind])
    E.Device.get(0).synchronize()
    assert got_kinds == [kind, kind]
    g0, g1 = cls[0].host_result(), cls[1].host_result()
    _check_fixed(g0, oracle.transient_ex(cls[0].circuit, g0.params, _ocfg(), 1e-12, 60,
                                         probes=cls[0].probes, method=int(M.TrBdf2),
                                         nthreads=16))
    _check_adaptive(g1, oracle.transient_ex(cls[1].circuit, g1.params, _ocfg(),
                                            probes=cls[1].probes, adaptive=ad, nthreads=16))


def test_graph_replays_with_methods_equal_eager():
    import torch
    ad = E.AdaptiveTransientParams(tstop=3e-10, h_init=1e-13, h_min=1e-16, h_max=2e-11,
                                   reltol=1e-5, abstol=1e-8, max_points=1000)
    cls = [E.TranClass(CK.ladder(24, seed=1), 40, 1e-12, 40, probes=[0, 12], seed=2,
                       method=M.Trapezoidal),
           E.TranClass(CK.ladder(96, seed=2), 20, 0.0, 0, probes=[0, 50], seed=4, adaptive=ad)]
    E.transient_mixed_device(cls, _cfg(), graph=False)
    E.Device.get(0).synchronize()
    ref = [{k: v.clone() for k, v in c.buffers.items()} for c in cls]
    for rep in range(3):
        for c in cls:
            for v in c.buffers.values():
                v.fill_(-7)
        torch.cuda.synchronize()
        E.transient_mixed_device(cls, _cfg(), graph=True)
        E.Device.get(0).synchronize()
        for c, r in zip(cls, ref):
            for k in r:
                if k in ("waveforms", "times"):  # rows past num_points are not written
                    continue
                assert torch.equal(c.buffers[k], r[k]), (rep, c.circuit.name, k)
        n = cls[1].buffers["num_points"]
        for i in range(n.numel()):
            m = int(n[i])
            assert torch.equal(cls[1].buffers["times"][i, :m], ref[1]["times"][i, :m])
            assert torch.equal(cls[1].buffers["waveforms"][i, :m], ref[1]["waveforms"][i, :m])
        assert torch.equal(cls[0].buffers["waveforms"], ref[0]["waveforms"])


@pytest.mark.parametrize("method", [M.Trapezoidal, M.TrBdf2])
def test_methods_ir_match_fp64_oracle(oracle, method):
    c = CK.ladder(32, seed=32)
    cfg = E.nr_config(limiter=E.Limiter.Pnjlim, precision=E.Precision.F32_IR)
    got = E.transient_sweep(c, 48, 1e-12, 100, seed=7, config=cfg, method=method)
    want = oracle.transient_ex(c, got.params, _ocfg(), 1e-12, 100, method=int(method),
                               nthreads=16)
    np.testing.assert_array_equal(got.status, want["status"])
    same = got.iterations == want["iterations"]
    assert same.mean() >= 0.999, np.count_nonzero(~same)
    assert np.all(close(got.waveforms[same], want["waveforms"][same]))
    assert np.all(spice_close(got.waveforms[~same], want["waveforms"][~same]))


def test_invalid_transient_settings():
    c = CK.ladder(8, seed=1)
    with pytest.raises(E.InvalidArgument):
        E.transient_sweep(c, 2, 1e-12, 5, method=7)
    for bad in (dict(h_max=1e-17), dict(h_min=0.0), dict(max_points=0), dict(reltol=-1.0),
                dict(tstop=float("inf"))):
        kw = dict(tstop=1e-9, h_init=1e-13, h_min=1e-16, h_max=2e-11, reltol=1e-3, abstol=1e-6,
                  max_points=100)
        kw.update(bad)
        with pytest.raises(E.InvalidArgument):
            E.transient_sweep(c, 2, 0.0, 0, adaptive=E.AdaptiveTransientParams(**kw))


def test_register_lu_pivot_ignores_negative_zero(oracle):
    # Trapezoidal matrices of a 64-unknown RC ladder at h = 1e-16 s: the
    # elimination's fma(-u, l, +0.0) underflows to -0.0 in a later column
    # for three of these four instances; the pivot search must rank it as
    # zero (it once ranked it above every positive value and picked it).
    c = CK.ladder(64, seed=192)
    n = c.size
    params = E.sample_params(c, 4, seed=5)
    ckt, keep = c.to_c()
    mats, rhss = [], []
    for p in params:
        J, _ = oracle.stamp(c, p, np.zeros(n))
        J = np.array(J).reshape(n, n)
        for e in range(ckt.num_elements):
            el = ckt.elements[e]
            if el.kind == A.SPX_EL_CAPACITOR:
                g = 2 * p[el.param[0]] / 1e-16
                a, b = el.node[0] - 1, el.node[1] - 1
                if a >= 0:
                    J[a, a] += g
                if b >= 0:
                    J[b, b] += g
                if a >= 0 and b >= 0:
                    J[a, b] -= g
                    J[b, a] -= g
        rhs = np.zeros(n)
        rhs[n - 1] = 1.8e-5
        mats.append(J)
        rhss.append(rhs)
    solver = E.BackendSelector().create_solver()
    cols = np.ascontiguousarray(np.stack(mats).transpose(0, 2, 1)).ravel()
    res = solver.solve_batch(cols, np.concatenate(rhss), n, 4)
    assert res.singular_indices == []
    x = np.asarray(res.solutions).reshape(4, n)
    for i in range(4):
        want = np.linalg.solve(mats[i], rhss[i])
        np.testing.assert_allclose(x[i], want, rtol=1e-12, atol=1e-20)
    # and the transient that met it
    got = E.transient_sweep(c, 4, 1e-16, 3, probes=[0, 62], seed=5, config=_cfg(),
                            method=M.Trapezoidal)
    assert np.all(got.status == A.SPX_CONVERGED)
