"""GPU parity of inductors (passive.rs:283-375; InductorState,
transient/companion.rs:158-302; solver.rs:60-75, 612-617) against the
oracle on identical seeded parameters, through the C ABI.  The engine keeps
each inductor's branch current as an unknown in every stage (its row
i - G (v0 - v1) = I); the oracle stamps the reference's conductance
companion with no branch current and forms G v + I afterwards, so the two
solve different matrices of the same circuit.  Bar: waveforms, branch
currents and final solutions within 1e-9 relative (1e-12 absolute); status,
flags and iteration counts equal for linear circuits, waveforms alone for
the diode-clamped ladder (its Newton tests converge on a branch current the
oracle's stages do not carry)."""
import math

import numpy as np
import pytest

from spicier_amd import _abi as A
from spicier_amd import circuit as CK
from spicier_amd import engine as E

from test_transient_methods_gpu import _check_adaptive, _check_fixed, close

pytestmark = pytest.mark.gpu

M = E.IntegrationMethod


def _cfg():
    return E.nr_config(limiter=E.Limiter.Pnjlim)


def _ocfg():
    return A.default_nr_config(limiter=A.SPX_LIMIT_PNJLIM)


def _probes(c):
    # s1, m1, s2 and the first inductor's branch current
    return [c.node_names["s1"] - 1, c.node_names["m1"] - 1,
            c.node_names.get("s2", c.node_names["s1"]) - 1, c.num_nodes + 1]


def test_dc_sweep_inductor_short(oracle):
    # parser/mod.rs:640-652's circuit with R varied: L carries 10 V / R
    c = CK.Circuit("ind")
    c.vsource("1", "0", 10.0)
    c.inductor("1", "2", 1e-3)
    c.resistor("2", "0", c.uniform(50.0, 200.0, name="R"))
    got = E.sweep_dcop(c, 64, seed=3)
    want = oracle.sweep_dcop(c, got.params)
    np.testing.assert_array_equal(got.status, want["status"])
    np.testing.assert_array_equal(got.iterations, want["iterations"])
    assert np.all(close(got.solutions, want["x"]))
    r = got.params[:, c.slot_names.index("R")]
    np.testing.assert_allclose(got.solutions[:, 3], 10.0 / r, rtol=1e-12)
    np.testing.assert_allclose(got.solutions[:, 1], 10.0, rtol=1e-12)


@pytest.mark.parametrize("method", [M.BackwardEuler, M.Trapezoidal, M.TrBdf2])
@pytest.mark.parametrize("sections", [2, 20, 42])
def test_rlc_ladder_methods_match_oracle(oracle, sections, method):
    # 8, 62 (register LU) and 128 (shared-symbolic sparse) unknowns
    c = CK.rlc_ladder(sections, seed=sections)
    B = 32 if sections <= 20 else 12
    probes = _probes(c)
    got = E.transient_sweep(c, B, 1e-12, 150, probes=probes, seed=7, config=_cfg(),
                            method=method)
    want = oracle.transient_ex(c, got.params, _ocfg(), 1e-12, 150, probes=probes,
                               method=int(method), nthreads=16)
    _check_fixed(got, want)
    assert np.all(got.accepted_mask())


@pytest.mark.parametrize("sections", [2, 20, 42])
def test_adaptive_rlc_ladder_matches_oracle(oracle, sections):
    c = CK.rlc_ladder(sections, seed=3 * sections)
    B = 32 if sections <= 20 else 12
    probes = _probes(c)
    ad = E.AdaptiveTransientParams(tstop=3e-10, h_init=1e-13, h_min=1e-16, h_max=2e-11,
                                   reltol=1e-3, abstol=1e-6, max_points=3000)
    got = E.transient_sweep(c, B, 0.0, 0, probes=probes, seed=5, config=_cfg(), adaptive=ad)
    want = oracle.transient_ex(c, got.params, _ocfg(), probes=probes, adaptive=ad, nthreads=16)
    _check_adaptive(got, want)
    assert np.all(got.status == A.SPX_CONVERGED)
    assert np.all(got.rejected_steps > 0)


@pytest.mark.parametrize("method", [M.BackwardEuler, M.Trapezoidal, M.TrBdf2])
def test_diode_clamped_rlc_ladder(oracle, method):
    c = CK.rlc_ladder(8, seed=1, diodes=True)
    probes = _probes(c)
    got = E.transient_sweep(c, 32, 1e-12, 150, probes=probes, seed=2, config=_cfg(),
                            method=method)
    want = oracle.transient_ex(c, got.params, _ocfg(), 1e-12, 150, probes=probes,
                               method=int(method), nthreads=16)
    np.testing.assert_array_equal(got.status, want["status"])
    assert np.all(got.status == A.SPX_CONVERGED)
    # Newton stops within its tolerances (1e-3 relative, 1e-6 V) on both sides
    assert np.all(close(got.waveforms, want["waveforms"], rtol=1e-6, atol=1e-9))


@pytest.mark.parametrize("kind", [E.KernelKind.Dense, E.KernelKind.Sparse64,
                                  E.KernelKind.Sparse256, E.KernelKind.Lanes])
def test_every_kernel_kind_with_inductors(oracle, kind):
    ad = E.AdaptiveTransientParams(tstop=2e-10, h_init=1e-13, h_min=1e-16, h_max=2e-11,
                                   reltol=1e-3, abstol=1e-6, max_points=2000)
    c0, c1 = CK.rlc_ladder(12, seed=4), CK.rlc_ladder(12, seed=6)
    cls = [E.TranClass(c0, 64, 1e-12, 60, probes=_probes(c0), seed=3, method=M.BackwardEuler),
           E.TranClass(c1, 64, 0.0, 0, probes=_probes(c1), seed=9, adaptive=ad)]
    got_kinds = E.transient_mixed_device(cls, _cfg(), graph=False, kernels=[kind, kind])
    E.Device.get(0).synchronize()
    assert got_kinds == [kind, kind]
    g0, g1 = cls[0].host_result(), cls[1].host_result()
    _check_fixed(g0, oracle.transient_ex(c0, g0.params, _ocfg(), 1e-12, 60, probes=cls[0].probes,
                                         method=int(M.BackwardEuler), nthreads=16))
    _check_adaptive(g1, oracle.transient_ex(c1, g1.params, _ocfg(), probes=cls[1].probes,
                                            adaptive=ad, nthreads=16))


def test_reference_lc_oscillation_on_gpu(oracle):
    # transient/mod.rs:451-540: C at 5 V, L with no current, trapezoidal,
    # 50 points per period, five periods; frequency within 5 %, amplitude
    # within 0.5 V; the tank's energy stays constant from the first step on
    ind, cap = 1e-3, 1e-6
    c = CK.Circuit("lc_tank")
    c.capacitor("top", "0", cap)
    c.inductor("top", "bot", ind)
    c.vpulse("bot", "0", 5.0, 0.0, 0.0, 0.0, 0.0, 1e6, 0.0)
    f0 = 1.0 / (2 * math.pi * math.sqrt(ind * cap))
    h = 1.0 / f0 / 50.0
    top, br = c.node_names["top"] - 1, c.num_nodes  # L is the first branch
    got = E.transient_sweep(c, 2, h, 250, probes=[top, br], seed=1, method=M.Trapezoidal)
    w, i = got.waveforms[0, :, 0], got.waveforms[0, :, 1]
    t = np.arange(w.size) * h
    k = np.nonzero((w[:-1] > 0) & (w[1:] <= 0))[0]
    cross = t[k] - w[k] * (t[k + 1] - t[k]) / (w[k + 1] - w[k])
    assert cross.size >= 2 and abs(1.0 / (cross[1] - cross[0]) - f0) / f0 < 0.05
    assert abs((w.max() - w.min()) / 2 - 5.0) < 0.5
    energy = 0.5 * cap * w ** 2 + 0.5 * ind * i ** 2
    assert np.ptp(energy[1:]) < 1e-12 * energy[1]
    want = oracle.transient_ex(c, got.params, A.default_nr_config(), h, 250, probes=[top, br],
                               method=int(M.Trapezoidal))
    _check_fixed(got, want)


def test_fp32_ir_with_inductors_matches_fp64_oracle(oracle):
    c = CK.rlc_ladder(10, seed=10)
    cfg = E.nr_config(limiter=E.Limiter.Pnjlim, precision=E.Precision.F32_IR)
    probes = _probes(c)
    got = E.transient_sweep(c, 48, 1e-12, 100, probes=probes, seed=7, config=cfg,
                            method=M.Trapezoidal)
    want = oracle.transient_ex(c, got.params, _ocfg(), 1e-12, 100, probes=probes,
                               method=int(M.Trapezoidal), nthreads=16)
    np.testing.assert_array_equal(got.status, want["status"])
    assert np.all(close(got.waveforms, want["waveforms"]))


@pytest.mark.parametrize("method", [M.BackwardEuler, M.Trapezoidal, M.TrBdf2])
def test_transmission_line_netlist_on_gpu(oracle, method):
    # the reference's 50-section LC expansion of a matched 50 Ohm, 5 ns line
    # (103 unknowns: the shared-symbolic sparse path), source resistance varied
    from spicier_amd.netlist import parse_netlist
    c = parse_netlist("t\nV1 1 0 PULSE(0 1 0 0 0 1 0)\nRS 1 in {aunif(50, 5)}\n"
                      "T1 in 0 out 0 Z0=50 TD=5n NL=50\nRL out 0 50\n.end\n").circuit
    probes = [c.node_names["in"] - 1, c.node_names["out"] - 1, c.num_nodes + 1]
    h, steps = 5e-9 / 200, 400
    got = E.transient_sweep(c, 24, h, steps, probes=probes, seed=3, method=method)
    want = oracle.transient_ex(c, got.params, A.default_nr_config(), h, steps, probes=probes,
                               method=int(method), nthreads=16)
    _check_fixed(got, want)
