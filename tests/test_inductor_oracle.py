"""Pins the oracle's inductor (passive.rs:283-375, InductorState in
transient/companion.rs:158-302, solver.rs:60-75 and 612-617): the DC short
with its branch current, the reference's LC-oscillation test
(transient/mod.rs:451-540), and what the mathematics fixes: the
Backward-Euler and trapezoidal RL step responses in closed form, the
trapezoidal LC's conserved energy and exact phase advance 2 atan(w h / 2),
Backward Euler's energy decay, the TR-BDF2 steady state, and the adaptive
solver's accuracy against the exact exponential."""
import math

import numpy as np
import pytest

from spicier_amd import _abi as A
from spicier_amd import circuit as CK

from test_transient_methods_oracle import Adaptive

R, L, V = 100.0, 1e-3, 5.0
TAU = L / R


def _rl(v=V, r=R, ind=L):
    """PULSE 0 -> v at t = 0+ through R into L to ground: i(t) = v/r (1 - e^(-t r / L))."""
    c = CK.Circuit("rl_step")
    c.vpulse("in", "0", 0.0, v, 0.0, 0.0, 0.0, 1e6, 0.0)
    c.resistor("in", "mid", r)
    c.inductor("mid", "0", ind)
    return c


def _lc(v0=5.0, ind=1e-3, cap=1e-6):
    """The reference's LC tank (mod.rs:451-470): C charged to v0 by a source
    in series with L that drops to 0 V at t = 0+, so that from the first
    step on C and L form a closed loop with v_C(0) = v0, i_L(0) = 0."""
    c = CK.Circuit("lc_tank")
    c.capacitor("top", "0", cap)
    c.inductor("top", "bot", ind)
    c.vpulse("bot", "0", v0, 0.0, 0.0, 0.0, 0.0, 1e6, 0.0)
    return c


def _run(oracle, c, method, h, steps, adaptive=None):
    p = CK.nominal_values(c).reshape(1, -1)
    out = oracle.transient_ex(c, p, A.default_nr_config(), h, steps, probes=None,
                              method=method, adaptive=adaptive)
    assert out["status"][0] == A.SPX_CONVERGED
    return out


def _branch(c):
    return c.num_nodes + [e[0] for e in c.elements
                          if e[0] in (A.SPX_EL_VSOURCE, A.SPX_EL_VPULSE, A.SPX_EL_VSIN,
                                      A.SPX_EL_INDUCTOR)].index(A.SPX_EL_INDUCTOR)


def test_dc_short_reference_parser_circuit(oracle):
    # parser/mod.rs:640-652's circuit: V1 1 0 10, L1 1 2 1m, R1 2 0 100
    c = CK.Circuit("ind")
    c.vsource("1", "0", 10.0)
    c.inductor("1", "2", 1e-3)
    c.resistor("2", "0", 100.0)
    out = oracle.sweep_dcop(c, CK.nominal_values(c).reshape(1, -1))
    assert out["status"][0] == A.SPX_CONVERGED
    x = out["x"][0]
    assert x.shape == (4,)
    np.testing.assert_allclose(x[:2], [10.0, 10.0], rtol=1e-12)
    assert x[3] == pytest.approx(0.1, rel=1e-12)       # L1 carries 10 V / 100 Ohm
    assert x[2] == pytest.approx(-0.1, rel=1e-12)      # V1's branch current


def test_reference_lc_oscillation_test(oracle):
    # mod.rs:451-540: trapezoidal, 50 points per period, five periods;
    # frequency within 5 %, amplitude within 0.5 V of 5 V
    ind, cap = 1e-3, 1e-6
    f0 = 1.0 / (2 * math.pi * math.sqrt(ind * cap))
    T = 1.0 / f0
    c = _lc(5.0, ind, cap)
    h = T / 50.0
    steps = int(math.ceil(5 * T / h))
    w = _run(oracle, c, A.SPX_TRAN_TRAP, h, steps)["waveforms"][0, :, c.node_names["top"] - 1]
    t = np.arange(w.size) * h
    k = np.nonzero((w[:-1] > 0) & (w[1:] <= 0))[0]
    cross = t[k] - w[k] * (t[k + 1] - t[k]) / (w[k + 1] - w[k])
    assert cross.size >= 2
    assert abs(1.0 / (cross[1] - cross[0]) - f0) / f0 < 0.05
    assert abs((w.max() - w.min()) / 2 - 5.0) < 0.5


@pytest.mark.parametrize("h", [1e-6, 1e-5, 1e-4])
def test_backward_euler_rl_closed_form(oracle, h):
    # i_k = i_k-1 + h/L (V - R i_k), i_0 = 0  =>  i_k = V/R (1 - (1 + hR/L)^-k)
    steps = 200
    c = _rl()
    out = _run(oracle, c, A.SPX_TRAN_BE, h, steps)
    i = out["waveforms"][0, :, _branch(c)]
    k = np.arange(steps + 1)
    np.testing.assert_allclose(i, V / R * (1 - (1 + h * R / L) ** (-k.astype(float))),
                               rtol=1e-11, atol=1e-15)
    vmid = out["waveforms"][0, 1:, c.node_names["mid"] - 1]
    np.testing.assert_allclose(vmid, V - R * i[1:], rtol=1e-10, atol=1e-13)


@pytest.mark.parametrize("h", [1e-6, 1e-5, 1e-4])
def test_trapezoidal_rl_closed_form(oracle, h):
    # (i_k - i_k-1) L/h = (v_k + v_k-1)/2 with v_0 = 0 (the DC short) and
    # v_k = V - R i_k: e_k = V/R - i_k = V/(R(1+b)) ((1-b)/(1+b))^(k-1), b = hR/(2L)
    steps = 200
    c = _rl()
    i = _run(oracle, c, A.SPX_TRAN_TRAP, h, steps)["waveforms"][0, :, _branch(c)]
    b = h * R / (2 * L)
    k = np.arange(1, steps + 1)
    e = V / (R * (1 + b)) * ((1 - b) / (1 + b)) ** (k - 1)
    np.testing.assert_allclose(i[1:], V / R - e, rtol=1e-11, atol=1e-14)
    assert i[0] == 0.0


def test_trapezoidal_lc_energy_and_phase(oracle):
    # from the first step on the trapezoidal LC conserves C v^2/2 + L i^2/2
    # exactly and every sample obeys v_k+1 + v_k-1 = 2 cos(theta) v_k with
    # theta = 2 atan(w h / 2)
    ind, cap = 1e-3, 1e-6
    w0 = 1.0 / math.sqrt(ind * cap)
    h = 2 * math.pi / w0 / 37.0
    c = _lc(5.0, ind, cap)
    out = _run(oracle, c, A.SPX_TRAN_TRAP, h, 400)
    v = out["waveforms"][0, :, c.node_names["top"] - 1]
    i = out["waveforms"][0, :, _branch(c)]
    E = 0.5 * cap * v ** 2 + 0.5 * ind * i ** 2
    assert np.ptp(E[1:]) < 1e-12 * E[1]
    th = 2 * math.atan(w0 * h / 2)
    res = v[3:] + v[1:-2] - 2 * math.cos(th) * v[2:-1]
    assert np.abs(res).max() < 1e-10 * 5.0


def test_backward_euler_lc_energy_decays(oracle):
    ind, cap = 1e-3, 1e-6
    h = 2 * math.pi * math.sqrt(ind * cap) / 40.0
    c = _lc(5.0, ind, cap)
    out = _run(oracle, c, A.SPX_TRAN_BE, h, 200)
    v = out["waveforms"][0, :, c.node_names["top"] - 1]
    i = out["waveforms"][0, :, _branch(c)]
    E = 0.5 * cap * v ** 2 + 0.5 * ind * i ** 2
    assert np.all(np.diff(E[1:]) < 0)
    # BE damps each step by |1 / (1 + i w h)|^2 = 1 / (1 + (w h)^2)
    wh2 = h * h / (ind * cap)
    np.testing.assert_allclose(E[2:] / E[1:-1], 1.0 / (1.0 + wh2), rtol=1e-9)


def test_trbdf2_rl_reaches_steady_state(oracle):
    # the reference's TR-BDF2 inductor update leaves v = 0, i = V/R fixed
    c = _rl()
    i = _run(oracle, c, A.SPX_TRAN_TRBDF2, TAU / 10, 1000)["waveforms"][0, :, _branch(c)]
    assert abs(i[-1] - V / R) < 1e-9 * V / R


def test_trbdf2_rl_diverges_for_stiff_steps(oracle):
    # the final update's history term reads v_prev twice (companion.rs:230-235):
    # the recursion's amplification exceeds 1 once h is a few L/R, and the
    # reference's current grows without bound where a consistent TR-BDF2
    # would settle at V/R
    c = _rl()
    i = _run(oracle, c, A.SPX_TRAN_TRBDF2, 10 * TAU, 400)["waveforms"][0, :, _branch(c)]
    assert abs(i[-1]) > 1e100
    i = _run(oracle, c, A.SPX_TRAN_TRBDF2, TAU, 400)["waveforms"][0, :, _branch(c)]
    assert abs(i[-1] - V / R) < 1e-4 * V / R  # h = L/R still settles


def test_trbdf2_rl_matches_scalar_restatement(oracle):
    # the stage formulas of InductorState for the scalar RL (companion.rs:197-262)
    g = 2.0 - math.sqrt(2.0)
    h, steps = 1e-5, 200
    al = (1 - g) / (g * (2 - g))
    h2, h1 = (1 - g) * h, g * h
    rho = h2 / h1
    den = 1 + 2 * rho
    a1, a2, b0 = (1 + rho) ** 2 / den, -rho * rho / den, (1 + rho) / den
    i, vp, ref = 0.0, 0.0, []
    for _ in range(steps):
        G1 = g * h / (2 * L)
        I1 = i + G1 * vp
        vg = (V / R - I1) / (1 / R + G1)
        ipp, i = i, i + G1 * (vg + vp)
        vp = vg
        G2 = b0 * h2 / L
        I2 = a1 * i + a2 * ipp
        vn = (V / R - I2) / (1 / R + G2)
        ref.append(G2 * vn + I2)
        i += h / L * ((1 + al) * vn - (1 + 2 * al) * vp + al * vp)
        vp = vn
    c = _rl()
    got = _run(oracle, c, A.SPX_TRAN_TRBDF2, h, steps)["waveforms"][0, 1:, _branch(c)]
    np.testing.assert_allclose(got, ref, rtol=1e-10, atol=1e-14)


def test_fixed_methods_converge_to_exact_rl(oracle):
    c = _rl()
    br = _branch(c)
    t = np.arange(101) * 1e-6
    exact = V / R * (1 - np.exp(-t / TAU))
    i_be = _run(oracle, c, A.SPX_TRAN_BE, 1e-7, 1000)["waveforms"][0, ::10, br]
    i_tr = _run(oracle, c, A.SPX_TRAN_TRAP, 1e-7, 1000)["waveforms"][0, ::10, br]
    # (the trapezoidal start averages the DC point's v = 0 into the first step:
    # an O(h) offset that decays with the response)
    assert np.abs(i_tr - exact)[1:].max() < 5e-3 * V / R
    assert np.abs(i_tr - exact)[50:].max() < 1e-4 * V / R
    assert np.abs(i_be - exact)[1:].max() < 5e-3 * V / R


def test_adaptive_rl_accuracy_and_inductor_lte(oracle):
    # the inductor's LTE drives the step control: an RL circuit alone (no
    # capacitor) still rejects and grows steps
    ad = Adaptive(5 * TAU, 1e-7, 1e-12, 2e-5, 1e-4, 1e-9, 100000)
    c = _rl()
    out = _run(oracle, c, A.SPX_TRAN_TRAP, 0.0, 0, adaptive=ad)
    n = out["num_points"][0]
    t = out["times"][0, :n]
    i = out["waveforms"][0, :n, _branch(c)]
    assert t[-1] == pytest.approx(ad.tstop, rel=1e-12)
    hs = np.diff(t)
    assert hs.max() > 10 * ad.h_init
    assert out["total_steps"][0] == (n - 1) + out["rejected_steps"][0]
    exact = V / R * (1 - np.exp(-t / TAU))
    assert np.abs(i - exact).max() < 1e-2 * V / R


def test_reference_transmission_line_dc(oracle):
    # parse_and_simulate.rs:597-638 and 641-680: a line is a short at DC; NL=3
    # gives three branch currents beside V1's
    from spicier_amd.netlist import parse_netlist
    c = parse_netlist("T DC\nV1 1 0 DC 10\nT1 1 0 2 0 Z0=50 TD=1n NL=3\nR1 2 0 100\n.end\n").circuit
    assert c.size - c.num_nodes == 4
    x = oracle.sweep_dcop(c, CK.nominal_values(c).reshape(1, -1))["x"][0]
    assert x[c.node_names["2"] - 1] == pytest.approx(10.0, rel=1e-12)
    c = parse_netlist("T matched\nV1 1 0 DC 10\nR_source 1 in 50\nT1 in 0 out 0 Z0=50 TD=5n NL=5\n"
                      "R_load out 0 50\n.end\n").circuit
    x = oracle.sweep_dcop(c, CK.nominal_values(c).reshape(1, -1))["x"][0]
    assert x[c.node_names["in"] - 1] == pytest.approx(5.0, rel=1e-12)
    assert x[c.node_names["out"] - 1] == pytest.approx(5.0, rel=1e-12)


def test_matched_line_delays_a_step(oracle):
    # a 50 Ohm line of TD = 5 ns, 50 sections, matched at both ends: the far
    # end stays near 0 until about TD, then settles at half the step
    # (trapezoidal, h = TD/200); the LC sections give L = Z0 TD / N and
    # C = TD / (Z0 N) per section
    from spicier_amd.netlist import parse_netlist
    c = parse_netlist("t\nV1 1 0 PULSE(0 1 0 0 0 1 0)\nRS 1 in 50\n"
                      "T1 in 0 out 0 Z0=50 TD=5n NL=50\nRL out 0 50\n.end\n").circuit
    ls = [CK.nominal_values(c)[e[2][0]] for e in c.elements if e[0] == A.SPX_EL_INDUCTOR]
    assert len(ls) == 50 and ls[0] == pytest.approx(50 * 5e-9 / 50)
    h, steps = 5e-9 / 200, 600
    out = _run(oracle, c, A.SPX_TRAN_TRAP, h, steps)
    w = out["waveforms"][0, :, c.node_names["out"] - 1]
    t = np.arange(steps + 1) * h
    assert np.abs(w[t < 0.8 * 5e-9]).max() < 0.02
    assert abs(w[-1] - 0.5) < 0.02
    k = np.argmax(w > 0.25)
    assert abs(t[k] - 5e-9) < 0.15 * 5e-9
