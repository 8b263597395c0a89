"""Sweeps of circuits read from SPICE netlists (spicier_amd/netlist.py) on
the GPU, against the oracle on the same sampled parameters, and against the
reference's own netlist checks (spicier-parser/tests/ngspice_validation.rs)."""
import numpy as np
import pytest

from spicier_amd import _abi as A
from spicier_amd import engine as E
from spicier_amd.netlist import parse_netlist

pytestmark = pytest.mark.gpu

RTOL, ATOL = 1e-9, 1e-12


def _close(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.all(np.abs(a - b) <= np.maximum(ATOL, RTOL * np.abs(b)))


def test_reference_diode_netlist(oracle):
    # ngspice_validation.rs:705-768: V(2) within 0.5..0.9 V, I = (5 - V)/1k in 4..5 mA
    p = parse_netlist("""
Diode Forward Bias
V1 1 0 DC 5
R1 1 2 1k
D1 2 0
.end
""")
    c = p.circuit
    assert c.size == 3 and c.nonlinear
    for cfg, acfg in ((E.nr_config(), A.default_nr_config()),
                      (E.nr_config(limiter=E.Limiter.Pnjlim),
                       A.default_nr_config(limiter=A.SPX_LIMIT_PNJLIM))):
        got = E.sweep_dcop(c, 4, seed=1, config=cfg)
        want = oracle.sweep_dcop(c, got.params, acfg, nthreads=4)
        np.testing.assert_array_equal(got.status, want["status"])
        assert _close(got.solutions, want["x"])
        v = got.solutions[:, 1]
        assert np.all((v > 0.5) & (v < 0.9)) and np.all(((5 - v) / 1e3 > 4e-3) & ((5 - v) / 1e3 < 5e-3))
    # with SPICE's junction limiter the KCL test accepts the Shockley root
    assert np.all(got.accepted_mask()) and abs(v[0] - 0.6928) < 5e-3


def test_monte_carlo_netlist_matches_oracle(oracle):
    p = parse_netlist("""CS amplifier, Monte Carlo
.param vto=agauss(0.7, 0.02, 1) kp=gauss(110u, 0.05, 1)
VDD vdd 0 DC 5
R1 vdd g {unif(100k, 0.01)}
R2 g 0 {unif(47k, 0.01)}
RD vdd d {unif(10k, 0.01)}
RS s 0 {unif(1k, 0.01)}
M1 d g s 0 nch W=10u L=1u
.model nch NMOS(VTO={vto} KP={kp} LAMBDA=0.02)
.end
""")
    c = p.circuit
    B = 1024
    got = E.sweep_dcop(c, B, seed=42)
    want = oracle.sweep_dcop(c, got.params, A.default_nr_config(), nthreads=16)
    np.testing.assert_array_equal(got.status, want["status"])
    np.testing.assert_array_equal(got.iterations, want["iterations"])
    assert _close(got.solutions, want["x"])
    assert np.all(got.accepted_mask())
    # the sampled VTO column follows N(0.7, 0.02)
    vto = got.params[:, p.params["VTO"].slot]
    assert abs(vto.mean() - 0.7) < 4 * 0.02 / np.sqrt(B) and abs(vto.std() - 0.02) < 0.002


def test_transient_netlist_rc_step():
    R, Cap = 2e3, 1.5e-12
    p = parse_netlist(f"""RC step
V1 in 0 PULSE(0 1 0 0 0 1)
R1 in out {R}
C1 out 0 {Cap}
.tran 1p 200p
.end
""")
    h, tstop = p.tran
    steps = int(round(tstop / h))
    got = E.transient_sweep(p.circuit, 3, h, steps, seed=1)
    k = np.arange(steps + 1, dtype=float)
    want = 1.0 * (1.0 - (1.0 + h / (R * Cap)) ** (-k))
    out = p.circuit.node_names["out"] - 1
    for i in range(3):
        np.testing.assert_allclose(got.waveforms[i, :, out], want, rtol=1e-12, atol=1e-15)


def test_subcircuit_inverter_chain_monte_carlo(oracle):
    # a self-biased CMOS inverter .SUBCKT instanced eight times, nested in a
    # pair, with sampled threshold voltages shared through .PARAM and
    # per-instance widths
    text = """subckt chain
.param vtn=agauss(0.5, 0.03, 1) vtp=agauss(-0.5, 0.03, 1)
.model nch NMOS(VTO={vtn} KP=120u LAMBDA=0.02)
.model pch PMOS(VTO={vtp} KP=40u LAMBDA=0.02)
.SUBCKT INV in out vdd PARAMS: WN=2u WP=6u
M1 out in 0 0 nch W={WN} L=1u
M2 out in vdd vdd pch W={WP} L=1u
RF out in 100k
.ENDS INV
.SUBCKT PAIR a b vdd PARAMS: K=1
X1 a m vdd INV PARAMS: WN={2u*K} WP={6u*K}
X2 m b vdd INV
.ENDS PAIR
VDD vdd 0 1.8
VIN vin 0 0.9
RIN vin in 100k
X1 in n1 vdd PAIR PARAMS: K=1
X2 n1 n2 vdd PAIR PARAMS: K=2
X3 n2 n3 vdd PAIR PARAMS: K=1.5
X4 n3 out vdd PAIR
.end
"""
    c = parse_netlist(text).circuit
    assert sum(e[0] in (A.SPX_EL_NMOS, A.SPX_EL_PMOS) for e in c.elements) == 16
    cfg, acfg = E.nr_config(limiter=E.Limiter.Pnjlim), A.default_nr_config(limiter=A.SPX_LIMIT_PNJLIM)
    got = E.sweep_dcop(c, 2048, seed=5, config=cfg)
    want = oracle.sweep_dcop(c, got.params, acfg, nthreads=16)
    np.testing.assert_array_equal(got.status, want["status"])
    # ~20 Newton iterations per instance through eight high-gain stages: a
    # rounding-level difference in the last update decides the count more
    # often than on the other workloads; instances with equal counts match to
    # 1e-9, all of them to the Newton tolerance
    same = got.iterations == want["iterations"]
    assert same.mean() >= 0.85
    assert _close(got.solutions[same], want["x"][same])
    assert np.all(np.abs(got.solutions - want["x"]) <= 1e-5)
    assert np.all(got.status == A.SPX_CONVERGED)
