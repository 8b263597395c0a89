"""PWL voltage sources (Waveform::Pwl, eval_pwl in waveforms.rs:116-118,
204-232; parse_pwl_waveform in parser/waveforms.rs:74-106): the oracle
pinned with the reference's test_pwl_waveform values (waveforms.rs:295-320)
read back through a circuit, hold before and after the breakpoints, the DC
value v1, a vertical step, and a Backward-Euler RC ramp response in closed
form; then GPU parity on an RC ladder driven by a many-point PWL on every
kernel kind."""
import numpy as np
import pytest

from spicier_amd import _abi as A
from spicier_amd import circuit as CK
from spicier_amd import engine as E
from spicier_amd.netlist import parse_netlist

REF_POINTS = [(0.0, 0.0), (1e-3, 5.0), (2e-3, 5.0), (3e-3, 0.0)]


def _source_trace(oracle, points, h, steps):
    c = CK.Circuit("pwl_r")
    c.vpwl("in", "0", points)
    c.resistor("in", "0", 1e3)
    p = CK.nominal_values(c).reshape(1, -1)
    out = oracle.transient_ex(c, p, A.default_nr_config(), h, steps, probes=[0],
                              method=A.SPX_TRAN_BE)
    assert out["status"][0] == A.SPX_CONVERGED
    return out["waveforms"][0, :, 0]


def test_reference_pwl_values(oracle):
    # waveforms.rs:295-320 at t = 0, 0.5m, 1m, 1.5m, 2m, 2.5m, 3m ... 5m
    w = _source_trace(oracle, REF_POINTS, 0.5e-3, 10)
    want = [0.0, 2.5, 5.0, 5.0, 5.0, 2.5, 0.0, 0.0, 0.0, 0.0, 0.0]
    np.testing.assert_allclose(w, want, atol=1e-10)


def test_hold_dc_value_and_vertical_step(oracle):
    # DC value v1 (waveforms.rs:153), hold before t1 and after tk; a vertical
    # edge (two points at 2 ns) takes the earlier interval's end value at 2 ns
    pts = [(1e-9, 0.3), (2e-9, 1.0), (2e-9, -1.0), (4e-9, -1.0)]
    w = _source_trace(oracle, pts, 0.5e-9, 10)
    np.testing.assert_allclose(w, [0.3, 0.3, 0.3, 0.65, 1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0],
                               atol=1e-12)


def test_parsed_pwl_matches_builder(oracle):
    c = parse_netlist("t\nV1 in 0 PWL(0 0 1m 5 2m 5 3m 0)\nR1 in 0 1k\n.end\n").circuit
    el = c.elements[0]
    assert el[0] == A.SPX_EL_VPWL and c.ctrl[0] == 4
    vals = CK.nominal_values(c)
    assert [vals[el[2][0] + j] for j in range(8)] == pytest.approx(
        [x for pt in REF_POINTS for x in pt])
    with pytest.raises(E.InvalidArgument):
        parse_netlist("t\n.param a=agauss(1,0.1,1)\nV1 in 0 PWL(0 0 1m {a})\nR1 in 0 1k\n")


def test_backward_euler_rc_ramp_closed_form(oracle):
    # a ramp of slope s into RC under BE: e_k = v_in(t_k) - v_k obeys
    # e_k = (e_k-1 + s h) / (1 + h/tau), e_0 = 0
    R, C, s, h, steps = 1e3, 1e-9, 1e6, 1e-7, 60
    c = CK.Circuit("pwl_rc")
    c.vpwl("in", "0", [(0.0, 0.0), (1.0, s * 1.0)])
    c.resistor("in", "out", R)
    c.capacitor("out", "0", C)
    p = CK.nominal_values(c).reshape(1, -1)
    out = oracle.transient_ex(c, p, A.default_nr_config(), h, steps,
                              probes=[c.node_names["out"] - 1], method=A.SPX_TRAN_BE)
    w = out["waveforms"][0, :, 0]
    a = 1.0 / (1.0 + h / (R * C))
    k = np.arange(steps + 1)
    e = s * h * a * (1 - a ** k) / (1 - a)
    np.testing.assert_allclose(w, s * h * k - e, rtol=1e-9, atol=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", [E.KernelKind.Dense, E.KernelKind.Sparse64,
                                  E.KernelKind.Sparse256, E.KernelKind.Lanes])
def test_pwl_driven_ladder_on_gpu(oracle, kind):
    # a 40-point staircase-and-ramp PWL into an RC ladder, on every kernel kind
    rng = np.random.default_rng(3)
    t = np.cumsum(rng.uniform(2e-12, 8e-12, 40))
    pts = list(zip(t, rng.uniform(0.0, 1.8, 40)))
    c = CK.Circuit("pwl_ladder")
    c.vpwl("in", "0", pts)
    prev = "in"
    for j in range(30):
        c.resistor(prev, f"n{j}", c.uniform(1e3, 5e3, name=f"R{j}"))
        c.capacitor(f"n{j}", "0", c.uniform(1e-15, 5e-15, name=f"C{j}"))
        prev = f"n{j}"
    probes = [0, 5, 29]
    cls = [E.TranClass(c, 64, 1e-12, 200, probes=probes, seed=2, method=E.IntegrationMethod.TrBdf2)]
    assert E.transient_mixed_device(cls, E.nr_config(), graph=False, kernels=[kind]) == [kind]
    E.Device.get(0).synchronize()
    got = cls[0].host_result()
    want = oracle.transient_ex(c, got.params, A.default_nr_config(), 1e-12, 200, probes=probes,
                               method=A.SPX_TRAN_TRBDF2, nthreads=16)
    np.testing.assert_array_equal(got.status, want["status"])
    np.testing.assert_array_equal(got.iterations, want["iterations"])
    err = np.abs(got.waveforms - want["waveforms"])
    assert np.all(err <= np.maximum(1e-12, 1e-9 * np.abs(want["waveforms"]))), err.max()
