"""Pins the CPU oracle (oracle/oracle.cpp) before it is trusted as the checker.

Each check is chosen so a plausible slip in the restatement (a dropped term,
a wrong sign or index, a transposed operand) fails: published constants,
the reference's own known-answer tests, LAPACK on the same systems, and
closed-form circuit solutions.
"""
import ctypes as C

import numpy as np
import pytest
from scipy.special import lambertw, wrightomega

from spicier_amd import _abi as A
from spicier_amd import circuit as CK


# ------------------------------------------------------------------ rng.rs
def test_splitmix64_published_sequence(oracle):
    # SplitMix64 (Steele, Lea, Flood 2014; Vigna's reference C) seeded with 0
    # outputs e220a8397b1dcdaf, 6e789e6aa1b965f4, 06c45d188009454f.  The
    # reference's splitmix64(x) is that generator's output for state x.
    g = 0x9E3779B97F4A7C15
    L = oracle.lib()
    assert L.orc_splitmix64(0) == 0xE220A8397B1DCDAF
    assert L.orc_splitmix64(g) == 0x6E789E6AA1B965F4
    assert L.orc_splitmix64((2 * g) % 2**64) == 0x06C45D188009454F


def test_combine_indices_separates_coordinates(oracle):
    L = oracle.lib()
    # rng.rs:51-56: XOR of the seed with odd multiples of the two indices.
    assert L.orc_combine_indices(42, 0, 0) == 42
    assert L.orc_combine_indices(0, 1, 0) == 0x517CC1B727220A95
    assert L.orc_combine_indices(0, 0, 1) == 0x5851F42D4C957F2D
    assert L.orc_combine_indices(7, 1, 1) == 7 ^ 0x517CC1B727220A95 ^ 0x5851F42D4C957F2D


def test_uniform_is_top_53_bits(oracle):
    L = oracle.lib()
    for s, p in ((0, 0), (5, 3), (1000, 9)):
        h = L.orc_splitmix64(L.orc_combine_indices(42, s, p))
        assert L.orc_uniform(42, s, p) == (h >> 11) / 2.0**53
        assert L.orc_uniform_f32(42, s, p) == np.float32((h >> 40) / 2.0**24)


def test_uniform_range_and_buckets(oracle):
    # rng.rs:349-399
    L = oracle.lib()
    vals = np.array([L.orc_uniform(42, s, p) for s in range(1000) for p in range(10)])
    assert vals.min() >= 0.0 and vals.max() < 1.0
    b = np.bincount(np.minimum((np.array([L.orc_uniform(42, i, 0) for i in range(10000)]) * 10)
                               .astype(int), 9), minlength=10)
    assert np.all((b > 800) & (b < 1200))
    assert L.orc_uniform(42, 0, 0) != L.orc_uniform(42, 1, 0)
    assert L.orc_uniform(42, 0, 0) != L.orc_uniform(43, 0, 0)


def test_gaussian_box_muller_and_moments(oracle):
    L = oracle.lib()
    # Box-Muller cos branch on params 2p, 2p+1 (rng.rs:120-131)
    for s, p in ((0, 0), (3, 1), (77, 5)):
        u1 = max(L.orc_uniform(42, s, 2 * p), 1e-10)
        u2 = L.orc_uniform(42, s, 2 * p + 1)
        ref = np.sqrt(-2 * np.log(u1)) * np.cos(2 * np.pi * u2)
        assert abs(L.orc_gaussian(42, s, p) - ref) <= 4e-16 * max(1.0, abs(ref))
    z = np.array([L.orc_gaussian(42, i, 0) for i in range(10000)])
    assert abs(z.mean()) < 0.05 and abs(z.std() - 1) < 0.05   # rng.rs:402-420
    zs = np.array([L.orc_gaussian_scaled(42, i, 0, 1000.0, 50.0) for i in range(10000)])
    assert abs(zs.mean() - 1000) < 5 and abs(zs.std() - 50) < 5  # rng.rs:423-457


def test_generate_params_layout_and_kinds(oracle):
    c = CK.inverter_chain()
    p = oracle.generate_params(c, 64)
    L = oracle.lib()
    assert p.shape == (64, c.num_params)
    # slot 0 = VTO_N0 ~ N(0.7, 0.03): gaussian_scaled at (seed, sweep, rng_index)
    assert p[5, 0] == L.orc_gaussian_scaled(42, 5, 0, 0.7, 0.03)
    lo, hi = 1e3, 10e3
    rl = p[:, 32:48]
    assert np.all((rl >= lo) & (rl < hi))
    # constants stay constant
    rs = c.slot_names.index("RS")
    assert np.all(p[:, rs] == 1e3)
    # sweep_offset shifts the sweep index
    q = oracle.generate_params(c, 8, sweep_offset=5)
    np.testing.assert_array_equal(q[0], p[5])


def test_lcg_monte_carlo_matches_sequential_generator(oracle):
    # MonteCarloGenerator (sweep.rs:170-200): one LCG stream over
    # (sample, variation); the sweep offset continues the same stream.
    c = CK.cs_amp(uniform_r=False)
    full = oracle.generate_params(c, 20, seed=12345, generator=A.SPX_GEN_LCG)
    tail = oracle.generate_params(c, 5, seed=12345, generator=A.SPX_GEN_LCG, sweep_offset=15)
    np.testing.assert_array_equal(full[15:], tail)
    # first draw by hand: s1 = seed*a + 1, u = (s1 >> 33) / 2^31
    a = 6364136223846793005
    s1 = (12345 * a + 1) % 2**64
    s2 = (s1 * a + 1) % 2**64
    u1 = max((s1 >> 33) / 2.0**31, 1e-10)
    u2 = (s2 >> 33) / 2.0**31
    z = np.sqrt(-2 * np.log(u1)) * np.cos(2 * np.pi * u2)
    nom, sig = 0.7, 0.02 / 0.7
    assert abs(full[0, 0] - np.clip(nom + z * sig * nom, 0.63, 0.77)) < 1e-15
    spec = c.specs
    for j in c.varied_slots():
        assert np.all(full[:, j] >= spec[j][4]) and np.all(full[:, j] <= spec[j][5])


def test_corner_and_linear_generators(oracle):
    c = CK.Circuit()
    c.uniform(1.0, 2.0)
    c.const(7.0)
    c.uniform(10.0, 20.0)
    p = oracle.generate_params(c, 4, generator=A.SPX_GEN_CORNER)
    np.testing.assert_array_equal(p, [[1, 7, 10], [2, 7, 10], [1, 7, 20], [2, 7, 20]])
    q = oracle.generate_params(c, 3, generator=A.SPX_GEN_LINEAR)
    np.testing.assert_array_equal(q, [[1, 7, 10], [1.5, 7, 15], [2, 7, 20]])


# ------------------------------------------------------------------ LU
def test_lu_known_answers(oracle):
    # solver.rs:408-460, faer_solver.rs:119-192, batched_lu.rs:581-618
    x, s = oracle.lu_solve(np.array([[2.0, 1.0], [1.0, 3.0]]), [5.0, 10.0])
    np.testing.assert_allclose(x, [1, 3], atol=1e-12)
    x, s = oracle.lu_solve(np.array([[2.0, 1.0], [1.0, 3.0]]), [5.0, 5.0])
    np.testing.assert_allclose(x, [2, 1], atol=1e-12)
    x, s = oracle.lu_solve(np.array([[1.0, 2.0], [1.0, 2.0]]), [1.0, 2.0])
    assert s and np.all(x == 0)
    # batch with identities + singular (solver.rs:407-460 column-major layout)
    m = np.array([1, 0, 0, 1, 1, 1, 2, 2], dtype=float)
    x, sing, ns = oracle.solve_batch(m, [1.0, 2.0, 1.0, 2.0], 2, 2)
    assert list(sing) == [0, 1] and ns == 1
    np.testing.assert_array_equal(x, [1, 2, 0, 0])


def test_lu_matches_lapack(oracle):
    rng = np.random.default_rng(0)
    for n in (1, 3, 5, 17, 64, 100):
        a = rng.standard_normal((n, n)) + n * 0.1 * np.eye(n)
        b = rng.standard_normal(n)
        x, s = oracle.lu_solve(a, b)
        assert not s
        np.testing.assert_allclose(x, np.linalg.solve(a, b), rtol=1e-10, atol=1e-12)


def test_lu_pivot_tie_takes_first_row(oracle):
    # Exact tie in column 0: nalgebra's icamax keeps the first maximum, so
    # rows are not swapped; the result is exact either way, the check is
    # that a permuted system gives the same answer.
    a = np.array([[1.0, 2.0, 0.0], [-1.0, 0.0, 1.0], [1.0, 1.0, 1.0]])
    b = np.array([3.0, 0.0, 3.0])
    x, s = oracle.lu_solve(a, b)
    np.testing.assert_allclose(x, [1, 1, 1], atol=1e-15)


# ------------------------------------------------------------------ devices
def _mos(oracle, pmos, vto, kp, lam, w, l, vgs, vds):
    d = C.c_double
    ids, gds, gm, reg = d(), d(), d(), C.c_int32()
    oracle.lib().orc_mosfet_evaluate(pmos, vto, kp, lam, w, l, vgs, vds, C.byref(ids),
                                     C.byref(gds), C.byref(gm), C.byref(reg))
    return ids.value, gds.value, gm.value, reg.value


def test_mosfet_level1_closed_forms(oracle):
    # level1.rs:342-397: default NMOS kp=2e-5, W/L=10
    ids, gds, gm, reg = _mos(oracle, 0, 0.7, 2e-5, 0.0, 10e-6, 1e-6, 0.3, 1.0)
    assert reg == 0 and ids == 0.0 and gm == 0.0 and gds == 1e-12
    beta = 2e-5 * 10
    ids, gds, gm, reg = _mos(oracle, 0, 0.7, 2e-5, 0.0, 10e-6, 1e-6, 2.0, 5.0)
    assert reg == 2 and abs(ids - 0.5 * beta * 1.3**2) < 1e-10 and abs(gm - beta * 1.3) < 1e-12
    ids, gds, gm, reg = _mos(oracle, 0, 0.7, 2e-5, 0.0, 10e-6, 1e-6, 2.0, 0.5)
    assert reg == 1 and abs(ids - beta * (1.3 * 0.5 - 0.125)) < 1e-10
    assert abs(gds - beta * (1.3 - 0.5)) < 1e-12 and abs(gm - beta * 0.5) < 1e-12
    # lambda terms: d/dvds of the linear-region current, by central difference
    f = lambda vds: _mos(oracle, 0, 0.7, 2e-5, 0.05, 10e-6, 1e-6, 2.0, vds)[0]
    g = _mos(oracle, 0, 0.7, 2e-5, 0.05, 10e-6, 1e-6, 2.0, 0.5)[1]
    assert abs(g - (f(0.5 + 1e-6) - f(0.5 - 1e-6)) / 2e-6) < 1e-9
    fs = lambda vds: _mos(oracle, 0, 0.7, 2e-5, 0.05, 10e-6, 1e-6, 2.0, vds)[0]
    gs = _mos(oracle, 0, 0.7, 2e-5, 0.05, 10e-6, 1e-6, 2.0, 3.0)[1]
    assert abs(gs - (fs(3.0 + 1e-6) - fs(3.0 - 1e-6)) / 2e-6) < 1e-10
    # PMOS saturation (level1.rs:454-486)
    ids, gds, gm, reg = _mos(oracle, 1, -0.7, 50e-6, 0.0, 10e-6, 1e-6, -2.0, -5.0)
    assert reg == 2 and abs(ids - (-0.5 * 500e-6 * 1.3**2)) < 1e-10 and abs(gm - 650e-6) < 1e-10


def test_pmos_stamp_matches_reference_entries(oracle):
    # level1.rs:564-620: PMOS d=1 g=2 s=3 at V = (0, 3, 5): gm=650uS,
    # gds=1e-12, ieq=877.5uA; G[0,0]=gds, G[0,1]=gm, G[0,2]=-gds-gm,
    # G[2,0]=-gds, G[2,1]=-gm, G[2,2]=gds+gm, rhs[0]=-ieq, rhs[2]=ieq.
    c = CK.Circuit()
    for n in ("1", "2", "3"):
        c.node(n)
    c.pmos("1", "2", "3", vto=-0.7, kp=50e-6, lam=0.0, w=10e-6, l=1e-6)
    J, rhs = oracle.stamp(c, CK.instance_params(c, 1)[0], [0.0, 3.0, 5.0])
    gm, gds = 650e-6, 1e-12
    ieq = -422.5e-6 - gds * (-5.0) - gm * (-2.0)   # ids - gds*vds - gm*vgs
    exp = np.array([[gds, gm, -gds - gm], [0, 0, 0], [-gds, -gm, gds + gm]])
    np.testing.assert_allclose(J, exp, atol=1e-15, rtol=0)
    np.testing.assert_allclose(rhs, [-ieq, 0, ieq], atol=1e-15, rtol=0)


def test_diode_model(oracle):
    d = C.c_double
    L = oracle.lib()
    vt = L.orc_thermal_voltage(300.15)
    assert vt == 1.380649e-23 * 300.15 / 1.602176634e-19 and abs(vt - 0.02585) < 1e-3
    i, g = d(), d()
    L.orc_diode_evaluate(1e-14, 1.0, -1.0, A.SPX_LIMIT_REFERENCE, C.byref(i), C.byref(g))
    assert i.value < 0 and abs(i.value) < 1e-12 and g.value == 1e-12
    L.orc_diode_evaluate(1e-14, 1.0, 0.0, A.SPX_LIMIT_REFERENCE, C.byref(i), C.byref(g))
    assert i.value == 0.0
    # below vcrit (~0.7297 V) the limiter is the identity: plain Shockley
    L.orc_diode_evaluate(1e-14, 1.0, 0.65, A.SPX_LIMIT_REFERENCE, C.byref(i), C.byref(g))
    assert abs(i.value - 1e-14 * np.expm1(0.65 / vt)) < 1e-12 * i.value
    vcrit = vt * np.log(vt / (np.sqrt(2) * 1e-14))
    assert L.orc_diode_limit(vcrit - 1e-3, vt) == vcrit - 1e-3
    lim = L.orc_diode_limit(100.0, 0.02585)
    assert 0 < lim < 100   # diode.rs:281-287
    # above vcrit: vcrit + nvt*(1 + ln(1 + arg)), continuous in arg but
    # jumping by nvt at vcrit
    v = vcrit + 0.1
    assert abs(L.orc_diode_limit(v, vt) - (vcrit + vt * (1 + np.log1p(0.1 / vt)))) < 1e-15


# ------------------------------------------------------------------ Newton
def tight(**kw):
    # tight enough that the last Newton update is at rounding level, loose
    # enough that rounding noise cannot keep the Δx test from passing
    return A.default_nr_config(v_reltol=1e-11, v_abstol=1e-13, i_abstol=1e-16, **kw)


def test_diode_resistor_lambert_w(oracle):
    # V - R - D: I = (V - Vd)/R, Vd = nVt ln(1 + I/IS), solved by Lambert W:
    # I = vt/R W(IS R/vt e^{(V + IS R)/vt}) - IS.  The SPICE pnjlim limiter
    # preserves the Shockley fixed point.
    vt = 1.380649e-23 * 300.15 / 1.602176634e-19
    IS = 1e-14
    for v, r in ((5.0, 1e3), (0.7, 100.0), (2.0, 10e3), (30.0, 100.0)):
        c = CK.diode_resistor(v, r)
        out = oracle.sweep_dcop(c, CK.instance_params(c, 1), tight(limiter=A.SPX_LIMIT_PNJLIM))
        assert out["status"][0] == A.SPX_CONVERGED
        assert out["flags"][0] == 0
        # W(e^z) = wrightomega(z), z = ln(IS R/vt) + (V + IS R)/vt (no overflow)
        z = np.log(IS * r / vt) + (v + IS * r) / vt
        i = vt / r * wrightomega(z).real - IS
        assert abs(out["x"][0, 1] - (v - i * r)) < 1e-10, (v, r, out["x"][0])
        if v < 10:
            assert abs(lambertw(np.exp(z)).real - wrightomega(z).real) < 1e-12 * wrightomega(z).real
        assert out["residual"][0] <= 1.0


def test_reference_limiter_fixed_point(oracle):
    # With the reference's memoryless limiter (diode.rs:113-148) the Newton
    # map's fixed point above vcrit solves
    #   (V - v)/R = IS(e^{L2/nvt} - 1) + gd(L2) (v - L1),  L1 = lim(v), L2 = lim(L1)
    # instead of Shockley's equation.  Pin that, and that the KCL test (which
    # evaluates the unlimited Shockley current) flags it.
    from scipy.optimize import brentq
    L = oracle.lib()
    vt = L.orc_thermal_voltage(300.15)
    IS, V, R = 1e-14, 5.0, 1e3

    def f(v):
        l1 = L.orc_diode_limit(v, vt)
        l2 = L.orc_diode_limit(l1, vt)
        e = np.exp(l2 / vt)
        return (V - v) / R - (IS * (e - 1) + max(IS * e / vt, 1e-12) * (v - l1))

    vcrit = vt * np.log(vt / (np.sqrt(2) * 1e-14))
    vstar = brentq(f, vcrit + 1e-9, 1.0, xtol=1e-15)
    c = CK.diode_resistor(V, R)
    out = oracle.sweep_dcop(c, CK.instance_params(c, 1),
                            tight(max_iterations=500, gmin_stepping=0, source_stepping=0))
    assert out["status"][0] == A.SPX_CONVERGED
    assert abs(out["x"][0, 1] - vstar) < 1e-8
    assert 0.5 < vstar < 0.9   # ngspice_validation.rs:705-768 bounds still hold
    assert out["flags"][0] & A.SPX_FLAG_KCL_FAIL
    # default tolerances stop early on the slowly converging map, as the
    # reference does; still inside the reference test's bounds
    d = oracle.sweep_dcop(c, CK.instance_params(c, 1))
    assert d["status"][0] == A.SPX_CONVERGED and 0.5 < d["x"][0, 1] < 0.9


def test_nmos_cs_closed_form(oracle):
    # ngspice_validation.rs:834-890 circuit, square law with lambda=0:
    # V(d) = 5 - RD * beta/2 * (VG - VTO)^2 = 5 - 0.169 = 4.831
    c = CK.nmos_cs_validation()
    out = oracle.sweep_dcop(c, CK.instance_params(c, 1), tight())
    d = c.node("d") - 1
    assert out["status"][0] == A.SPX_CONVERGED
    assert abs(out["x"][0, d] - (5 - 1e3 * 0.5 * 2e-4 * 1.3**2)) < 1e-12


def test_pmos_rd_closed_form(oracle):
    # level1.rs:520-560 circuit: PMOS s=VDD 5 V, g=3 V, d -> 1k -> 0;
    # saturation, lambda 0: V(d) = 1k * beta/2 * 1.3^2 = 0.4225 V.
    c = CK.Circuit()
    for n in ("1", "2", "3"):
        c.node(n)
    c.resistor("1", "0", 1e3)
    c.vsource("3", "0", 5.0)
    c.vsource("2", "0", 3.0)
    c.pmos("1", "2", "3", vto=-0.7, kp=50e-6, lam=0.0, w=10e-6, l=1e-6)
    out = oracle.sweep_dcop(c, CK.instance_params(c, 1), tight())
    assert abs(out["x"][0, 0] - 0.4225) < 1e-12
    # branch current of VDD = -Id (current flows out of + terminal into the circuit)
    assert abs(out["x"][0, 3] - (-0.4225e-3)) < 1e-15


def test_cs_amp_cubic_closed_form(oracle):
    # CS amp with degeneration and lambda: with Vg = 5*R2/(R1+R2) fixed,
    # Id solves Id = beta/2 (Vg - Id*RS - VTO)^2 (1 + lambda (5 - Id (RD+RS)))
    # in saturation; bisection on that scalar equation.
    from scipy.optimize import brentq
    c = CK.cs_amp()
    p = CK.instance_params(c, 1)[0]
    out = oracle.sweep_dcop(c, p[None], tight())
    beta = 110e-6 * 10
    vg = 5 * 47e3 / 147e3
    f = lambda i: i - beta / 2 * (vg - i * 1e3 - 0.7) ** 2 * (1 + 0.02 * (5 - i * 11e3))
    i = brentq(f, 0.0, (vg - 0.7) / 1e3, xtol=1e-18, rtol=1e-15)
    x = out["x"][0]
    assert abs(x[c.node("g") - 1] - vg) < 1e-12
    assert abs(x[c.node("d") - 1] - (5 - i * 10e3)) < 1e-9
    assert abs(x[c.node("s") - 1] - i * 1e3) < 1e-12


def test_symmetric_inverter_trips_at_midrail(oracle):
    # One self-biased CMOS stage with matched devices sits exactly at VDD/2:
    # pins the PMOS sign conventions against the NMOS ones.
    c = CK.Circuit()
    c.vsource("vdd", "0", 1.8)
    c.nmos("o", "g", "0", vto=0.7, kp=100e-6, lam=0.02, w=10e-6, l=1e-6)
    c.pmos("o", "g", "vdd", vto=-0.7, kp=100e-6, lam=0.02, w=10e-6, l=1e-6)
    c.resistor("o", "g", 100e3)
    out = oracle.sweep_dcop(c, CK.instance_params(c, 1), tight())
    x = out["x"][0]
    assert out["status"][0] == A.SPX_CONVERGED
    assert abs(x[c.node("o") - 1] - 0.9) < 1e-12 and abs(x[c.node("g") - 1] - 0.9) < 1e-12


def test_convergence_check_semantics(oracle):
    # A linear circuit takes the sweep's single solve (sweep.rs:159-217);
    # Newton (newton.rs:134-159) stops on the first iteration whose update
    # passes the Δx test, so it needs one more than the last big update.
    c = CK.voltage_divider()
    out = oracle.sweep_dcop(c, CK.instance_params(c, 1))
    assert out["iterations"][0] == 1 and abs(out["x"][0, 1] - 5.0) < 1e-12
    c = CK.nmos_cs_validation()
    o = oracle.sweep_dcop(c, CK.instance_params(c, 1))
    t = oracle.sweep_dcop(c, CK.instance_params(c, 1), A.default_nr_config(max_iterations=o["iterations"][0] - 1, gmin_stepping=0, source_stepping=0))
    assert o["status"][0] == A.SPX_CONVERGED and t["status"][0] == A.SPX_FAILED


def test_gmin_and_source_stepping_reach_the_same_point(oracle):
    c = CK.cs_amp()
    p = CK.instance_params(c, 1)
    plain = oracle.sweep_dcop(c, p, tight())["x"][0]
    # force the fallbacks: a plain-NR budget of 1 iteration cannot converge
    cfg = tight(max_iterations=1)
    cfg.max_iterations = 1
    o = oracle.sweep_dcop(c, p, cfg)
    assert o["status"][0] == A.SPX_FAILED
    cfg2 = A.default_nr_config(max_iterations=50, source_stepping=0)
    o2 = oracle.sweep_dcop(c, p, cfg2)
    assert o2["status"][0] == A.SPX_CONVERGED
    o3 = oracle.sweep_dcop(c, p, A.default_nr_config(gmin_stepping=0))
    assert o3["status"][0] == A.SPX_CONVERGED
    np.testing.assert_allclose(o2["x"][0], plain, rtol=1e-5, atol=1e-9)


def test_stepping_counts_every_attempt_of_a_failed_level(oracle):
    # newton.rs:300-315 / 411-426: a level whose Newton solve does not
    # converge is attempted max_attempts_per_level times, and every attempt's
    # iterations go into total_iterations.  With a 1-iteration budget no solve
    # from zero can pass the dx test (x moves by volts), so by hand: plain NR
    # 1, the first Gmin level 5 x 1 and the stage gives up, then source
    # stepping at factors 0.1, 0.05, 0.025, 0.0125, 0.00625 (the step halves
    # while it exceeds 0.01) 5 x 1 each: 1 + 5 + 25 = 31.
    c = CK.cs_amp()
    p = CK.instance_params(c, 1)
    o = oracle.sweep_dcop(c, p, A.default_nr_config(max_iterations=1))
    assert o["status"][0] == A.SPX_FAILED and o["iterations"][0] == 31
    o = oracle.sweep_dcop(c, p, A.default_nr_config(max_iterations=1, gmin_max_attempts=2,
                                                    src_max_attempts=3))
    assert o["iterations"][0] == 1 + 2 + 5 * 3
    # a 3-iteration budget fails plain NR and converges later: the attempts
    # change only the count, by (attempts - 1) x 3 per failed level, never the
    # solution (the retries are replays)
    base = oracle.sweep_dcop(c, p, A.default_nr_config(max_iterations=3, gmin_max_attempts=1,
                                                       src_max_attempts=1))
    five = oracle.sweep_dcop(c, p, A.default_nr_config(max_iterations=3))
    assert base["status"][0] == five["status"][0]
    np.testing.assert_array_equal(base["x"], five["x"])
    d = int(five["iterations"][0] - base["iterations"][0])
    assert d > 0 and d % (4 * 3) == 0


def test_source_stepping_overshoot_is_flagged(oracle):
    # newton.rs:312-350: after a halved step the factor can jump past 1.0
    # (0.975 + 0.05) and the loop ends "converged" without a level at 1.0.
    # The restatement keeps that; the FP64 KCL test at full sources flags it.
    c = CK.diode_resistor(5.0, 1e3)
    cfg = A.default_nr_config(v_reltol=1e-11, v_abstol=1e-13, i_abstol=1e-16)
    out = oracle.sweep_dcop(c, CK.instance_params(c, 1), cfg)
    assert out["status"][0] == A.SPX_CONVERGED
    assert out["flags"][0] & A.SPX_FLAG_SOURCE_STEPPING
    assert abs(out["x"][0, 0] - 0.975 * 5.0) < 1e-12
    assert out["flags"][0] & A.SPX_FLAG_KCL_FAIL and out["residual"][0] > 1


def test_workload_sweeps_converge_and_pass_kcl(oracle):
    for mk, B in ((CK.cs_amp, 256), (CK.inverter_chain, 64)):
        c = mk()
        out = oracle.sweep_dcop(c, oracle.generate_params(c, B), nthreads=4)
        assert np.all(out["status"] == A.SPX_CONVERGED)
        assert np.all(out["residual"] <= 1.0)
        assert np.all((out["flags"] & A.SPX_FLAG_KCL_FAIL) == 0)


# ------------------------------------------------------------------ statistics
def test_statistics_known_answers(oracle):
    # statistics.rs:760-800
    s = oracle.stats_accumulate(np.arange(1.0, 6.0)[:, None])[0]
    assert list(s) == [5, 15, 55, 1, 5]
    o = np.zeros(3)
    oracle.lib().orc_stats_from_moments(5, 15.0, 55.0, C.c_void_p(o.ctypes.data))
    assert abs(o[0] - 3) < 1e-12 and abs(o[1] - 2) < 1e-12 and abs(o[2] - 2**0.5) < 1e-12
    s1 = oracle.stats_accumulate(np.array([[42.0]]))[0]
    assert list(s1) == [1, 42, 42 * 42, 42, 42]
    # mask + merge: masked rows are skipped; merge is associative
    v = np.random.default_rng(1).standard_normal((100, 3))
    m = (np.arange(100) % 3 != 0).astype(np.uint8)
    full = oracle.stats_accumulate(v, m)
    a = oracle.stats_accumulate(v[:40], m[:40])
    b = oracle.stats_accumulate(v[40:], m[40:])
    np.testing.assert_allclose(full[:, :3], a[:, :3] + b[:, :3], rtol=1e-13)
    np.testing.assert_array_equal(full[:, 3], np.minimum(a[:, 3], b[:, 3]))
    np.testing.assert_array_equal(full[:, 4], np.maximum(a[:, 4], b[:, 4]))


# ------------------------------------------------------------------ GMRES
def _sparse_system(rng, n, extra=3, dom=1.0):
    rows, cols, vals = [], [], []
    for i in range(n):
        js = rng.choice(n, size=extra, replace=False)
        for j in js:
            if j != i:
                rows.append(i); cols.append(j); vals.append(rng.standard_normal())
        rows.append(i); cols.append(i); vals.append(dom * (extra + 1) + rng.random())
    rows.append(0); cols.append(0); vals.append(0.25)          # a duplicate, summed
    return np.array(rows), np.array(cols), np.array(vals)


def _dense_of(n, rows, cols, vals):
    m = np.zeros((n, n))
    np.add.at(m, (rows, cols), vals)
    return m


def test_gmres_converges_to_lapack_solution(oracle):
    rng = np.random.default_rng(11)
    for n, restart in ((60, 30), (200, 30), (200, 5)):
        rows, cols, vals = _sparse_system(rng, n, dom=0.6)
        b = rng.standard_normal(n)
        x, conv, it, res = oracle.gmres(n, rows, cols, vals, b, restart=restart, tol=1e-12)
        m = _dense_of(n, rows, cols, vals)
        want = np.linalg.solve(m, b)
        assert conv and res < 1e-12 and it > 1
        assert np.linalg.norm(b - m @ x) / np.linalg.norm(b) < 1e-11
        np.testing.assert_allclose(x, want, rtol=1e-9, atol=1e-11)


def test_gmres_jacobi_solves_diagonal_in_one_step_and_edge_cases(oracle):
    n = 50
    d = np.linspace(1.0, 1e6, n)
    idx = np.arange(n)
    b = np.linspace(-3.0, 5.0, n)
    x, conv, it, res = oracle.gmres(n, idx, idx, d, b, tol=1e-14)
    assert conv and it == 1
    np.testing.assert_allclose(x, b / d, rtol=1e-14)
    # without Jacobi the Krylov space needs more than one vector
    x2, conv2, it2, _ = oracle.gmres(n, idx, idx, d, b, tol=1e-14, jacobi=False)
    assert it2 > 1
    # zero right-hand side: converged, no iterations, x = 0 (real.rs:38-45)
    x, conv, it, res = oracle.gmres(n, idx, idx, d, np.zeros(n))
    assert conv and it == 0 and np.all(x == 0)
    # iteration budget exhausted: reported not converged; the reference counts
    # the iteration that trips the budget check (real.rs:289-292), so the
    # count is max_iter + 1 when the budget runs out inside a cycle
    rng = np.random.default_rng(2)
    rows, cols, vals = _sparse_system(rng, 120, dom=0.3)
    x, conv, it, res = oracle.gmres(120, rows, cols, vals, rng.standard_normal(120),
                                    restart=4, max_iter=6, tol=1e-14)
    assert not conv and it == 7 and res > 1e-14


# ------------------------------------------------------------- BE transient
def test_transient_rc_step_is_the_backward_euler_recurrence(oracle):
    # v_k = V (1 - (1 + h/tau)^-k): the exact Backward-Euler recurrence of an
    # RC charged from a step, with the source branch current (V - v_k)/R
    R, Cap, V = 2e3, 1.5e-12, 1.0
    c = CK.rc_step(R, Cap, V)
    h, steps = 1e-12, 200
    out = oracle.transient(c, CK.instance_params(c, 1), A.default_nr_config(), h, steps)
    k = np.arange(steps + 1)
    want = V * (1.0 - (1.0 + h / (R * Cap)) ** (-k.astype(float)))
    # unknowns: in, out, branch
    np.testing.assert_allclose(out["waveforms"][0, :, 1], want, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(out["waveforms"][0, 1:, 0], V, rtol=0, atol=1e-15)
    np.testing.assert_allclose(out["waveforms"][0, 1:, 2], -(V - want[1:]) / R, rtol=1e-10)
    assert out["status"][0] == A.SPX_CONVERGED and out["fail_step"][0] == -1
    assert out["iterations"][0] == steps + 1 and out["residual"][0] <= 1.0


def test_transient_pulse_waveform_and_ladder(oracle):
    # the driven node follows eval_pulse exactly (waveforms.rs:160-180)
    c = CK.ladder(16, seed=3)
    h, steps = 1e-12, 40
    out = oracle.transient(c, CK.instance_params(c, 2), A.default_nr_config(), h, steps)
    t = np.arange(steps + 1) * h
    ramp = np.clip(t / 10e-12, 0, 1) * 1.8
    np.testing.assert_allclose(out["waveforms"][0, :, 0], ramp, rtol=1e-12, atol=1e-15)
    assert np.all(out["status"] == A.SPX_CONVERGED)
    # the ladder charges monotonically from the source side for the first steps
    w = out["waveforms"][0]
    assert np.all(w[5:, 1] > 0) and w[20, 1] > w[20, 5]
