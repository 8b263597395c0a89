"""Circuit programs for the sweep engine, and the sweep workloads' netlists.

A :class:`Circuit` is the host-side builder of ``spx_circuit`` (the
device-resident circuit program of ``include/spicier_hip.h``): a list of
elements whose values live in per-instance *parameter slots*, each slot with a
distribution (``spx_param_spec``).  It plays the part the reference's
``SweepStamperFactory`` closures play (crates/spicier-solver/src/sweep.rs:272-326),
in a form that can cross a C ABI and be evaluated on the GPU.

Element stamping follows spicier-core's MNA conventions
(crates/spicier-core/src/mna.rs:87-169); node 0 is ground and node k>0 is
unknown k-1, voltage sources add branch unknowns after the nodes in order of
appearance (crates/spicier-core/src/netlist.rs:243-326).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _abi as A


# element kinds that carry a branch current unknown, numbered in element order
BRANCH_KINDS = (A.SPX_EL_VSOURCE, A.SPX_EL_VPULSE, A.SPX_EL_VSIN, A.SPX_EL_VPWL,
                A.SPX_EL_INDUCTOR, A.SPX_EL_VCVS, A.SPX_EL_CCVS)


@dataclass(frozen=True)
class Param:
    """Handle of a parameter slot."""
    slot: int


@dataclass
class Circuit:
    name: str = "circuit"
    elements: list = field(default_factory=list)
    specs: list = field(default_factory=list)   # (kind, rng_index, a, b, c, d)
    node_names: dict = field(default_factory=lambda: {"0": 0})
    slot_names: list = field(default_factory=list)
    ctrl: dict = field(default_factory=dict)    # element index -> controlling element index

    # ---------------------------------------------------------------- nodes
    def node(self, name) -> int:
        if isinstance(name, int):
            return name
        if name in ("0", "gnd", "GND"):
            return 0
        if name not in self.node_names:
            self.node_names[name] = len(self.node_names)
        return self.node_names[name]

    @property
    def num_nodes(self) -> int:
        return len(self.node_names) - 1

    @property
    def num_vsources(self) -> int:
        return sum(e[0] in BRANCH_KINDS for e in self.elements)

    @property
    def size(self) -> int:
        return self.num_nodes + self.num_vsources

    def branch_unknown(self, element: int) -> int:
        """Unknown index of a branch element's current (V, L, E, H sources;
        branch rows follow the nodes in element order, mna.rs:124-145)."""
        if self.elements[element][0] not in BRANCH_KINDS:
            raise ValueError(f"element {element} has no branch current")
        return self.num_nodes + sum(e[0] in BRANCH_KINDS for e in self.elements[:element])

    @property
    def num_params(self) -> int:
        return len(self.specs)

    @property
    def nonlinear(self) -> bool:
        return any(e[0] in (A.SPX_EL_DIODE, A.SPX_EL_NMOS, A.SPX_EL_PMOS) for e in self.elements)

    # ---------------------------------------------------------------- slots
    def _slot(self, kind, a, b=0.0, c=0.0, d=0.0, name=None) -> Param:
        s = len(self.specs)
        self.specs.append((kind, s, float(a), float(b), float(c), float(d)))
        self.slot_names.append(name or f"p{s}")
        return Param(s)

    def const(self, v, name=None) -> Param:
        return self._slot(A.SPX_P_CONST, v, name=name)

    def gauss_rel(self, nominal, sigma=0.05, vmin=None, vmax=None, name=None) -> Param:
        """ParameterVariation semantics: (nom + z*sigma*nom).clamp(min, max), default
        bounds +-10% (crates/spicier-solver/src/sweep.rs:31-40,193)."""
        vmin = nominal * 0.9 if vmin is None else vmin
        vmax = nominal * 1.1 if vmax is None else vmax
        return self._slot(A.SPX_P_GAUSS_REL, nominal, sigma, vmin, vmax, name)

    def gauss(self, mean, sigma, name=None) -> Param:
        """mean + z*sigma (rng.rs:148-150); corner/linear sweeps use mean -+ 3 sigma."""
        return self._slot(A.SPX_P_GAUSS, mean, sigma, mean - 3 * sigma, mean + 3 * sigma, name)

    def uniform(self, lo, hi, name=None) -> Param:
        return self._slot(A.SPX_P_UNIFORM, lo, hi, lo, hi, name)

    def loguniform(self, lo, hi, name=None) -> Param:
        return self._slot(A.SPX_P_LOGUNIFORM, lo, hi, lo, hi, name)

    def _p(self, v) -> int:
        if isinstance(v, Param):
            return v.slot
        return self.const(v).slot

    # ------------------------------------------------------------- elements
    def _add(self, kind, nodes, params, ctrl=None) -> int:
        """Appends an element; returns its index."""
        n = [self.node(x) for x in nodes] + [0] * (4 - len(nodes))
        p = [self._p(v) for v in params] + [-1] * (A.SPX_MAX_ELEMENT_PARAMS - len(params))
        self.elements.append((kind, n, p))
        if ctrl is not None:
            self.ctrl[len(self.elements) - 1] = ctrl
        return len(self.elements) - 1

    def _branch_element(self, ctrl) -> int:
        if not (isinstance(ctrl, int) and 0 <= ctrl < len(self.elements)
                and self.elements[ctrl][0] in BRANCH_KINDS):
            raise ValueError("a current-controlled source needs an earlier element with a "
                             "branch current (voltage source, inductor, VCVS or CCVS)")
        return ctrl

    def vcvs(self, op, on, cp, cn, gain):
        """E: v(op) - v(on) = gain (v(cp) - v(cn)) (controlled.rs:58-90)."""
        return self._add(A.SPX_EL_VCVS, (op, on, cp, cn), (gain,))

    def vccs(self, op, on, cp, cn, gm):
        """G: gm (v(cp) - v(cn)) enters op and leaves on (controlled.rs:172-201)."""
        return self._add(A.SPX_EL_VCCS, (op, on, cp, cn), (gm,))

    def cccs(self, op, on, ctrl: int, gain):
        """F: gain times element `ctrl`'s branch current, KCL +at op, -at on
        (controlled.rs:267-285)."""
        return self._add(A.SPX_EL_CCCS, (op, on), (gain,), ctrl=self._branch_element(ctrl))

    def ccvs(self, op, on, ctrl: int, gain):
        """H: v(op) - v(on) = gain times element `ctrl`'s branch current
        (controlled.rs:351-375)."""
        return self._add(A.SPX_EL_CCVS, (op, on), (gain,), ctrl=self._branch_element(ctrl))

    def resistor(self, a, b, r):
        return self._add(A.SPX_EL_RESISTOR, (a, b), (r,))

    def capacitor(self, a, b, c):
        return self._add(A.SPX_EL_CAPACITOR, (a, b), (c,))

    def inductor(self, a, b, l):
        """Inductor of l henries from a to b (passive.rs:283-375); its branch
        current (a -> b) is an unknown, after the node voltages in element
        order with the voltage sources."""
        return self._add(A.SPX_EL_INDUCTOR, (a, b), (l,))

    def vsource(self, pos, neg, v):
        return self._add(A.SPX_EL_VSOURCE, (pos, neg), (v,))

    def vpulse(self, pos, neg, v1, v2, td, tr, tf, pw, per):
        return self._add(A.SPX_EL_VPULSE, (pos, neg), (v1, v2, td, tr, tf, pw, per))

    def vsin(self, pos, neg, vo, va, freq, td=0.0, theta=0.0, phase=0.0):
        """SIN(vo va freq td theta phase) source (waveforms.rs:34-49, 185-201);
        phase in degrees; DC value vo."""
        return self._add(A.SPX_EL_VSIN, (pos, neg), (vo, va, freq, td, theta, phase))

    def vpwl(self, pos, neg, points):
        """PWL((t1, v1), (t2, v2), ...) source (waveforms.rs:116-118, 204-232):
        the 2k constant values take consecutive parameter slots; DC value v1."""
        pts = [(float(t), float(v)) for t, v in points]
        if not pts:
            raise ValueError("PWL needs at least one point")
        first = len(self.specs)
        for j, (t, v) in enumerate(pts):
            self.const(t, name=f"PWL{first}_T{j}")
            self.const(v, name=f"PWL{first}_V{j}")
        i = self._add(A.SPX_EL_VPWL, (pos, neg), ())
        self.elements[i][2][0] = first
        self.ctrl[i] = len(pts)
        return i

    def isource(self, frm, to, i):
        """Current i flows from `frm` through the source into `to`."""
        return self._add(A.SPX_EL_ISOURCE, (frm, to), (i,))

    def diode(self, anode, cathode, is_=1e-14, n=1.0):
        return self._add(A.SPX_EL_DIODE, (anode, cathode), (is_, n))

    def nmos(self, d, g, s, vto=0.7, kp=2e-5, lam=0.0, w=10e-6, l=1e-6):
        return self._add(A.SPX_EL_NMOS, (d, g, s), (vto, kp, lam, w, l))

    def pmos(self, d, g, s, vto=-0.7, kp=1e-5, lam=0.0, w=10e-6, l=1e-6):
        return self._add(A.SPX_EL_PMOS, (d, g, s), (vto, kp, lam, w, l))

    # ------------------------------------------------------------ C structs
    def to_c(self):
        """Returns (SpxCircuit, keepalive) — keep `keepalive` referenced while
        the struct is in use."""
        arr = (A.SpxElement * max(1, len(self.elements)))()
        for i, (k, n, p) in enumerate(self.elements):
            arr[i].kind = k
            for j in range(4):
                arr[i].node[j] = n[j]
            for j in range(A.SPX_MAX_ELEMENT_PARAMS):
                arr[i].param[j] = p[j]
            arr[i].ctrl = self.ctrl.get(i, -1)
        ckt = A.SpxCircuit(num_nodes=self.num_nodes, num_elements=len(self.elements),
                           elements=C.cast(arr, C.POINTER(A.SpxElement)),
                           num_params=self.num_params)
        return ckt, arr

    def param_specs_c(self):
        arr = (A.SpxParamSpec * max(1, len(self.specs)))()
        for i, (k, r, a, b, c, d) in enumerate(self.specs):
            arr[i] = A.SpxParamSpec(kind=k, rng_index=r, a=a, b=b, c=c, d=d)
        return arr

    def sweep_spec(self, batch, seed=42, generator=A.SPX_GEN_HASH, sweep_offset=0,
                   explicit_params=None):
        """Returns (SpxSweepSpec, keepalive)."""
        specs = self.param_specs_c()
        sp = A.SpxSweepSpec(generator=generator, seed=seed, batch=batch,
                            sweep_offset=sweep_offset,
                            params=C.cast(specs, C.POINTER(A.SpxParamSpec)),
                            explicit_params=A.ptr(explicit_params))
        return sp, (specs, explicit_params)

    def varied_slots(self):
        return [i for i, s in enumerate(self.specs) if s[0] != A.SPX_P_CONST]


# --------------------------------------------------------------------------
# Workload netlists
# --------------------------------------------------------------------------

def cs_amp(uniform_r=True) -> Circuit:
    """Common-source NMOS amplifier, 4 nodes + 1 branch = 5 unknowns.

    VDD=5 V; R1 100k VDD-G; R2 47k G-0; RD 10k VDD-D; RS 1k S-0; M1 D G S with
    VTO~N(0.7,0.02) V, KP~N(110e-6, 5%) A/V^2, LAMBDA=0.02, W/L=10; resistors
    uniform +-1% (the 1,024-instance Monte Carlo case).  Topology as the
    reference's NMOS common-source validation circuit
    (crates/spicier-parser/tests/ngspice_validation.rs:834-842).
    With ``uniform_r=False`` every varied quantity is a ParameterVariation
    (Gaussian relative), the only form MonteCarloGenerator draws.
    """
    c = Circuit("cs_amp")
    rvals = (("R1", 100e3), ("R2", 47e3), ("RD", 10e3), ("RS", 1e3))
    if uniform_r:
        vto = c.gauss(0.7, 0.02, name="VTO")
        kp = c.gauss(110e-6, 0.05 * 110e-6, name="KP")
        r = {k: c.uniform(0.99 * v, 1.01 * v, name=k) for k, v in rvals}
    else:
        vto = c.gauss_rel(0.7, 0.02 / 0.7, name="VTO")
        kp = c.gauss_rel(110e-6, 0.05, name="KP")
        r = {k: c.gauss_rel(v, 0.01 / 3, name=k) for k, v in rvals}
    c.vsource("vdd", "0", 5.0)
    c.resistor("vdd", "g", r["R1"])
    c.resistor("g", "0", r["R2"])
    c.resistor("vdd", "d", r["RD"])
    c.resistor("s", "0", r["RS"])
    c.nmos("d", "g", "s", vto=vto, kp=kp, lam=0.02, w=10e-6, l=1e-6)
    return c


INV_STAGES = 12
INV_RC_NODES = 16
INV_VDD = 1.8


def inverter_chain() -> Circuit:
    """12-stage self-biased CMOS inverter chain with a 16-node RC load: 64 unknowns.

    Stage k (k = 1..12) has nodes g_k (gate), o_k (output), n_k (NMOS source)
    and p_k (PMOS source):
      * NMOS  d=o_k g=g_k s=n_k, W/L = 10, KP_N in group (k-1)//3;
      * PMOS  d=o_k g=g_k s=p_k, W/L = 22, KP_P in group (k-1)//3;
      * R_S = 1k from n_k to ground; p_k reaches the ideal 1.8 V rail through
        R_S = 1k, written as its Norton equivalent (1k to ground plus
        1.8 mA injected into p_k), so the rail adds no unknown;
      * feedback R_F = 100k o_k-g_k biases the stage at its trip point
        (midrail); coupling R_C = 100k o_{k-1}-g_k, stage 1 driven through
        R_C from an ideal 0.9 V (Norton again).
    RC load: r_1..r_16 in a ladder from o_12, R_j~U(1,10) kOhm, C_j~U(1,10) fF
    to ground (open at DC).
    Varied: VTO_N[12]~N(0.7,0.03), VTO_P[12]~N(-0.7,0.03), KP_N[4]~N(110e-6,5%),
    KP_P[4]~N(50e-6,5%) — slots 0..31, the 32 DC-relevant parameters — then
    R_j (slots 32..47) and C_j (48..63).  LAMBDA = 0.02.
    """
    c = Circuit("inverter_chain")
    vton = [c.gauss(0.7, 0.03, name=f"VTO_N{k}") for k in range(INV_STAGES)]
    vtop = [c.gauss(-0.7, 0.03, name=f"VTO_P{k}") for k in range(INV_STAGES)]
    kpn = [c.gauss(110e-6, 0.05 * 110e-6, name=f"KP_N{g}") for g in range(4)]
    kpp = [c.gauss(50e-6, 0.05 * 50e-6, name=f"KP_P{g}") for g in range(4)]
    rl = [c.uniform(1e3, 10e3, name=f"RL{j}") for j in range(INV_RC_NODES)]
    cl = [c.uniform(1e-15, 10e-15, name=f"CL{j}") for j in range(INV_RC_NODES)]
    rs, rf, rc = c.const(1e3, "RS"), c.const(100e3, "RF"), c.const(100e3, "RC")
    i_rail = c.const(INV_VDD / 1e3, "I_RAIL")
    i_in = c.const(0.9 / 100e3, "I_IN")
    lam = c.const(0.02, "LAMBDA")
    wn, wp, l = c.const(10e-6, "WN"), c.const(22e-6, "WP"), c.const(1e-6, "L")
    # node numbering: stage k -> g, o, n, p; then the RC ladder
    for k in range(1, INV_STAGES + 1):
        for t in "gonp":
            c.node(f"{t}{k}")
    for j in range(1, INV_RC_NODES + 1):
        c.node(f"r{j}")
    for k in range(1, INV_STAGES + 1):
        grp = (k - 1) // 3
        c.nmos(f"o{k}", f"g{k}", f"n{k}", vto=vton[k - 1], kp=kpn[grp], lam=lam, w=wn, l=l)
        c.pmos(f"o{k}", f"g{k}", f"p{k}", vto=vtop[k - 1], kp=kpp[grp], lam=lam, w=wp, l=l)
        c.resistor(f"n{k}", "0", rs)
        c.resistor(f"p{k}", "0", rs)
        c.isource("0", f"p{k}", i_rail)
        c.resistor(f"o{k}", f"g{k}", rf)
        if k == 1:
            c.resistor("g1", "0", rc)
            c.isource("0", "g1", i_in)
        else:
            c.resistor(f"o{k - 1}", f"g{k}", rc)
    prev = f"o{INV_STAGES}"
    for j in range(1, INV_RC_NODES + 1):
        c.resistor(prev, f"r{j}", rl[j - 1])
        c.capacitor(f"r{j}", "0", cl[j - 1])
        prev = f"r{j}"
    assert c.size == 64
    return c


GRID_N = 32
GRID_DIODE_EVERY = 8


def grid_diodes(N: int = GRID_N, every: int = GRID_DIODE_EVERY) -> Circuit:
    """N x N resistive grid with diodes: the sparse FP64 sweep workload.

    Node (r, c) is unknown r*N + c.  Every horizontal and vertical edge is a
    resistor R~U(0.9, 1.1) kOhm (2N(N-1) slots, row edges first); the four
    corner nodes (0,0), (0,N-1), (N-1,0), (N-1,N-1) are driven by ideal
    sources of 1.0, 1.7667, 2.5333 and 3.3 V (1.0 + k * 2.3/3, adding four
    branch unknowns); every `every`-th node, r*N + c = every/2 (mod every),
    has a diode to ground with IS~logU(1e-15, 1e-13) A, N = 1.0.  The offset
    keeps diodes off the corner nodes, where a diode straight across an ideal
    1-3.3 V source would carry hundreds of amperes to exp-overflow currents.
    At N=32: 1,024 nodes + 4 branches = 1,028 unknowns, 1,984 resistors,
    128 diodes.
    """
    c = Circuit(f"grid_diodes_{N}")
    for r in range(N):
        for col in range(N):
            c.node(f"n{r}_{col}")
    redges = [((r, col), (r, col + 1)) for r in range(N) for col in range(N - 1)]
    cedges = [((r, col), (r + 1, col)) for r in range(N - 1) for col in range(N)]
    rs = [c.uniform(0.9e3, 1.1e3, name=f"R{k}") for k in range(len(redges) + len(cedges))]
    dn = [k for k in range(N * N) if k % every == every // 2]
    isl = [c.loguniform(1e-15, 1e-13, name=f"IS{k}") for k in dn]
    nd = c.const(1.0, "N")
    name = lambda rc: f"n{rc[0]}_{rc[1]}"
    for (a, b), r in zip(redges + cedges, rs):
        c.resistor(name(a), name(b), r)
    corners = [(0, 0), (0, N - 1), (N - 1, 0), (N - 1, N - 1)]
    for k, rc in enumerate(corners):
        c.vsource(name(rc), "0", c.const(1.0 + k * 2.3 / 3, f"V{k}"))
    for k, s in zip(dn, isl):
        c.diode(name((k // N, k % N)), "0", s, nd)
    return c


# Thermal voltage at 300.15 K (crates/spicier-devices/src/diode.rs:41-45)
THERMAL_VOLTAGE = 1.380649e-23 * 300.15 / 1.602176634e-19


def newton_system(circuit: Circuit, params, x):
    """The linear systems Newton steps solve at operating points x:
    resistors, DC voltage and current sources, and diodes linearised at x
    without junction limiting (the companion conductance
    gd = max(IS e^(vd/(N Vt)) / (N Vt), 1e-12) and current
    ieq = IS (e^(vd/(N Vt)) - 1) - gd vd of diode.rs:91-131, as the device
    kernels stamp them), no Gmin.  params [B][P] and x [B][n] are a sweep's
    parameters and solutions; the solution of system b is x[b] up to the
    Newton tolerance when x[b] converged.  Returns (n, rows, cols,
    vals[B][nnz], rhs[B][n]): one triplet pattern for the batch, duplicates
    kept (the solvers sum them).  Other element kinds raise ValueError."""
    params = np.asarray(params, dtype=np.float64)
    x = np.asarray(x, dtype=np.float64)
    B, n = x.shape[0], circuit.size
    rows, cols, vals = [], [], []
    rhs = np.zeros((B, n))

    def v(node):  # node voltage per instance (ground 0)
        return x[:, node - 1] if node > 0 else np.zeros(B)

    def add(i, j, val):  # MNA entry (i, j) of node indices, ground skipped
        if i > 0 and j > 0:
            rows.append(i - 1); cols.append(j - 1); vals.append(val)

    branch = circuit.num_nodes
    for kind, nd, p in circuit.elements:
        a, b = nd[0], nd[1]
        if kind == A.SPX_EL_RESISTOR:
            g = 1.0 / params[:, p[0]]
            add(a, a, g); add(b, b, g); add(a, b, -g); add(b, a, -g)
        elif kind == A.SPX_EL_VSOURCE:
            one = np.ones(B)
            if a > 0:
                rows += [a - 1, branch]; cols += [branch, a - 1]; vals += [one, one]
            if b > 0:
                rows += [b - 1, branch]; cols += [branch, b - 1]; vals += [-one, -one]
            rhs[:, branch] = params[:, p[0]]
            branch += 1
        elif kind == A.SPX_EL_ISOURCE:
            i = params[:, p[0]]
            if a > 0:
                rhs[:, a - 1] -= i
            if b > 0:
                rhs[:, b - 1] += i
        elif kind == A.SPX_EL_DIODE:
            is_, nvt = params[:, p[0]], params[:, p[1]] * THERMAL_VOLTAGE
            vd = v(a) - v(b)
            e = np.exp(vd / nvt)
            gd = np.maximum(is_ * e / nvt, 1e-12)
            ieq = is_ * (e - 1.0) - gd * vd
            add(a, a, gd); add(b, b, gd); add(a, b, -gd); add(b, a, -gd)
            if a > 0:
                rhs[:, a - 1] -= ieq
            if b > 0:
                rhs[:, b - 1] += ieq
        else:
            raise ValueError(f"newton_system: element kind {kind} is not linearised here")
    return n, np.array(rows, dtype=np.int32), np.array(cols, dtype=np.int32), \
        np.stack(vals, axis=1), rhs


def grid_linear_batch(N: int = GRID_N, batch: int = 4096, seed: int = 0, g_diode: float = 0.04):
    """Triplets of `batch` linear systems of the N x N grid workload's
    Jacobian family: edge conductances 1/R with R~U(0.9, 1.1) kOhm, the four
    corner sources as MNA branch rows (zero diagonal), and at every diode
    node the junction conductance g_diode (about a conducting diode's
    1 mA / 26 mV).  Right-hand sides are the corner voltages.  Returns
    (n, rows, cols, values[batch][nnz], rhs[batch][n]); duplicates are
    stamped as separate triplets (summed by the solvers)."""
    rng = np.random.default_rng(seed)
    nn = N * N
    n = nn + 4
    rows, cols, sign, which = [], [], [], []
    e = 0
    for r in range(N):
        for c in range(N):
            k = r * N + c
            for nb in ((k + 1) if c + 1 < N else None, (k + N) if r + 1 < N else None):
                if nb is None:
                    continue
                for (i, j, sg) in ((k, k, 1.0), (nb, nb, 1.0), (k, nb, -1.0), (nb, k, -1.0)):
                    rows.append(i); cols.append(j); sign.append(sg); which.append(e)
                e += 1
    g = 1.0 / rng.uniform(0.9e3, 1.1e3, size=(batch, e))
    vals = g[:, which] * np.array(sign)
    extra_r, extra_c, extra_v = [], [], []
    for k in range(nn):
        if k % GRID_DIODE_EVERY == GRID_DIODE_EVERY // 2:
            extra_r.append(k); extra_c.append(k); extra_v.append(g_diode)
    corners = (0, N - 1, N * (N - 1), N * N - 1)
    for s_, node in enumerate(corners):
        extra_r += [node, nn + s_]; extra_c += [nn + s_, node]; extra_v += [1.0, 1.0]
    rows = np.array(rows + extra_r, dtype=np.int32)
    cols = np.array(cols + extra_c, dtype=np.int32)
    vals = np.concatenate([vals, np.tile(np.array(extra_v), (batch, 1))], axis=1)
    rhs = np.zeros((batch, n))
    rhs[:, nn:] = [1.0 + k * 2.3 / 3 for k in range(4)]
    return n, rows, cols, vals, rhs


TRAN_PULSE = dict(v1=0.0, v2=1.8, td=0.0, tr=10e-12, tf=10e-12, pw=1e-9, per=2e-9)


def ladder(n: int, seed: int = 0) -> Circuit:
    """Random R/C/diode/MOSFET ladder with exactly n unknowns, driven by a
    PULSE source (0 -> 1.8 V, 10 ps rise, 1 ns wide): the mixed-size
    transient workload.

    Nodes in, l1 .. l(n-2) plus the source branch.  Rung k joins l(k-1) (or
    in) to l(k) through R~U(1,10) kOhm and loads l(k) with C~U(1,10) fF to
    ground; a structure seed picks for each rung (from k = 2 on) nothing, a
    diode l(k) -> 0 (IS~logU(1e-15,1e-13) A, N = 1), or an NMOS pull-down
    d=l(k), g=l(k-1), s=0 (VTO~N(0.7,0.03) V, KP~N(110e-6, 5%), LAMBDA
    0.02, W/L = 10), with probabilities 0.6 / 0.2 / 0.2.  The structure is
    fixed by `seed`; instances differ only in parameter values.
    """
    if n < 3:
        raise ValueError("a ladder needs n >= 3 unknowns")
    rng = np.random.default_rng(seed)
    c = Circuit(f"ladder_{n}_{seed}")
    c.node("in")
    for k in range(1, n - 1):
        c.node(f"l{k}")
    pulse = [c.const(TRAN_PULSE[k], k.upper()) for k in ("v1", "v2", "td", "tr", "tf", "pw", "per")]
    c.vpulse("in", "0", *pulse)
    lam, w, l, nd = c.const(0.02, "LAMBDA"), c.const(10e-6, "W"), c.const(1e-6, "L"), c.const(1.0, "N")
    prev = "in"
    for k in range(1, n - 1):
        node = f"l{k}"
        c.resistor(prev, node, c.uniform(1e3, 10e3, name=f"R{k}"))
        c.capacitor(node, "0", c.uniform(1e-15, 10e-15, name=f"C{k}"))
        pick = rng.random() if k >= 2 else 0.0
        if 0.6 <= pick < 0.8:
            c.diode(node, "0", c.loguniform(1e-15, 1e-13, name=f"IS{k}"), nd)
        elif pick >= 0.8:
            c.nmos(node, f"l{k - 1}", "0", vto=c.gauss(0.7, 0.03, name=f"VTO{k}"),
                   kp=c.gauss(110e-6, 0.05 * 110e-6, name=f"KP{k}"), lam=lam, w=w, l=l)
        prev = node
    assert c.size == n
    return c


def rlc_ladder(sections: int, seed: int = 0, diodes: bool = False) -> Circuit:
    """Low-pass RLC ladder driven by the ladder's PULSE source: section k
    runs s(k-1) (or in) -> R~U(10,100) Ohm -> m(k) -> L~U(1,10) nH -> s(k),
    with C~U(0.1,1) pF from s(k) to ground; a 1 kOhm load ends the line.
    With `diodes`, every other section's output also clamps to ground
    through a diode (IS~logU(1e-15,1e-13) A).  2 + 3 * sections unknowns:
    in, m(k), s(k), the source branch and one branch per inductor.  `seed`
    names the circuit; instances differ only in parameter values."""
    if sections < 1:
        raise ValueError("an RLC ladder needs at least one section")
    c = Circuit(f"rlc_ladder_{sections}_{seed}{'_d' if diodes else ''}")
    pulse = [c.const(TRAN_PULSE[k], k.upper()) for k in ("v1", "v2", "td", "tr", "tf", "pw", "per")]
    c.vpulse("in", "0", *pulse)
    nd = c.const(1.0, "N")
    prev = "in"
    for k in range(1, sections + 1):
        c.resistor(prev, f"m{k}", c.uniform(10.0, 100.0, name=f"R{k}"))
        c.inductor(f"m{k}", f"s{k}", c.uniform(1e-9, 10e-9, name=f"L{k}"))
        c.capacitor(f"s{k}", "0", c.uniform(0.1e-12, 1e-12, name=f"C{k}"))
        if diodes and k % 2 == 0:
            c.diode(f"s{k}", "0", c.loguniform(1e-15, 1e-13, name=f"IS{k}"), nd)
        prev = f"s{k}"
    c.resistor(prev, "0", c.const(1e3, "RLOAD"))
    assert c.size == 2 + 3 * sections
    return c


def controlled_ladder(sections: int, seed: int = 0) -> Circuit:
    """Cascade of buffered RC sections with every controlled-source kind,
    driven by the ladder's PULSE source.  Section k: a 0 V sense source
    from the previous buffer output (or in) to s(k); R~U(1,10) kOhm from
    s(k) to a(k) and C~U(0.1,1) pF from a(k) to ground; F injects 0.1 times
    the sense current into a(k); E buffers b(k) = g a(k), g~U(0.5,0.9), with
    a 10 kOhm load; H drives h(k) = 1 kOhm times the sense current into
    10 kOhm; G sinks gm a(k), gm~U(0.1,1) mS, from node q(k) loaded by 1 kOhm.
    8 * sections + 2 unknowns; instances differ only in parameter values."""
    if sections < 1:
        raise ValueError("a controlled ladder needs at least one section")
    c = Circuit(f"controlled_ladder_{sections}_{seed}")
    pulse = [c.const(TRAN_PULSE[k], k.upper()) for k in ("v1", "v2", "td", "tr", "tf", "pw", "per")]
    c.vpulse("in", "0", *pulse)
    zero, rl, rm, fk = c.const(0.0, "ZERO"), c.const(10e3, "RL"), c.const(1e3, "RM"), c.const(0.1, "FK")
    prev = "in"
    for k in range(1, sections + 1):
        vs = c.vsource(prev, f"s{k}", zero)
        c.resistor(f"s{k}", f"a{k}", c.uniform(1e3, 10e3, name=f"R{k}"))
        c.capacitor(f"a{k}", "0", c.uniform(0.1e-12, 1e-12, name=f"C{k}"))
        c.cccs("0", f"a{k}", vs, fk)
        c.vcvs(f"b{k}", "0", f"a{k}", "0", c.uniform(0.5, 0.9, name=f"E{k}"))
        c.resistor(f"b{k}", "0", rl)
        c.ccvs(f"h{k}", "0", vs, rm)
        c.resistor(f"h{k}", "0", rl)
        c.vccs("0", f"q{k}", f"a{k}", "0", c.uniform(1e-4, 1e-3, name=f"G{k}"))
        c.resistor(f"q{k}", "0", rm)
        prev = f"b{k}"
    assert c.size == 8 * sections + 2
    return c


def rc_step(r=1e3, cap=1e-12, v=1.0) -> Circuit:
    """PULSE 0 -> v (no rise) through R into C: the Backward-Euler step
    response v_k = v (1 - (1 + h/RC)^-k) (unknowns in, out, source branch)."""
    c = Circuit("rc_step")
    c.vpulse("in", "0", 0.0, v, 0.0, 0.0, 0.0, 1.0, 0.0)
    c.resistor("in", "out", r)
    c.capacitor("out", "0", cap)
    return c


def diode_resistor(v=5.0, r=1e3, is_=1e-14, n=1.0) -> Circuit:
    """V - R - D to ground (the reference's 5 V / 1 kOhm diode check,
    crates/spicier-parser/tests/ngspice_validation.rs:705-768)."""
    c = Circuit("diode_resistor")
    c.vsource("in", "0", v)
    c.resistor("in", "a", r)
    c.diode("a", "0", is_, n)
    return c


def voltage_divider(r1=1e3, r2=1e3, v=10.0, sigma=0.05) -> Circuit:
    """The reference's SimpleDividerFactory sweep fixture
    (crates/spicier-batched-sweep/src/sweep.rs:236-284): V - R1 - out - R2 - 0,
    R2 varied as a ParameterVariation."""
    c = Circuit("voltage_divider")
    r2p = c.gauss_rel(r2, sigma, name="R2")
    c.vsource("in", "0", v)
    c.resistor("in", "out", r1)
    c.resistor("out", "0", r2p)
    return c


def nmos_cs_validation() -> Circuit:
    """NMOS common-source circuit of ngspice_validation.rs:834-890 (VDD 5 V,
    RD 1k, VG 2 V, default-ish L1 NMOS); the reference asserts V(d) in (3,5)."""
    c = Circuit("nmos_cs_validation")
    c.vsource("vdd", "0", 5.0)
    c.vsource("g", "0", 2.0)
    c.resistor("vdd", "d", 1e3)
    c.nmos("d", "g", "0", vto=0.7, kp=2e-5, lam=0.0, w=10e-6, l=1e-6)
    return c


def nominal_values(circuit: Circuit) -> np.ndarray:
    """Nominal value of every slot: a for constants and Gaussians, the
    midpoint for uniform slots, the geometric mean for log-uniform ones."""
    out = []
    for kind, _, a, b, _, _ in circuit.specs:
        if kind == A.SPX_P_UNIFORM:
            out.append(0.5 * (a + b))
        elif kind == A.SPX_P_LOGUNIFORM:
            out.append(float(np.sqrt(a * b)))
        else:
            out.append(a)
    return np.array(out, dtype=np.float64)


def instance_params(circuit: Circuit, batch: int) -> np.ndarray:
    """Parameter matrix [batch][P] with every slot at its nominal value."""
    return np.tile(nominal_values(circuit), (batch, 1))
