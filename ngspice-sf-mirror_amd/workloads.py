"""Synthetic workloads with seeded, counter-based random streams (numpy's
Philox), so the engine and the oracle see identical values.

device_eval_workload: 2**20 isolated BSIM4 instances of the 45 nm base card
(NMOS vth0 = 0.40 V, PMOS vth0 = -0.40 V, toxe = 1.8 nm, other parameters at
their BSIM4 4.8.3 defaults), Vgs, Vds ~ U[0, 1] V, Vbs ~ U[-0.3, 0] V,
W log-uniform on [100 nm, 10 um], L log-uniform on [45 nm, 1 um], a vth0
offset (instance parameter delvto) ~ N(0, 30 mV); the first half NMOS, the
second half PMOS.  Each instance's 16 Jacobian entries land in a random
2**21-row sparse matrix: its four terminals map to four distinct random rows.
"""
import numpy as np

BASE_NMOS = {'nmos': 1, 'vth0': 0.40, 'toxe': 1.8e-9}
BASE_PMOS = {'pmos': 1, 'vth0': -0.40, 'toxe': 1.8e-9}

EVAL_N = 1 << 20
EVAL_ROWS = 1 << 21
H_STEP = 1e-11          # 10 ps backward-Euler step used by the evaluation workload


def rng(config, seed):
    return np.random.Generator(np.random.Philox(key=[config, seed]))


def device_eval_workload(n=EVAL_N, seed=1, rows=EVAL_ROWS, with_matrix=True):
    """Instance parameters, terminal voltages and the element map.

    Returns a dict with
      model_idx[n] (0 NMOS, 1 PMOS), params {l, w, delvto},
      vterm[4, n] (node voltages d g s b in circuit polarity, source at 0),
      vds, vgs, vbs (device polarity), and, with_matrix, elements
      rows_e[16 n], cols_e[16 n] where element e = slot * n + i is stamp slot
      `slot` (row-major over d g s b) of instance i, and rhs_rows[4 n].
    """
    g = rng(1, seed)
    vgs = g.uniform(0.0, 1.0, n)
    vds = g.uniform(0.0, 1.0, n)
    vbs = g.uniform(-0.3, 0.0, n)
    w = np.exp(g.uniform(np.log(100e-9), np.log(10e-6), n))
    l = np.exp(g.uniform(np.log(45e-9), np.log(1e-6), n))
    delvto = g.normal(0.0, 0.030, n)
    model_idx = np.zeros(n, dtype=np.int32)
    model_idx[n // 2:] = 1
    typ = np.where(model_idx == 0, 1.0, -1.0)
    vterm = np.stack([typ * vds, typ * vgs, np.zeros(n), typ * vbs])
    out = dict(n=n, model_idx=model_idx, params={'l': l, 'w': w, 'delvto': delvto},
               vterm=np.ascontiguousarray(vterm), vds=vds, vgs=vgs, vbs=vbs, type=typ)
    if with_matrix:
        # four distinct rows per instance: draw, then resolve collisions by offset
        term = g.integers(0, rows, size=(4, n), dtype=np.int64)
        for a in range(1, 4):
            for _ in range(8):
                clash = np.zeros(n, dtype=bool)
                for b in range(a):
                    clash |= term[a] == term[b]
                if not clash.any():
                    break
                term[a, clash] = (term[a, clash] + 1 + a) % rows
        slot_r = np.repeat(np.arange(4), 4)
        slot_c = np.tile(np.arange(4), 4)
        out['rows_e'] = term[slot_r].reshape(-1).astype(np.int32)
        out['cols_e'] = term[slot_c].reshape(-1).astype(np.int32)
        out['rhs_rows'] = term.reshape(-1).astype(np.int64)
        out['nrows'] = rows
        out['terms'] = term
    return out


def eval_states(wl):
    """State levels 0..2 (29 x n each) for the evaluation workload: level 0
    holds the previous iterate's terminal voltages equal to this one's (so
    Newton limiting leaves them unchanged), charge history at zero."""
    n = wl['n']
    s0 = np.zeros((29, n))
    s0[3] = wl['vds']          # vds
    s0[2] = wl['vgs']          # vgs
    s0[1] = wl['vbs']          # vbs
    s0[0] = wl['vbs'] - wl['vds']   # vbd
    return [s0, np.zeros((29, n)), np.zeros((29, n))]


# ---------------------------------------------------------------------------
# Transient circuits.  A circuit is a plain dict (the same description feeds
# the engine and the oracle):
#   nnodes, models {key: card}, mos [(key, d, g, s, b, {inst params})],
#   caps [(p, n, C)], res [(p, n, R)], vsrc [(p, n, dc, pulse or None)],
#   ic {node: v}, tstep, tstop.
# Node 0 is ground; nodes are numbered 1..nnodes.
# ---------------------------------------------------------------------------

RING_STAGES = 101
RING_WN, RING_WP, RING_L = 200e-9, 400e-9, 45e-9
VDD = 1.0


def ring_oscillator(stages=RING_STAGES, cload=1e-15, ic_node1=0.51, tstep=10e-12, tstop=200e-9):
    """CMOS inverter ring: stage k drives node k+1 from node k (node
    stages+1 wraps to node 1), 1 fF from every ring node to ground, VDD on
    node stages+1; node 1 starts from .ic v(1) = 0.51 V."""
    vdd = stages + 1
    mos, caps = [], []
    for k in range(1, stages + 1):
        a = k
        y = k + 1 if k < stages else 1
        mos.append(('n', y, a, 0, 0, {'w': RING_WN, 'l': RING_L}))
        mos.append(('p', y, a, vdd, vdd, {'w': RING_WP, 'l': RING_L}))
        caps.append((k, 0, cload))
    return {'nnodes': vdd, 'models': {'n': dict(BASE_NMOS), 'p': dict(BASE_PMOS)}, 'mos': mos,
            'caps': caps, 'res': [], 'vsrc': [(vdd, 0, VDD, None)], 'ic': {1: ic_node1},
            'tstep': tstep, 'tstop': tstop}


UNEVEN_RINGS = 2048
UNEVEN_TSTOP = 10e-9


def uneven_rings(count=UNEVEN_RINGS, seed=3, tstop=UNEVEN_TSTOP):
    """The uneven batch: ring oscillators whose stage counts are drawn
    uniformly from the odd numbers 11..301 (Philox stream (3, seed)), each
    with its own sparsity pattern; .tran 10p 10n."""
    g = rng(3, seed)
    stages = 11 + 2 * g.integers(0, (301 - 11) // 2 + 1, size=count)
    return [ring_oscillator(stages=int(k), tstop=tstop) for k in stages]


def ring_with_divider(stages=11, tstop=0.3e-9):
    """A ring oscillator with a floating 0.3 V source between two extra
    nodes: node a to ground through 10 uOhm, node b to ground through
    100 kOhm and to ring node 1 through 10 kOhm.  The rows of the source's
    two nodes differ by ten orders in scale, so KLU's threshold pivoting
    (klu_kernel.c:370-530) takes an off-diagonal pivot at every fresh
    factorization on KLU's own ordering."""
    d = ring_oscillator(stages=stages, tstop=tstop)
    a, b = d['nnodes'] + 1, d['nnodes'] + 2
    d['nnodes'] = b
    d['vsrc'] = d['vsrc'] + [(a, b, 0.3, None)]
    d['res'] = [(a, 0, 1e-5), (b, 0, 1e5), (b, 1, 1e4)]
    return d


def inverter_chain(stages=5, cload=1e-15, tstep=10e-12, tstop=1e-9):
    """CMOS inverters in a chain: node 1 is driven by a source pulsing from
    0 to 1 V at 0.1 ns (rise and fall 20 ps, width 0.3 ns), stage k drives
    node k+1, 1 fF from every node to ground, VDD on node stages+2."""
    vdd = stages + 2
    mos, caps = [], []
    for k in range(1, stages + 1):
        mos.append(('n', k + 1, k, 0, 0, {'w': RING_WN, 'l': RING_L}))
        mos.append(('p', k + 1, k, vdd, vdd, {'w': RING_WP, 'l': RING_L}))
        caps.append((k + 1, 0, cload))
    pulse = (0.0, VDD, 0.1e-9, 20e-12, 20e-12, 0.3e-9, 1.0)
    return {'nnodes': vdd, 'models': {'n': dict(BASE_NMOS), 'p': dict(BASE_PMOS)}, 'mos': mos,
            'caps': caps, 'res': [], 'vsrc': [(vdd, 0, VDD, None), (1, 0, 0.0, pulse)], 'ic': {},
            'tstep': tstep, 'tstop': tstop}


def with_floating_node(desc, node=2, gshunt=0.0):
    """The circuit with one more node tied to `node` by a 2 fF capacitor
    alone: at DC its row and column are empty and the operating point's
    matrix is singular (E_SINGULAR) unless .options gshunt puts a
    conductance on every diagonal; gmin and source stepping then fail and
    CKTop ends with optran, the transient operating point."""
    d = dict(desc)
    f = d['nnodes'] + 1
    d['nnodes'] = f
    d['caps'] = list(d['caps']) + [(f, node, 2e-15)]
    d['options'] = dict(d.get('options', {}), gshunt=gshunt)
    return d


def rc_charge(v=1.0, r=1e3, c=1e-12, ic=None, uic=False, tstep=20e-12, tstop=5e-9):
    """A source of v volts on node 1 charging c through r into node 2;
    ic: {node: volts} for .ic; uic: .tran ... uic."""
    return {'nnodes': 2, 'models': {}, 'mos': [], 'caps': [(2, 0, c)], 'res': [(1, 2, r)],
            'vsrc': [(1, 0, v, None)], 'ic': dict(ic or {}), 'tstep': tstep, 'tstop': tstop,
            'options': {'uic': int(uic)}}


def with_uic(desc, ic=None):
    """desc under .tran ... uic, with its .ic values replaced by ic when
    given"""
    d = dict(desc)
    if ic is not None:
        d['ic'] = dict(ic)
    d['options'] = dict(desc.get('options', {}), uic=1)
    return d


def mirror_adder_topology(bits=16):
    """Node map and transistor list of a ripple-carry adder of 28-transistor
    mirror full adders (carry: 10 T, sum: 14 T, two output inverters).
    Returns (nnodes, mos list with model keys 'n'/'p' and no params, named
    nodes dict)."""
    nodes = {'vdd': 1}
    nxt = [2]

    def new(name):
        nodes[name] = nxt[0]
        nxt[0] += 1
        return nodes[name]

    A = [new('a%d' % i) for i in range(bits)]
    B = [new('b%d' % i) for i in range(bits)]
    cin = new('cin')
    vdd = 1
    mos = []
    ci = cin
    for i in range(bits):
        a, b = A[i], B[i]
        cob, sb = new('cob%d' % i), new('sb%d' % i)
        co, s = new('co%d' % i), new('s%d' % i)
        na, nx, pa, px = new('na%d' % i), new('nx%d' % i), new('pa%d' % i), new('px%d' % i)
        n1, n2, ny = new('n1_%d' % i), new('n2_%d' % i), new('ny%d' % i)
        p1, p2, py = new('p1_%d' % i), new('p2_%d' % i), new('py%d' % i)
        # carry: Co_b = not(a b + ci (a + b))
        mos += [('n', cob, a, na, 0), ('n', na, b, 0, 0),
                ('n', cob, ci, nx, 0), ('n', nx, a, 0, 0), ('n', nx, b, 0, 0),
                ('p', cob, a, pa, vdd), ('p', pa, b, vdd, vdd),
                ('p', cob, ci, px, vdd), ('p', px, a, vdd, vdd), ('p', px, b, vdd, vdd)]
        # sum: S_b = not(a b ci + Co_b (a + b + ci))
        mos += [('n', sb, a, n1, 0), ('n', n1, b, n2, 0), ('n', n2, ci, 0, 0),
                ('n', sb, cob, ny, 0), ('n', ny, a, 0, 0), ('n', ny, b, 0, 0), ('n', ny, ci, 0, 0),
                ('p', sb, a, p1, vdd), ('p', p1, b, p2, vdd), ('p', p2, ci, vdd, vdd),
                ('p', sb, cob, py, vdd), ('p', py, a, vdd, vdd), ('p', py, b, vdd, vdd),
                ('p', py, ci, vdd, vdd)]
        # output inverters
        mos += [('n', co, cob, 0, 0), ('p', co, cob, vdd, vdd),
                ('n', s, sb, 0, 0), ('p', s, sb, vdd, vdd)]
        ci = co
    return nxt[0] - 1, mos, nodes


ADDER_BITS = 16
ADDER_TSTEP, ADDER_TSTOP = 5e-12, 20e-9
ADDER_SWITCH = 1e-9
ADDER_EDGE = 20e-12


def mc_adder(seed, bits=ADDER_BITS, tstep=ADDER_TSTEP, tstop=ADDER_TSTOP):
    """One Monte Carlo instance of the ripple-carry adder: its own model card
    pair (toxe ~ N(1.8 nm, 2 %)), per transistor delvto ~ N(0, 30 mV) and
    L ~ N(45 nm, 3 %), W = 200 nm (NMOS) / 400 nm (PMOS), and its own input
    vector: every input moves from a random bit to a random bit at 1 ns
    (PULSE v1 v2 1n 20p 20p 1 2)."""
    nn, topo, names = mirror_adder_topology(bits)
    g = rng(2, seed)
    toxe = 1.8e-9 * (1.0 + 0.02 * g.standard_normal())
    models = {'n': dict(BASE_NMOS, toxe=toxe), 'p': dict(BASE_PMOS, toxe=toxe)}
    nm = len(topo)
    dv = g.normal(0.0, 0.030, nm)
    ll = 45e-9 * (1.0 + 0.03 * g.standard_normal(nm))
    mos = []
    for k, (key, d, gg, s, b) in enumerate(topo):
        w = RING_WN if key == 'n' else RING_WP
        mos.append((key, d, gg, s, b, {'w': w, 'l': float(ll[k]), 'delvto': float(dv[k])}))
    bits0 = g.integers(0, 2, 2 * bits + 1)
    bits1 = g.integers(0, 2, 2 * bits + 1)
    inputs = [names['a%d' % i] for i in range(bits)] + [names['b%d' % i] for i in range(bits)] \
        + [names['cin']]
    vsrc = [(1, 0, VDD, None)]
    for k, node in enumerate(inputs):
        v1, v2 = VDD * bits0[k], VDD * bits1[k]
        vsrc.append((node, 0, v1, (v1, v2, ADDER_SWITCH, ADDER_EDGE, ADDER_EDGE, 1.0, 2.0)))
    return {'nnodes': nn, 'models': models, 'mos': mos, 'caps': [], 'res': [], 'vsrc': vsrc,
            'ic': {}, 'tstep': tstep, 'tstop': tstop, 'names': names,
            'bits0': bits0, 'bits1': bits1}


def deck(desc, title='engine workload'):
    """A SPICE deck of a circuit description (node k named "k"; source j
    named "v<j>", written in reverse order so that the deck's branch
    numbering, the reverse of deck order, puts source j at unknown
    nnodes + 1 + j as the description does)."""
    out = [title]
    for key, card in desc['models'].items():
        typ = 'nmos' if card.get('nmos') else 'pmos'
        ps = ' '.join('%s=%r' % (k, float(v)) for k, v in card.items() if k not in ('nmos', 'pmos'))
        out.append('.model %s %s level=54 %s' % (key, typ, ps))
    for k, (key, d, g, s, b, ip) in enumerate(desc['mos']):
        ps = ' '.join('%s=%r' % (n, float(v)) for n, v in ip.items())
        out.append('m%d %d %d %d %d %s %s' % (k, d, g, s, b, key, ps))
    for k, (p, n, c) in enumerate(desc.get('caps', [])):
        out.append('c%d %d %d %r' % (k, p, n, float(c)))
    for k, (p, n, r) in enumerate(desc.get('res', [])):
        out.append('r%d %d %d %r' % (k, p, n, float(r)))
    for k in reversed(range(len(desc['vsrc']))):
        p, n, dc, pulse = desc['vsrc'][k]
        line = 'v%d %d %d dc %r' % (k, p, n, float(dc))
        if pulse is not None:
            line += ' pulse(%s)' % ' '.join(repr(float(x)) for x in pulse)
        out.append(line)
    ics = ' '.join('v(%d)=%r' % (k, float(v)) for k, v in desc.get('ic', {}).items())
    if ics:
        out.append('.ic ' + ics)
    o = {k: v for k, v in desc.get('options', {}).items() if k not in ('uic', 'optran')}
    if o:
        out.append('.options ' + ' '.join('%s=%s' % (k, ('gear' if v == 2 else 'trap')
                                                      if k == 'method' else repr(v))
                                          for k, v in o.items()))
    out.append('.tran %r %r%s' % (float(desc['tstep']), float(desc['tstop']),
                                  ' uic' if desc.get('options', {}).get('uic') else ''))
    out.append('.end')
    return '\n'.join(out) + '\n'


MESH_SIDE = 1024
MESH_LOADS = 16384
MESH_PAD_EVERY = 64
MESH_G = 10.0           # 0.1 ohm segments
MESH_C = 10e-15         # per node
MESH_LOAD_C = 1e-15     # inverter output load


def power_grid_jacobian(side=MESH_SIDE, loads=MESH_LOADS, pad_every=MESH_PAD_EVERY, seed=0, h=10e-12,
                        gmin=1e-12):
    """The Newton matrix of the power-grid mesh at one trapezoidal step h:
    a side x side RC mesh (0.1 ohm segments, 10 fF per node to ground),
    VDD pads (voltage sources to ground) every pad_every nodes in both
    directions, and `loads` CMOS inverters at seeded random mesh nodes
    (Philox stream (4, seed)), each PMOS drawing from its mesh node, its
    input driven by its own clock source, 1 fF on its output.  MOSFET
    small-signal values are seeded draws in the ranges a 45 nm inverter
    takes (gds 1e-6..1e-4 S, gm 0..2e-4 S, gate capacitances 0.05..0.2
    fF), stamped as BSIM4load stamps a drain/source row (b4ld.c:5056-5389);
    capacitors enter as (2/h) C, gmin on every node diagonal
    (LoadGmin_CSC).

    Unknowns (0-based, ground removed): mesh node (i, j) at i*side + j,
    inverter k's input at side^2 + 2k and output at side^2 + 2k + 1, then
    the pad branches, then the clock-source branches.  Returns (n, Ap, Ai,
    Ax) in CSC with ascending rows, Ap int64, Ai int32."""
    import scipy.sparse as sp
    g = rng(4, seed)
    nm = side * side
    ag0 = 2.0 / h
    pads = [(a * pad_every) * side + b * pad_every for a in range(side // pad_every)
            for b in range(side // pad_every)]
    npad = len(pads)
    n = nm + 2 * loads + npad + loads
    rows, cols, vals = [], [], []

    def add(r, c, v):
        rows.append(np.asarray(r, dtype=np.int64))
        cols.append(np.asarray(c, dtype=np.int64))
        vals.append(np.broadcast_to(np.asarray(v, dtype=np.float64), np.shape(r)).copy())

    def two_terminal(a, b, gv):
        add(a, a, gv), add(b, b, gv), add(a, b, -gv), add(b, a, -gv)

    ij = np.arange(nm).reshape(side, side)
    two_terminal(ij[:, :-1].ravel(), ij[:, 1:].ravel(), MESH_G)
    two_terminal(ij[:-1, :].ravel(), ij[1:, :].ravel(), MESH_G)
    node_diag = np.arange(nm + 2 * loads)
    add(node_diag, node_diag, gmin)
    add(np.arange(nm), np.arange(nm), ag0 * MESH_C)
    pads = np.array(pads, dtype=np.int64)
    br = nm + 2 * loads + np.arange(npad)
    add(pads, br, 1.0), add(br, pads, 1.0)
    k = np.arange(loads)
    tap = g.integers(0, nm, size=loads)
    vin = nm + 2 * k
    out = vin + 1
    cb = nm + 2 * loads + npad + k
    add(vin, cb, 1.0), add(cb, vin, 1.0)
    add(out, out, ag0 * MESH_LOAD_C)
    gds_n, gds_p = g.uniform(1e-6, 1e-4, loads), g.uniform(1e-6, 1e-4, loads)
    gm_n, gm_p = g.uniform(0, 2e-4, loads), g.uniform(0, 2e-4, loads)
    cgd_n, cgd_p = g.uniform(0.05e-15, 0.2e-15, loads), g.uniform(0.05e-15, 0.2e-15, loads)
    cgs_n, cgs_p = g.uniform(0.05e-15, 0.2e-15, loads), g.uniform(0.05e-15, 0.2e-15, loads)
    # NMOS d=out g=in s=b=ground: only the drain row remains
    add(out, out, gds_n), add(out, vin, gm_n)
    # PMOS d=out g=in s=b=tap (gmbs cancels with the bulk on the source)
    add(out, out, gds_p), add(out, vin, gm_p), add(out, tap, -(gds_p + gm_p))
    add(tap, out, -gds_p), add(tap, vin, -gm_p), add(tap, tap, gds_p + gm_p)
    # gate capacitances as (2/h) C two-terminal stamps
    two_terminal(vin, out, ag0 * (cgd_n + cgd_p))
    add(vin, vin, ag0 * cgs_n)
    two_terminal(vin, tap, ag0 * cgs_p)
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    v = np.concatenate(vals)
    A = sp.csc_matrix((v, (r, c)), shape=(n, n))
    A.sum_duplicates()
    A.sort_indices()
    return n, A.indptr.astype(np.int64), A.indices.astype(np.int32), A.data.astype(np.float64)


MESH_CLOCK_PERIOD = 1e-9
MESH_CLOCK_EDGE = 20e-12


def power_grid(side=MESH_SIDE, loads=None, pad_every=MESH_PAD_EVERY, seed=0, tstep=10e-12, tstop=10e-9):
    """The power-grid mesh as a circuit description: side x side nodes on a
    0.1 ohm segment grid (node (i, j) = 1 + i*side + j), 10 fF from every
    mesh node to ground, 1 V pads (voltage sources to ground) every
    pad_every nodes in both directions, and `loads` CMOS inverters at seeded
    mesh nodes (Philox stream (4, seed), the taps power_grid_jacobian uses):
    inverter k's PMOS draws from its mesh node, its input (node
    side^2 + 2k + 1) is driven by its own clock PULSE(0 1 td 20p 20p 480p 1n)
    with td uniform in [0, 1 ns], its output (node side^2 + 2k + 2) carries
    1 fF.  loads defaults to 16384 scaled with the mesh area."""
    if loads is None:
        loads = MESH_LOADS * side * side // (MESH_SIDE * MESH_SIDE)
    g = rng(4, seed)
    nm = side * side
    tap = g.integers(0, nm, size=loads)
    td = rng(6, seed).uniform(0.0, MESH_CLOCK_PERIOD, loads)
    ij = np.arange(nm).reshape(side, side) + 1
    res = [(int(a), int(b), 1.0 / MESH_G) for a, b in zip(ij[:, :-1].ravel(), ij[:, 1:].ravel())]
    res += [(int(a), int(b), 1.0 / MESH_G) for a, b in zip(ij[:-1, :].ravel(), ij[1:, :].ravel())]
    caps = [(int(u), 0, MESH_C) for u in range(1, nm + 1)]
    vsrc = [(int(ij[a * pad_every, b * pad_every]), 0, VDD, None)
            for a in range(side // pad_every) for b in range(side // pad_every)]
    mos = []
    for k in range(loads):
        vin, out, t = nm + 2 * k + 1, nm + 2 * k + 2, int(tap[k]) + 1
        mos.append(('n', out, vin, 0, 0, {'w': RING_WN, 'l': RING_L}))
        mos.append(('p', out, vin, t, t, {'w': RING_WP, 'l': RING_L}))
        caps.append((out, 0, MESH_LOAD_C))
        edge = MESH_CLOCK_EDGE
        vsrc.append((vin, 0, 0.0, (0.0, VDD, float(td[k]), edge, edge, MESH_CLOCK_PERIOD / 2 - edge,
                                   MESH_CLOCK_PERIOD)))
    return {'nnodes': nm + 2 * loads, 'models': {'n': dict(BASE_NMOS), 'p': dict(BASE_PMOS)}, 'mos': mos,
            'caps': caps, 'res': res, 'vsrc': vsrc, 'ic': {}, 'tstep': tstep, 'tstop': tstop}
