"""The deck front end (host only): MNA numbering as ngspice assigns it,
subcircuit flattening, values and expressions, and error reporting.

The numbering rules are restated from the reference's sources and checked
here on decks whose expected numbering follows from those rules by hand:
nodes by first appearance in the flattened deck (INPpas2 -> INPtermInsert
-> CKTnewNode), subcircuit-local nodes named "<instance without x>.<node>"
(subckt.c:1107-1127), source branches after the nodes in reverse deck
order (VSRCsetup walks the instance list, cktcrte.c:62-64 puts new
instances at its head).  No ngspice binary runs here, so the numbering is
pinned by these hand-derived cases, not by the reference's own output.
"""
import math

import pytest

from engine_path import ngmi, workloads

INV_CHAIN = """inverter chain
* comment line
.subckt inv a y vdd
mp y a vdd vdd pch l=45n w=400n
mn y a mid 0 nch l=45n w=200n
rm mid 0 1   ; local node
.ends
v1 in 0 pulse(0 1 1n 10p 10p 1n 2n)
vdd vdd 0 dc 1.0
x1 in n2 vdd inv
x2 n2 out vdd inv
c1 out 0 1f
.model nch nmos level=54 toxe=1.8n vth0=0.4
.model pch pmos level=54
+ toxe=1.8n vth0=-0.4
.tran 1p 1n
.end
"""


def test_numbering_follows_first_appearance_and_reverse_branches():
    nl = ngmi.Netlist(INV_CHAIN)
    assert nl.title == 'inverter chain'
    names = nl.names()
    # v1: in; vdd: vdd; x1/mp: n2 (y), in, vdd; x1/mn: n2, in, 1.mid; x1/rm: 1.mid;
    # x2/mp: out, n2, vdd; x2/mn: out, n2, 2.mid; c1: out
    assert names[:7] == ['0', 'in', 'vdd', 'n2', '1.mid', 'out', '2.mid']
    assert nl.size() == {'nodes': 6, 'unknowns': 8, 'elements': 9}
    # branches in reverse deck order: vdd (second V card) first
    assert nl.find('vdd#branch') == 7
    assert nl.find('v1#branch') == 8
    assert nl.find('VDD') == 2          # names are case-insensitive
    els = nl.elements()
    assert [e[0] for e in els] == ['v1', 'vdd', 'm.1.mp', 'm.1.mn', 'r.1.rm', 'm.2.mp', 'm.2.mn',
                                   'r.2.rm', 'c1']
    assert els[2][2] == [3, 1, 2, 2]
    assert els[3][2] == [3, 1, 4, 0]
    assert els[8] == ('c1', 'c', [5, 0], pytest.approx(1e-15, rel=1e-15))
    assert nl.tran()['tstop'] == pytest.approx(1e-9)


def test_values_suffixes_and_expressions():
    deck = """values
.param cl=0.5f rbase=1.5k k2={2*2}
r1 a 0 1.5k
r2 a 0 2meg
r3 a 0 {sqrt(4)*1k}
r4 a 0 'rbase*k2'
c1 a 0 {2*cl}
c2 a 0 10pF
r5 a 0 3.3e2ohm
r6 a 0 {max(3, 2^3) - -1}
vdd a 0 {k2/4}
.tran 1p 1n
.end
"""
    nl = ngmi.Netlist(deck)
    vals = [e[3] for e in nl.elements()]
    expect = [1.5e3, 2e6, 2e3, 6e3, 1e-15, 10e-12, 330.0, 9.0, 1.0]
    for v, e in zip(vals, expect):
        assert v == pytest.approx(e, rel=1e-14), (vals, expect)


def test_monte_carlo_functions_follow_the_seed():
    deck = """mc
.param dv=agauss(0, 30m, 1) u=aunif(1, 0.1)
c1 a 0 {1f + dv*1f}
r1 a 0 {u}
vdd a 0 1
.tran 1p 1n
.end
"""
    a = [e[3] for e in ngmi.Netlist(deck, seed=7).elements()]
    b = [e[3] for e in ngmi.Netlist(deck, seed=7).elements()]
    c = [e[3] for e in ngmi.Netlist(deck, seed=8).elements()]
    assert a == b
    assert a != c
    assert 0.9 <= a[1] <= 1.1


def test_ring_deck_matches_the_description():
    d = workloads.ring_oscillator(stages=11, tstop=2e-9)
    nl = ngmi.Netlist(workloads.deck(d))
    sz = nl.size()
    assert sz['nodes'] == d['nnodes'] and sz['unknowns'] == d['nnodes'] + len(d['vsrc'])
    # the source written last-first lands at unknown nnodes + 1 + its index
    assert nl.find('v0#branch') == d['nnodes'] + 1
    a = nl.circuit().analyze()
    b = ngmi.Circuit(d).analyze()
    assert a['unknowns'] == b['unknowns'] and a['lu_nonzeros'] == b['lu_nonzeros']
    # element values survive the text round trip exactly
    caps = [e[3] for e in nl.elements() if e[1] == 'c']
    assert caps == [c[2] for c in d['caps']]


def test_adder_deck_matches_the_description():
    d = workloads.mc_adder(0)
    nl = ngmi.Netlist(workloads.deck(d))
    assert nl.size()['unknowns'] == d['nnodes'] + len(d['vsrc'])
    assert nl.circuit().analyze()['unknowns'] == ngmi.Circuit(d).analyze()['unknowns']


@pytest.mark.parametrize('deck,msg', [
    ('t\nd1 a 0 dmod\n.end\n', "line 2: element type 'd'"),
    ('t\nm1 a b 0 0 nope\n.end\n', 'line 2: unknown model nope'),
    ('t\n.subckt s a\nr1 a 0 1\n', 'has no .ends'),
    ('t\n.model q npn\n.end\n', "model type 'npn'"),
    ('t\nr1 a 0 {1 +}\n.end\n', 'line 2'),
    ('t\nx1 a b s\n.subckt s a\nr1 a 0 1\n.ends\n.end\n', 'has 1 ports, instance gives 2'),
    ('t\nv1 a 0 sin(0 1 1g)\n.end\n', "source function 'sin'"),
])
def test_errors_name_the_line(deck, msg):
    with pytest.raises(ngmi.EngineError):
        ngmi.Netlist(deck)
    assert msg in ngmi.last_error()


def test_number_parser_corner_cases():
    deck = """n
r1 a 0 1meg
r2 a 0 1m
r3 a 0 1mil
r4 a 0 .5k
r5 a 0 1e-3k
r6 a 0 2.5E+1
vdd a 0 1
.end
"""
    vals = [e[3] for e in ngmi.Netlist(deck).elements()][:6]
    expect = [1e6, 1e-3, 25.4e-6, 500.0, 1.0, 25.0]
    assert all(math.isclose(v, e, rel_tol=1e-14) for v, e in zip(vals, expect)), vals


def test_tran_uic_reaches_the_circuit():
    """'.tran ... uic' sets MODEUIC on the circuit (TRANuic, transetp.c:51-53)"""
    d = workloads.with_uic(workloads.rc_charge(), {2: 0.5})
    text = workloads.deck(d)
    assert '.tran 2e-11 5e-09 uic' in text
    nl = ngmi.Netlist(text)
    assert nl.tran()['uic'] == 1
    assert nl.circuit().analyze()['unknowns'] == 3
