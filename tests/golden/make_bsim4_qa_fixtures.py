#!/usr/bin/env python3
"""Write tests/golden/bsim4_qa.json from the reference's BSIM4 CMC QA data.

Reads (in a container that has the reference tree):
  tests/bsim4/{nmos,pmos}/qaSpec                 test definitions
  tests/bsim4/{nmos,pmos}/parameters/*Parameters  model cards
  tests/bsim4/{nmos,pmos}/reference/*.standard    expected results
and stores, per DC test: the model card, extra model/instance parameters,
the bias grid and the expected terminal currents.  Only data is written.

usage: make_bsim4_qa_fixtures.py <reference_root> <out.json>
"""
import json
import os
import re
import sys


def parse_params(path):
    out = {}
    for line in open(path):
        m = re.match(r'\+\s*(\w+)\s*=\s*\(([^)]*)\)', line.strip())
        if m:
            out[m.group(1).lower()] = float(m.group(2))
    return out


def parse_spec(path):
    tests = []
    cur = None
    glob = {}
    for line in open(path):
        line = line.strip()
        if not line or line.startswith('//') or line.startswith('`'):
            continue
        key, _, rest = line.partition(' ')
        rest = rest.strip()
        if key == 'test':
            cur = {'name': rest, 'model_extra': {}, 'model_files': []}
            tests.append(cur)
            continue
        target = cur if cur is not None else glob
        if key in ('temperature', 'biases', 'biasList', 'biasSweep', 'outputs',
                   'instanceParameters', 'freq'):
            target[key] = rest
        elif key == 'modelParameters' and cur is not None:
            if '=' in rest:
                for kv in rest.split():
                    k, v = kv.split('=')
                    cur['model_extra'][k.lower()] = float(v)
            else:
                cur['model_files'].append(rest)
    return glob, tests


def parse_bias(s):
    out = {}
    for tok in s.split():
        m = re.match(r'V\((\w+)\)=(.*)', tok)
        out[m.group(1)] = m.group(2)
    return out


def main():
    ref, outp = sys.argv[1], sys.argv[2]
    data = {'source': 'tests/bsim4/{nmos,pmos} of ngspice 46 (CMC QA, reltol 1e-6 '
                      'clip 1e-13 A per tests/bin/modelQaTestRoutines.pm:105-118)',
            'devices': {}}
    for pol in ('nmos', 'pmos'):
        d = os.path.join(ref, 'tests/bsim4', pol)
        glob, tests = parse_spec(os.path.join(d, 'qaSpec'))
        dev = {'tests': [], 'ac_tests': []}
        for t in tests:
            if t['name'].startswith('ac'):
                temps = [float(x) for x in t.get('temperature', glob.get('temperature', '27')).split()]
                biases = parse_bias(t['biases'])
                inst = {}
                for kv in t['instanceParameters'].split():
                    k, v = kv.split('=')
                    inst[k.lower()] = float(v)
                model = {}
                for f in t['model_files']:
                    model.update(parse_params(os.path.join(d, f)))
                model.update(t['model_extra'])
                lines = open(os.path.join(d, 'reference', t['name'] + '.standard')).read().split('\n')
                header = lines[0].split()
                rows = [[float(x) for x in l.split()] for l in lines[1:] if l.strip()]
                entry = {'name': t['name'], 'model': model, 'instance': inst,
                         'columns': header, 'temps': temps, 'rows': rows,
                         'bias': {p: float(v) for p, v in biases.items()}}
                if 'biasSweep' in t:
                    sname, svals = list(parse_bias(t['biasSweep']).items())[0]
                    entry['sweep_pin'] = sname
                    assert len(rows) % len(temps) == 0
                if 'freq' in t:
                    entry['freq'] = t['freq']
                dev['ac_tests'].append(entry)
                continue
            if not t['name'].startswith('dcSweep'):
                continue
            temps = [float(x) for x in t.get('temperature', glob.get('temperature', '27')).split()]
            biases = parse_bias(t['biases'])
            lname, lvals = list(parse_bias(t['biasList']).items())[0]
            sname, svals = list(parse_bias(t['biasSweep']).items())[0]
            start, stop, step = [float(x) for x in svals.split(',')]
            nsweep = int(round((stop - start) / step)) + 1
            inst = {}
            for kv in t['instanceParameters'].split():
                k, v = kv.split('=')
                inst[k.lower()] = float(v)
            model = {}
            for f in t['model_files']:
                model.update(parse_params(os.path.join(d, f)))
            model.update(t['model_extra'])
            rows = []
            for line in open(os.path.join(d, 'reference', t['name'] + '.standard')):
                parts = line.split()
                try:
                    rows.append([float(x) for x in parts])
                except ValueError:
                    continue
            lv = [float(x) for x in lvals.split(',')]
            assert len(rows) == len(temps) * len(lv) * nsweep, t['name']
            points = []
            k = 0
            for T in temps:
                for v in lv:
                    for i in range(nsweep):
                        r = rows[k]
                        k += 1
                        bias = {p: float(val) for p, val in biases.items()}
                        bias[lname] = v
                        bias[sname] = r[0]
                        points.append({'temp': T, 'bias': bias, 'I': r[1]})
            dev['tests'].append({'name': t['name'], 'model': model, 'instance': inst,
                                 'output': t['outputs'], 'points': points})
        data['devices'][pol] = dev
    with open(outp, 'w') as fh:
        json.dump(data, fh, separators=(',', ':'))
    print('wrote', outp, sum(len(t['points']) for d in data['devices'].values() for t in d['tests']), 'points')


if __name__ == '__main__':
    main()
