#!/usr/bin/env python3
"""Generate the BSIM4 4.8.3 parameter tables used by the engine and by the oracle.

The tables are data, not logic: the netlist names of the model and instance
parameters, the struct field each one sets, its value type, and the default
value ngspice applies when the parameter is not given.  They are read from the
reference sources (run this in a container that has them) and written as an
X-macro header so that the C/C++/HIP code can expand them into enums, name
lookup tables and the defaulting pass:

  src/spicelib/devices/bsim4/b4.c      BSIM4pTable / BSIM4mPTable (names -> ids)
  src/spicelib/devices/bsim4/b4par.c   instance id -> field
  src/spicelib/devices/bsim4/b4mpar.c  model id -> field
  src/spicelib/devices/bsim4/b4set.c:79-2300  model defaults, in order

Defaults that are not a literal, a copy of another field, or an NMOS/PMOS or
mobMod choice (the *sd gate-tunnelling families, dlcigd, njts*d, tnjts*d,
oxide trap densities) are listed as SPECIAL and implemented by hand next to
the defaulting pass.

usage: gen_bsim4_tables.py <reference_root> <out_header> [<out_header> ...]
"""
import re
import sys
import os


def strip_comments(s):
    s = re.sub(r'/\*.*?\*/', '', s, flags=re.S)
    return re.sub(r'//[^\n]*', '', s)


# Fields whose defaults depend on other Given flags: handled by hand.
SPECIAL_FAMILIES = set()
for pre in ('', 'l', 'w', 'p'):
    for q in ('a', 'b', 'c'):
        for sfx in ('igs', 'igd', 'igsd'):
            SPECIAL_FAMILIES.add(pre + q + sfx)
SPECIAL_FAMILIES |= {'dlcigd', 'njtsd', 'njtsswd', 'njtsswgd', 'tnjtsd',
                     'tnjtsswd', 'tnjtsswgd', 'oxideTrapDensityA',
                     'oxideTrapDensityB'}


def main():
    ref = sys.argv[1]
    outs = sys.argv[2:]
    d = os.path.join(ref, 'src/spicelib/devices/bsim4')
    b4 = strip_comments(open(os.path.join(d, 'b4.c')).read())
    par = strip_comments(open(os.path.join(d, 'b4par.c')).read())
    mpar = strip_comments(open(os.path.join(d, 'b4mpar.c')).read())
    setc = strip_comments(open(os.path.join(d, 'b4set.c')).read())

    ipart = b4[b4.index('BSIM4pTable[]'):b4.index('BSIM4mPTable[]')]
    mpart = b4[b4.index('BSIM4mPTable[]'):]
    entry = re.compile(r'\b(IOP|IOPR|IP|OP)\s*\(\s*"([^"]+)"\s*,\s*(\w+)\s*,\s*(\w+)')

    def table(part):
        out = []
        for k, name, pid, typ in entry.findall(part):
            if k == 'OP':
                continue
            out.append((name, pid, typ))
        return out

    inst_names = table(ipart)
    mod_names = table(mpart)

    case = re.compile(r'case\s+(\w+)\s*:\s*(.*?)break;', re.S)

    def id2field(src, obj):
        m = {}
        for pid, body in case.findall(src):
            f = re.search(obj + r'->BSIM4(\w+)\s*=\s*value->([ris])Value', body)
            if f:
                m[pid] = (f.group(1), f.group(2))
        return m

    mfield = id2field(mpar, 'mod')
    ifield = id2field(par, 'here')

    # model defaults in reference order
    sec = setc[setc.index('if (!model->BSIM4typeGiven)'):
               setc.index('DMCGeff = model->BSIM4dmcg - model->BSIM4dmcgt;')]
    # The engine implements BSIM4 4.8.1 and later; drop the 4.8.0-only branch
    # of the mobility defaults (b4set.c:358-366).
    v480 = re.search(r'if\s*\(\s*model->BSIM4v48intVersion\s*<=\s*BSIM4v48V480\s*\)\s*\{[^{}]*\}\s*else', sec)
    if v480:
        sec = sec[:v480.start()] + sec[v480.end():]
    pat = re.compile(r'if\s*\(\s*!\s*model->BSIM4(\w+)Given\s*\)\s*'
                     r'model->BSIM4(\w+)\s*=\s*([^;]+);')
    defaults = []
    seen = set()
    for m in pat.finditer(sec):
        g, f, e = m.group(1), m.group(2), ' '.join(m.group(3).split())
        if g != f or f in SPECIAL_FAMILIES or f in seen:
            continue
        seen.add(f)
        try:
            float(e)
            defaults.append(('DEF_CONST', f, e))
            continue
        except ValueError:
            pass
        mm = re.fullmatch(r'model->BSIM4(\w+)', e)
        if mm:
            defaults.append(('DEF_COPY', f, mm.group(1)))
            continue
        mm = re.fullmatch(r'\(?\s*model->BSIM4type\s*==\s*NMOS\s*\)?\s*\?\s*([-+.\deE]+)\s*:\s*([-+.\deE]+)', e)
        if mm:
            defaults.append(('DEF_TYPE', f, mm.group(1) + ', ' + mm.group(2)))
            continue
        mm = re.fullmatch(r'\(+\s*model->BSIM4mobMod\s*==\s*(\d)\s*\|\|\s*model->BSIM4mobMod\s*==\s*(\d)\s*\)+\s*\?\s*([-+.\deE]+)\s*:\s*([-+.\deE]+)', e)
        if mm:
            defaults.append(('DEF_MOBMOD', f, '%s, %s, %s, %s' % mm.groups()))
            continue
        if f == 'tnom':
            defaults.append(('DEF_TNOM', f, '0'))
            continue
        if f in ('type', 'version'):
            continue
        raise SystemExit('unhandled default %s = %s' % (f, e))

    # every field named anywhere
    fields = []
    fseen = set()

    def addf(f):
        if f not in fseen:
            fseen.add(f)
            fields.append(f)
    for _, pid, _ in mod_names:
        if pid in mfield:
            addf(mfield[pid][0])
    for k, f, _ in defaults:
        addf(f)
    for f in sorted(SPECIAL_FAMILIES):
        addf(f)
    addf('type')

    lines = []
    lines.append('/* Generated by gen_bsim4_tables.py from the ngspice 46 BSIM4 4.8.3 sources')
    lines.append(' * (b4.c parameter tables, b4par.c/b4mpar.c setters, b4set.c defaults).')
    lines.append(' * Data only: parameter names, target fields, value kinds and defaults. */')
    lines.append('#ifndef BSIM4_TABLES_INCLUDED')
    lines.append('#define BSIM4_TABLES_INCLUDED')
    lines.append('')
    lines.append('/* BSIM4_MODEL_FIELD(field) : every model field that a parameter sets */')
    lines.append('#define BSIM4_MODEL_FIELDS(X) \\')
    for f in fields:
        lines.append('  X(%s) \\' % f)
    lines.append('')
    lines.append('/* BSIM4_MODEL_PARAM(name, field, kind) kind: r real, i integer, f flag */')
    lines.append('#define BSIM4_MODEL_PARAMS(X) \\')
    for name, pid, typ in mod_names:
        if pid == 'BSIM4_MOD_NMOS':
            lines.append('  X("%s", type, n) \\' % name)
        elif pid == 'BSIM4_MOD_PMOS':
            lines.append('  X("%s", type, p) \\' % name)
        elif pid in mfield:
            f, k = mfield[pid]
            if k == 's':
                continue
            lines.append('  X("%s", %s, %s) \\' % (name, f, k))
    lines.append('')
    lines.append('/* defaults in the order ngspice applies them */')
    lines.append('#define BSIM4_MODEL_DEFAULTS(C, Y, T, M, N) \\')
    for k, f, e in defaults:
        mac = {'DEF_CONST': 'C', 'DEF_COPY': 'Y', 'DEF_TYPE': 'T',
               'DEF_MOBMOD': 'M', 'DEF_TNOM': 'N'}[k]
        lines.append('  %s(%s, %s) \\' % (mac, f, e))
    lines.append('')
    # size-dependent (binned) parameters: p = p0 + pl*InvL + pw*InvW + pp*InvLW
    temp = strip_comments(open(os.path.join(d, 'b4temp.c')).read())
    binpat = re.compile(r'pParam->BSIM4(\w+)\s*=\s*model->BSIM4(\w+)\s*\+\s*model->BSIM4l(\w+)\s*\*\s*Inv_L\s*'
                        r'\+\s*model->BSIM4w(\w+)\s*\*\s*Inv_W\s*\+\s*model->BSIM4p(\w+)\s*\*\s*Inv_LW\s*;')
    binned = []
    for m in binpat.finditer(temp):
        f, a, l, w, p = m.groups()
        if not (f == a == l == w == p):
            raise SystemExit('irregular binning for %s' % f)
        for pre in ('', 'l', 'w', 'p'):
            if pre + f not in fseen:
                raise SystemExit('binned field %s%s has no setter' % (pre, f))
        binned.append(f)
    lines.append('/* BSIM4_BINNED(field): size-dependent parameters, in b4temp.c order */')
    lines.append('#define BSIM4_BINNED(X) \\')
    for f in binned:
        lines.append('  X(%s) \\' % f)
    lines.append('')
    lines.append('/* BSIM4_INST_PARAM(name, field, kind) */')
    lines.append('#define BSIM4_INST_PARAMS(X) \\')
    for name, pid, typ in inst_names:
        if pid in ifield:
            f, k = ifield[pid]
            lines.append('  X("%s", %s, %s) \\' % (name, f, k))
    lines.append('')
    lines.append('#endif')
    text = '\n'.join(lines) + '\n'
    for o in outs:
        with open(o, 'w') as fh:
            fh.write(text)
    print('fields %d, model params %d, defaults %d, instance params %d' % (
        len(fields), len(mod_names), len(defaults), len(inst_names)))


if __name__ == '__main__':
    main()
