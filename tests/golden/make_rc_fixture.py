"""Extract the v(2) column of the reference's RC transient golden
(tests/general/rc.out: .tran 0.1 7 of r 1 2 1.0, c 2 0 1.0, vin pulse(0 1))
into rc_golden.json.  Run with the reference at /root/reference."""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else '/root/reference'
src = os.path.join(REF, 'tests', 'general', 'rc.out')
rows = []
rows_reported = None
for line in open(src):
    m = re.match(r'\s*([0-9.]+e[+-]\d+)\s+([-0-9.]+e[+-]\d+)\s', line)
    if m:
        rows.append([float(m.group(1)), float(m.group(2))])
    m = re.match(r'No\. of Data Rows : (\d+)', line)
    if m:
        rows_reported = int(m.group(1))
out = {'source': 'tests/general/rc.out of ngspice (reference test suite)',
       'netlist': 'r 1 2 1.0; c 2 0 1.0; vin 1 0 pulse(0 1); .tran 0.1 7.0',
       'note': 'printed at 4 significant digits on the 0.1 s output grid',
       'data_rows': rows_reported, 't_v2': rows}
json.dump(out, open(os.path.join(os.path.dirname(os.path.abspath(__file__)), 'rc_golden.json'), 'w'),
          indent=0)
print(len(rows), rows_reported)
