"""Writes state_diagram.json from the reference's tests/state_diagram.xlsx.

The workbook (sheets 'Without waning' and 'With waning') holds, for each pair of
states (s1, s2), whether s1=True implies s2=True (1), s2=False (-1) or either (0)
(tests/test_immunity.py:25-85). Parsed with the standard library's zipfile/XML
readers (openpyxl is not installed). Run in the build container:
    python tests/golden/make_state_diagram.py
"""
import json
import os
import re
import xml.etree.ElementTree as ET
import zipfile

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = '/root/reference/tests/state_diagram.xlsx'
NS = {'m': 'http://schemas.openxmlformats.org/spreadsheetml/2006/main'}


def col_index(ref):
    letters = re.match(r'[A-Z]+', ref).group(0)
    n = 0
    for ch in letters:
        n = n * 26 + (ord(ch) - 64)
    return n - 1


def read_sheet(z, shared, path):
    root = ET.fromstring(z.read(path))
    rows = []
    for row in root.iter('{%s}row' % NS['m']):
        cells = {}
        for c in row.findall('m:c', NS):
            v = c.find('m:v', NS)
            if v is None:
                continue
            val = v.text
            if c.get('t') == 's':
                val = shared[int(val)]
            elif c.get('t') not in ('str', 'inlineStr'):
                val = float(val)
            cells[col_index(c.get('r'))] = val
        if cells:
            rows.append([cells.get(i) for i in range(max(cells) + 1)])
    return rows


def main():
    z = zipfile.ZipFile(SRC)
    sroot = ET.fromstring(z.read('xl/sharedStrings.xml'))
    shared = [''.join(t.text or '' for t in si.iter('{%s}t' % NS['m'])) for si in sroot.findall('m:si', NS)]
    out = {}
    for name, path in [('without_waning', 'xl/worksheets/sheet2.xml'), ('with_waning', 'xl/worksheets/sheet3.xml')]:
        rows = read_sheet(z, shared, path)
        header = [h for h in rows[0][1:] if h]
        matrix = {}
        for r in rows[1:]:
            if not r or r[0] is None or r[0] not in header:
                continue
            matrix[r[0]] = {header[j]: int(r[j + 1]) for j in range(len(header)) if r[j + 1] is not None}
        out[name] = dict(states=header, relation=matrix)
    with open(os.path.join(HERE, 'state_diagram.json'), 'w') as f:
        json.dump(out, f, indent=1)
    print({k: len(v['states']) for k, v in out.items()})


if __name__ == '__main__':
    main()
