"""Write the polynomial force-law coefficients as the reference's sources
spell them, for the constant pins of the oracle and the device library.

Reads (as text) the literals of
  BGQStep16.c:167        const float ma0 = 0.269327, ... ma5 = ...  (the
                         scalar tail of Step16_int; the QPX constants at
                         :40-45 repeat them)
  ForceLaw.cxx:96-116    m_a[k] = ...;  under POLY_ORDER 6 and 5
and writes ref_poly_coeffs.json with each literal kept as its decimal
string.  Runs only where the reference tree is mounted; the JSON is the
fixture the tests read.
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/src/halo_finder"


def step16_scalar(text):
    m = re.search(r"const float ma0 = ([^,]+), ma1 = ([^,]+), ma2 = ([^,]+), "
                  r"ma3 = ([^,]+), ma4 = ([^,]+), ma5 = ([^;]+);", text)
    return [v.strip() for v in m.groups()]


def step16_qpx(text):
    out = []
    for k in range(6):
        m = re.search(r"vector4double a1%d = \(vector4double\)\{ *([^,]+)," % k, text)
        out.append(m.group(1).strip())
    return out


def forcelaw_order(text, order):
    block = re.search(r"#if POLY_ORDER == %d\n(.*?)#endif" % order, text, re.S).group(1)
    vals = dict((int(i), v.strip()) for i, v in re.findall(r"m_a\[(\d)\] = *([^;]+);", block))
    return [vals[k] for k in range(7)]


def main():
    if not os.path.isdir(SRC):
        sys.exit("the reference tree is not mounted")
    s16 = open(os.path.join(SRC, "BGQStep16.c")).read()
    fl = open(os.path.join(SRC, "ForceLaw.cxx")).read()
    out = {
        "step16_scalar": step16_scalar(s16),
        "step16_qpx": step16_qpx(s16),
        "forcelaw_poly5": forcelaw_order(fl, 5),
        "forcelaw_poly6": forcelaw_order(fl, 6),
        "note": "decimal literals as written; each is a double constant that the "
                "reference converts to float (float ma0 = ..., float m_a[7])",
    }
    with open(os.path.join(HERE, "ref_poly_coeffs.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
