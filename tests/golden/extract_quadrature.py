"""Writes quadrature_reference.json: the numeric tables of the reference's
Bem/Integration/quadrature.hpp, read as text (nothing of the reference is
run or imported).

  gauss_7      lines 30-38:   quadrature_1d(0.5*(1.0 +- t), 0.5*w)  -> t, w
  quadrature_7 lines 88-96:   quadrature_2d(x, y, w*0.5)            -> x, y, w
  quadrature_19 lines 126-146: quadrature_2d(x, y, w*0.5)           -> x, y, w

Usage: python tests/golden/extract_quadrature.py [/path/to/reference]
"""

import json
import os
import re
import sys

NUM = r"([-+]?[0-9.]+(?:e[-+]?[0-9]+)?)"


def table(text, name):
    m = re.search(r"\b" + name + r"\s*=\s*\{(.*?)\};", text, re.S)
    return m.group(1)


def main(ref):
    path = os.path.join(ref, "Bem", "Integration", "quadrature.hpp")
    text = open(path).read()
    g_t, g_w = [], []
    for sign, t, w in re.findall(r"quadrature_1d\(0\.5\*\(1\.0([-+])" + NUM + r"\),0\.5\*" + NUM + r"\)",
                                 table(text, "gauss_7")):
        g_t.append(float(t) * (-1.0 if sign == "-" else 1.0))
        g_w.append(float(w))
    out = {"source": "Bem/Integration/quadrature.hpp (reference): gauss_7 at lines 30-38 as (0.5*(1+t), 0.5*w); "
                     "quadrature_7 at lines 88-96 and quadrature_19 at lines 126-146 as (x, y, w*0.5)",
           "gauss_7_t": g_t, "gauss_7_w": g_w}
    for name in ("quadrature_7", "quadrature_19"):
        rows = re.findall(r"quadrature_2d\(\s*" + NUM + r"\s*,\s*" + NUM + r"\s*,\s*" + NUM + r"\s*\*\s*0\.5\s*\)",
                          table(text, name))
        out[name + "_x"] = [float(r[0]) for r in rows]
        out[name + "_y"] = [float(r[1]) for r in rows]
        out[name + "_w_unhalved"] = [float(r[2]) for r in rows]
    assert len(g_t) == 7 and len(out["quadrature_7_x"]) == 7 and len(out["quadrature_19_x"]) == 19
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "quadrature_reference.json")
    json.dump(out, open(dst, "w"), indent=1)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
