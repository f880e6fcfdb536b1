"""Writes tests/golden/gauss_legendre_reference.json: the 30-digit
Gauss-Legendre nodes and weights tabulated by the reference
(gsGaussRule::lookupReference, src/gsAssembler/gsGaussRule.hpp:218-547),
copied as decimal strings for n = 1..10.  Run in the build container only
(it reads the reference tree); the JSON is the committed fixture."""
import json
import os
import re
import sys

SRC = "/root/reference/src/gsAssembler/gsGaussRule.hpp"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gauss_legendre_reference.json")


def main():
    text = open(SRC).read()
    start = text.index("lookupReference")
    body = text[start:]
    out = {}
    for m in re.finditer(r"case\s+(\d+)\s*:\s*\{(.*?)return true;", body, re.S):
        n = int(m.group(1))
        if n > 10:
            continue
        blk = m.group(2)
        xs = re.search(r"x\s*<<(.*?);", blk, re.S).group(1)
        ws = re.search(r"w\s*<<(.*?);", blk, re.S).group(1)
        out[str(n)] = {"x": [s.strip() for s in xs.split(",")],
                       "w": [s.strip() for s in ws.split(",")]}
    if len(out) < 5:
        sys.exit("table not found")
    json.dump({"source": "src/gsAssembler/gsGaussRule.hpp lookupReference", "rules": out},
              open(OUT, "w"), indent=1)
    print("wrote", OUT, sorted(int(k) for k in out))


if __name__ == "__main__":
    main()
