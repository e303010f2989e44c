"""Extract the reference's codon, CAI-weight and codon-pair tables as data.

Reads include/info.cuh of the reference checkout as text (never compiles or
imports it) and writes the package's data/reference_tables.json:
  codons   -- 64 codons, 3 letters each, in the reference's order
              (info.cuh:11-31: amino-acid blocks A C D E F G H I K L M N P Q R
              S T V W Y Z, synonyms by ascending weight, stops last)
  weights  -- 64 float32 CAI weights (info.cuh:32-52), evaluated in float32
              exactly as the C expressions `a / b.0f` are
  cps      -- 61*61 float32 codon-pair scores (info.cuh:54)
  gc3_ref  -- ref_GC3_percent (info.cuh:7)
Only this container has the reference; the JSON is what ships.
"""
import json
import os
import re
import sys

import numpy as np

SRC = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/include/info.cuh"
ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
OUT = os.path.join(ROOT, "protein_encoding_v2_nsga-iii_amd", "data", "reference_tables.json")


def main():
    text = open(SRC).read()
    live = text.split("#if 0")[0]  # the alternate tables after '#if 0' are compiled out
    m = re.search(r'char codons\[[^\]]*\]\s*=\s*"(.*?)";', live, re.S)
    codons = re.sub(r'[\\\s"]', "", m.group(1))
    assert len(codons) == 192, len(codons)
    m = re.search(r"float codons_weight\[[^\]]*\]\s*=\s*\{(.*?)\};", live, re.S)
    weights = []
    for term in m.group(1).split(","):
        term = term.strip()
        if "/" in term:
            a, b = term.split("/")
            v = np.float32(float(a.strip())) / np.float32(float(b.strip().rstrip("f")))
        else:
            v = np.float32(float(term.rstrip("f")))
        weights.append(float(np.float32(v)))
    assert len(weights) == 64
    m = re.search(r"const float cps\[[^\]]*\]\s*=\s*\{(.*?)\};", live, re.S)
    cps = [float(np.float32(t.strip().rstrip("f"))) for t in m.group(1).split(",") if t.strip()]
    assert len(cps) == 61 * 61
    m = re.search(r"const float ref_GC3_percent\s*=\s*([0-9.]+)f", live)
    gc3 = float(np.float32(m.group(1)))
    json.dump({"source": "include/info.cuh:7-54", "codons": codons, "weights": weights,
               "cps": cps, "gc3_ref": gc3}, open(OUT, "w"))
    print("wrote", OUT)


if __name__ == "__main__":
    main()
