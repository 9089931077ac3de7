"""Writes tests/golden/gamma_tables.json: SHA-256 digests of the reference's
two gamma exponent tables (src/nvmath/Gamma.cpp, pow_5_11_table and
pow_11_5_table, 512 floats each, as little-endian float32 bytes), parsed from
the reference source text.  Gamma.cpp cannot be compiled here on its own (its
header pulls in the cmake-generated nvconfig.h), so the digests are how the
oracle's closed-form tables are pinned to the reference's literal ones.  Run in
the container that holds the reference tree:

    python tests/golden/make_gamma_fixture.py
"""
import hashlib
import json
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/src/nvmath/Gamma.cpp"


def parse(text, name):
    m = re.search(r"static const float %s\[512\] = \{(.*?)\};" % name, text, re.S)
    body = re.sub(r"//[^\n]*", "", m.group(1))
    vals = [np.inf if v == "INFINITE_RESULT" else float(v.rstrip("f"))
            for v in re.findall(r"INFINITE_RESULT|[-+]?[0-9.]+(?:[eE][-+]?[0-9]+)?f?", body)]
    assert len(vals) == 512, (name, len(vals))
    return np.array(vals, dtype="<f4")


def main():
    text = open(SRC).read()
    out = {"source": "src/nvmath/Gamma.cpp (castano/nvidia-texture-tools 2.1.2)", "encoding": "512 x float32 little-endian"}
    for key, name in (("pow_5_11", "pow_5_11_table"), ("pow_11_5", "pow_11_5_table")):
        t = parse(text, name)
        out[key] = {"sha256": hashlib.sha256(t.tobytes()).hexdigest(), "nonzero": int(np.count_nonzero(t)),
                    "inf": int(np.isinf(t).sum())}
    with open(os.path.join(HERE, "gamma_tables.json"), "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
