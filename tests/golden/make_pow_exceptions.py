"""Writes tests/golden/pow_exceptions.json: the result of checking the C
library's pow against the correctly rounded powers of the kernels on every
float r^2 in [0, fl(h^2)] (see tests/test_pow_tables.py): how many inputs
differ for pow(h^2 - r^2, 3) and pow(h - sqrt(r^2), 2), the first entries of
each table, and fingerprints of all 990,437,133 pow results of each kind.

Run from the repository root after building the library and tests/hip:
    python tests/golden/make_pow_exceptions.py
"""
import json
import os
import platform
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import test_pow_tables as t  # noqa: E402


def main():
    tab3, tab2, libc = t.load_tables()
    out = t.host_check(tab3, tab2, threads=os.cpu_count() or 1)
    assert out[0] == 0 and out[1] == 0, out
    rec = {
        "what": "inputs r^2 (every float in [0, fl(h^2)]) where glibc pow differs from the correctly rounded power",
        "libc": libc,
        "machine": platform.machine(),
        "inputs": t.HSQ_BITS + 1,
        "pow3_exceptions": len(tab3),
        "pow2_exceptions": len(tab2),
        "pow3_first": [int(v) for v in tab3[:4]],
        "pow2_first": [int(v) for v in tab2[:4]],
        "entry_format": "float bits of r^2, bit 31 set when pow is the next double above the correctly rounded power",
        "fingerprint_pow3": out[4],
        "fingerprint_pow2": out[5],
    }
    with open(os.path.join(HERE, "pow_exceptions.json"), "w") as f:
        json.dump(rec, f, indent=1)
        f.write("\n")
    print(json.dumps(rec, indent=1))


if __name__ == "__main__":
    main()
