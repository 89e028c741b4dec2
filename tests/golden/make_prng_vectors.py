"""Writes tests/golden/sitmo_prng.json from the reference PRNG build.

The driver oracle/_ref/prng_driver is our own program compiled against the
reference's prng_engine.hpp (recipe: `make -C oracle ref`).  It prints, per
script, the seed, the steps ("d<n>" = n draws, "s<z>" = discard(z)) and the
drawn 32-bit outputs.  This script runs it once and stores the text as JSON,
so the tests need neither the reference nor the driver.
"""
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def main():
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"])
    out = subprocess.check_output([os.path.join(ROOT, "oracle", "_ref", "prng_driver")], text=True)
    cases = []
    for line in out.strip().splitlines():
        seed, script, values = line.split("|")
        cases.append({"seed": int(seed), "steps": script.split(","),
                      "values": [int(v) for v in values.split()]})
    dst = os.path.join(HERE, "sitmo_prng.json")
    with open(dst, "w") as f:
        json.dump({"source": "prng_engine.hpp (sitmo::prng_engine) via oracle/ref_prng_driver.cpp",
                   "cases": cases}, f, indent=0)
    print(f"wrote {len(cases)} cases to {dst}", file=sys.stderr)


if __name__ == "__main__":
    main()
