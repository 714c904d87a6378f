"""Extract the reference's C0 input track and hits into a JSON fixture.

Reads src/input_track.h of cerati/p2r-tests (the active, uncommented block:
`inputtrk` and `inputhit00` .. `inputhit21`, input_track.h:1-96) and writes
input_track_c0.json next to this script.  The values are data (the reference's
own input fixture); the JSON is what the tests compare the oracle's and the
product's embedded C0 tables against.

Usage: python tests/golden/make_input_c0.py [/root/reference]
"""
import json
import os
import re
import sys


def parse(text):
    # Only the block before the first commented-out alternative set.
    active = text.split("/*")[0]
    num = r"[-+]?\d*\.?\d+(?:[eE][-+]?\d+)?"
    trk = re.search(r"ATRK\s+inputtrk\s*=\s*\{\s*\{([^}]*)\}\s*,\s*\{([^}]*)\}\s*,\s*(" + num + r")\s*\}", active)
    par = [float(v) for v in re.findall(num, trk.group(1))]
    cov = [float(v) for v in re.findall(num, trk.group(2))]
    q = int(trk.group(3))
    hits = {}
    for m in re.finditer(r"AHIT\s+inputhit(\d\d)\s*=\s*\{\s*\{([^}]*)\}\s*,\s*\{([^}]*)\}", active):
        hits[int(m.group(1))] = {
            "pos": [float(v) for v in re.findall(num, m.group(2))],
            "cov": [float(v) for v in re.findall(num, m.group(3))],
        }
    assert len(par) == 6 and len(cov) == 21 and sorted(hits) == list(range(22))
    return {"track": {"par": par, "cov": cov, "q": q},
            "hits": [hits[i] for i in range(22)],
            "layer_order": list(range(21, -1, -1)),
            "source": "src/input_track.h:1-96; layer l uses inputhit[21-l] (propagate-tor-test_tbb.cpp:818-820)"}


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    with open(os.path.join(ref, "src", "input_track.h")) as f:
        data = parse(f.read())
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "input_track_c0.json")
    with open(out, "w") as f:
        json.dump(data, f, indent=1)
    print("wrote", out)


if __name__ == "__main__":
    main()
