"""Writes four_domain_enumeration.json: the exact distribution of the number
of fully bound domain pairs of the enumerable four-domain system (at most two
staples) at 330 and 345 K, from the oracle's enumeration (which the
enumeration tests pin), as the reference of the learned proposal's accuracy
protocol.

    python tests/golden/make_four_domain_enumeration.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

TEMPS = (330.0, 345.0)


def exact_numfulldomains(temp):
    import pyoracle
    from latticednaorigami_amd import systems
    s = systems.four_domain_system(max_total_staples=2)
    o = pyoracle.Oracle(s, temps=[temp])
    probs, _, _ = o.enumerate()
    o.close()
    out = [0.0] * 5
    for (ns, nfull, nmis, nstack), p in probs.items():
        out[nfull] += p
    return out


def main():
    table = {f"{t:g}": exact_numfulldomains(t) for t in TEMPS}
    with open(os.path.join(HERE, "four_domain_enumeration.json"), "w") as f:
        json.dump({"system": "four_domain_system(max_total_staples=2)",
                   "order_parameter": "num_fully_bound_domain_pairs",
                   "p_by_temperature": table}, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
