"""Writes tests/golden/golden_matrix_quotient.json: the rows of reference/golden_matrix.txt with
CanonicalizeStates -> Full, for every corpus case of golden_matrix_none.json and every event
identity mode (None, Full, Automatic): '<case> <mode>' holds the quotient row ('<case> Full <mode>
1': [states, events, causal triples, causal pairs, branchial pairs]), '<case> <mode> full' the
full-capture row ('<case> Full <mode> 0'). Run in the build container:
python tests/golden/make_golden_quotient.py
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))
MATRIX = "/root/reference/reference/golden_matrix.txt"


def main():
    names = {c["name"] for c in json.load(open(os.path.join(HERE, "golden_matrix_none.json")))["cases"]}
    rows = {}
    for line in open(MATRIX):
        t = line.split()
        if not t or t[0].startswith("#") or len(t) < 11 or t[0] not in names or t[1] != "Full":
            continue
        key = f"{t[0]} {t[2]}" + ("" if t[3] == "1" else " full")
        rows[key] = [int(v) for v in t[5:10]]
    doc = {"source": "reference/golden_matrix.txt rows '<case> Full <event mode> <quotient>': "
                     "states, events, causal triples, causal pairs, branchial pairs",
           "rows": dict(sorted(rows.items()))}
    with open(os.path.join(HERE, "golden_matrix_quotient.json"), "w") as f:
        json.dump(doc, f, indent=0)
    print(len(rows), "rows")


if __name__ == "__main__":
    main()
