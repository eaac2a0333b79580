"""Writes tests/golden/golden_matrix_full.json: for every corpus case of golden_matrix_none.json,
the reference's 'Full None 0' row of reference/golden_matrix.txt -- the number of exact (IR)
canonical states and the sum of mix64 over their distinct IR hashes (the fingerprint of
tools/generate_golden_matrix.cpp:42-59) -- and the event counts of its 'None Full 0' and
'None Automatic 0' rows (event canonicalization by signature), and the state count of its
'Automatic None 0' row (content identity). Full multiway expands every raw state in every mode,
so these are the distinct IR hashes over the None-mode run's raw states. Run in the build
container: python tests/golden/make_golden_full.py
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))
MATRIX = "/root/reference/reference/golden_matrix.txt"


def main():
    names = [c["name"] for c in json.load(open(os.path.join(HERE, "golden_matrix_none.json")))["cases"]]
    rows, rest = {}, {}
    for line in open(MATRIX):
        t = line.split()
        if not t or t[0].startswith("#") or len(t) < 11:
            continue
        case, smode, emode, quotient = t[0], t[1], t[2], t[3]
        if smode == "Full" and emode == "None" and quotient == "0":
            rows.setdefault(case, {}).update(ir_states=int(t[5]), events=int(t[6]), ir_fingerprint=int(t[10]))
        if smode == "None" and emode in ("Full", "Automatic") and quotient == "0":
            rows.setdefault(case, {})["events_" + emode.lower()] = int(t[6])
        if quotient == "0" and smode in ("None", "Automatic"):
            rest[(case, smode, emode)] = t[6:10]
        if smode == "Automatic" and emode == "None" and quotient == "0":
            rows.setdefault(case, {})["content_states"] = int(t[5])
    # under Automatic state identity the events, causal edges, pairs and branchial edges are
    # the None rows' in every event mode (only the state count changes): checked, not assumed
    for (case, smode, emode), v in rest.items():
        if smode == "Automatic":
            assert v == rest[(case, "None", emode)], (case, emode)
    doc = {"source": "reference/golden_matrix.txt rows '<case> Full None 0': states = distinct IR hashes, "
                     "fingerprint = sum of mix64 over them; rows '<case> None Full|Automatic 0': events; "
                     "rows '<case> Automatic None 0': content_states (Automatic rows' other columns equal None's)",
           "cases": {n: rows[n] for n in names if n in rows}}
    with open(os.path.join(HERE, "golden_matrix_full.json"), "w") as f:
        json.dump(doc, f, indent=1)
    print(len(doc["cases"]), "cases")


if __name__ == "__main__":
    main()
