"""Writes tests/golden/ir_vectors.json from the REFERENCE exact canonicalizer (oracle/_ref/irref,
built from the reference sources by oracle/Makefile). Run in the build container:

    make -C oracle ref && python tests/golden/make_ir_vectors.py

Each vector is a state (edge lists over arbitrary vertex labels) with the reference's
IRCanonicalizer::compute_canonical_hash, canonicalize_edges() canonical form and
compute_canonical_hash_with_edge_rank() per-edge ranks. The states mix
random hypergraphs (arity 1-4, repeated vertices, unary edges) with symmetric ones (cycles,
circulants, disjoint copies, complete digraphs, stars, permutation hypergraphs), where the
search and its automorphism pruning run.
"""
from __future__ import annotations

import itertools
import json
import os
import random
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
IRREF = os.path.join(ROOT, "oracle", "_ref", "irref")


def states(seed: int = 20261202):
    rnd = random.Random(seed)
    out = []
    for n in range(2, 13):
        for a in (2, 3):
            out.append([[(i + k) % n for k in range(a)] for i in range(n)])              # cycles
    for n in (8, 12, 15):
        out.append([[i, (i + 1) % n] for i in range(n)] + [[i, (i + 3) % n] for i in range(n)])
    for c in (2, 3, 5, 8):
        base = [[0, 1], [1, 2], [2, 0], [0, 0]]
        out.append([[v + 10 * k for v in e] for k in range(c) for e in base])            # copies
    for n in range(2, 7):
        out.append([[i, j] for i in range(n) for j in range(n) if i != j])              # K_n
    for k in (5, 9, 17):
        out.append([[0, i] for i in range(1, k)])                                        # stars
    out.append([list(p) for p in itertools.permutations(range(4), 3)])
    out.append([[v, v] for v in range(6)])
    out.append([[i] for i in range(5)])
    while len(out) < 600:
        nv = rnd.randint(1, 14)
        out.append([[rnd.randrange(nv) * 7 + 3 for _ in range(rnd.choice([1, 2, 2, 2, 3, 4]))]
                    for _ in range(rnd.randint(1, 24))])
    return out


def main():
    sts = states()
    with tempfile.NamedTemporaryFile("w", suffix=".txt", delete=False) as f:
        f.write(f"{len(sts)}\n")
        for st in sts:
            f.write(f"{len(st)} " + " ".join(f"{len(e)} " + " ".join(map(str, e)) for e in st) + "\n")
        path = f.name
    try:
        lines = subprocess.run([IRREF, path], capture_output=True, text=True, check=True).stdout.splitlines()
    finally:
        os.unlink(path)
    vectors = []
    for st, line in zip(sts, lines):
        t = list(map(int, line.split()))
        h, nv, ne, rest = t[0], t[1], t[2], t[3:]
        form, w = [], 0
        for _ in range(ne):
            a = rest[w]
            form.append(rest[w + 1:w + 1 + a])
            w += 1 + a
        rank = rest[w:w + len(st)]
        vectors.append({"edges": st, "hash": h, "vertex_count": nv, "form": form, "rank": rank})
    doc = {"source": "oracle/_ref/irref: IRCanonicalizer::compute_canonical_hash and canonicalize_edges "
                     "(hypergraph/src/ir_canonicalization.cpp over common/include/hgcommon/ir_core.hpp)",
           "vectors": vectors}
    with open(os.path.join(HERE, "ir_vectors.json"), "w") as f:
        json.dump(doc, f, separators=(",", ":"))
    print(len(vectors), "vectors")


if __name__ == "__main__":
    main()
