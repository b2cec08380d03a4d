"""Writes the golden fixtures under tests/golden/.

readme_partition.json -- the 9-nonzero example tensor of the reference's
README (README.md:87-95, written out as readme_example.tns), partitioned for
several P_n.  Expected values come from oracle/brute.py's independent Python
statement (mixed-radix block id + Python's stable sort), not from oracle.c,
and the P=(2,1,1) case is also spelled out by hand in
tests/test_oracle_partition.py.

Run:  python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import brute  # noqa: E402


def read_tns(path, order):
    idx, vals = [], []
    with open(path) as f:
        for line in f:
            tok = line.split()
            if not tok:
                continue
            idx.append([int(t) - 1 for t in tok[:order]])
            vals.append(float(tok[order]))
    return np.array(idx, np.uint32).T.copy(), np.array(vals)


def main():
    idx, vals = read_tns(os.path.join(HERE, "readme_example.tns"), 3)
    dims = [int(idx[k].max()) + 1 for k in range(3)]
    cases = []
    for part in ([1, 1, 1], [2, 1, 1], [3, 1, 1], [1, 2, 1], [2, 2, 2], [3, 2, 2]):
        bd, counts, order, bids = brute.partition_python(idx, dims, part)
        starts = np.concatenate([[0], np.cumsum(counts)]).tolist()
        csr = []
        for n in range(3):
            per_block = []
            for b in range(len(counts)):
                members = order[starts[b]:starts[b + 1]]
                local = [int(idx[n, e]) - (bids[e] // int(np.prod(part[n + 1:])) % part[n]) * bd[n]
                         for e in members]
                rp, wh = brute.csr_python(local, bd[n])
                per_block.append({"row_ptr": [int(x) for x in rp], "where": [int(x) for x in wh]})
            csr.append(per_block)
        cases.append({"part": part, "block_dims": bd, "block_nnz": counts, "perm": order,
                      "csr": csr})
    out = {"source": "README.md:87-95", "dims": dims, "nnz": int(len(vals)),
           "sum_sq": float(np.sum(vals ** 2)), "cases": cases}
    with open(os.path.join(HERE, "readme_partition.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
